"""GPU tests of the device-drained ring server (sigops_server mode SIGOPS_SERVE_RING): requests are
packed by bucket into tiles that mix requests and handed to one running persistent launch.

Every result must equal the oracle's, and a request's results must be the same alone and in the
mix: concurrent submitters; the mixed workload's request sizes (1..4096 items, log-uniform) at
three arrival rates through the native replay client (also for the slot-batch server); the recovery and Ed25519 edge cases (the
exceptional additions end the launch and go through its careful pass); a ring of 64 positions,
so positions are reused many times and the packer waits for free ones;
launches that end when the ring idles and the next request starts another; refusals at submit."""
import time

import numpy as np
import pytest

import edge_cases
import oracle_lib as O
from sigops_pkg import sigops
from test_gpu_serve import _check_pool_results, _pool, _requests, _slice

pytestmark = pytest.mark.gpu
W = __import__("importlib").import_module("wgpu-sigops_amd.workload")


def _say(*a):
    print(*a, flush=True)


def _ring(**kw):
    return sigops.Server(mode=sigops.SERVE_RING, **kw)


def _gather(pool, idx):
    """items idx of the pool laid out back to back (the replay client's layout)"""
    schemes, sigs, aux, msgs, off = pool[:5]
    lens = (off[idx + 1] - off[idx]).astype(np.int64)
    g_off = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    ed = idx[schemes[idx] == 2]
    g_msgs = np.concatenate([msgs[off[i]:off[i + 1]] for i in ed]) if len(ed) else np.zeros(0, np.uint8)
    return schemes[idx], sigs[idx], aux[idx], g_msgs, g_off


def _expect_at(pool, idx, out, st):
    """results of the gathered items idx against the pool's oracle results"""
    schemes, _, _, _, _, out_exp, st_exp, v, pidx = pool
    sec = schemes[idx] < 2
    assert np.array_equal(out[sec], out_exp[idx[sec]]) and np.array_equal(st[sec], st_exp[idx[sec]])
    ed_pos = np.full(len(schemes), -1, np.int64)
    ed_pos[pidx[2]] = np.arange(len(pidx[2]))
    e = ~sec
    assert np.array_equal(st[e] == 0, v.astype(bool)[ed_pos[idx[e]]])
    assert not out[e].any()


def test_ring_alone_and_mixed_are_identical():
    n = 6000
    _say(f"ring alone/mixed n={n}")
    pool = _pool(n, 31)
    req = _requests(n, 5)
    with _ring() as srv:
        alone_out = np.zeros((n, 64), np.uint8)
        alone_st = np.zeros(n, np.uint8)
        for r in range(len(req) - 1):
            a, b = req[r], req[r + 1]
            o, s, _, _ = srv.wait(srv.submit(*_slice(pool, a, b)))
            alone_out[a:b], alone_st[a:b] = o, s
        _check_pool_results(pool, alone_out, alone_st)
        tickets = [srv.submit(*_slice(pool, req[r], req[r + 1])) for r in range(len(req) - 1)]
        mix_out = np.zeros((n, 64), np.uint8)
        mix_st = np.zeros(n, np.uint8)
        for r, t in enumerate(tickets):
            o, s, ts, td = srv.wait(t)
            assert td >= ts
            mix_out[req[r]:req[r + 1]], mix_st[req[r]:req[r + 1]] = o, s
        batches, items = srv.stats()
    assert np.array_equal(mix_out, alone_out) and np.array_equal(mix_st, alone_st)
    assert items == 2 * n


@pytest.mark.parametrize("mode", ["ring", "slots"])
@pytest.mark.parametrize("rate", [2000, 10000, 40000])
def test_mixed_request_sizes_at_three_rates(rate, mode):
    """requests of 1..4096 items (log-uniform) drawn from a pool, Poisson arrivals at `rate` per
    second, through either server; every request's results equal the oracle's and the same
    request submitted alone"""
    n = 8192
    pool = _pool(n, 61)
    rng = np.random.default_rng(rate)
    nreq = 40
    sizes = np.floor(np.exp(rng.random(nreq) * np.log(4096))).astype(np.int64).clip(1, 4096)
    starts = rng.integers(0, n - 4096, nreq)
    idx = np.concatenate([np.arange(a, a + k) for a, k in zip(starts, sizes)])
    req_off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    arrivals = np.cumsum(rng.exponential(1e9 / rate, nreq)).astype(np.uint64)
    _say(f"{mode} replay rate {rate}/s: {nreq} requests, {len(idx)} items")
    arrs = _gather(pool, idx)
    with (_ring() if mode == "ring" else sigops.Server(slots=3)) as srv:
        lat, out, st = srv.replay(*arrs, req_off, arrivals)
        _expect_at(pool, idx, out, st)
        assert (lat > 0).all() and (lat < 10e9).all()
        # three of the requests again, alone
        for r in (0, nreq // 2, nreq - 1):
            a, b = int(req_off[r]), int(req_off[r + 1])
            g = _gather(pool, idx[a:b])
            o, s, _, _ = srv.wait(srv.submit(*g))
            assert np.array_equal(o, out[a:b]) and np.array_equal(s, st[a:b]), r


def test_ring_edge_cases_and_careful_pass():
    """every recovery and Ed25519 edge case among valid items, in requests of a few hundred"""
    items = []
    for c in (0, 1):
        cases = edge_cases.secp_cases(c)
        for (name, sig, dig), e in zip(cases, edge_cases.secp_expected(c, cases)):
            items.append((c, sig, dig, b"", e))
    ecases = edge_cases.ed_cases()
    for (name, sig, pk, m), e in zip(ecases, edge_cases.ed_expected(ecases)):
        items.append((2, sig, pk, m, e))
    rng = np.random.default_rng(17)
    order = rng.permutation(len(items) * 6)
    items = [items[k % len(items)] for k in order]
    n = len(items)
    _say(f"ring edge cases n={n}")
    sch = np.array([it[0] for it in items], np.uint8)
    sigs = np.frombuffer(b"".join(it[1] for it in items), np.uint8).reshape(n, 64)
    aux = np.frombuffer(b"".join(it[2] for it in items), np.uint8).reshape(n, 32)
    off = np.concatenate([[0], np.cumsum([len(it[3]) for it in items])]).astype(np.uint64)
    msgs = np.frombuffer(b"".join(it[3] for it in items), np.uint8)
    cuts = list(range(0, n, 300)) + [n]
    with _ring() as srv:
        tickets = [srv.submit(sch[a:b], sigs[a:b], aux[a:b], msgs, off[a:b + 1]) for a, b in zip(cuts, cuts[1:])]
        res = [srv.wait(t) for t in tickets]
        batches = srv.stats()[0]
    out = np.concatenate([r[0] for r in res])
    st = np.concatenate([r[1] for r in res])
    for i, it in enumerate(items):
        if it[0] < 2:
            assert (int(st[i]), bytes(out[i])) == it[4], i
        else:
            assert (st[i] == 0) == it[4], i
    assert batches >= 1


def test_ring_small_ring_wraps():
    """64 positions with 4 KiB of messages each: the 20000 items reuse every position many times"""
    n = 20000
    _say(f"ring wrap n={n}")
    pool = _pool(n, 71)
    req = _requests(n, 19)
    with _ring(ring_tiles=64, ring_msg_bytes=1 << 18) as srv:
        tickets = [srv.submit(*_slice(pool, req[r], req[r + 1])) for r in range(len(req) - 1)]
        out = np.zeros((n, 64), np.uint8)
        st = np.zeros(n, np.uint8)
        for r, t in enumerate(tickets):
            o, s, _, _ = srv.wait(t)
            out[req[r]:req[r + 1]], st[req[r]:req[r + 1]] = o, s
        # a message longer than a position's message slot is refused
        sch = np.array([2], np.uint8)
        with pytest.raises(sigops.SigopsError):
            srv.submit(sch, np.zeros((1, 64), np.uint8), np.zeros((1, 32), np.uint8), np.zeros(20000, np.uint8),
                       np.array([0, 20000], np.uint64))
    _check_pool_results(pool, out, st)


def test_ring_launches_end_when_idle():
    n = 2000
    _say(f"ring idle epochs n={n}")
    pool = _pool(n, 81)
    with _ring(idle_close_us=500) as srv:
        o1, s1, _, _ = srv.wait(srv.submit(*_slice(pool, 0, 1000)))
        b1 = srv.stats()[0]
        time.sleep(0.05)  # the launch ends after 0.5 ms without work
        o2, s2, _, _ = srv.wait(srv.submit(*_slice(pool, 1000, 2000)))
        b2 = srv.stats()[0]
        assert b2 > b1
        # empty request completes at once
        t0 = srv.submit(np.zeros(0, np.uint8), np.zeros((0, 64), np.uint8), np.zeros((0, 32), np.uint8),
                        np.zeros(0, np.uint8), np.zeros(1, np.uint64))
        assert srv.wait(t0)[0].shape == (0, 64)
    _check_pool_results(pool, np.concatenate([o1, o2]), np.concatenate([s1, s2]))


def test_ring_budget_queues_requests():
    n = 3000
    _say(f"ring budget n={n}")
    pool = _pool(n, 41)
    req = _requests(n, 9)
    with _ring(hbm_budget=400_000, max_batch=2048) as srv:
        with pytest.raises(sigops.SigopsError):
            srv.submit(*_slice(pool, 0, 2500))  # larger than max_batch
        kept = [r for r in range(len(req) - 1) if (req[r + 1] - req[r]) * 250 + 20000 < 400_000]
        tickets = [srv.submit(*_slice(pool, req[r], req[r + 1])) for r in kept]
        out = np.zeros((n, 64), np.uint8)
        st = np.full(n, 0xEE, np.uint8)
        for r, t in zip(kept, tickets):
            o, s, _, _ = srv.wait(t)
            out[req[r]:req[r + 1]], st[req[r]:req[r + 1]] = o, s
    mask = np.zeros(n, bool)
    for r in kept:
        mask[req[r]:req[r + 1]] = True
    schemes, sigs, aux, msgs, off, out_exp, st_exp, v, idx = pool
    sec = np.nonzero(mask & (schemes < 2))[0]
    assert np.array_equal(out[sec], out_exp[sec]) and np.array_equal(st[sec], st_exp[sec])
    ed_pos = {int(i): j for j, i in enumerate(idx[2])}
    for i in np.nonzero(mask & (schemes == 2))[0]:
        assert (st[i] == 0) == bool(v[ed_pos[int(i)]])
