"""GPU tests of the continuous-batching server: every request's results are identical whether it
runs alone or mixed with concurrent requests (exact comparison), admission by HBM budget holds
requests back without losing them, and bad requests are rejected at submit."""
import threading

import numpy as np
import pytest

import oracle_lib as O
from sigops_pkg import sigops

pytestmark = pytest.mark.gpu
W = __import__("importlib").import_module("wgpu-sigops_amd.workload")


def _pool(n, seed):
    """a mixed item pool built with the oracle generators, plus the oracle's expected outputs"""
    rng = np.random.default_rng(seed)
    schemes = rng.choice(3, size=n, p=[0.5, 0.3, 0.2]).astype(np.uint8)
    idx = [np.nonzero(schemes == s)[0] for s in range(3)]
    sigs = np.zeros((n, 64), np.uint8)
    aux = np.zeros((n, 32), np.uint8)
    out_exp = np.zeros((n, 64), np.uint8)
    st_exp = np.zeros(n, np.uint8)
    lens = np.zeros(n, np.int64)
    for c in (0, 1):
        s, d, p, cor = O.gen_secp(c, seed + c, len(idx[c]), corrupt_ppm=50000)
        sigs[idx[c]], aux[idx[c]] = s, d
        out_exp[idx[c]], st_exp[idx[c]] = O.recover_batch(c, s, d)
    le = W.loguniform_lengths(len(idx[2]), 2000, seed)
    lens[idx[2]] = le
    off_e = W.msg_offsets(le).astype(np.uint64)
    s, p, m, cor = O.gen_ed25519(seed + 2, len(idx[2]), off_e, corrupt_ppm=50000)
    sigs[idx[2]], aux[idx[2]] = s, p
    v = O.verify_batch(s, p, m, off_e)
    off = W.msg_offsets(lens).astype(np.uint64)
    msgs = np.zeros(int(off[-1]), np.uint8)
    for j, i in enumerate(idx[2]):
        msgs[off[i]:off[i + 1]] = m[off_e[j]:off_e[j + 1]]
    return schemes, sigs, aux, msgs, off, out_exp, st_exp, v, idx


def _requests(n, seed):
    rng = np.random.default_rng(seed)
    sizes, total = [], 0
    while total < n:
        k = int(min(n - total, max(1, np.floor(np.exp(rng.random() * np.log(600))))))
        sizes.append(k)
        total += k
    return np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)


def _check_pool_results(pool, out, st):
    schemes, sigs, aux, msgs, off, out_exp, st_exp, v, idx = pool
    sec = np.concatenate([idx[0], idx[1]])
    assert np.array_equal(out[sec], out_exp[sec]) and np.array_equal(st[sec], st_exp[sec])
    assert np.array_equal(st[idx[2]] == 0, v.astype(bool))


def _slice(pool, a, b):
    schemes, sigs, aux, msgs, off = pool[:5]
    return schemes[a:b], sigs[a:b], aux[a:b], msgs, off[a:b + 1]


def test_requests_alone_and_mixed_are_identical():
    n = 6000
    print(f"serve alone/mixed n={n}", flush=True)
    pool = _pool(n, 31)
    req = _requests(n, 5)
    with sigops.Server(slots=3) as srv:
        alone_out = np.zeros((n, 64), np.uint8)
        alone_st = np.zeros(n, np.uint8)
        for r in range(len(req) - 1):
            a, b = req[r], req[r + 1]
            t = srv.submit(*_slice(pool, a, b))
            o, s, _, _ = srv.wait(t)
            alone_out[a:b], alone_st[a:b] = o, s
        _check_pool_results(pool, alone_out, alone_st)
        batches_alone = srv.stats()[0]
        # mixed: all requests in flight at once from two submitting threads
        tickets = {}

        def submitter(rs):
            for r in rs:
                a, b = req[r], req[r + 1]
                tickets[r] = srv.submit(*_slice(pool, a, b))

        rs = list(range(len(req) - 1))
        th = [threading.Thread(target=submitter, args=(rs[k::2],)) for k in range(2)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        mix_out = np.zeros((n, 64), np.uint8)
        mix_st = np.zeros(n, np.uint8)
        for r in rs:
            o, s, ts, td = srv.wait(tickets[r])
            assert td >= ts
            mix_out[req[r]:req[r + 1]], mix_st[req[r]:req[r + 1]] = o, s
        batches_total = srv.stats()[0]
    assert np.array_equal(mix_out, alone_out) and np.array_equal(mix_st, alone_st)
    # the concurrent requests were batched together (fewer launches than requests)
    assert batches_total - batches_alone < len(rs)


def test_budget_admission_and_rejection():
    n = 3000
    print(f"serve budget n={n}", flush=True)
    pool = _pool(n, 41)
    req = _requests(n, 9)
    # a budget that fits ~2 requests forces queueing; every request still completes exactly
    with sigops.Server(slots=2, hbm_budget=400_000, max_batch=2048) as srv:
        with pytest.raises(sigops.SigopsError):
            srv.submit(*_slice(pool, 0, 2500))  # larger than max_batch
        tickets = [srv.submit(*_slice(pool, req[r], req[r + 1])) for r in range(len(req) - 1)
                   if (req[r + 1] - req[r]) * 250 + 20000 < 400_000]
        kept = [r for r in range(len(req) - 1) if (req[r + 1] - req[r]) * 250 + 20000 < 400_000]
        out = np.zeros((n, 64), np.uint8)
        st = np.full(n, 0xEE, np.uint8)
        for r, t in zip(kept, tickets):
            o, s, _, _ = srv.wait(t)
            out[req[r]:req[r + 1]], st[req[r]:req[r + 1]] = o, s
        # empty request completes at once
        t0 = srv.submit(np.zeros(0, np.uint8), np.zeros((0, 64), np.uint8), np.zeros((0, 32), np.uint8),
                        np.zeros(0, np.uint8), np.zeros(1, np.uint64))
        assert srv.wait(t0)[0].shape == (0, 64)
    mask = np.zeros(n, bool)
    for r in kept:
        mask[req[r]:req[r + 1]] = True
    schemes, sigs, aux, msgs, off, out_exp, st_exp, v, idx = pool
    sec = np.nonzero(mask & (schemes < 2))[0]
    assert np.array_equal(out[sec], out_exp[sec]) and np.array_equal(st[sec], st_exp[sec])
    ed_pos = {int(i): j for j, i in enumerate(idx[2])}
    for i in np.nonzero(mask & (schemes == 2))[0]:
        assert (st[i] == 0) == bool(v[ed_pos[int(i)]])


def test_replay_trace_latencies():
    n = 4000
    print(f"serve replay n={n}", flush=True)
    pool = _pool(n, 51)
    req = _requests(n, 13)
    nreq = len(req) - 1
    arrivals = (np.cumsum(np.random.default_rng(3).exponential(2e5, nreq))).astype(np.uint64)  # 5k req/s
    with sigops.Server(slots=2) as srv:
        lat, out, st = srv.replay(*pool[:5], req.astype(np.uint64), arrivals)
    assert (lat > 0).all() and (lat < 10e9).all()
    _check_pool_results(pool, out, st)
