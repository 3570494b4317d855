"""GPU tests of the mixed-scheme batch (one persistent kernel, graph-captured) and of the device
workload generator. Expected values come from the oracle, the closed forms and the generators'
round-trip property; every comparison is exact."""
import numpy as np
import pytest
import torch

import edge_cases
import oracle_lib as O
from sigops_pkg import sigops

pytestmark = pytest.mark.gpu
W = __import__("importlib").import_module("wgpu-sigops_amd.workload")


@pytest.fixture(scope="module")
def ctx():
    c = sigops.Context()
    yield c
    c.close()


@pytest.fixture(scope="module")
def ctx_eager():
    c = sigops.Context(use_graph=False)
    yield c
    c.close()


def _say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("curve", [0, 1])
def test_device_generator_matches_oracle_secp(ctx, curve):
    n = 2048
    _say(f"gen secp curve={curve} n={n}")
    it = W.secp(ctx, curve, n, seed=321, corrupt_ppm=100000)
    torch.cuda.synchronize()
    sigs, dig, pubs, cor = O.gen_secp(curve, 321, n, corrupt_ppm=100000)
    assert np.array_equal(it.sigs.cpu().numpy(), sigs)
    assert np.array_equal(it.aux.cpu().numpy(), dig)
    assert np.array_equal(it.pubs.cpu().numpy(), pubs)
    assert np.array_equal(it.corrupted.cpu().numpy(), cor)


def test_device_generator_matches_oracle_ed25519(ctx):
    n = 1024
    lens = W.uniform_lengths(n, 700, 5)
    _say(f"gen ed25519 n={n}")
    it = W.ed25519(ctx, n, lens, seed=654, corrupt_ppm=100000)
    torch.cuda.synchronize()
    off = it.msg_off_host
    sigs, pks, msgs, cor = O.gen_ed25519(654, n, off.astype(np.uint64), corrupt_ppm=100000)
    assert np.array_equal(it.sigs.cpu().numpy(), sigs)
    assert np.array_equal(it.aux.cpu().numpy(), pks)
    assert np.array_equal(it.msgs.cpu().numpy()[: int(off[-1])], msgs)
    assert np.array_equal(it.corrupted.cpu().numpy(), cor)


def _oracle_mixed(n, seed):
    """mixed batch assembled from the oracle generators; returns arrays and expected outputs"""
    rng = np.random.default_rng(seed)
    schemes = rng.choice(3, size=n, p=[0.5, 0.3, 0.2]).astype(np.uint8)
    idx = [np.nonzero(schemes == s)[0] for s in range(3)]
    sigs = np.zeros((n, 64), np.uint8)
    aux = np.zeros((n, 32), np.uint8)
    out_exp = np.zeros((n, 64), np.uint8)
    st_exp = np.zeros(n, np.uint8)
    lens = np.zeros(n, np.int64)
    for curve in (0, 1):
        s, d, p, c = O.gen_secp(curve, seed + curve, len(idx[curve]), corrupt_ppm=50000)
        sigs[idx[curve]] = s
        aux[idx[curve]] = d
        o, st = O.recover_batch(curve, s, d)
        out_exp[idx[curve]] = o
        st_exp[idx[curve]] = st
    le = W.loguniform_lengths(len(idx[2]), 4096, seed)
    lens[idx[2]] = le
    off_e = W.msg_offsets(le).astype(np.uint64)
    s, p, m, c = O.gen_ed25519(seed + 2, len(idx[2]), off_e, corrupt_ppm=50000)
    sigs[idx[2]] = s
    aux[idx[2]] = p
    valid = O.verify_batch(s, p, m, off_e)
    off = W.msg_offsets(lens).astype(np.uint64)
    msgs = np.zeros(int(off[-1]), np.uint8)
    for j, i in enumerate(idx[2]):
        msgs[off[i]:off[i + 1]] = m[off_e[j]:off_e[j + 1]]
    return schemes, sigs, aux, msgs, off, out_exp, st_exp, idx, valid


def test_batch_matches_oracle(ctx, ctx_eager):
    n = 12000
    _say(f"batch mixed n={n}")
    schemes, sigs, aux, msgs, off, out_exp, st_exp, idx, valid = _oracle_mixed(n, 900)
    for c in (ctx, ctx_eager):
        out, st = c.batch(schemes, sigs, aux, msgs, off)
        sec = np.concatenate([idx[0], idx[1]])
        assert np.array_equal(out[sec], out_exp[sec])
        assert np.array_equal(st[sec], st_exp[sec])
        assert np.array_equal(st[idx[2]] == 0, valid.astype(bool))
        assert not out[idx[2]].any()


def test_batch_graph_replays_match_eager(ctx, ctx_eager):
    """the captured graph replayed several times on different data, with eager work in between"""
    data = [_oracle_mixed(3000, s) for s in (11, 12)]
    for rep in range(3):
        d = data[rep % 2]
        out, st = ctx.batch(*d[:5])
        out_e, st_e = ctx_eager.batch(*d[:5])
        assert np.array_equal(out, out_e) and np.array_equal(st, st_e), rep
        sec = np.concatenate([d[7][0], d[7][1]])
        assert np.array_equal(out[sec], d[5][sec]) and np.array_equal(st[sec], d[6][sec])
        # eager, non-batch work between the replays
        ctx.secp256k1_recover(np.asarray(d[1][:64]), np.asarray(d[2][:64]))
        torch.cuda.synchronize()


def test_batch_stage_timing_path(ctx):
    """the eager, event-timed launch path gives the same results and sane stage durations"""
    schemes, sigs, aux, msgs, off, out_exp, st_exp, idx, valid = _oracle_mixed(4000, 31)
    ref_out, ref_st = ctx.batch(schemes, sigs, aux, msgs, off)
    ctx.set_stage_timing(True)
    try:
        out, st = ctx.batch(schemes, sigs, aux, msgs, off)
        ms = ctx.batch_stage_ms()
    finally:
        ctx.set_stage_timing(False)
    assert np.array_equal(out, ref_out) and np.array_equal(st, ref_st)
    sec = np.concatenate([idx[0], idx[1]])
    assert np.array_equal(out[sec], out_exp[sec]) and np.array_equal(st[sec], st_exp[sec])
    assert set(ms) == set(sigops.Context.STAGES)
    assert all(v >= 0 for v in ms.values()) and ms["verify_kernel"] > 0
    # graph path again after switching timing off
    out2, st2 = ctx.batch(schemes, sigs, aux, msgs, off)
    assert np.array_equal(out2, ref_out) and np.array_equal(st2, ref_st)


def test_autotune_variants_agree():
    """every variant the autotuner tries (waves per SIMD, tile rounds, bucket order) gives the same
    outputs (it raises otherwise), and a decision is recorded"""
    at = __import__("importlib").import_module("wgpu-sigops_amd.autotune")
    entry = at.tune(n=8192, steps=1, rounds=2, write=False, log=_say)
    assert len(entry["results"]) == len(at.CANDIDATES)
    assert all(r["ms_median"] > 0 for r in entry["results"])
    assert {k: entry[k] for k in at.DEFAULT} in list(at.CANDIDATES)


@pytest.fixture(scope="module")
def ctx9():
    c = sigops.Context(tuned=False, g_window=9)
    yield c
    c.close()


@pytest.fixture(scope="module")
def ctx8():
    c = sigops.Context(tuned=False, g_window=8)
    yield c
    c.close()


@pytest.mark.parametrize("g_window", [8, 9])
def test_batch_windows_match_oracle(ctx8, ctx9, g_window):
    """the mixed-batch variants with 8- and 9-bit fixed-base windows against the oracle"""
    n = 12000
    _say(f"batch mixed, {g_window}-bit G windows, n={n}")
    schemes, sigs, aux, msgs, off, out_exp, st_exp, idx, valid = _oracle_mixed(n, 901)
    out, st = (ctx9 if g_window == 9 else ctx8).batch(schemes, sigs, aux, msgs, off)
    sec = np.concatenate([idx[0], idx[1]])
    assert np.array_equal(out[sec], out_exp[sec])
    assert np.array_equal(st[sec], st_exp[sec])
    assert np.array_equal(st[idx[2]] == 0, valid.astype(bool))


@pytest.mark.parametrize("g_window", [8, 9])
def test_batch_edge_cases(ctx8, ctx9, g_window):
    """secp edge cases of both curves and the Ed25519 edge cases in one interleaved batch (fixed-base
    windows of 8 and 9 bits)"""
    ctx = ctx9 if g_window == 9 else ctx8
    items = []
    for curve in (0, 1):
        cases = edge_cases.secp_cases(curve)
        for (name, sig, d), exp in zip(cases, edge_cases.secp_expected(curve, cases)):
            items.append((curve, sig, d, b"", exp))
    ec = edge_cases.ed_cases()
    for (name, sig, pk, m), e in zip(ec, edge_cases.ed_expected(ec)):
        items.append((2, sig, pk, m, e))
    order = np.random.default_rng(3).permutation(len(items))
    items = [items[i] for i in order]
    n = len(items)
    _say(f"batch edge cases n={n}")
    schemes = np.array([it[0] for it in items], np.uint8)
    sigs = np.frombuffer(b"".join(it[1] for it in items), np.uint8).reshape(n, 64)
    aux = np.frombuffer(b"".join(it[2] for it in items), np.uint8).reshape(n, 32)
    off = W.msg_offsets([len(it[3]) for it in items]).astype(np.uint64)
    msgs = np.frombuffer(b"".join(it[3] for it in items), np.uint8)
    out, st = ctx.batch(schemes, sigs, aux, msgs, off)
    for i, it in enumerate(items):
        if it[0] < 2:
            assert (int(st[i]), bytes(out[i])) == it[4], i
        else:
            assert (st[i] == 0) == it[4], i


def test_batch_rejects_unknown_scheme(ctx):
    with pytest.raises(sigops.SigopsError):
        ctx.batch(np.array([0, 7], np.uint8), np.zeros((2, 64), np.uint8), np.zeros((2, 32), np.uint8),
                  np.zeros(0, np.uint8), np.zeros(3, np.uint64))


def test_batch_full_size_mixed_properties(ctx):
    """2^20 mixed items generated on the device: every uncorrupted secp item recovers the
    generator's key, every uncorrupted Ed25519 item verifies, corrupted items do not reproduce the
    original; 4096 sampled items are compared with the oracle exactly."""
    n = 1 << 20
    _say(f"batch full size n={n}")
    mx = W.mixed(ctx, n, seed=5)
    out = torch.zeros((n, 64), dtype=torch.uint8, device="cuda:0")
    st = torch.full((n,), 0xEE, dtype=torch.uint8, device="cuda:0")
    ctx.batch_dev(mx.schemes.data_ptr(), mx.sigs.data_ptr(), mx.aux.data_ptr(), mx.msgs.data_ptr(),
                  mx.msg_off.data_ptr(), n, out.data_ptr(), st.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    out_h, st_h, cor = out.cpu().numpy(), st.cpu().numpy(), mx.corrupted.cpu().numpy()
    sch = mx.schemes_host
    assert (st_h != 0xEE).all()
    for s in (0, 1):
        ii = mx.idx[s]
        pubs = mx.parts[s].pubs.cpu().numpy()
        ok = cor[ii] == 0
        assert np.array_equal(out_h[ii][ok], pubs[ok]) and not st_h[ii][ok].any()
        bad = ~ok
        assert not (out_h[ii][bad] == pubs[bad]).all(axis=1).any()
    ii = mx.idx[2]
    assert np.array_equal(st_h[ii] == 0, cor[ii] == 0)
    # exact comparison with the oracle on a sample (corrupted items included)
    rng = np.random.default_rng(0)
    sample = np.sort(rng.choice(n, 4096, replace=False))
    sigs_h = mx.sigs.cpu().numpy()
    aux_h = mx.aux.cpu().numpy()
    msgs_h = mx.msgs.cpu().numpy()
    for s in (0, 1):
        ss = sample[sch[sample] == s]
        o, t = O.recover_batch(s, sigs_h[ss], aux_h[ss])
        assert np.array_equal(o, out_h[ss]) and np.array_equal(t, st_h[ss])
    ss = sample[sch[sample] == 2]
    offs = mx.msg_off_host
    lens = offs[ss + 1] - offs[ss]
    sub_off = W.msg_offsets(lens).astype(np.uint64)
    sub_msgs = np.concatenate([msgs_h[offs[i]:offs[i + 1]] for i in ss]) if len(ss) else np.zeros(0, np.uint8)
    v = O.verify_batch(sigs_h[ss], aux_h[ss], sub_msgs, sub_off)
    assert np.array_equal(v.astype(bool), st_h[ss] == 0)


def test_batch_max_size_properties(ctx):
    """a batch of the context's full capacity (max_batch = 2^22 items, messages up to 256 bytes):
    every uncorrupted secp item recovers the generator's key, every uncorrupted Ed25519 item
    verifies and no corrupted item passes; 2048 sampled items match the oracle exactly. One item
    more is refused before anything runs."""
    n = 1 << 22
    _say(f"batch max size n={n}")
    mx = W.mixed(ctx, n, seed=9, max_len=256)
    out = torch.zeros((n, 64), dtype=torch.uint8, device="cuda:0")
    st = torch.full((n,), 0xEE, dtype=torch.uint8, device="cuda:0")
    s = torch.cuda.current_stream().cuda_stream
    ctx.batch_dev(mx.schemes.data_ptr(), mx.sigs.data_ptr(), mx.aux.data_ptr(), mx.msgs.data_ptr(),
                  mx.msg_off.data_ptr(), n, out.data_ptr(), st.data_ptr(), s)
    torch.cuda.synchronize()
    out_h, st_h, cor = out.cpu().numpy(), st.cpu().numpy(), mx.corrupted.cpu().numpy()
    assert (st_h != 0xEE).all()
    for sc in (0, 1):
        ii = mx.idx[sc]
        pubs = mx.parts[sc].pubs.cpu().numpy()
        ok = cor[ii] == 0
        assert np.array_equal(out_h[ii][ok], pubs[ok]) and not st_h[ii][ok].any()
        assert not (out_h[ii][~ok] == pubs[~ok]).all(axis=1).any()
    ii = mx.idx[2]
    assert np.array_equal(st_h[ii] == 0, cor[ii] == 0)
    rng = np.random.default_rng(1)
    sample = np.sort(rng.choice(n, 2048, replace=False))
    sch = mx.schemes_host
    d_sample = torch.from_numpy(sample).cuda()
    sigs_h, aux_h = mx.sigs[d_sample].cpu().numpy(), mx.aux[d_sample].cpu().numpy()
    for sc in (0, 1):
        sel = sch[sample] == sc
        o, t = O.recover_batch(sc, sigs_h[sel], aux_h[sel])
        assert np.array_equal(o, out_h[sample[sel]]) and np.array_equal(t, st_h[sample[sel]])
    with pytest.raises(sigops.SigopsError):
        ctx.batch_dev(mx.schemes.data_ptr(), mx.sigs.data_ptr(), mx.aux.data_ptr(), mx.msgs.data_ptr(),
                      mx.msg_off.data_ptr(), n + 1, out.data_ptr(), st.data_ptr(), s)


def test_long_messages_in_batch_and_plain(ctx):
    """Ed25519 messages from 0 bytes up to 200 KB (the batch's last bucket holds 40 or more SHA-512
    blocks, so its lanes hash very different lengths), interleaved with secp items, through the
    mixed batch and through the plain entry point; every result equals the oracle's"""
    lens = np.array([0, 1, 111, 112, 4990, 5000, 5100, 8192, 20000, 65537, 100000, 200000] * 6
                    + list(np.random.default_rng(5).integers(0, 300, 60)), np.int64)
    ne = len(lens)
    off_e = W.msg_offsets(lens).astype(np.uint64)
    _say(f"long messages n_ed={ne} bytes={int(off_e[-1])}")
    esigs, epks, emsgs, ecor = O.gen_ed25519(4040, ne, off_e, corrupt_ppm=200000)
    ref_e = O.verify_batch(esigs, epks, emsgs, off_e)
    valid, st = ctx.ed25519_verify(esigs, epks, emsgs, off_e)
    assert np.array_equal(valid, ref_e.astype(bool))
    assert np.array_equal(valid, ecor == 0)
    # the same items in a mixed batch with secp256k1 items
    ssigs, sdig, spubs, scor = O.gen_secp(0, 4041, 100, corrupt_ppm=100000)
    n = ne + 100
    order = np.random.default_rng(6).permutation(n)
    is_ed = order < ne
    schemes = np.where(is_ed, 2, 0).astype(np.uint8)
    sigs = np.where(is_ed[:, None], esigs[np.minimum(order, ne - 1)], ssigs[np.maximum(order - ne, 0)])
    aux = np.where(is_ed[:, None], epks[np.minimum(order, ne - 1)], sdig[np.maximum(order - ne, 0)])
    blens = np.where(is_ed, lens[np.minimum(order, ne - 1)], 0)
    boff = W.msg_offsets(blens).astype(np.uint64)
    bmsgs = np.concatenate([emsgs[off_e[j]:off_e[j + 1]] for j in order[is_ed]])
    out, bst = ctx.batch(schemes, sigs, aux, bmsgs, boff)
    for i in range(n):
        j = order[i]
        if is_ed[i]:
            assert (bst[i] == 0) == bool(ref_e[j]), i
        else:
            o, t = O.recover_batch(0, ssigs[j - ne:j - ne + 1], sdig[j - ne:j - ne + 1])
            assert bst[i] == t[0] and bytes(out[i]) == bytes(o[0]), i
