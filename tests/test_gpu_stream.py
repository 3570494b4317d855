"""GPU tests of the streamed host batch (sigops_batch with streaming on, the default for n >= 4096):
the host packs chunks into pinned memory while one persistent launch drains them. Every result must
equal the copy-then-launch path's and the oracle's: mixed batches with corrupted items and long
messages, the recovery and Ed25519 edge cases (exceptional additions go through the careful pass
of the streamed launch), batches of one scheme, consecutive batches of different sizes, and the
canary error path."""
import numpy as np
import pytest

import edge_cases
import oracle_lib as O
from sigops_pkg import sigops

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def ctx():
    c = sigops.Context()
    yield c
    c.close()


@pytest.fixture(scope="module", params=[8, 9], ids=["g8", "g9"])
def ctxw(request):
    """a context with either fixed-base window of the secp streams (the streamed kernel is built
    for both)"""
    c = sigops.Context(tuned=False, g_window=request.param)
    yield c
    c.close()


def _mixed(n, seed, corrupt_ppm=100000, max_len=600, long_every=0):
    """host arrays of a mixed batch made by the oracle's generators (one third per scheme)"""
    rng = np.random.default_rng(seed)
    sch = rng.integers(0, 3, n).astype(np.uint8)
    sigs = np.zeros((n, 64), np.uint8)
    aux = np.zeros((n, 32), np.uint8)
    lens = np.zeros(n, np.int64)
    ed = np.nonzero(sch == 2)[0]
    lens[ed] = rng.integers(0, max_len + 1, len(ed))
    if long_every:
        lens[ed[::long_every]] = rng.integers(4000, 20000, len(ed[::long_every]))
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    for c in (0, 1):
        ii = np.nonzero(sch == c)[0]
        s, d, _, _ = O.gen_secp(c, seed * 7 + c, len(ii), corrupt_ppm=corrupt_ppm)
        sigs[ii], aux[ii] = s, d
    e_off = np.concatenate([[0], np.cumsum(lens[ed])]).astype(np.uint64)
    s, p, m, _ = O.gen_ed25519(seed * 7 + 2, len(ed), e_off, corrupt_ppm=corrupt_ppm)
    sigs[ed], aux[ed] = s, p
    msgs = np.zeros(int(off[-1]), np.uint8)
    for k, i in enumerate(ed):
        msgs[off[i]:off[i + 1]] = m[e_off[k]:e_off[k + 1]]
    return sch, sigs, aux, msgs, off


def _expect(sch, sigs, aux, msgs, off):
    out = np.zeros((len(sch), 64), np.uint8)
    st = np.zeros(len(sch), np.uint8)
    for c in (0, 1):
        ii = np.nonzero(sch == c)[0]
        o, t = O.recover_batch(c, sigs[ii], aux[ii])
        out[ii], st[ii] = o, t
    ii = np.nonzero(sch == 2)[0]
    sub_off = np.concatenate([[0], np.cumsum(off[ii + 1] - off[ii])]).astype(np.uint64)
    sub = np.concatenate([msgs[off[i]:off[i + 1]] for i in ii]) if len(ii) else np.zeros(0, np.uint8)
    v = O.verify_batch(sigs[ii], aux[ii], sub, sub_off)
    return out, st, ii, v.astype(bool)


def _both(ctx, *arrs):
    ctx.set_streaming(True)
    a = ctx.batch(*arrs)
    ctx.set_streaming(False)
    b = ctx.batch(*arrs)
    ctx.set_streaming(True)
    return a, b


def test_streamed_mixed_matches_copy_path_and_oracle(ctxw):
    n = 60000
    _say(f"streamed mixed n={n}, {ctxw.g_window}-bit G windows")
    arrs = _mixed(n, 3, long_every=97)
    (so, ss), (co, cs) = _both(ctxw, *arrs)
    assert np.array_equal(so, co) and np.array_equal(ss, cs)
    out, st, ii, v = _expect(*arrs)
    secp = arrs[0] < 2
    assert np.array_equal(so[secp], out[secp]) and np.array_equal(ss[secp], st[secp])
    assert np.array_equal(ss[ii] == 0, v) and not so[ii].any()


def test_streamed_edge_cases(ctxw):
    """every recovery and Ed25519 edge case, repeated to a streamed size among valid items"""
    ctx = ctxw
    items = []
    for c in (0, 1):
        cases = edge_cases.secp_cases(c)
        for (name, sig, dig), e in zip(cases, edge_cases.secp_expected(c, cases)):
            items.append((c, sig, dig, b"", e))
    ecases = edge_cases.ed_cases()
    for (name, sig, pk, m), e in zip(ecases, edge_cases.ed_expected(ecases)):
        items.append((2, sig, pk, m, e))
    reps = 4096 // len(items) + 2
    rng = np.random.default_rng(9)
    order = rng.permutation(len(items) * reps)
    items = [items[k % len(items)] for k in order]
    n = len(items)
    _say(f"streamed edge cases n={n}, {ctx.g_window}-bit G windows")
    sch = np.array([it[0] for it in items], np.uint8)
    sigs = np.frombuffer(b"".join(it[1] for it in items), np.uint8).reshape(n, 64)
    aux = np.frombuffer(b"".join(it[2] for it in items), np.uint8).reshape(n, 32)
    off = np.concatenate([[0], np.cumsum([len(it[3]) for it in items])]).astype(np.uint64)
    msgs = np.frombuffer(b"".join(it[3] for it in items), np.uint8)
    ctx.set_streaming(True)
    out, st = ctx.batch(sch, sigs, aux, msgs, off)
    for i, it in enumerate(items):
        if it[0] < 2:
            assert (int(st[i]), bytes(out[i])) == it[4], i
        else:
            assert (st[i] == 0) == it[4], i


@pytest.mark.parametrize("scheme", [0, 1, 2])
def test_streamed_single_scheme(ctx, scheme):
    n = 15000
    _say(f"streamed single scheme {scheme} n={n}")
    sch, sigs, aux, msgs, off = _mixed(n, 10 + scheme)
    keep = sch == scheme
    sch, sigs, aux = sch[keep], sigs[keep], aux[keep]
    lens = (off[1:] - off[:-1])[keep]
    idx = np.nonzero(keep)[0]
    new_off = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    new_msgs = np.concatenate([msgs[off[i]:off[i + 1]] for i in idx]) if scheme == 2 else np.zeros(0, np.uint8)
    arrs = (sch, sigs, aux, new_msgs, new_off)
    if len(sch) < 4096:
        pytest.skip("too few items for the streamed path")
    (so, ss), (co, cs) = _both(ctx, *arrs)
    assert np.array_equal(so, co) and np.array_equal(ss, cs)


def test_streamed_consecutive_sizes(ctx):
    """batches of different sizes in a row reuse the pinned layout (a new batch tag each time)"""
    for k, n in enumerate((30000, 5000, 45000, 4096)):
        _say(f"streamed consecutive {k}: n={n}")
        arrs = _mixed(n, 20 + k, max_len=300)
        (so, ss), (co, cs) = _both(ctx, *arrs)
        assert np.array_equal(so, co) and np.array_equal(ss, cs), k


def test_streamed_canary(ctx):
    arrs = _mixed(6000, 31)
    ctx.set_fault(sigops.FAULT_SKIP_COMPUTE)
    try:
        with pytest.raises(sigops.ShaderFailureError):
            ctx.batch(*arrs)
    finally:
        ctx.set_fault(0)
    out, st = ctx.batch(*arrs)
    (so, ss), _ = _both(ctx, *arrs)
    assert np.array_equal(out, so) and np.array_equal(st, ss)
