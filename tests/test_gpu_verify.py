"""GPU parity of the predicates next to the path, through the C ABI: ECDSA verification with a
known public key (fuel Signature::verify; libsecp256k1 rules for secp256k1, p256 rules for
secp256r1) and Ed25519 verify_strict (ed25519-dalek 2.1). Every status code equals the oracle's,
on generated signatures with corruption, wrong keys and high-S variants, and on the hand-built
edge cases whose expected values come from the pure-Python closed forms."""
import numpy as np
import pytest

import edge_cases
import oracle_lib as O
import pyref
from sigops_pkg import sigops

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = sigops.Context()
    yield c
    c.close()


@pytest.fixture(scope="module", params=[8, 9], ids=["g8", "g9"])
def ctxw(request):
    """a context with either fixed-base window (the per-scheme secp kernels are built for both)"""
    c = sigops.Context(tuned=False, g_window=request.param)
    yield c
    c.close()


def _say(*a):
    print(*a, flush=True)


def _mutate(curve, sigs, pubs, seed):
    """wrong keys for 5% of the items and, where fuel's format can hold it, s -> n - s for 5%"""
    rng = np.random.default_rng(seed)
    n = sigs.shape[0]
    sigs, pubs = sigs.copy(), pubs.copy()
    wrong = rng.random(n) < 0.05
    pubs[wrong] = np.roll(pubs, 1, axis=0)[wrong]
    order = pyref.CURVES[curve]["n"]
    for i in np.nonzero(rng.random(n) < 0.05)[0]:
        s = int.from_bytes(bytes([sigs[i, 32] & 0x7F]) + sigs[i, 33:].tobytes(), "big")
        t = order - s
        if t < 2**255:
            sigs[i, 32:] = np.frombuffer(t.to_bytes(32, "big"), np.uint8)
    return sigs, pubs


@pytest.mark.parametrize("curve,n", [(0, 4096), (1, 8192)])
def test_secp_verify_matches_oracle(ctxw, curve, n):
    _say(f"secp verify curve={curve} n={n}, {ctxw.g_window}-bit G windows")
    sigs, dig, pubs, cor = O.gen_secp(curve, 2000 + curve, n, corrupt_ppm=100000)
    sigs, pubs = _mutate(curve, sigs, pubs, 5 + curve)
    fn = ctxw.secp256k1_verify if curve == 0 else ctxw.secp256r1_verify
    valid, st = fn(sigs, dig, pubs)
    ref = O.secp_verify_batch(curve, sigs, dig, pubs)
    assert np.array_equal(st, ref)
    assert np.array_equal(valid, ref == 0)
    assert len(set(ref.tolist())) >= 2 and (ref == 0).mean() > 0.7


@pytest.mark.parametrize("curve", [0, 1])
def test_secp_verify_edge_cases(ctxw, curve):
    ctx = ctxw
    cases = edge_cases.secp_verify_cases(curve)
    exp = edge_cases.secp_verify_expected(curve, cases)
    _say(f"secp verify edge cases curve={curve} n={len(cases)}, {ctx.g_window}-bit G windows")
    sigs = np.frombuffer(b"".join(c[1] for c in cases), np.uint8).reshape(-1, 64)
    digs = np.frombuffer(b"".join(c[2] for c in cases), np.uint8).reshape(-1, 32)
    pubs = np.frombuffer(b"".join(c[3] for c in cases), np.uint8).reshape(-1, 64)
    fn = ctx.secp256k1_verify if curve == 0 else ctx.secp256r1_verify
    valid, st = fn(sigs, digs, pubs)
    for i, (name, _, _, _) in enumerate(cases):
        assert int(st[i]) == exp[i], name
        assert bool(valid[i]) == (exp[i] == 0), name


@pytest.mark.parametrize("curve", [0, 1])
def test_secp_verify_dev_65536(ctx, curve):
    """device-generated signatures at the reference's N = 65536, verified on the device; every
    uncorrupted item is valid and the corrupted ones agree with the oracle"""
    import torch

    n = 65536
    _say(f"secp verify_dev curve={curve} n={n}")
    d = torch.device("cuda:0")
    sigs = torch.empty((n, 64), dtype=torch.uint8, device=d)
    digs = torch.empty((n, 32), dtype=torch.uint8, device=d)
    pubs = torch.empty((n, 64), dtype=torch.uint8, device=d)
    cor = torch.empty(n, dtype=torch.uint8, device=d)
    st = torch.full((n,), 0xEE, dtype=torch.uint8, device=d)
    ctx.gen_secp_dev(curve, 4242 + curve, n, 50000, sigs.data_ptr(), digs.data_ptr(), pubs.data_ptr(), cor.data_ptr())
    fn = ctx.secp256k1_verify_dev if curve == 0 else ctx.secp256r1_verify_dev
    fn(sigs.data_ptr(), digs.data_ptr(), pubs.data_ptr(), n, st.data_ptr())
    torch.cuda.synchronize()
    st_h, cor_h = st.cpu().numpy(), cor.cpu().numpy()
    assert (st_h[cor_h == 0] == 0).all()
    idx = np.nonzero(cor_h)[0]
    ref = O.secp_verify_batch(curve, sigs.cpu().numpy()[idx], digs.cpu().numpy()[idx], pubs.cpu().numpy()[idx])
    assert np.array_equal(st_h[idx], ref)


def test_ed25519_verify_strict_cases(ctx):
    cases = edge_cases.ed_strict_cases()
    exp = edge_cases.ed_strict_expected(cases)
    _say(f"ed25519 verify_strict edge cases n={len(cases)}")
    sigs = np.frombuffer(b"".join(c[1] for c in cases), np.uint8).reshape(-1, 64)
    pks = np.frombuffer(b"".join(c[2] for c in cases), np.uint8).reshape(-1, 32)
    msgs = [c[3] for c in cases]
    off = np.zeros(len(cases) + 1, np.uint64)
    off[1:] = np.cumsum([len(m) for m in msgs])
    data = np.frombuffer(b"".join(msgs), np.uint8)
    valid, st = ctx.ed25519_verify_strict(sigs, pks, data, off)
    plain, _ = ctx.ed25519_verify(sigs, pks, data, off)
    for i, (name, sig, pk, m) in enumerate(cases):
        assert bool(valid[i]) == exp[i], name
        assert bool(plain[i]) == O.ed25519_verify(sig, pk, m), name
    assert {0, 1, 2, 3, 4, 5} <= set(st.tolist())
    assert (plain & ~valid).sum() >= 8  # cases only the strict rules reject


def test_ed25519_verify_strict_matches_oracle(ctx):
    n = 8192
    _say(f"ed25519 verify_strict n={n}")
    rng = np.random.default_rng(31)
    lens = rng.integers(0, 600, n)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    sigs, pks, msgs, cor = O.gen_ed25519(3131, n, off, corrupt_ppm=100000)
    valid, st = ctx.ed25519_verify_strict(sigs, pks, msgs, off)
    ref = O.verify_strict_batch(sigs, pks, msgs, off)
    assert np.array_equal(valid, ref.astype(bool))
    assert np.array_equal(valid, cor == 0)  # honest keys and R are never of small order


def test_mirrors(ctx):
    sigs, dig, pubs, cor = O.gen_secp(0, 55, 64, corrupt_ppm=200000)
    got = sigops.secp256k1_ecdsa.verify(list(sigs), list(dig), list(pubs), ctx=ctx)
    assert got == [O.secp_verify(0, bytes(sigs[i]), bytes(dig[i]), bytes(pubs[i])) == 0 for i in range(64)]
    sigs, dig, pubs, cor = O.gen_secp(1, 56, 64, corrupt_ppm=200000)
    ok, st = sigops.secp256r1_ecdsa.verify_with_status(sigs, dig, pubs, ctx=ctx)
    assert list(st) == [O.secp_verify(1, bytes(sigs[i]), bytes(dig[i]), bytes(pubs[i])) for i in range(64)]
    assert sigops.secp256k1_ecdsa.verify([], [], [], ctx=ctx) == []
    with pytest.raises(ValueError):
        sigops.secp256k1_ecdsa.verify(list(sigs), list(dig), [], ctx=ctx)
    cases = edge_cases.ed_strict_cases()[:40]
    got = sigops.ed25519_eddsa.ecverify_strict([c[1] for c in cases], [c[3] for c in cases], [c[2] for c in cases],
                                               ctx=ctx)
    assert got == edge_cases.ed_strict_expected(cases)
