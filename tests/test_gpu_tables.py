"""GPU tests of caller-supplied fixed-base tables in the reference's layout (`table_limbs`,
src/precompute.rs:12-69), through the C ABI: the reference's own tables give exactly the plain
results; a table of multiples of another point P gives u1 * P + u2 * R (secp) or accepts exactly
the signatures made against base P (Ed25519), as the reference's table-driven shaders compute;
malformed tables are refused. Expected values come from the affine closed forms in pyref."""
import hashlib

import numpy as np
import pytest

import oracle_lib as O
import pyref
from sigops_pkg import sigops

pytestmark = pytest.mark.gpu
P = sigops.precompute


@pytest.fixture(scope="module")
def ctx():
    c = sigops.Context()
    yield c
    c.close()


def _say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("curve", [0, 1])
def test_reference_tables_equal_plain(ctx, curve):
    sigs, dig, pubs, cor = O.gen_secp(curve, 600 + curve, 2048, corrupt_ppm=100000)
    plain_out, plain_st = (ctx.secp256k1_recover if curve == 0 else ctx.secp256r1_recover)(sigs, dig)
    for lbits in (11, 13, 15):
        _say(f"reference table curve={curve} l={lbits}")
        tab = (P.secp256k1_bases if curve == 0 else P.secp256r1_bases)(lbits)
        out, st = ctx.secp_recover_table(curve, sigs, dig, tab, lbits)
        assert np.array_equal(out, plain_out) and np.array_equal(st, plain_st)


def _lift(c, r, odd):
    p = c["p"]
    rhs = (r**3 + c["a"] * r + c["b"]) % p
    y = pow(rhs, (p + 1) // 4, p)
    return (r, y if y % 2 == odd else p - y) if y * y % p == rhs else None


@pytest.mark.parametrize("curve", [0, 1])
def test_table_of_another_point(ctx, curve):
    """u1 * P + u2 * R for the table's point P = 12345 G, on valid and corrupted signatures"""
    c = pyref.CURVES[curve]
    n = c["n"]
    Pt = pyref.w_mul(c, c["G"], 12345)
    sigs, dig, pubs, cor = O.gen_secp(curve, 700 + curve, 48, corrupt_ppm=200000)
    _say(f"table of 12345 G curve={curve}")
    out, st = ctx.secp_recover_table(curve, sigs, dig, P.secp_table_for_point(curve, Pt, 13), 13)
    for i in range(48):
        sig = bytes(sigs[i])
        r = int.from_bytes(sig[:32], "big")
        sb = bytearray(sig[32:])
        odd = sb[0] >> 7
        sb[0] &= 0x7F
        s = int.from_bytes(bytes(sb), "big")
        ref_st, _ = pyref.secp_recover(curve, sig, bytes(dig[i]))  # same acceptance rules
        assert int(st[i]) == ref_st, i
        if ref_st:
            continue
        R = _lift(c, r, odd)
        z = int.from_bytes(bytes(dig[i]), "big") % n
        rinv = pow(r, -1, n)
        Q = pyref.w_add(c, pyref.w_mul(c, Pt, (-z * rinv) % n), pyref.w_mul(c, R, s * rinv % n))
        want = bytes(64) if Q is None else Q[0].to_bytes(32, "big") + Q[1].to_bytes(32, "big")
        assert bytes(out[i]) == want, i


def test_ed25519_tables(ctx):
    """the reference's table reproduces plain verification; a table of B' = 5 B accepts exactly the
    signatures made against B' (s = r + k a, A = a B', R = r B')"""
    rng = np.random.default_rng(9)
    lens = rng.integers(0, 200, 512)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    sigs, pks, msgs, cor = O.gen_ed25519(901, 512, off, corrupt_ppm=100000)
    valid, st = ctx.ed25519_verify(sigs, pks, msgs, off)
    v2, st2 = ctx.ed25519_verify_table(sigs, pks, msgs, off, P.ed25519_bases(13), 13)
    assert np.array_equal(valid, v2) and np.array_equal(st, st2)
    Bp = pyref.e_mul(pyref.EB, 5)
    items = []
    for i in range(12):
        a = int.from_bytes(hashlib.sha256(b"a%d" % i).digest(), "little") % pyref.EL
        r = int.from_bytes(hashlib.sha256(b"r%d" % i).digest(), "little") % pyref.EL
        A = pyref.e_compress(pyref.e_mul(Bp, a))
        R = pyref.e_compress(pyref.e_mul(Bp, r))
        m = b"base %d" % i
        k = int.from_bytes(hashlib.sha512(R + A + m).digest(), "little") % pyref.EL
        S = (r + k * a) % pyref.EL
        sig = bytearray(R + S.to_bytes(32, "little"))
        if i % 4 == 3:
            sig[40] ^= 4  # corrupted: rejected under either base
        items.append((bytes(sig), A, m, i % 4 != 3))
    s_ = np.frombuffer(b"".join(t[0] for t in items), np.uint8).reshape(-1, 64)
    a_ = np.frombuffer(b"".join(t[1] for t in items), np.uint8).reshape(-1, 32)
    o_ = np.zeros(len(items) + 1, np.uint64)
    o_[1:] = np.cumsum([len(t[2]) for t in items])
    m_ = np.frombuffer(b"".join(t[2] for t in items), np.uint8)
    vb, _ = ctx.ed25519_verify_table(s_, a_, m_, o_, P.ed25519_table_for_point(Bp, 13), 13)
    assert list(vb) == [t[3] for t in items]
    vplain, _ = ctx.ed25519_verify(s_, a_, m_, o_)
    assert not vplain.any()  # the same signatures do not verify against B


def test_malformed_tables_refused(ctx):
    sigs, dig, pubs, cor = O.gen_secp(0, 5, 4)
    good = np.array(P.secp256k1_bases(13), np.uint32)
    for name, t in (("zero", np.zeros_like(good)), ("limb", good.copy()), ("swap", good.copy())):
        if name == "limb":
            t[3] |= 1 << 13
        if name == "swap":
            t[80:120], t[120:160] = good[120:160], good[80:120]
        with pytest.raises(sigops.SigopsError):
            ctx.secp_recover_table(0, sigs, dig, t, 13)
    with pytest.raises(sigops.SigopsError):
        ctx.secp_recover_table(0, sigs, dig, good[:-1], 13)  # wrong length
    with pytest.raises(sigops.SigopsError):
        ctx.secp_recover_table(0, sigs, dig, good, 10)  # limb size outside 11..15
    T = pyref.ed25519_torsion()[-1]
    with pytest.raises(sigops.SigopsError):
        ctx.ed25519_verify_table(np.zeros((1, 64), np.uint8), np.zeros((1, 32), np.uint8), np.zeros(0, np.uint8),
                                 np.zeros(2, np.uint64), P.ed25519_table_for_point(pyref.e_add(pyref.EB, T), 13), 13)
    # the context still works after refusals
    out, st = ctx.secp_recover_table(0, sigs, dig, good, 13)
    assert np.array_equal(out, pubs)
