"""The device algorithms (wgpu-sigops_amd/csrc/*.hpp) compiled for the host with g++ and compared
with the oracle and the closed forms, so that a logic error is caught on the CPU before any GPU
run. The GPU tests (test_gpu_*.py) then check the HIP build of the same code."""
import hashlib
import random

import numpy as np
import pytest

import devcheck_lib as D
import edge_cases
import oracle_lib as O
import pyref

MODULI = [pyref.CURVES[0]["p"], pyref.CURVES[0]["n"], pyref.CURVES[1]["p"], pyref.CURVES[1]["n"], pyref.EP, pyref.EL]


@pytest.mark.parametrize("mod", range(6))
def test_field_ops(mod):
    m = MODULI[mod]
    rnd = random.Random(mod)
    for a, b in [(0, 0), (m - 1, m - 1), (1, m - 1)] + [(rnd.randrange(m), rnd.randrange(m)) for _ in range(60)]:
        assert D.field(0, mod, a, b) == a * b % m
        assert D.field(1, mod, a) == a * a % m
        assert D.field(3, mod, a, b) == (a + b) % m
        assert D.field(4, mod, a, b) == (a - b) % m
        if a:
            assert D.field(2, mod, a) == pow(a, -1, m)


@pytest.mark.parametrize("mod", range(6))
def test_inversion_divsteps_edge_values(mod):
    """the divstep inversion (op 2) and the Fermat chain (op 5) against pow(x, -1, m) on values that
    stress the recurrence: tiny, near m, powers of two, halves, and random"""
    m = MODULI[mod]
    rnd = random.Random(50 + mod)
    xs = [1, 2, 3, m - 1, m - 2, (m + 1) // 2, (m - 1) // 2, 2**255 % m, 2**128, 2**64 + 1, 3**100 % m]
    xs += [pow(2, k, m) for k in range(0, 256, 17)] + [rnd.randrange(1, m) for _ in range(100)]
    for x in xs:
        want = pow(x, -1, m)
        assert D.field(2, mod, x) == want, (mod, x)
        assert D.field(5, mod, x) == want, (mod, x)


def _val(limbs):
    return sum(v << (29 * i) for i, v in enumerate(limbs))


def _limbs(x):
    return [(x >> (29 * i)) & (2**29 - 1) for i in range(9)]


@pytest.mark.parametrize("mod", range(6))
def test_montgomery_raw_limbs(mod):
    """mont_mul / mont_sqr on raw limb vectors: the output is congruent to a*b/2^261, normalised,
    and below 2M whenever a, b < 2M. Includes the column-sum extremes (every limb 2^29-1), where
    an overflowing or sign-confused column accumulator would show."""
    m = MODULI[mod]
    R = 2**261
    rinv = pow(R, -1, m)
    rnd = random.Random(100 + mod)
    cases = [(2 * m - 1, 2 * m - 1), (0, 2 * m - 1), (1, 1), (m, m), (m - 1, 2 * m - 1)]
    cases += [(rnd.randrange(2 * m), rnd.randrange(2 * m)) for _ in range(200)]
    # sparse operands (rows of zero limbs leave the upper columns to the reduction terms and the
    # carries alone): products with 1 and small powers of two, as from_mont and doublings make
    cases += [(x, y) for x in (2 * m - 1, m + 1, (m + 1) // 2, 2**255 % (2 * m), rnd.randrange(2 * m))
              for y in (1, 2, 2**28, 2**29, 2**58 + 1)]
    # limbs all at the maximum: value 2^261 - 1 (outside [0, 2M), checked for congruence only)
    top = [2**29 - 1] * 9
    raw = [(_limbs(a), _limbs(b), True) for a, b in cases] + [(top, top, False), (top, _limbs(2 * m - 1), False)]
    for al, bl, in_range in raw:
        a, b = _val(al), _val(bl)
        for op in (0, 1):
            rb = al if op == 1 else bl
            out = D.mont_raw(op, mod, al, rb)
            v = _val(out)
            assert v % m == a * _val(rb) * rinv % m, (mod, op, al, rb)
            assert all(x < 2**29 for x in out[:8])
            if in_range:
                assert v < 2 * m and out[8] < 2**29


def test_glv_split_secp256k1():
    """k = k1 + k2*lambda (mod n) with both halves odd and below 2^129 in magnitude. lambda is
    pinned independently: lambda^3 = 1 (mod n) and lambda*G = (beta*Gx, Gy) with beta^3 = 1 (mod p),
    checked with the oracle's own point multiplication."""
    c = pyref.CURVES[0]
    n, p = c["n"], c["p"]
    lam = 0x5363AD4CC05C30E0A5261C028812645A122E22EA20816678DF02967C1B23BD72
    beta = 0x7AE96A2B657C07106E64479EAC3434E99CF0497512F58995C1396C28719501EE
    assert pow(lam, 3, n) == 1 and pow(beta, 3, p) == 1 and lam != 1 and beta != 1
    G = c["G"][0].to_bytes(32, "big") + c["G"][1].to_bytes(32, "big")
    rc, lamG = O.secp_mul(0, G, lam)
    assert rc == 0 and lamG == (beta * c["G"][0] % p).to_bytes(32, "big") + c["G"][1].to_bytes(32, "big")
    rnd = random.Random(7)
    ks = [0, 1, 2, 3, n - 1, n - 2, lam, n - lam, (n - 1) // 2, 2**128, 2**128 - 1, 2**255]
    ks += [rnd.randrange(n) for _ in range(3000)]
    for k in ks:
        k1, k2 = D.glv_split(k)
        assert (k1 + k2 * lam - k) % n == 0, k
        assert k1 & 1 and k2 & 1, k
        assert abs(k1) < 2**129 and abs(k2) < 2**129, k


def test_sha256_message_new_and_address_forms():
    """the device SHA-256 (built for the host) against hashlib on every length 0..300 and a few
    long ones, at every byte alignment of the start, with the word reader and with the 16-byte
    reader (every start offset within a 16-byte chunk: every placement of the block's words),
    each reading nothing outside the message's words or chunks; and the fixed 64-byte form used
    for addresses"""
    rnd = random.Random(9)
    for L in list(range(0, 301)) + [1000, 4095, 4096, 4097]:
        msg = bytes(rnd.randrange(256) for _ in range(L))
        want = hashlib.sha256(msg).digest()
        for at in (0, 1, 2, 3) if L < 80 or L > 999 else (L % 4,):
            assert D.sha256(msg, at) == want, (L, at)
        for at in range(16) if L < 80 or L > 999 else (L % 16, 16 + (L * 7) % 16):
            assert D.sha256(msg, at, x4=True) == want, (L, at, "x4")
    for _ in range(50):
        pk = bytes(rnd.randrange(256) for _ in range(64))
        assert D.sha256_64(pk) == hashlib.sha256(pk).digest()


@pytest.mark.parametrize("g_window", [8, 9])
@pytest.mark.parametrize("curve", [0, 1])
def test_verify_random_and_edges(curve, g_window):
    """device verification (host build, bound checker on) against the oracle: random signatures
    with corruption, every edge case, and the careful path forced on ordinary inputs; fixed-base
    windows of 8 and 9 bits"""
    sigs, dig, pubs, cor = O.gen_secp(curve, 12, 48, corrupt_ppm=300000)
    for i in range(48):
        st, _ = D.secp_verify(curve, sigs[i], dig[i], pubs[i], g_window=g_window)
        assert st == O.secp_verify(curve, bytes(sigs[i]), bytes(dig[i]), bytes(pubs[i]))
    hits = 0
    for (name, sig, d, pub), exp in zip(edge_cases.secp_verify_cases(curve),
                                        edge_cases.secp_verify_expected(curve, edge_cases.secp_verify_cases(curve))):
        st, careful = D.secp_verify(curve, sig, d, pub, g_window=g_window)
        hits += careful
        assert st == exp, name
    if g_window == 8:
        assert hits > 0  # the careful path was exercised (the crafted cases aim at the 8-bit schedule)
    for i in range(6):
        st, _ = D.secp_verify(curve, sigs[i], dig[i], pubs[i], force_careful=True, g_window=g_window)
        assert st == O.secp_verify(curve, bytes(sigs[i]), bytes(dig[i]), bytes(pubs[i]))


def test_sha512_preimage_stream():
    """SHA-512 of a 64-byte (R || A) or 32-byte header and a message of every length 0..259 and a
    few long ones, the message at every byte alignment; the reader touches no word outside the
    message"""
    rnd = random.Random(2)
    for L in list(range(0, 260)) + [1000, 4095, 4096]:
        msg = bytes(rnd.randrange(256) for _ in range(L))
        for hl in (64, 32):
            hdr = bytes(rnd.randrange(256) for _ in range(hl))
            want = hashlib.sha512(hdr + msg).digest()
            for at in (0, 1, 2, 3) if L < 140 or L > 999 else (L % 4,):
                assert D.sha512(hdr, msg, at) == want, (L, hl, at)


@pytest.mark.parametrize("g_window", [8, 9])
@pytest.mark.parametrize("curve", [0, 1])
def test_recover_random_and_edges(curve, g_window):
    sigs, dig, pubs, cor = O.gen_secp(curve, 11, 64, corrupt_ppm=150000)
    for i in range(64):
        st, out = D.secp_recover(curve, bytes(sigs[i]), bytes(dig[i]), g_window=g_window)
        assert (st & 0xFF, out) == O.secp_recover(curve, bytes(sigs[i]), bytes(dig[i]))
    cases = edge_cases.secp_cases(curve)
    hits = 0
    for (name, sig, d), exp in zip(cases, edge_cases.secp_expected(curve, cases)):
        st, out = D.secp_recover(curve, sig, d, g_window=g_window)
        hits += st >> 8
        assert (st & 0xFF, out) == exp, name
    if g_window == 8:
        assert hits > 0  # the careful path was exercised (the crafted cases aim at the 8-bit schedule)
    for i in range(8):  # careful path forced on ordinary inputs
        st, out = D.secp_recover(curve, bytes(sigs[i]), bytes(dig[i]), force_careful=True, g_window=g_window)
        assert (st & 0xFF, out) == O.secp_recover(curve, bytes(sigs[i]), bytes(dig[i]))


@pytest.mark.parametrize("lattice", [True, False])
def test_ed25519_random_and_edges(lattice):
    """both forms of the check: lattice-reduced (Ed25519-only kernels) and full-length (mixed)"""
    cases = edge_cases.ed_cases()
    for (name, sig, pk, m), exp in zip(cases, edge_cases.ed_expected(cases)):
        assert (D.ed25519_verify(sig, pk, m, lattice=lattice) == 0) == exp, name
    rng = np.random.default_rng(4)
    lens = rng.integers(0, 300, 48)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    sigs, pks, msgs, cor = O.gen_ed25519(8, 48, off, corrupt_ppm=200000)
    for i in range(48):
        m = bytes(msgs[off[i]:off[i + 1]])
        assert (D.ed25519_verify(bytes(sigs[i]), bytes(pks[i]), m, lattice=lattice) == 0) == (not cor[i])


def test_ed25519_lattice_pairs():
    """the lattice reduction of the Ed25519 check: every pair it returns has c0 = c1 k (mod 8l), c1 odd
    and positive, both below 2^136 (checked with Python integers); it finds one for every k below
    2^128 (the pair (k, 1)) and for all but a few in 10^5 random k (about 3 in 10^6 have none)"""
    N8 = 8 * pyref.EL
    rnd = random.Random(21)
    ks = [0, 1, 2, 3, 8, 2**127, 2**128 - 1, pyref.EL - 1, pyref.EL - 2, (pyref.EL - 1) // 2, 2**252 - 1,
          N8 // 16, N8 // 24, 3 * 2**250]
    ks += [rnd.randrange(pyref.EL) for _ in range(40000)]
    ks += [rnd.randrange(2**rnd.randrange(1, 253)) for _ in range(2000)]
    random_misses = 0
    for j, k in enumerate(ks):
        ok, c0, c1 = D.ed_lattice(k)
        if ok:
            assert c1 > 0 and c1 & 1 and c1 < 2**136 and abs(c0) < 2**136, k
            assert (c0 - c1 * k) % N8 == 0, k
        else:
            assert k >= 2**128, k
            random_misses += 14 <= j < 40014
    assert random_misses <= 3


def test_ed25519_full_length_check():
    """the full-length check that takes over when the reduction finds no pair: forced for every item,
    and reached on its own by the fixture's challenges"""
    cases = edge_cases.ed_cases()
    exp = edge_cases.ed_expected(cases)
    D.set_ed_force_full(True)
    try:
        for (name, sig, pk, m), e in zip(cases, exp):
            assert (D.ed25519_verify(sig, pk, m) == 0) == e, name
    finally:
        D.set_ed_force_full(False)
    long = [(c, e) for c, e in zip(cases, exp) if c[0].startswith("long_lattice_")]
    assert len(long) >= 8
    for (name, sig, pk, m), e in long:
        k = int.from_bytes(hashlib.sha512(sig[:32] + pk + m).digest(), "little") % pyref.EL
        assert not D.ed_lattice(k)[0], name
        assert (D.ed25519_verify(sig, pk, m) == 0) == e, name


def test_ed25519_strict_cases():
    """device verify_strict (host build) against the oracle's restatement on the strict edge cases:
    small-order keys and R, undecodable and non-canonical R, and the ordinary cases"""
    cases = edge_cases.ed_strict_cases()
    codes = set()
    for (name, sig, pk, m), exp in zip(cases, edge_cases.ed_strict_expected(cases)):
        st = D.ed25519_verify(sig, pk, m, strict=True)
        codes.add(st)
        assert (st == 0) == exp == O.ed25519_verify_strict(sig, pk, m), name
        assert D.ed25519_verify(sig, pk, m, strict=True, lattice=False) == st, name
        assert (D.ed25519_verify(sig, pk, m) == 0) == O.ed25519_verify(sig, pk, m), name
    assert {0, 1, 2, 3, 4, 5} <= codes


def _mont_limbs(v, p):
    """an engine table word group: v * 2^261 mod p as nine 29-bit limbs"""
    m = v * 2**261 % p
    return [(m >> (29 * k)) & (2**29 - 1) for k in range(9)]


@pytest.mark.parametrize("lbits", [11, 12, 13, 14, 15])
def test_caller_tables_canonical(lbits):
    """the reference's own tables (precompute::*_bases) convert to exactly the resident tables"""
    from sigops_pkg import sigops
    P = sigops.precompute
    for which, tab in ((0, P.secp256k1_bases(lbits)), (1, P.secp256r1_bases(lbits)), (2, P.ed25519_bases(lbits))):
        bad, out = D.table(which, tab, lbits)
        assert bad == 0, (which, lbits)
        assert np.array_equal(out, D.resident_table(which)[:len(out)]), (which, lbits)


def test_caller_tables_other_points_and_bad_tables():
    from sigops_pkg import sigops
    P = sigops.precompute
    for curve in (0, 1):
        c = pyref.CURVES[curve]
        Q = pyref.w_mul(c, c["G"], 12345)
        bad, out = D.table(curve, P.secp_table_for_point(curve, Q, 13), 13)
        assert bad == 0
        for j in (0, 1, 17, 127):  # odd multiple (2j+1) Q, closed form
            x, y = pyref.w_mul(c, Q, 2 * j + 1)
            assert list(out[j, :9]) == _mont_limbs(x, c["p"]) and list(out[j, 9:18]) == _mont_limbs(y, c["p"])
        good = np.array(P.secp_table_for_point(curve, Q, 13), np.uint32)
        t = good.copy()
        t[5] ^= 1  # one limb of entry 0's x: off the curve
        assert D.table(curve, t, 13)[0] & 1
        t = good.copy()
        t[40 * 3:40 * 4], t[40 * 4:40 * 5] = good[40 * 4:40 * 5], good[40 * 3:40 * 4]  # entries 3 and 4 swapped
        assert D.table(curve, t, 13)[0] == 2
        t = good.copy()
        t[7] = 1 << 13  # a limb out of range
        assert D.table(curve, t, 13)[0] & 1
    # Ed25519: another point of the prime-order subgroup, and one with a torsion component
    B5 = pyref.e_mul(pyref.EB, 5)
    bad, out = D.table(2, P.ed25519_table_for_point(B5, 12), 12)
    assert bad == 0
    x, y = pyref.e_mul(B5, 7)  # entry j = 3 holds 7 * B5 as (y+x, y-x, 2d x y)
    p = pyref.EP
    assert list(out[3, :9]) == _mont_limbs((y + x) % p, p) and list(out[3, 9:18]) == _mont_limbs((y - x) % p, p)
    assert list(out[3, 18:27]) == _mont_limbs(2 * pyref.ED * x * y % p, p)
    T = pyref.ed25519_torsion()[-1]
    mixed = pyref.e_add(B5, T)
    assert D.table(2, P.ed25519_table_for_point(mixed, 12), 12)[0] == 8
