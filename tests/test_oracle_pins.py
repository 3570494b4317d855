"""Independent pins of the oracle's hot-path rules (CPU only).

The edge-case lists of edge_cases.py take their expectations from pyref.py, a second restatement
written alongside the oracle. The tests here check the same cases against code that was not
written here, and against the reference's own golden vectors:
  * OpenSSL libcrypto's Ed25519 verification on every Ed25519 edge case (its rules coincide with
    ed25519-dalek 2.1 `verify`: s < l, cofactorless equation, byte comparison with R);
  * OpenSSL's ECDSA_do_verify on every verification edge case, with the two places where the
    rules differ asserted as such: libsecp256k1's lower-S rule (secp256k1 only) and keys OpenSSL
    refuses to construct;
  * recovery with OpenSSL doing the curve arithmetic (point decompression for R, EC_POINT_mul for
    u1 G + u2 R) on every recovery edge case and on corrupted generated signatures;
  * the reference's nonce KAT (src/curve_algos/secp256k1_ecdsa.rs:146-158, :206-214) through the
    oracle's signer, and its Strauss-Shamir vector (src/tests/secp256k1_curve.rs:683-740) through
    the oracle's double-scalar multiplication and recovery.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import edge_cases
import oracle_lib as O
import openssl_lib as S
import pyref

HERE = os.path.dirname(os.path.abspath(__file__))

pytestmark = pytest.mark.skipif(not S.available(), reason="OpenSSL libcrypto not found")


def _r_s(sig: bytes):
    return int.from_bytes(sig[:32], "big"), int.from_bytes(bytes([sig[32] & 0x7F]) + sig[33:], "big")


def test_group_orders_from_openssl():
    for curve in (0, 1):
        assert S.group_order(curve) == pyref.CURVES[curve]["n"]


def test_ed25519_edge_cases_vs_openssl():
    cases = edge_cases.ed_cases()
    agree = [O.ed25519_verify(sig, pk, m) == S.ed25519_verify(sig, pk, m) for (_, sig, pk, m) in cases]
    assert all(agree), [c[0] for c, a in zip(cases, agree) if not a]
    outcomes = {O.ed25519_verify(sig, pk, m) for (_, sig, pk, m) in cases}
    assert outcomes == {True, False}


def test_ed25519_strict_cases_plain_verify_vs_openssl():
    """the verify_strict list run through plain verify (small-order R with s = 0, undecodable and
    non-canonical R): OpenSSL's verify makes the same decision as dalek's `verify`"""
    for name, sig, pk, m in edge_cases.ed_strict_cases():
        assert O.ed25519_verify(sig, pk, m) == S.ed25519_verify(sig, pk, m), name


def test_ed25519_generated_vs_openssl():
    rnd = np.random.default_rng(4)
    lens = rnd.integers(0, 300, 200)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    sigs, pks, msgs, cor = O.gen_ed25519(1717, 200, off, corrupt_ppm=300000, nthreads=4)
    valid = O.verify_batch(sigs, pks, msgs, off, nthreads=4)
    for i in range(200):
        m = bytes(msgs[off[i]:off[i + 1]])
        assert bool(valid[i]) == S.ed25519_verify(bytes(sigs[i]), bytes(pks[i]), m), i
    assert 0 < cor.sum() < 200


@pytest.mark.parametrize("curve", [0, 1])
def test_secp_verify_edge_cases_vs_openssl(curve):
    n = pyref.CURVES[curve]["n"]
    half = (n - 1) // 2
    high_s, refused = [], []
    for name, sig, dig, pub in edge_cases.secp_verify_cases(curve):
        st = O.secp_verify(curve, sig, dig, pub)
        r, s = _r_s(sig)
        ossl = S.ecdsa_verify(curve, pub, dig, r, s)
        if ossl is None:  # OpenSSL refuses the key: not a curve point with coordinates below p
            assert st == 6, name
            refused.append(name)
        elif st == 5:  # libsecp256k1 verifies lower-S signatures only; OpenSSL accepts either s
            assert curve == 0 and s > half and ossl is True, name
            high_s.append(name)
        else:
            assert (st == 0) == ossl, name
    assert sorted(refused) == ["key_not_on_curve", "key_x_ge_p", "key_y_plus_p", "key_zero"]
    assert sorted(high_s) == (["s_half_plus_1", "s_max"] if curve == 0 else [])


@pytest.mark.parametrize("curve", [0, 1])
def test_secp_verify_generated_vs_openssl(curve):
    """lower-S generated signatures, 30% corrupted, with their keys and with the next item's key"""
    sigs, dig, pubs, cor = O.gen_secp(curve, 515, 96, corrupt_ppm=300000, nthreads=4)
    for i in range(96):
        for pub in (pubs[i].tobytes(), pubs[(i + 1) % 96].tobytes()):
            sig = sigs[i].tobytes()
            st = O.secp_verify(curve, sig, dig[i].tobytes(), pub)
            r, s = _r_s(sig)
            ossl = S.ecdsa_verify(curve, pub, dig[i].tobytes(), r, s)
            if st == 5:
                assert curve == 0 and s > (pyref.CURVES[0]["n"] - 1) // 2
            else:
                assert (st == 0) == ossl, i


@pytest.mark.parametrize("curve", [0, 1])
def test_recover_edge_cases_vs_openssl(curve):
    statuses = set()
    for name, sig, dig in edge_cases.secp_cases(curve):
        want = S.ecdsa_recover(curve, sig, dig)
        assert O.secp_recover(curve, sig, dig) == want, name
        statuses.add(want[0])
        if want[0] == 0:  # the recovered key verifies the signature (OpenSSL accepts either s)
            r, s = _r_s(sig)
            assert S.ecdsa_verify(curve, want[1], dig, r, s) is True, name
    assert statuses == {0, 1, 2, 3, 4}


@pytest.mark.parametrize("curve", [0, 1])
def test_recover_generated_vs_openssl(curve):
    sigs, dig, pubs, cor = O.gen_secp(curve, 616, 128, corrupt_ppm=400000, nthreads=4)
    out, st = O.recover_batch(curve, sigs, dig, nthreads=4)
    for i in range(128):
        want = S.ecdsa_recover(curve, sigs[i].tobytes(), dig[i].tobytes())
        assert (int(st[i]), out[i].tobytes()) == want, i
        if not cor[i]:
            assert want == (0, pubs[i].tobytes())


@pytest.mark.parametrize("curve", [0, 1])
def test_recover_agrees_with_verification_under_the_signers_key(curve):
    """On corrupted signatures, OpenSSL's ECDSA verification under the signer's key accepts exactly
    the signatures from which recovery returns the signer's key, except those whose recovery bit
    (bit 7 of s's first byte) was flipped: verification ignores that bit, and recovery with it
    restored returns the signer. (The benchmark checks its OpenSSL baseline by the same rule.)"""
    sigs, digs, pubs, cor = O.gen_secp(curve, 77 + curve, 12000, corrupt_ppm=250000, nthreads=8)
    out, st = O.recover_batch(curve, sigs, digs, nthreads=8)
    _, valid = S.time_ecdsa_verify(curve, sigs, digs, pubs, 8)
    signer = (st == 0) & (out == pubs).all(axis=1)
    assert signer[cor == 0].all()
    odd = np.nonzero(valid.astype(bool) != signer)[0]
    fixed = sigs[odd].copy()
    fixed[:, 32] ^= 0x80
    o2, t2 = O.recover_batch(curve, fixed, digs[odd])
    assert (valid.astype(bool)[odd] & (t2 == 0) & (o2 == pubs[odd]).all(axis=1)).all()
    assert int(cor.sum()) > 2000


def test_secp256k1_reference_nonce_kat():
    """the reference signs with sk = 1 and its fixed nonce k() and asserts the signature; the
    oracle's signer with the same inputs gives it byte for byte, with recovery parity 0"""
    kat = json.load(open(os.path.join(HERE, "golden", "secp256k1_kat.json")))
    k = sum(w << (64 * i) for i, w in enumerate(kat["nonce_words_le"]))
    digest = hashlib.sha256(kat["message_utf8"].encode()).digest()
    assert digest.hex() == kat["digest"]
    sig = O.secp_sign(0, bytes.fromhex(kat["secret_key"]), digest, k.to_bytes(32, "big"))
    assert sig is not None and sig.hex() == kat["signature"]
    assert sig[32] >> 7 == kat["recid"]
    # OpenSSL: the signature verifies under G (sk = 1) and recovers G
    r, s = _r_s(sig)
    G = bytes.fromhex(kat["pubkey"])
    assert S.ecdsa_verify(0, G, digest, r, s) is True
    assert S.ecdsa_recover(0, sig, digest) == (0, G)


def _strauss():
    v = json.load(open(os.path.join(HERE, "golden", "secp256k1_strauss.json")))
    B = int(v["bx_decimal"]).to_bytes(32, "big") + int(v["by_decimal"]).to_bytes(32, "big")
    return int(v["x"], 16), int(v["y"], 16), B


def test_secp256k1_reference_strauss_vector():
    c = pyref.CURVES[0]
    n = c["n"]
    x, y, B = _strauss()
    Bp = (int.from_bytes(B[:32], "big"), int.from_bytes(B[32:], "big"))
    want = S.point_mul2(0, B, x % n, y % n)
    closed = pyref.w_add(c, pyref.w_mul(c, c["G"], x % n), pyref.w_mul(c, Bp, y % n))
    assert want == closed[0].to_bytes(32, "big") + closed[1].to_bytes(32, "big")
    assert O.secp_mul2(0, B, x % n, y % n) == (0, want)
    # the vector's point: a joint left-to-right double-and-add over (x, y) reaches an addition
    # of a point and its negative (the case the reference's comment describes)
    T = {(1, 0): c["G"], (0, 1): Bp, (1, 1): pyref.w_add(c, c["G"], Bp)}
    acc, hit = None, False
    for i in range(255, -1, -1):
        acc = pyref.w_add(c, acc, acc) if acc else None
        d = ((x >> i) & 1, (y >> i) & 1)
        if d != (0, 0):
            P = T[d]
            if acc is not None and acc[0] == P[0] and acc[1] != P[1]:
                hit = True
            acc = pyref.w_add(c, acc, P)
    assert hit


def test_secp256k1_strauss_vector_as_recovery():
    """the same vector reached through recovery (R = B, u1 = x, u2 = y), the form the GPU edge
    cases take (edge_cases.secp_cases lists it as ref_strauss_vector)"""
    cases = {name: (sig, dig) for name, sig, dig in edge_cases.secp_cases(0)}
    sig, dig = cases["ref_strauss_vector"]
    x, y, B = _strauss()
    n = pyref.CURVES[0]["n"]
    want = S.point_mul2(0, B, x % n, y % n)
    assert O.secp_recover(0, sig, dig) == (0, want)
    assert S.ecdsa_recover(0, sig, dig) == (0, want)


def test_openssl_cpu_baseline_harness():
    """the threaded OpenSSL timing harness (bench.py's tuned-library CPU baseline) verifies every
    uncorrupted item and rejects every corrupted one, for any thread count"""
    if S.baseline_lib() is None:
        pytest.skip("oracle/libossl_baseline.so is not built")
    for curve in (0, 1):
        sigs, dig, pubs, cor = O.gen_secp(curve, 31, 300, corrupt_ppm=200000)
        for threads in (1, 3):
            t, valid = S.time_ecdsa_verify(curve, sigs, dig, pubs, threads)
            assert t > 0 and np.array_equal(valid, (cor == 0).astype(np.uint8))
    lens = np.arange(300) % 200
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    sigs, pks, msgs, cor = O.gen_ed25519(32, 300, off, corrupt_ppm=200000)
    for threads in (1, 4):
        t, valid = S.time_ed25519_verify(sigs, pks, msgs, off, threads)
        assert t > 0 and np.array_equal(valid, (cor == 0).astype(np.uint8))
