"""GPU parity: the HIP path, called through the C ABI (libsigops.so), against the oracle, the
closed forms and the golden vectors. Bit-exact on every output byte and status code.

Sizes follow the reference's configurations where the oracle finishes in seconds: secp256k1 at
N = 1024 (its plain-method configuration), secp256r1 and Ed25519 at N = 8192 here (the 65536-item
runs are in test_gpu_full.py, checked by properties and by the oracle on the box's cores).
"""
import hashlib
import json
import os

import numpy as np
import pytest

import edge_cases
import oracle_lib as O
from sigops_pkg import sigops

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


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


@pytest.mark.parametrize("curve,n", [(0, 1024), (1, 8192)])
def test_secp_recover_matches_oracle(ctxw, curve, n):
    _say(f"secp recover curve={curve} n={n}, {ctxw.g_window}-bit G windows")
    sigs, dig, pubs, cor = O.gen_secp(curve, 1000 + curve, n, corrupt_ppm=100000)
    fn = ctxw.secp256k1_recover if curve == 0 else ctxw.secp256r1_recover
    out, st = fn(sigs, dig)
    ref_out, ref_st = O.recover_batch(curve, sigs, dig)
    assert np.array_equal(st, ref_st)
    assert np.array_equal(out, ref_out)
    ok = cor == 0
    assert np.array_equal(out[ok], pubs[ok])  # round trip recover(sign(sk)) == sk*G


@pytest.mark.parametrize("curve", [0, 1])
def test_secp_edge_cases(ctxw, curve):
    cases = edge_cases.secp_cases(curve)
    exp = edge_cases.secp_expected(curve, cases)
    _say(f"secp edge cases curve={curve} n={len(cases)}, {ctxw.g_window}-bit G windows")
    sigs = np.frombuffer(b"".join(c[1] for c in cases), np.uint8).reshape(-1, 64)
    digs = np.frombuffer(b"".join(c[2] for c in cases), np.uint8).reshape(-1, 32)
    fn = ctxw.secp256k1_recover if curve == 0 else ctxw.secp256r1_recover
    out, st = fn(sigs, digs)
    for i, (name, _, _) in enumerate(cases):
        assert (int(st[i]), bytes(out[i])) == exp[i], name


def test_secp256k1_reference_kat(ctx):
    kat = json.load(open(os.path.join(HERE, "golden", "secp256k1_kat.json")))
    sig = bytes.fromhex(kat["signature"])
    dig = hashlib.sha256(kat["message_utf8"].encode()).digest()
    res = sigops.secp256k1_ecdsa.ecrecover([sig], [dig], sigops.precompute.secp256k1_bases(13), 13, ctx=ctx)
    assert res[0].hex() == kat["pubkey"]
    res = sigops.secp256k1_ecdsa.ecrecover_single_shader([sig], [dig], 13, ctx=ctx)
    assert res[0].hex() == kat["pubkey"]


def test_ed25519_verify_matches_oracle(ctx):
    n = 8192
    rng = np.random.default_rng(7)
    lens = rng.integers(0, 1025, n)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    _say(f"ed25519 verify n={n} bytes={int(off[-1])}")
    sigs, pks, msgs, cor = O.gen_ed25519(77, n, off, corrupt_ppm=100000)
    valid, st = ctx.ed25519_verify(sigs, pks, msgs, off)
    ref = O.verify_batch(sigs, pks, msgs, off)
    assert np.array_equal(valid, ref.astype(bool))
    assert np.array_equal(valid, cor == 0)


def test_ed25519_edge_cases(ctx):
    cases = edge_cases.ed_cases()
    exp = edge_cases.ed_expected(cases)
    _say(f"ed25519 edge cases n={len(cases)}")
    res = sigops.ed25519_eddsa.ecverify_bytes([c[1] for c in cases], [c[3] for c in cases], [c[2] for c in cases],
                                              ctx=ctx)
    for (name, *_), r, e in zip(cases, res, exp):
        assert r == e, name


def test_ed25519_rfc8032_and_fixed_32_byte_api(ctx):
    vec = json.load(open(os.path.join(HERE, "golden", "rfc8032_ed25519.json")))["vectors"]
    res = sigops.ed25519_eddsa.ecverify_bytes([bytes.fromhex(v["signature"]) for v in vec],
                                              [bytes.fromhex(v["message"]) for v in vec],
                                              [bytes.fromhex(v["public"]) for v in vec], ctx=ctx)
    assert res == [True, True, True]
    # the reference API: 32-byte messages (fuel Message = SHA-256 of the data), as its tests do
    seeds = [bytes([i]) * 32 for i in range(1, 6)]
    msgs = [hashlib.sha256(bytes([i]) * 100).digest() for i in range(5)]
    sigs = [O.ed25519_sign(s, m) for s, m in zip(seeds, msgs)]
    pks = [O.ed25519_pubkey(s) for s in seeds]
    bad = bytearray(sigs[2])
    bad[40] ^= 4
    sigs[2] = bytes(bad)
    got = sigops.ed25519_eddsa.ecverify(sigs, msgs, pks, sigops.precompute.ed25519_bases(13), 13, ctx=ctx)
    assert got == [True, True, False, True, True]
    assert sigops.ed25519_eddsa.ecverify_single(sigs, msgs, pks, 13, ctx=ctx) == got


def test_empty_and_ragged_inputs(ctx):
    out, st = ctx.secp256k1_recover(np.zeros((0, 64), np.uint8), np.zeros((0, 32), np.uint8))
    assert out.shape == (0, 64)
    v, st = ctx.ed25519_verify(np.zeros((0, 64), np.uint8), np.zeros((0, 32), np.uint8), np.zeros(0, np.uint8),
                               np.zeros(1, np.uint64))
    assert v.shape == (0,)
    # sizes that are not multiples of a wavefront or a workgroup
    for n in (1, 63, 65, 257):
        sigs, dig, pubs, cor = O.gen_secp(0, 50 + n, n)
        out, st = ctx.secp256k1_recover(sigs, dig)
        assert np.array_equal(out, pubs) and not st.any()
    # all-empty messages
    n = 70
    off = np.zeros(n + 1, np.uint64)
    sigs, pks, msgs, cor = O.gen_ed25519(5, n, off)
    valid, st = ctx.ed25519_verify(sigs, pks, msgs, off)
    assert valid.all()
