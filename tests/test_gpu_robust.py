"""GPU tests of the error path and of calls the round-trip tests do not make:
  * the status canary: with the fault switch set, no compute kernel runs, and every host entry
    point and reference-API mirror reports ShaderFailureError (the reference's only error,
    src/lib.rs:12-14, raised when the `success` canary is missing, src/secp256k1_ecdsa.rs:203-205);
    with the switch off again the same context gives correct results;
  * one captured graph per context serves a batch whatever buffers each call passes: device
    buffers from the caching allocator that move between calls, each call checked against the
    oracle;
  * the device entry points refuse message buffers that are not 4-byte aligned;
  * inputs larger than the context's max_batch are split into several calls (the mirrors accept up
    to 2^30 items, src/secp256k1_ecdsa.rs:22)."""
import numpy as np
import pytest
import torch

import oracle_lib as O
from sigops_pkg import sigops

pytestmark = pytest.mark.gpu
W = __import__("importlib").import_module("wgpu-sigops_amd.workload")
DEV = "cuda:0"


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def ctx():
    c = sigops.Context()
    yield c
    c.close()


def _ed_inputs(n, seed, maxlen=300):
    lens = np.arange(n) % maxlen
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    sigs, pks, msgs, cor = O.gen_ed25519(seed, n, off, corrupt_ppm=200000)
    return sigs, pks, msgs, off


def test_canary_reports_shader_failure(ctx):
    _say("canary: fault switch on")
    sigs, dig, pubs, cor = O.gen_secp(0, 71, 300, corrupt_ppm=0)
    esigs, pks, msgs, off = _ed_inputs(200, 72)
    ctx.set_fault(sigops.FAULT_SKIP_COMPUTE)
    try:
        with pytest.raises(sigops.ShaderFailureError):
            ctx.secp256k1_recover(sigs, dig)
        with pytest.raises(sigops.ShaderFailureError):
            ctx.secp256r1_recover(sigs, dig)
        with pytest.raises(sigops.ShaderFailureError):
            ctx.secp256k1_verify(sigs, dig, pubs)
        with pytest.raises(sigops.ShaderFailureError):
            ctx.ed25519_verify(esigs, pks, msgs, off)
        with pytest.raises(sigops.ShaderFailureError):
            ctx.ed25519_verify_strict(esigs, pks, msgs, off)
        sch = np.zeros(300, np.uint8)
        with pytest.raises(sigops.ShaderFailureError):
            ctx.batch(sch, sigs, dig, np.zeros(0, np.uint8), np.zeros(301, np.uint64))
        # the reference-API mirrors raise the reference's error type
        with pytest.raises(sigops.ShaderFailureError):
            sigops.secp256k1_ecdsa.ecrecover_single_shader([bytes(s) for s in sigs[:8]],
                                                           [bytes(d) for d in dig[:8]], 13, ctx=ctx)
        with pytest.raises(sigops.ShaderFailureError):
            sigops.ed25519_eddsa.ecverify_bytes([bytes(s) for s in esigs[:8]],
                                                [bytes(msgs[off[i]:off[i + 1]]) for i in range(8)],
                                                [bytes(p) for p in pks[:8]], ctx=ctx)
    finally:
        ctx.set_fault(0)
    _say("canary: fault switch off")
    out, st = ctx.secp256k1_recover(sigs, dig)
    assert not st.any() and np.array_equal(out, pubs)
    valid, _ = ctx.ed25519_verify(esigs, pks, msgs, off)
    assert np.array_equal(valid, O.verify_batch(esigs, pks, msgs, off).astype(bool))
    with pytest.raises(sigops.SigopsError):
        ctx.set_fault(2)


def test_batch_graph_serves_moving_buffers(ctx):
    """fresh device buffers on every call (as the torch caching allocator or an FFI caller hands
    them out): every call's output equals the oracle's"""
    s = torch.cuda.current_stream()
    for k in range(6):
        n = 700 + 301 * k
        _say(f"graph reuse call {k}, n={n}")
        mx = W.mixed(ctx, n, seed=1000 + k, corrupt_ppm=100000)
        pad = torch.zeros(4096 * (k + 1), dtype=torch.uint8, device=DEV)  # moves the next allocations
        out = torch.zeros((n, 64), dtype=torch.uint8, device=DEV)
        st = torch.full((n,), 0xEE, dtype=torch.uint8, device=DEV)
        ctx.batch_dev(mx.schemes.data_ptr(), mx.sigs.data_ptr(), mx.aux.data_ptr(), mx.msgs.data_ptr(),
                      mx.msg_off.data_ptr(), n, out.data_ptr(), st.data_ptr(), s.cuda_stream)
        torch.cuda.synchronize()
        del pad
        sch, sigs, aux = mx.schemes_host, mx.sigs.cpu().numpy(), mx.aux.cpu().numpy()
        out_h, st_h = out.cpu().numpy(), st.cpu().numpy()
        for c in (0, 1):
            ii = np.nonzero(sch == c)[0]
            o, t = O.recover_batch(c, sigs[ii], aux[ii])
            assert np.array_equal(out_h[ii], o) and np.array_equal(st_h[ii], t), (k, c)
        ii = np.nonzero(sch == 2)[0]
        off = mx.msg_off_host
        msgs = mx.msgs.cpu().numpy()
        lens = off[ii + 1] - off[ii]
        sub_off = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
        sub = np.concatenate([msgs[off[i]:off[i + 1]] for i in ii]) if len(ii) else np.zeros(0, np.uint8)
        v = O.verify_batch(sigs[ii], aux[ii], sub, sub_off)
        assert np.array_equal(st_h[ii] == 0, v.astype(bool)), k


def test_device_entry_points_refuse_unaligned_messages(ctx):
    n = 4
    sigs = torch.zeros((n, 64), dtype=torch.uint8, device=DEV)
    aux = torch.zeros((n, 32), dtype=torch.uint8, device=DEV)
    msgs = torch.zeros(64, dtype=torch.uint8, device=DEV)
    off = torch.zeros(n + 1, dtype=torch.int64, device=DEV)
    st = torch.zeros(n, dtype=torch.uint8, device=DEV)
    out = torch.zeros((n, 64), dtype=torch.uint8, device=DEV)
    sch = torch.full((n,), 2, dtype=torch.uint8, device=DEV)
    bad = msgs.data_ptr() + 1
    for call in (lambda: ctx.ed25519_verify_dev(sigs.data_ptr(), aux.data_ptr(), bad, off.data_ptr(), n,
                                                st.data_ptr()),
                 lambda: ctx.ed25519_verify_strict_dev(sigs.data_ptr(), aux.data_ptr(), bad, off.data_ptr(), n,
                                                       st.data_ptr()),
                 lambda: ctx.batch_dev(sch.data_ptr(), sigs.data_ptr(), aux.data_ptr(), bad, off.data_ptr(), n,
                                       out.data_ptr(), st.data_ptr())):
        with pytest.raises(sigops.SigopsError, match="aligned"):
            call()


def test_inputs_above_max_batch_are_split():
    """a context of max_batch 4096 and 3 x 4096 + 5 items through the mirrors and the array API:
    identical to the oracle"""
    with sigops.Context(max_batch=4096) as small:
        n = 3 * 4096 + 5
        _say(f"split: n={n} with max_batch 4096")
        sigs, dig, pubs, cor = O.gen_secp(1, 81, n, corrupt_ppm=50000)
        keys = sigops.secp256r1_ecdsa.ecrecover_single_shader(sigs, dig, 13, ctx=small)
        o, t = O.recover_batch(1, sigs, dig)
        assert [bytes(r) for r in o] == keys
        esigs, pks, msgs, off = _ed_inputs(n, 82, maxlen=97)
        valid, st = small.ed25519_verify(esigs, pks, msgs, off)
        assert np.array_equal(valid, O.verify_batch(esigs, pks, msgs, off).astype(bool))
        sch = np.full(n, 2, np.uint8)
        sch[: n // 2] = 0
        bsigs = esigs.copy()
        baux = pks.copy()
        ks, kd, kp, kc = O.gen_secp(0, 83, n // 2, corrupt_ppm=50000)
        bsigs[: n // 2], baux[: n // 2] = ks, kd
        lens = np.diff(off).astype(np.int64)
        lens[: n // 2] = 0
        boff = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
        bmsgs = np.concatenate([msgs[off[i]:off[i + 1]] for i in range(n // 2, n)])
        out, bst = small.batch(sch, bsigs, baux, bmsgs, boff)
        ko, kt = O.recover_batch(0, ks, kd)
        assert np.array_equal(out[: n // 2], ko) and np.array_equal(bst[: n // 2], kt)
        assert np.array_equal(bst[n // 2:] == 0, valid[n // 2:])


def test_context_refuses_unknown_variants():
    """sigops_create refuses a fixed-base window other than 8 or 9 bits, and 9-bit windows with the
    one-wave kernel (that variant is not built); the valid combinations create contexts, and a
    window left at 0 means 9 bits with two waves per SIMD, 8 with one"""
    for kw in ({"g_window": 7}, {"g_window": 10}, {"g_window": 9, "waves": 1}, {"waves": 3}, {"bucket_order": 3}):
        with pytest.raises(sigops.SigopsError):
            sigops.Context(tuned=False, max_batch=1024, **kw)
    for kw, want in (({"g_window": 8}, 8), ({"g_window": 9}, 9), ({"g_window": 8, "waves": 1}, 8), ({}, 9),
                     ({"waves": 1}, 8), ({"waves": 2}, 9)):
        c = sigops.Context(tuned=False, max_batch=1024, **kw)
        assert c.g_window == want
        c.close()
