"""GPU parity on every item at 2^20 signatures per scheme: secp256k1 and secp256r1 recovery and
Ed25519 verification through the host entry points, and the mixed 2^20 batch through the device
entry point. Inputs come from the seeded device generators with 5 % of the items corrupted (so
rejections are compared too); every output and status is compared bit for bit with the oracle,
run on 16 host threads (the CPU share of one GPU). Each test takes about 10 to 25 s of oracle
time."""
import importlib

import numpy as np
import pytest
import torch

import oracle_lib as O
from sigops_pkg import sigops

pytestmark = pytest.mark.gpu
W = importlib.import_module("wgpu-sigops_amd.workload")
THREADS = 16
N = 1 << 20


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def ctx():
    c = sigops.Context()
    yield c
    c.close()


@pytest.mark.parametrize("curve", [0, 1], ids=["secp256k1", "secp256r1"])
def test_recover_every_item_2p20(ctx, curve):
    _say(f"recover curve={curve} n={N}")
    it = W.secp(ctx, curve, N, seed=4040 + curve, corrupt_ppm=50000)
    torch.cuda.synchronize()
    sigs, digs = it.sigs.cpu().numpy(), it.aux.cpu().numpy()
    pubs, cor = it.pubs.cpu().numpy(), it.corrupted.cpu().numpy()
    fn = ctx.secp256k1_recover if curve == 0 else ctx.secp256r1_recover
    out, st = fn(sigs, digs)
    ref_out, ref_st = O.recover_batch(curve, sigs, digs, nthreads=THREADS)
    assert np.array_equal(st, ref_st)
    assert np.array_equal(out, ref_out)
    # the uncorrupted items recover the generator's keys
    ok = cor == 0
    assert 0.03 < 1 - ok.mean() < 0.07
    assert np.array_equal(out[ok], pubs[ok]) and not st[ok].any()


def test_ed25519_verify_every_item_2p20(ctx):
    lens = W.loguniform_lengths(N, 4096, 5050)
    _say(f"ed25519 verify n={N} bytes={int(lens.sum())}")
    it = W.ed25519(ctx, N, lens, seed=5051, corrupt_ppm=50000)
    torch.cuda.synchronize()
    off = it.msg_off_host.astype(np.uint64)
    sigs, pks = it.sigs.cpu().numpy(), it.aux.cpu().numpy()
    msgs, cor = it.msgs.cpu().numpy()[: int(off[-1])], it.corrupted.cpu().numpy()
    valid, st = ctx.ed25519_verify(sigs, pks, msgs, off)
    ref = O.verify_batch(sigs, pks, msgs, off, nthreads=THREADS).astype(bool)
    assert np.array_equal(valid, ref)
    assert np.array_equal(valid, st == 0)
    assert np.array_equal(valid, cor == 0)


def test_mixed_batch_every_item_2p20(ctx):
    """the 2^20 mixed batch of the headline measurement (50 % secp256k1, 30 % secp256r1, 20 %
    Ed25519 with messages of 0..4096 bytes), every item against the oracle"""
    _say(f"mixed batch n={N}")
    mx = W.mixed(ctx, N, seed=6060)
    out = torch.zeros((N, 64), dtype=torch.uint8, device="cuda:0")
    st = torch.full((N,), 0xEE, dtype=torch.uint8, device="cuda:0")
    ctx.batch_dev(mx.schemes.data_ptr(), mx.sigs.data_ptr(), mx.aux.data_ptr(), mx.msgs.data_ptr(),
                  mx.msg_off.data_ptr(), N, out.data_ptr(), st.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    out_h, st_h = out.cpu().numpy(), st.cpu().numpy()
    sigs, aux = mx.sigs.cpu().numpy(), mx.aux.cpu().numpy()
    msgs, off = mx.msgs.cpu().numpy(), mx.msg_off_host.astype(np.uint64)
    for s in (0, 1):
        ii = mx.idx[s]
        o, t = O.recover_batch(s, sigs[ii], aux[ii], nthreads=THREADS)
        assert np.array_equal(t, st_h[ii]), f"scheme {s} status"
        assert np.array_equal(o, out_h[ii]), f"scheme {s} keys"
    ii = mx.idx[2]
    # secp items carry no message bytes, so the Ed25519 messages are contiguous in item order
    sub_off = np.concatenate([off[ii], off[ii[-1] + 1:ii[-1] + 2]])
    ref = O.verify_batch(sigs[ii], aux[ii], msgs, sub_off, nthreads=THREADS).astype(bool)
    assert np.array_equal(st_h[ii] == 0, ref)
    assert not out_h[ii].any()


def test_mixed_batch_both_windows_2p20(ctx):
    """the same 2^20 mixed batch through the variants with 8- and 9-bit fixed-base windows: every
    output and status is the same (the test above compares the tuned variant with the oracle)"""
    _say(f"mixed batch, 8- and 9-bit G windows, n={N}")
    mx = W.mixed(ctx, N, seed=6060)
    res = []
    ctxs = [sigops.Context(tuned=False, g_window=w) for w in (8, 9)]
    try:
        for c in ctxs:
            out = torch.zeros((N, 64), dtype=torch.uint8, device="cuda:0")
            st = torch.full((N,), 0xEE, dtype=torch.uint8, device="cuda:0")
            c.batch_dev(mx.schemes.data_ptr(), mx.sigs.data_ptr(), mx.aux.data_ptr(), mx.msgs.data_ptr(),
                        mx.msg_off.data_ptr(), N, out.data_ptr(), st.data_ptr(), torch.cuda.current_stream().cuda_stream)
            torch.cuda.synchronize()
            res.append((out.cpu(), st.cpu()))
    finally:
        for c in ctxs:
            c.close()
    assert torch.equal(res[0][1], res[1][1]) and torch.equal(res[0][0], res[1][0])
    assert int((res[1][1] == 0xEE).sum()) == 0
