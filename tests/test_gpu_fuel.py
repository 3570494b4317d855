"""GPU tests of Fuel's hashing steps either side of the path: Message::new (SHA-256 of raw data)
and the address of a recovered key (SHA-256 of X || Y), against hashlib and the oracle, plus the
whole chain data -> digest -> recovery -> address on the device."""
import hashlib

import numpy as np
import pytest
import torch

import oracle_lib as O
from sigops_pkg import sigops

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = sigops.Context()
    yield c
    c.close()


def _say(*a):
    print(*a, flush=True)


def test_message_new_lengths(ctx):
    rng = np.random.default_rng(21)
    lens = [0, 1, 3, 55, 56, 63, 64, 65, 119, 120, 127, 128, 1000, 4096] + list(rng.integers(0, 4097, 2000))
    datas = [rng.integers(0, 256, int(L), dtype=np.uint8).tobytes() for L in lens]
    _say(f"message_new n={len(datas)}")
    got = ctx.message_new(datas)
    want = np.array([np.frombuffer(hashlib.sha256(d).digest(), np.uint8) for d in datas])
    assert np.array_equal(got, want)
    assert sigops.fuel.message_new(datas[:5], ctx=ctx) == [hashlib.sha256(d).digest() for d in datas[:5]]


def test_message_new_single_block_count(ctx):
    """messages of one SHA-256 block count keep their index order (the longest-first sort is
    skipped): lengths 0..55 (one block each), all 64-byte messages (two), and a single message"""
    rng = np.random.default_rng(23)
    for lens in (rng.integers(0, 56, 5000), np.full(3000, 64), np.array([200])):
        datas = [rng.integers(0, 256, int(L), dtype=np.uint8).tobytes() for L in lens]
        _say(f"message_new one block count n={len(datas)} len0={len(datas[0])}")
        got = ctx.message_new(datas)
        want = np.array([np.frombuffer(hashlib.sha256(d).digest(), np.uint8) for d in datas])
        assert np.array_equal(got, want)


def test_message_new_dev_unaligned_offsets(ctx):
    rng = np.random.default_rng(22)
    n = 777
    lens = rng.integers(0, 300, n)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64) + 3  # every start unaligned
    data = rng.integers(0, 256, int(off[-1]) + 16, dtype=np.uint8)
    d_data = torch.from_numpy(np.concatenate([data, np.zeros(4, np.uint8)])).cuda()
    d_off = torch.from_numpy(off.astype(np.int64)).cuda()
    d_out = torch.zeros((n, 32), dtype=torch.uint8, device="cuda")
    ctx.message_new_dev(d_data.data_ptr(), d_off.data_ptr(), n, d_out.data_ptr())
    torch.cuda.synchronize()
    got = d_out.cpu().numpy()
    for i in range(n):
        assert got[i].tobytes() == hashlib.sha256(data[off[i]:off[i + 1]].tobytes()).digest(), i


def test_message_new_dev_buffer_4_byte_aligned(ctx):
    """a data pointer 4 bytes past a 16-byte boundary (the documented minimum alignment), with
    log-uniform lengths up to 3000 bytes: the longest-first order over many block counts"""
    rng = np.random.default_rng(25)
    n = 3000
    lens = np.exp(rng.uniform(0, np.log(3000), n)).astype(np.int64)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    data = rng.integers(0, 256, int(off[-1]) + 16, dtype=np.uint8)
    d_buf = torch.from_numpy(np.concatenate([np.zeros(4, np.uint8), data])).cuda()
    d_off = torch.from_numpy(off).cuda()
    d_out = torch.zeros((n, 32), dtype=torch.uint8, device="cuda")
    assert d_buf.data_ptr() % 16 == 0
    _say(f"message_new_dev 4-byte-aligned buffer n={n}")
    ctx.message_new_dev(d_buf.data_ptr() + 4, d_off.data_ptr(), n, d_out.data_ptr())
    torch.cuda.synchronize()
    got = d_out.cpu().numpy()
    for i in range(n):
        assert got[i].tobytes() == hashlib.sha256(data[off[i]:off[i + 1]].tobytes()).digest(), i


def test_message_new_queue_many_messages_twice(ctx):
    """far more messages than resident lanes, a heavy-tailed length mix (empty, 64-byte and
    log-uniform up to 8 KiB), then a second, smaller call: every lane takes many messages from the
    work counter, which each call clears"""
    rng = np.random.default_rng(24)
    n = 300000
    kind = rng.random(n)
    lens = np.where(kind < 0.6, 0, np.where(kind < 0.85, 64, np.exp(rng.uniform(0, np.log(8192), n)).astype(np.int64)))
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    data = rng.integers(0, 256, int(off[-1]) + 16, dtype=np.uint8)
    d_data = torch.from_numpy(data).cuda()
    d_off = torch.from_numpy(off).cuda()
    want = [hashlib.sha256(data[off[i]:off[i + 1]].tobytes()).digest() for i in range(n)]
    for m in (n, 12345):
        _say(f"message_new_dev queue n={m}")
        d_out = torch.zeros((m, 32), dtype=torch.uint8, device="cuda")
        ctx.message_new_dev(d_data.data_ptr(), d_off.data_ptr(), m, d_out.data_ptr())
        torch.cuda.synchronize()
        got = d_out.cpu().numpy()
        bad = [i for i in range(m) if got[i].tobytes() != want[i]]
        assert not bad, (m, bad[:5])


@pytest.mark.parametrize("curve", [0, 1])
def test_recover_address_matches_hash_of_recovered_key(ctx, curve):
    n = 3000
    _say(f"recover_address curve={curve} n={n}")
    sigs, dig, pubs, cor = O.gen_secp(curve, 70 + curve, n, corrupt_ppm=100000)
    addr, st = ctx.recover_address(curve, sigs, dig)
    out, st2 = (ctx.secp256k1_recover if curve == 0 else ctx.secp256r1_recover)(sigs, dig)
    assert np.array_equal(st, st2)
    for i in range(n):
        want = hashlib.sha256(out[i].tobytes()).digest() if st[i] == 0 else bytes(32)
        assert addr[i].tobytes() == want, i
    assert (st != 0).any() and (st == 0).any()


def test_address_from_pubkeys_dev_without_status(ctx):
    rng = np.random.default_rng(23)
    pubs = rng.integers(0, 256, (500, 64), dtype=np.uint8)
    d_p = torch.from_numpy(pubs).cuda()
    d_a = torch.zeros((500, 32), dtype=torch.uint8, device="cuda")
    ctx.address_from_pubkeys_dev(d_p.data_ptr(), None, 500, d_a.data_ptr())
    torch.cuda.synchronize()
    got = d_a.cpu().numpy()
    for i in range(500):
        assert got[i].tobytes() == hashlib.sha256(pubs[i].tobytes()).digest()


def test_chain_data_to_address():
    """raw transaction bytes -> Message::new -> recovery -> address, against the oracle's signer"""
    ctx = sigops.default_context()
    rng = np.random.default_rng(24)
    n = 256
    datas = [rng.integers(0, 256, int(L), dtype=np.uint8).tobytes() for L in rng.integers(0, 2000, n)]
    sks = [int.from_bytes(rng.integers(0, 256, 32, dtype=np.uint8).tobytes(), "big") % (2**255) + 1 for _ in range(n)]
    digests = ctx.message_new(datas)
    sigs = []
    for i in range(n):
        sk = sks[i].to_bytes(32, "big")
        nonce = hashlib.sha256(b"nonce" + sk + digests[i].tobytes()).digest()
        sigs.append(O.secp_sign(0, sk, digests[i].tobytes(), nonce))
    addrs, st = sigops.fuel.recover_address(sigs, [d.tobytes() for d in digests], ctx=ctx)
    assert not st.any()
    for i in range(n):
        pk = O.secp_pubkey(0, sks[i].to_bytes(32, "big"))
        assert addrs[i] == hashlib.sha256(pk).digest(), i
