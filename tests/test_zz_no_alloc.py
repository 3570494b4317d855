"""GPU test of sigops_message_new_noalloc (the entry point that allocates nothing): through the C ABI
its digests equal sigops_message_new_dev's on the same messages, bit for bit, on calls in a row --
mixed lengths, then fewer messages (the fixed scratch still holds the first call's order), then
messages of a single length (the order is not used) -- and a sample is checked against hashlib.
More messages than the fixed scratch holds are refused."""
import hashlib

import numpy as np
import pytest
import torch

from sigops_pkg import sigops

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _messages(n, seed, max_len, fixed=None):
    rng = np.random.default_rng(seed)
    lens = np.full(n, fixed, np.int64) if fixed is not None else rng.integers(0, max_len + 1, n)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    data = rng.integers(0, 256, int(off[-1]), dtype=np.uint8)
    # 16-byte aligned with 16 readable bytes after the last message
    buf = torch.zeros(((int(off[-1]) + 16 + 15) // 16) * 16, dtype=torch.uint8, device=DEV)
    if off[-1]:
        buf[: int(off[-1])] = torch.from_numpy(data).to(DEV)
    return buf, torch.from_numpy(off).to(DEV), data, off


def test_message_new_noalloc_matches_dev_entry_point():
    with sigops.Context() as ctx:
        s = torch.cuda.current_stream().cuda_stream
        for n, seed, max_len, fixed in ((300000, 1, 300, None), (1000, 2, 5000, None), (777, 3, 0, 100)):
            print(f"noalloc call: n={n} seed={seed}", flush=True)
            buf, off, data, off_h = _messages(n, seed, max_len, fixed)
            ref = torch.zeros((n, 32), dtype=torch.uint8, device=DEV)
            got = torch.full((n, 32), 0xAB, dtype=torch.uint8, device=DEV)
            ctx.message_new_dev(buf.data_ptr(), off.data_ptr(), n, ref.data_ptr(), s)
            ctx.message_new_noalloc_dev(buf.data_ptr(), off.data_ptr(), n, got.data_ptr(), s)
            torch.cuda.synchronize()
            ref_h, got_h = ref.cpu().numpy(), got.cpu().numpy()
            assert np.array_equal(got_h, ref_h), n
            for i in np.random.default_rng(seed).choice(n, 200, replace=False):
                assert got_h[i].tobytes() == hashlib.sha256(data[off_h[i]:off_h[i + 1]].tobytes()).digest()
        with pytest.raises(sigops.SigopsError, match="SIGOPS_NOALLOC_MAX_MESSAGES"):
            ctx.message_new_noalloc_dev(buf.data_ptr(), off.data_ptr(), sigops.NOALLOC_MAX_MESSAGES + 1,
                                        got.data_ptr(), s)
