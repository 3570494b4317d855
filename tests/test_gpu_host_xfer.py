"""GPU tests of the host entry points' transfers: caller buffers are packed into pinned staging by
a pool of host threads, in 8 MiB parts whose DMA starts as each is packed, and results come back
the same way. The host path must give exactly the device path's results for batches of many
parts, for offsets that start at a nonzero base (rebased while packing), and for any number of
pool threads."""
import hashlib

import numpy as np
import pytest
import torch

from sigops_pkg import sigops

pytestmark = pytest.mark.gpu
W = __import__("importlib").import_module("wgpu-sigops_amd.workload")


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def mixed():
    """300,000 mixed items (about 60 MB of inputs: several staging parts) and their device-path
    results"""
    n = 300000
    with sigops.Context() as c:
        mx = W.mixed(c, n, seed=77)
        out = torch.zeros((n, 64), dtype=torch.uint8, device="cuda:0")
        st = torch.full((n,), 0xEE, dtype=torch.uint8, device="cuda:0")
        c.batch_dev(mx.schemes.data_ptr(), mx.sigs.data_ptr(), mx.aux.data_ptr(), mx.msgs.data_ptr(),
                    mx.msg_off.data_ptr(), n, out.data_ptr(), st.data_ptr(), torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        host = (mx.schemes_host, mx.sigs.cpu().numpy(), mx.aux.cpu().numpy(),
                mx.msgs.cpu().numpy()[: int(mx.msg_off_host[-1])], mx.msg_off_host.astype(np.uint64))
        yield host, out.cpu().numpy(), st.cpu().numpy()


@pytest.mark.parametrize("threads", ["1", "3", "8"])
def test_host_batch_equals_device_path(mixed, threads, monkeypatch):
    (sch, sigs, aux, msgs, off), out_d, st_d = mixed
    monkeypatch.setenv("SIGOPS_HOST_THREADS", threads)
    _say(f"host batch n={len(sch)} threads={threads}")
    with sigops.Context() as c:
        out, st = c.batch(sch, sigs, aux, msgs, off)
        assert np.array_equal(out, out_d) and np.array_equal(st, st_d)
        # a sub-range whose offsets start at a nonzero base, with the whole message buffer
        a, b = 12345, 171717
        assert off[a] > 0
        out2, st2 = c.batch(sch[a:b], sigs[a:b], aux[a:b], msgs, off[a:b + 1])
        assert np.array_equal(out2, out_d[a:b]) and np.array_equal(st2, st_d[a:b])


def test_host_ed25519_and_message_new_rebased_offsets(mixed):
    (sch, sigs, aux, msgs, off), out_d, st_d = mixed
    ed = np.nonzero(sch == 2)[0]
    # the Ed25519 items' messages in a buffer of their own whose first message starts at byte 1000
    lens = (off[ed + 1] - off[ed]).astype(np.int64)
    body = np.concatenate([msgs[off[i]:off[i + 1]] for i in ed])
    pad = np.random.default_rng(3).integers(0, 256, 1000, dtype=np.uint8)
    buf = np.concatenate([pad, body])
    offs = (1000 + np.concatenate([[0], np.cumsum(lens)])).astype(np.uint64)
    _say(f"host ed25519 verify rebased n={len(ed)}")
    with sigops.Context() as c:
        valid, st = c.ed25519_verify(sigs[ed], aux[ed], buf, offs)
        assert np.array_equal(st, st_d[ed]) and np.array_equal(valid, st_d[ed] == 0)
        dig = c.message_new(buf, offs)
    for k in np.random.default_rng(4).choice(len(ed), 300, replace=False):
        assert dig[k].tobytes() == hashlib.sha256(buf[offs[k]:offs[k + 1]].tobytes()).digest(), k
