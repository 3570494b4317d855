"""GPU parity at the full sizes of the per-scheme configurations: secp256r1 recovery of 65 536
valid signatures and Ed25519 verification of 65 536 signatures with message lengths uniform on
0..1024 bytes and 10 % corrupted. Every output is compared bit for bit with the oracle, run on 16
host threads (the box's CPU share for one GPU)."""
import numpy as np
import pytest

import oracle_lib as O
from sigops_pkg import sigops

pytestmark = pytest.mark.gpu
THREADS = 16


@pytest.fixture(scope="module")
def ctx():
    c = sigops.Context()
    yield c
    c.close()


def _say(*a):
    print(*a, flush=True)


def test_secp256r1_recover_65536(ctx):
    n = 65536
    _say(f"secp256r1 recover n={n}")
    sigs, dig, pubs, cor = O.gen_secp(1, 2002, n, corrupt_ppm=0, nthreads=THREADS)
    out, st = ctx.secp256r1_recover(sigs, dig)
    assert not st.any()
    assert np.array_equal(out, pubs)
    ref_out, ref_st = O.recover_batch(1, sigs, dig, nthreads=THREADS)
    assert np.array_equal(out, ref_out) and np.array_equal(st, ref_st)


def test_ed25519_verify_65536(ctx):
    n = 65536
    rng = np.random.default_rng(3003)
    lens = rng.integers(0, 1025, n)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    _say(f"ed25519 verify n={n} bytes={int(off[-1])}")
    sigs, pks, msgs, cor = O.gen_ed25519(3003, n, off, corrupt_ppm=100000, nthreads=THREADS)
    valid, st = ctx.ed25519_verify(sigs, pks, msgs, off)
    ref = O.verify_batch(sigs, pks, msgs, off, nthreads=THREADS)
    assert np.array_equal(valid, ref.astype(bool))
    assert np.array_equal(valid, cor == 0)
    assert 0.08 < (cor != 0).mean() < 0.12
