"""CPU test: threads that share a Context (the reference-API mirrors share the default context) never
enter the native library at the same time, and the default context is created once.

The native library is replaced by a stub whose entry points record whether another call is in
progress, so the test runs without a GPU."""
import threading
import time

import numpy as np

from sigops_pkg import sigops


class _StubLib:
    def __init__(self):
        self.inside = 0
        self.overlaps = 0
        self.calls = 0
        self.created = 0
        self.guard = threading.Lock()

    def sigops_create(self, cfg):
        with self.guard:
            self.created += 1
        time.sleep(0.01)  # a slow create widens the window for a second default context
        return 1

    def sigops_destroy(self, h):
        pass

    def sigops_last_error(self, h):
        return b""

    def _native(self, *args):
        with self.guard:
            self.inside += 1
            self.calls += 1
            if self.inside > 1:
                self.overlaps += 1
        time.sleep(0.002)
        with self.guard:
            self.inside -= 1
        return 0

    sigops_secp256k1_recover = sigops_secp256r1_recover = sigops_ed25519_verify = _native
    sigops_secp256k1_verify = sigops_batch = _native


def test_mirrors_serialize_native_calls(monkeypatch):
    stub = _StubLib()
    monkeypatch.setattr(sigops, "_lib", stub)
    monkeypatch.setattr(sigops, "_default_ctx", None)
    monkeypatch.setattr(sigops, "_tuned", lambda device: None)
    sigs = [bytes(64)] * 3
    digs = [bytes(32)] * 3
    errors = []

    def worker(k):
        try:
            for _ in range(25):
                if k % 2:
                    sigops.secp256k1_ecdsa.ecrecover_single_shader(sigs, digs, 13)
                else:
                    sigops.ed25519_eddsa.ecverify_bytes(sigs, [b"m"] * 3, [bytes(32)] * 3)
        except Exception as e:  # noqa: BLE001 -- reported below
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(k,)) for k in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors
    assert stub.calls == 100
    assert stub.overlaps == 0
    assert stub.created == 1


def test_chunks_cover_inputs_above_max_batch(monkeypatch):
    """host arrays above max_batch go to the library as consecutive calls of at most max_batch
    items, with the matching slices of every array"""
    seen = []

    class Lib(_StubLib):
        def sigops_secp256k1_recover(self, h, ps, pd, n, po, pst):
            seen.append(n)
            return 0

        def sigops_ed25519_verify(self, h, ps, pk, pm, poff, n, pv, pst):
            seen.append(n)
            return 0

    monkeypatch.setattr(sigops, "_lib", Lib())
    monkeypatch.setattr(sigops, "_tuned", lambda device: None)
    ctx = sigops.Context(max_batch=1000)
    ctx.secp256k1_recover(np.zeros((2500, 64), np.uint8), np.zeros((2500, 32), np.uint8))
    assert seen == [1000, 1000, 500]
    seen.clear()
    ctx.ed25519_verify(np.zeros((1001, 64), np.uint8), np.zeros((1001, 32), np.uint8), np.zeros(10, np.uint8),
                       np.zeros(1002, np.uint64))
    assert seen == [1000, 1]
