"""Fuel's hashing steps either side of the recovery path, on the device.

``message_new(datas)`` is fuel ``Message::new`` (SHA-256 of the data), the step that makes the
digests the reference's tests and benchmarks sign (src/tests/ed25519_eddsa.rs:19,
src/benchmarks/secp256k1_ecdsa.rs:162). ``recover_address(signatures, messages)`` recovers the
secp256k1 public keys and hashes each 64-byte X || Y with SHA-256 to the Fuel address without
bringing the keys back to the host (32 zero bytes where recovery fails, the reason in the status).
"""
import numpy as np

from ._secp_common import _as_bytes_array


def message_new(datas, ctx=None):
    """SHA-256 digests (32 bytes each) of a list of byte strings"""
    from . import default_context

    if len(datas) == 0:
        return []
    out = (ctx or default_context()).message_new(list(datas))
    return [bytes(r) for r in out]


def recover_address(signatures, messages, ctx=None, curve=0):
    """addresses (32 bytes each) and per-item status of the keys recovered from the signatures"""
    from . import default_context

    if len(signatures) != len(messages):
        raise ValueError("signatures and messages differ in length")
    if len(signatures) == 0:
        return [], np.zeros(0, np.uint8)
    sigs = _as_bytes_array(signatures, 64, "signature")
    digs = _as_bytes_array(messages, 32, "message")
    addr, st = (ctx or default_context()).recover_address(curve, sigs, digs)
    return [bytes(r) for r in addr], st


__all__ = ["message_new", "recover_address"]
