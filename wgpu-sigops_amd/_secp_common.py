"""Shared body of the secp256k1 / secp256r1 mirrors."""
import numpy as np


def _as_bytes_array(items, width, what):
    if isinstance(items, np.ndarray):
        arr = np.ascontiguousarray(items, dtype=np.uint8).reshape(-1, width)
        return arr
    buf = bytearray()
    for x in items:
        b = bytes(x)
        if len(b) != width:
            raise ValueError(f"each {what} must be {width} bytes")
        buf += b
    return np.frombuffer(bytes(buf), dtype=np.uint8).reshape(-1, width)


def ecrecover_impl(curve, signatures, messages, ctx, table=None):
    """table: None (the resident generator table) or (table_limbs, log_limb_size) in the
    reference's layout, converted and used as the fixed base on the device"""
    from . import default_context

    if len(signatures) != len(messages):
        raise ValueError("signatures and messages differ in length")
    if len(signatures) > 256 * 256 * 256 * 64:
        raise ValueError("too many signatures")
    if len(signatures) == 0:
        return [], np.zeros(0, np.uint8)
    sigs = _as_bytes_array(signatures, 64, "signature")
    digs = _as_bytes_array(messages, 32, "message")
    ctx = ctx or default_context()
    if table is not None:
        out, st = ctx.secp_recover_table(curve, sigs, digs, table[0], table[1])
    else:
        fn = ctx.secp256k1_recover if curve == 0 else ctx.secp256r1_recover
        out, st = fn(sigs, digs)
    return [bytes(row) for row in out], st


def verify_impl(curve, signatures, messages, public_keys, ctx):
    """fuel Signature::verify over a batch -> (list of bool, numpy status codes)"""
    from . import default_context

    if not (len(signatures) == len(messages) == len(public_keys)):
        raise ValueError("signatures, messages and public keys differ in length")
    if len(signatures) > 256 * 256 * 256 * 64:
        raise ValueError("too many signatures")
    if len(signatures) == 0:
        return [], np.zeros(0, np.uint8)
    sigs = _as_bytes_array(signatures, 64, "signature")
    digs = _as_bytes_array(messages, 32, "message")
    pubs = _as_bytes_array(public_keys, 64, "public key")
    ctx = ctx or default_context()
    fn = ctx.secp256k1_verify if curve == 0 else ctx.secp256r1_verify
    valid, st = fn(sigs, digs, pubs)
    return [bool(v) for v in valid], st
