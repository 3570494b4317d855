"""Mirror of the reference's Ed25519 verification API (src/ed25519_eddsa.rs).

``ecverify(signatures, messages, verifying_keys, table_limbs, log_limb_size)`` and
``ecverify_single(signatures, messages, verifying_keys, log_limb_size)`` take 64-byte R||s
signatures, 32-byte messages (fuel ``Message`` digests, which is what the reference hashes,
src/ed25519_eddsa.rs:39) and 32-byte verifying keys, and return a list of bools under
ed25519-dalek ``verify`` rules. ``ecverify_bytes`` is the variable-length extension (messages of
any length). ``ecverify_strict`` applies dalek ``verify_strict`` rules instead (R must decompress,
neither R nor the key may have small order), the adjacent predicate the reference names
(src/curve_algos/ed25519_eddsa.rs:197-198). Errors as in the reference: ShaderFailureError when results are missing, [] for no
input, IndexError when fewer keys than signatures are given (the reference indexes, :38).
"""
import numpy as np

from . import precompute
from ._secp_common import _as_bytes_array


def _verify(signatures, messages, verifying_keys, ctx, fixed_len, strict=False, table=None):
    from . import default_context

    if len(signatures) != len(messages):
        raise ValueError("signatures and messages differ in length")
    n = len(signatures)
    if n == 0:
        return []
    if len(verifying_keys) < n:
        raise IndexError("fewer verifying keys than signatures")
    sigs = _as_bytes_array(signatures, 64, "signature")
    pks = _as_bytes_array(list(verifying_keys)[:n] if not isinstance(verifying_keys, np.ndarray)
                          else verifying_keys[:n], 32, "verifying key")
    if fixed_len:
        msgs = _as_bytes_array(messages, 32, "message").reshape(-1)
        off = np.arange(n + 1, dtype=np.uint64) * 32
    else:
        parts = [bytes(m) for m in messages]
        off = np.zeros(n + 1, np.uint64)
        off[1:] = np.cumsum([len(p) for p in parts])
        msgs = np.frombuffer(b"".join(parts), dtype=np.uint8)
    ctx = ctx or default_context()
    if table is not None:
        valid, _ = ctx.ed25519_verify_table(sigs, pks, msgs, off, table[0], table[1])
    else:
        valid, _ = (ctx.ed25519_verify_strict if strict else ctx.ed25519_verify)(sigs, pks, msgs, off)
    return [bool(v) for v in valid]


def ecverify(signatures, messages, verifying_keys, table_limbs, log_limb_size, ctx=None):
    precompute.check_table(table_limbs, 3, log_limb_size)
    return _verify(signatures, messages, verifying_keys, ctx, True, table=(table_limbs, log_limb_size))


def ecverify_single(signatures, messages, verifying_keys, log_limb_size, ctx=None):
    precompute.check_log_limb_size(log_limb_size)
    return _verify(signatures, messages, verifying_keys, ctx, True)


def ecverify_bytes(signatures, messages, verifying_keys, ctx=None):
    """Variable-length messages (each message is any bytes object)."""
    return _verify(signatures, messages, verifying_keys, ctx, False)


def ecverify_strict(signatures, messages, verifying_keys, ctx=None):
    """ed25519-dalek ``verify_strict`` for each item (messages of any length): a list of bools."""
    return _verify(signatures, messages, verifying_keys, ctx, False, strict=True)
