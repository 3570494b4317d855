"""Mirror of the reference's secp256k1 recovery API (src/secp256k1_ecdsa.rs).

``ecrecover(signatures, messages, table_limbs, log_limb_size)`` and
``ecrecover_single_shader(signatures, messages, log_limb_size)`` keep the reference's argument
meaning and results: fuel-format 64-byte signatures (recovery parity in bit 7 of byte 32), 32-byte
message digests, and a list of 64-byte X||Y public keys. Like the reference they raise
``ShaderFailureError`` when the GPU did not produce the results, fail on a length mismatch
(src/secp256k1_ecdsa.rs:21) and return [] for no input (:71-73). An invalid signature yields 64 zero
bytes (the reference has no per-item error channel; ``recover_with_status`` exposes the codes).
"""
import numpy as np

from . import precompute
from ._secp_common import ecrecover_impl, verify_impl

CURVE = 0


def ecrecover(signatures, messages, table_limbs, log_limb_size, ctx=None):
    precompute.check_table(table_limbs, 2, log_limb_size)
    return ecrecover_impl(CURVE, signatures, messages, ctx, table=(table_limbs, log_limb_size))[0]


def ecrecover_single_shader(signatures, messages, log_limb_size, ctx=None):
    precompute.check_log_limb_size(log_limb_size)
    return ecrecover_impl(CURVE, signatures, messages, ctx)[0]


ecrecover_single = ecrecover_single_shader


def recover_with_status(signatures, messages, ctx=None):
    """(list of 64-byte keys, numpy status codes)"""
    return ecrecover_impl(CURVE, signatures, messages, ctx)


def verify(signatures, messages, public_keys, ctx=None):
    """fuel ``Signature::verify(&public_key, &message)`` for each item: a list of bools.
    Rules of libsecp256k1 (lower-S signatures only); the recovery parity bit is ignored.
    No counterpart among the reference's entry points."""
    return verify_impl(CURVE, signatures, messages, public_keys, ctx)[0]


def verify_with_status(signatures, messages, public_keys, ctx=None):
    """(list of bools, numpy status codes SIGOPS_ST_*)"""
    return verify_impl(CURVE, signatures, messages, public_keys, ctx)


__all__ = ["ecrecover", "ecrecover_single_shader", "ecrecover_single", "recover_with_status", "verify",
           "verify_with_status", "np"]
