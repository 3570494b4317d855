"""ctypes binding of the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is the checker: tests compare the HIP path against it. It is built by
``make -C oracle`` (``__graft_entry__.build()`` does this).
"""
import ctypes
import os
import subprocess

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(REPO, "oracle")
# SIGOPS_ORACLE_SO names another build of the same sources (the sanitizer test runs the tests on
# an AddressSanitizer / UBSan build through it)
ORACLE_SO = os.environ.get("SIGOPS_ORACLE_SO") or os.path.join(ORACLE_DIR, "liboracle.so")

K1, R1 = 0, 1
ST_OK, ST_BAD_R, ST_BAD_S, ST_NO_POINT, ST_INFINITY = 0, 1, 2, 3, 4

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
        L = ctypes.CDLL(ORACLE_SO)
        c_p = ctypes.c_char_p
        vp = ctypes.c_void_p
        L.or_sha512.argtypes = [c_p, ctypes.c_size_t, c_p]
        L.or_sha256.argtypes = [c_p, ctypes.c_size_t, c_p]
        L.or_secp_recover.argtypes = [ctypes.c_int, c_p, c_p, c_p]
        L.or_secp_pubkey.argtypes = [ctypes.c_int, c_p, c_p]
        L.or_secp_sign.argtypes = [ctypes.c_int, c_p, c_p, c_p, c_p]
        L.or_ed25519_verify.argtypes = [c_p, c_p, c_p, ctypes.c_size_t]
        L.or_ed25519_pubkey.argtypes = [c_p, c_p]
        L.or_ed25519_sign.argtypes = [c_p, c_p, ctypes.c_size_t, c_p]
        L.or_mod_mul.argtypes = [ctypes.c_int, c_p, c_p, c_p]
        L.or_mod_inv.argtypes = [ctypes.c_int, c_p, c_p]
        L.or_mod_pow.argtypes = [ctypes.c_int, c_p, c_p, c_p]
        L.or_mod_sqrt.argtypes = [ctypes.c_int, c_p, c_p]
        L.or_reduce512_l.argtypes = [c_p, c_p]
        L.or_secp_mul.argtypes = [ctypes.c_int, c_p, c_p, c_p]
        L.or_secp_mul2.argtypes = [ctypes.c_int, c_p, c_p, c_p, c_p]
        L.or_ed25519_decompress.argtypes = [c_p, c_p]
        L.or_secp_recover_batch.argtypes = [ctypes.c_int, vp, vp, ctypes.c_size_t, vp, vp, ctypes.c_int]
        L.or_ed25519_verify_batch.argtypes = [vp, vp, vp, vp, ctypes.c_size_t, vp, ctypes.c_int]
        L.or_ed25519_verify_strict_batch.argtypes = [vp, vp, vp, vp, ctypes.c_size_t, vp, ctypes.c_int]
        L.or_secp_verify.argtypes = [ctypes.c_int, c_p, c_p, c_p]
        L.or_secp_verify_batch.argtypes = [ctypes.c_int, vp, vp, vp, ctypes.c_size_t, vp, ctypes.c_int]
        L.or_ed25519_verify_strict.argtypes = [c_p, c_p, c_p, ctypes.c_size_t]
        L.or_ed25519_is_small_order.argtypes = [c_p]
        L.or_gen_secp.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_size_t, ctypes.c_uint32,
                                  vp, vp, vp, vp, ctypes.c_int]
        L.or_gen_ed25519.argtypes = [ctypes.c_uint64, ctypes.c_size_t, vp, ctypes.c_uint32,
                                     vp, vp, vp, vp, ctypes.c_int]
        L.or_item_random.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, c_p, ctypes.c_size_t]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def sha512(data: bytes) -> bytes:
    out = ctypes.create_string_buffer(64)
    lib().or_sha512(data, len(data), out)
    return out.raw


def sha256(data: bytes) -> bytes:
    out = ctypes.create_string_buffer(32)
    lib().or_sha256(data, len(data), out)
    return out.raw


def secp_recover(curve: int, sig: bytes, digest: bytes):
    out = ctypes.create_string_buffer(64)
    st = lib().or_secp_recover(curve, sig, digest, out)
    return st, out.raw


def secp_pubkey(curve: int, sk: bytes) -> bytes:
    out = ctypes.create_string_buffer(64)
    assert lib().or_secp_pubkey(curve, sk, out) == 0
    return out.raw


def secp_sign(curve: int, sk: bytes, digest: bytes, nonce: bytes):
    out = ctypes.create_string_buffer(64)
    rc = lib().or_secp_sign(curve, sk, digest, nonce, out)
    return None if rc else out.raw


def ed25519_verify(sig: bytes, pk: bytes, msg: bytes) -> bool:
    return bool(lib().or_ed25519_verify(sig, pk, msg, len(msg)))


def ed25519_verify_strict(sig: bytes, pk: bytes, msg: bytes) -> bool:
    return bool(lib().or_ed25519_verify_strict(sig, pk, msg, len(msg)))


def ed25519_is_small_order(enc: bytes) -> bool:
    return bool(lib().or_ed25519_is_small_order(enc))


def secp_verify(curve: int, sig: bytes, digest: bytes, pub: bytes) -> int:
    """status of fuel Signature::verify(pub, digest): 0 valid, else a SIGOPS_ST_* code"""
    return lib().or_secp_verify(curve, sig, digest, pub)


def ed25519_pubkey(seed: bytes) -> bytes:
    out = ctypes.create_string_buffer(32)
    lib().or_ed25519_pubkey(seed, out)
    return out.raw


def ed25519_sign(seed: bytes, msg: bytes) -> bytes:
    out = ctypes.create_string_buffer(64)
    lib().or_ed25519_sign(seed, msg, len(msg), out)
    return out.raw


def mod_op(name, mod, *args):
    out = ctypes.create_string_buffer(32)
    ints = [a.to_bytes(32, "big") for a in args]
    rc = getattr(lib(), "or_mod_" + name)(mod, *ints, out)
    return rc, int.from_bytes(out.raw, "big")


def reduce512_l(h: bytes) -> bytes:
    out = ctypes.create_string_buffer(32)
    lib().or_reduce512_l(h, out)
    return out.raw


def secp_mul(curve, P: bytes, k: int):
    out = ctypes.create_string_buffer(64)
    rc = lib().or_secp_mul(curve, P, k.to_bytes(32, "big"), out)
    return rc, out.raw


def secp_mul2(curve, P: bytes, u1: int, u2: int):
    out = ctypes.create_string_buffer(64)
    rc = lib().or_secp_mul2(curve, P, u1.to_bytes(32, "big"), u2.to_bytes(32, "big"), out)
    return rc, out.raw


def ed25519_decompress(b: bytes):
    out = ctypes.create_string_buffer(64)
    ok = lib().or_ed25519_decompress(b, out)
    return ok, out.raw


def recover_batch(curve, sigs: np.ndarray, digests: np.ndarray, nthreads=0):
    n = sigs.shape[0]
    out = np.zeros((n, 64), np.uint8)
    st = np.zeros(n, np.uint8)
    lib().or_secp_recover_batch(curve, _ptr(np.ascontiguousarray(sigs)), _ptr(np.ascontiguousarray(digests)),
                                n, _ptr(out), _ptr(st), nthreads)
    return out, st


def verify_batch(sigs, pks, msgs, msg_off, nthreads=0):
    n = sigs.shape[0]
    valid = np.zeros(n, np.uint8)
    msgs = np.ascontiguousarray(msgs) if msgs.size else np.zeros(1, np.uint8)
    lib().or_ed25519_verify_batch(_ptr(np.ascontiguousarray(sigs)), _ptr(np.ascontiguousarray(pks)), _ptr(msgs),
                                  _ptr(np.ascontiguousarray(msg_off, dtype=np.uint64)), n, _ptr(valid), nthreads)
    return valid


def verify_strict_batch(sigs, pks, msgs, msg_off, nthreads=0):
    n = sigs.shape[0]
    valid = np.zeros(n, np.uint8)
    msgs = np.ascontiguousarray(msgs) if msgs.size else np.zeros(1, np.uint8)
    lib().or_ed25519_verify_strict_batch(_ptr(np.ascontiguousarray(sigs)), _ptr(np.ascontiguousarray(pks)),
                                         _ptr(msgs), _ptr(np.ascontiguousarray(msg_off, dtype=np.uint64)), n,
                                         _ptr(valid), nthreads)
    return valid


def secp_verify_batch(curve, sigs, digests, pubs, nthreads=0):
    n = sigs.shape[0]
    st = np.zeros(n, np.uint8)
    lib().or_secp_verify_batch(curve, _ptr(np.ascontiguousarray(sigs)), _ptr(np.ascontiguousarray(digests)),
                               _ptr(np.ascontiguousarray(pubs)), n, _ptr(st), nthreads)
    return st


def gen_secp(curve, seed, n, corrupt_ppm=0, nthreads=0):
    sigs = np.zeros((n, 64), np.uint8)
    digests = np.zeros((n, 32), np.uint8)
    pubs = np.zeros((n, 64), np.uint8)
    corrupted = np.zeros(n, np.uint8)
    lib().or_gen_secp(curve, seed, n, corrupt_ppm, _ptr(sigs), _ptr(digests), _ptr(pubs), _ptr(corrupted), nthreads)
    return sigs, digests, pubs, corrupted


def gen_ed25519(seed, n, msg_off, corrupt_ppm=0, nthreads=0):
    msg_off = np.ascontiguousarray(msg_off, dtype=np.uint64)
    sigs = np.zeros((n, 64), np.uint8)
    pks = np.zeros((n, 32), np.uint8)
    msgs = np.zeros(max(int(msg_off[-1]), 1), np.uint8)
    corrupted = np.zeros(n, np.uint8)
    lib().or_gen_ed25519(seed, n, _ptr(msg_off), corrupt_ppm, _ptr(sigs), _ptr(pks), _ptr(msgs), _ptr(corrupted),
                         nthreads)
    return sigs, pks, msgs[: int(msg_off[-1])], corrupted
