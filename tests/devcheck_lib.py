"""ctypes binding of tests/devcheck/libdevcheck.so: the device headers built for the host."""
import ctypes
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
SRC = os.path.join(HERE, "devcheck", "devcheck.cpp")
SO = os.path.join(HERE, "devcheck", "libdevcheck.so")
CSRC = os.path.join(REPO, "wgpu-sigops_amd", "csrc")

_lib = None


def build():
    if os.environ.get("SIGOPS_DEVCHECK_SO"):  # another build of the same sources (the sanitizer test)
        return os.environ["SIGOPS_DEVCHECK_SO"]
    deps = [SRC] + [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hpp")]
    if os.path.exists(SO) and all(os.path.getmtime(SO) >= os.path.getmtime(d) for d in deps):
        return SO
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wno-unknown-pragmas", "-DSIGOPS_HOST_CHECK=1", "-I" + CSRC,
                    "-o", SO, SRC], check=True)
    return SO


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(build())
        c_p = ctypes.c_char_p
        L.dc_secp_recover.argtypes = [ctypes.c_int, c_p, c_p, c_p, ctypes.c_int]
        L.dc_ed25519_verify.argtypes = [c_p, c_p, c_p, ctypes.c_uint, ctypes.c_int]
        L.dc_field.argtypes = [ctypes.c_int, ctypes.c_int, c_p, c_p, c_p]
        L.dc_sha512.argtypes = [c_p, ctypes.c_int, c_p, ctypes.c_uint, ctypes.c_uint, c_p]
        u32p = ctypes.POINTER(ctypes.c_uint32)
        L.dc_mont_raw.argtypes = [ctypes.c_int, ctypes.c_int, u32p, u32p, u32p]
        L.dc_glv_split.argtypes = [u32p, u32p, u32p]
        L.dc_sha256.argtypes = [c_p, ctypes.c_uint, ctypes.c_uint, c_p, ctypes.c_int]
        L.dc_secp_verify.argtypes = [ctypes.c_int, c_p, c_p, c_p, ctypes.c_int]
        L.dc_table.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_uint, ctypes.c_void_p]
        L.dc_table.restype = ctypes.c_uint
        L.dc_resident_table.argtypes = [ctypes.c_int, ctypes.c_void_p]
        L.dc_sha256_64.argtypes = [c_p, c_p]
        L.dc_ed_lattice.argtypes = [u32p, u32p, u32p, ctypes.POINTER(ctypes.c_int)]
        L.dc_set_ed_force_full.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def secp_recover(curve, sig, dig, force_careful=False, g_window=8):
    out = ctypes.create_string_buffer(64)
    st = lib().dc_secp_recover(curve, sig, dig, out, int(force_careful) | (2 if g_window == 9 else 0))
    return st, out.raw


def ed25519_verify(sig, pk, msg, strict=False, lattice=True):
    """status of the device verification built for the host: the lattice-reduced check of the
    Ed25519-only kernels, or (lattice=False) the full-length check of the mixed-scheme kernels"""
    return lib().dc_ed25519_verify(sig, pk, msg, len(msg), int(strict) | (0 if lattice else 2))


def ed_lattice(k):
    """the device's lattice reduction for challenge k: (found, c0, c1) with c0 signed"""
    kw = (ctypes.c_uint32 * 8)(*[(k >> (32 * i)) & 0xFFFFFFFF for i in range(8)])
    c0 = (ctypes.c_uint32 * 5)()
    c1 = (ctypes.c_uint32 * 5)()
    neg = ctypes.c_int(0)
    ok = lib().dc_ed_lattice(kw, c0, c1, ctypes.byref(neg))
    v0 = sum(int(c0[i]) << (32 * i) for i in range(5))
    v1 = sum(int(c1[i]) << (32 * i) for i in range(5))
    return bool(ok), -v0 if neg.value else v0, v1


def set_ed_force_full(on):
    lib().dc_set_ed_force_full(int(on))


def field(op, mod, a, b=0):
    out = ctypes.create_string_buffer(32)
    lib().dc_field(op, mod, a.to_bytes(32, "big"), b.to_bytes(32, "big"), out)
    return int.from_bytes(out.raw, "big")


def sha512(hdr, msg, at=0):
    """SHA-512(hdr || msg) by the device code built for the host, hdr of 32 or 64 bytes, msg at byte
    offset `at` of its buffer; raises if the hash read memory outside the message's words"""
    out = ctypes.create_string_buffer(64)
    rc = lib().dc_sha512(bytes(hdr), len(hdr), bytes(msg), len(msg), at, out)
    if rc != 0:
        raise AssertionError(f"SHA-512 of {len(msg)} bytes at offset {at} read outside the message's words")
    return out.raw


def sha512_ed(hdr, msg):
    return sha512(hdr, msg)


def mont_raw(op, mod, a_limbs, b_limbs):
    """mont_mul (op 0) or mont_sqr (op 1) of raw 9-limb vectors; returns the 9 output limbs"""
    A = (ctypes.c_uint32 * 9)(*a_limbs)
    B = (ctypes.c_uint32 * 9)(*b_limbs)
    O = (ctypes.c_uint32 * 9)()
    lib().dc_mont_raw(op, mod, A, B, O)
    return list(O)


def glv_split(k):
    """device GLV split of k < n (secp256k1): returns signed halves (k1, k2)"""
    K = (ctypes.c_uint32 * 8)(*[(k >> (32 * i)) & 0xFFFFFFFF for i in range(8)])
    H1 = (ctypes.c_uint32 * 5)()
    H2 = (ctypes.c_uint32 * 5)()
    sg = lib().dc_glv_split(K, H1, H2)
    h1 = sum(v << (32 * i) for i, v in enumerate(H1))
    h2 = sum(v << (32 * i) for i, v in enumerate(H2))
    return (-h1 if sg & 1 else h1), (-h2 if sg & 2 else h2)


def secp_verify(curve, sig, dig, pub, force_careful=False, g_window=8):
    """(status, careful path ran) of the device verification built for the host"""
    st = lib().dc_secp_verify(curve, bytes(sig), bytes(dig), bytes(pub), int(force_careful) | (2 if g_window == 9 else 0))
    return st & 0xFF, bool(st >> 8)


def table(which, limbs, lbits):
    """caller-table conversion (0 k1, 1 r1, 2 ed25519): (TAB_BAD_* bits, (128, 20) uint32 table, or
    (256, 28) for Ed25519: the odd multiples of the point, then those of [2^136] times it)"""
    import numpy as np
    limbs = np.ascontiguousarray(limbs, dtype=np.uint32)
    out = np.zeros((256, 28) if which == 2 else (128, 20), np.uint32)
    bad = lib().dc_table(which, limbs.ctypes.data, lbits, out.ctypes.data)
    return bad, out


def resident_table(which):
    import numpy as np
    out = np.zeros((256, 28) if which == 2 else (256, 20), np.uint32)
    lib().dc_resident_table(which, out.ctypes.data)
    return out


def sha256(msg, at=0, x4=False):
    """x4: the reader with 16-byte loads (from a 16-byte-aligned buffer)"""
    out = ctypes.create_string_buffer(32)
    rc = lib().dc_sha256(bytes(msg), len(msg), at, out, int(x4))
    if rc != 0:
        unit = "16-byte chunks" if x4 else "words"
        raise AssertionError(f"SHA-256 of {len(msg)} bytes at offset {at} read outside the message's {unit}")
    return out.raw


def sha256_64(data):
    out = ctypes.create_string_buffer(32)
    lib().dc_sha256_64(bytes(data), out)
    return out.raw
