"""OpenSSL libcrypto (system library, called through ctypes) as an independent cross-check of the
oracle: ECDSA signing and EC point multiplication on secp256k1 / P-256, and Ed25519 keys and
signatures. Used only by the CPU tests."""
import ctypes
import ctypes.util

NID = {0: 714, 1: 415}  # NID_secp256k1, NID_X9_62_prime256v1
EVP_PKEY_ED25519 = 1087

_c = None


def available():
    try:
        crypto()
        return True
    except OSError:
        return False


def crypto():
    global _c
    if _c is None:
        name = ctypes.util.find_library("crypto") or "libcrypto.so.3"
        C = ctypes.CDLL(name)
        vp = ctypes.c_void_p
        for fn, res, args in [
            ("EC_KEY_new_by_curve_name", vp, [ctypes.c_int]),
            ("EC_KEY_free", None, [vp]),
            ("EC_KEY_get0_group", vp, [vp]),
            ("EC_KEY_set_private_key", ctypes.c_int, [vp, vp]),
            ("EC_KEY_set_public_key", ctypes.c_int, [vp, vp]),
            ("EC_POINT_new", vp, [vp]),
            ("EC_POINT_free", None, [vp]),
            ("EC_POINT_mul", ctypes.c_int, [vp, vp, vp, vp, vp, vp]),
            ("EC_POINT_oct2point", ctypes.c_int, [vp, vp, ctypes.c_char_p, ctypes.c_size_t, vp]),
            ("EC_POINT_point2oct", ctypes.c_size_t, [vp, vp, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t, vp]),
            ("EC_POINT_is_at_infinity", ctypes.c_int, [vp, vp]),
            ("EC_GROUP_get_order", ctypes.c_int, [vp, vp, vp]),
            ("BN_new", vp, []),
            ("BN_bin2bn", vp, [ctypes.c_char_p, ctypes.c_int, vp]),
            ("BN_bn2binpad", ctypes.c_int, [vp, ctypes.c_char_p, ctypes.c_int]),
            ("BN_free", None, [vp]),
            ("ECDSA_do_sign", vp, [ctypes.c_char_p, ctypes.c_int, vp]),
            ("ECDSA_SIG_get0", None, [vp, ctypes.POINTER(vp), ctypes.POINTER(vp)]),
            ("ECDSA_SIG_free", None, [vp]),
            ("ECDSA_SIG_new", vp, []),
            ("ECDSA_SIG_set0", ctypes.c_int, [vp, vp, vp]),
            ("ECDSA_do_verify", ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, vp, vp]),
            ("EVP_PKEY_new_raw_private_key", vp, [ctypes.c_int, vp, ctypes.c_char_p, ctypes.c_size_t]),
            ("EVP_PKEY_new_raw_public_key", vp, [ctypes.c_int, vp, ctypes.c_char_p, ctypes.c_size_t]),
            ("EVP_PKEY_get_raw_public_key", ctypes.c_int, [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_size_t)]),
            ("EVP_PKEY_free", None, [vp]),
            ("EVP_MD_CTX_new", vp, []),
            ("EVP_MD_CTX_free", None, [vp]),
            ("EVP_DigestSignInit", ctypes.c_int, [vp, vp, vp, vp, vp]),
            ("EVP_DigestSign", ctypes.c_int, [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_size_t), ctypes.c_char_p,
                                              ctypes.c_size_t]),
            ("EVP_DigestVerifyInit", ctypes.c_int, [vp, vp, vp, vp, vp]),
            ("EVP_DigestVerify", ctypes.c_int, [vp, ctypes.c_char_p, ctypes.c_size_t, ctypes.c_char_p,
                                                ctypes.c_size_t]),
        ]:
            f = getattr(C, fn)
            f.restype = res
            f.argtypes = args
        _c = C
    return _c


def _bn(x: int):
    b = x.to_bytes(32, "big")
    return crypto().BN_bin2bn(b, 32, None)


def pubkey(curve: int, sk: int) -> bytes:
    C = crypto()
    key = C.EC_KEY_new_by_curve_name(NID[curve])
    grp = C.EC_KEY_get0_group(key)
    pt = C.EC_POINT_new(grp)
    k = _bn(sk)
    assert C.EC_POINT_mul(grp, pt, k, None, None, None) == 1
    buf = ctypes.create_string_buffer(65)
    assert C.EC_POINT_point2oct(grp, pt, 4, buf, 65, None) == 65
    C.BN_free(k)
    C.EC_POINT_free(pt)
    C.EC_KEY_free(key)
    return buf.raw[1:]


def point_mul(curve: int, P: bytes, k: int):
    """k*P for P = X||Y (64 bytes); returns None for infinity."""
    C = crypto()
    key = C.EC_KEY_new_by_curve_name(NID[curve])
    grp = C.EC_KEY_get0_group(key)
    pt = C.EC_POINT_new(grp)
    res = C.EC_POINT_new(grp)
    assert C.EC_POINT_oct2point(grp, pt, b"\x04" + P, 65, None) == 1
    kb = _bn(k)
    assert C.EC_POINT_mul(grp, res, None, pt, kb, None) == 1
    out = None
    if not C.EC_POINT_is_at_infinity(grp, res):
        buf = ctypes.create_string_buffer(65)
        assert C.EC_POINT_point2oct(grp, res, 4, buf, 65, None) == 65
        out = buf.raw[1:]
    C.BN_free(kb)
    C.EC_POINT_free(pt)
    C.EC_POINT_free(res)
    C.EC_KEY_free(key)
    return out


def ecdsa_sign(curve: int, sk: int, digest: bytes):
    """OpenSSL ECDSA signature (random nonce) -> (r, s) integers."""
    C = crypto()
    key = C.EC_KEY_new_by_curve_name(NID[curve])
    grp = C.EC_KEY_get0_group(key)
    k = _bn(sk)
    assert C.EC_KEY_set_private_key(key, k) == 1
    pt = C.EC_POINT_new(grp)
    assert C.EC_POINT_mul(grp, pt, k, None, None, None) == 1
    assert C.EC_KEY_set_public_key(key, pt) == 1
    sig = C.ECDSA_do_sign(digest, len(digest), key)
    assert sig
    r = ctypes.c_void_p()
    s = ctypes.c_void_p()
    C.ECDSA_SIG_get0(sig, ctypes.byref(r), ctypes.byref(s))
    rb = ctypes.create_string_buffer(32)
    sb = ctypes.create_string_buffer(32)
    C.BN_bn2binpad(r, rb, 32)
    C.BN_bn2binpad(s, sb, 32)
    C.ECDSA_SIG_free(sig)
    C.EC_POINT_free(pt)
    C.BN_free(k)
    C.EC_KEY_free(key)
    return int.from_bytes(rb.raw, "big"), int.from_bytes(sb.raw, "big")


def ecdsa_verify(curve: int, pub: bytes, digest: bytes, r: int, s: int):
    """OpenSSL ECDSA_do_verify of (r, s) on digest under pub = X||Y: True / False, or None when
    OpenSSL refuses the key itself (not a curve point)"""
    C = crypto()
    key = C.EC_KEY_new_by_curve_name(NID[curve])
    grp = C.EC_KEY_get0_group(key)
    pt = C.EC_POINT_new(grp)
    try:
        if C.EC_POINT_oct2point(grp, pt, b"\x04" + pub, 65, None) != 1 or C.EC_KEY_set_public_key(key, pt) != 1:
            return None
        sig = C.ECDSA_SIG_new()
        rb, sb = crypto().BN_bin2bn(r.to_bytes(32, "big"), 32, None), crypto().BN_bin2bn(s.to_bytes(32, "big"), 32, None)
        assert C.ECDSA_SIG_set0(sig, rb, sb) == 1  # the signature object now owns rb, sb
        rc = C.ECDSA_do_verify(digest, len(digest), sig, key)
        C.ECDSA_SIG_free(sig)
        return rc == 1
    finally:
        C.EC_POINT_free(pt)
        C.EC_KEY_free(key)


def ed25519_pubkey(seed: bytes) -> bytes:
    C = crypto()
    pk = C.EVP_PKEY_new_raw_private_key(EVP_PKEY_ED25519, None, seed, 32)
    out = ctypes.create_string_buffer(32)
    ln = ctypes.c_size_t(32)
    assert C.EVP_PKEY_get_raw_public_key(pk, out, ctypes.byref(ln)) == 1
    C.EVP_PKEY_free(pk)
    return out.raw


def ed25519_sign(seed: bytes, msg: bytes) -> bytes:
    C = crypto()
    pk = C.EVP_PKEY_new_raw_private_key(EVP_PKEY_ED25519, None, seed, 32)
    ctx = C.EVP_MD_CTX_new()
    assert C.EVP_DigestSignInit(ctx, None, None, None, pk) == 1
    out = ctypes.create_string_buffer(64)
    ln = ctypes.c_size_t(64)
    assert C.EVP_DigestSign(ctx, out, ctypes.byref(ln), msg, len(msg)) == 1
    C.EVP_MD_CTX_free(ctx)
    C.EVP_PKEY_free(pk)
    return out.raw


def ed25519_verify(sig: bytes, pub: bytes, msg: bytes) -> bool:
    C = crypto()
    pk = C.EVP_PKEY_new_raw_public_key(EVP_PKEY_ED25519, None, pub, 32)
    if not pk:
        return False
    ctx = C.EVP_MD_CTX_new()
    assert C.EVP_DigestVerifyInit(ctx, None, None, None, pk) == 1
    rc = C.EVP_DigestVerify(ctx, sig, 64, msg, len(msg))
    C.EVP_MD_CTX_free(ctx)
    C.EVP_PKEY_free(pk)
    return rc == 1


def group_order(curve: int) -> int:
    """the order n of the curve's group, as OpenSSL holds it"""
    C = crypto()
    key = C.EC_KEY_new_by_curve_name(NID[curve])
    grp = C.EC_KEY_get0_group(key)
    bn = C.BN_new()
    assert C.EC_GROUP_get_order(grp, bn, None) == 1
    buf = ctypes.create_string_buffer(32)
    assert C.BN_bn2binpad(bn, buf, 32) == 32
    C.BN_free(bn)
    C.EC_KEY_free(key)
    return int.from_bytes(buf.raw, "big")


def ecdsa_recover(curve: int, sig: bytes, digest: bytes):
    """Public-key recovery of a fuel-format signature (r || s, recovery parity in bit 7 of byte 32)
    with OpenSSL doing the curve arithmetic: R is lifted from r by OpenSSL's point decompression
    (prefix 02 / 03 by the parity; it fails when r is not an x-coordinate or r >= p) and
    Q = u1 G + u2 R by EC_POINT_mul, u1 = -z / r, u2 = s / r mod n. The acceptance rules around
    them are libsecp256k1's secp256k1_ecdsa_recover (r, s in [1, n-1], z = digest mod n, R.x = r,
    Q not at infinity), which fuel-crypto 0.49 calls for both curves' recover paths. Returns
    (status, X||Y) with the oracle's status codes (1 bad r, 2 bad s, 3 no point, 4 infinity) and
    64 zero bytes on failure."""
    C = crypto()
    n = group_order(curve)
    r = int.from_bytes(sig[:32], "big")
    odd = sig[32] >> 7
    s = int.from_bytes(bytes([sig[32] & 0x7F]) + sig[33:], "big")
    if not 0 < r < n:
        return 1, bytes(64)
    if not 0 < s < n:
        return 2, bytes(64)
    key = C.EC_KEY_new_by_curve_name(NID[curve])
    grp = C.EC_KEY_get0_group(key)
    R = C.EC_POINT_new(grp)
    Q = C.EC_POINT_new(grp)
    try:
        if C.EC_POINT_oct2point(grp, R, bytes([2 + odd]) + r.to_bytes(32, "big"), 33, None) != 1:
            return 3, bytes(64)
        z = int.from_bytes(digest, "big") % n
        rinv = pow(r, -1, n)
        u1, u2 = (-z * rinv) % n, s * rinv % n
        b1, b2 = _bn(u1), _bn(u2)
        assert C.EC_POINT_mul(grp, Q, b1, R, b2, None) == 1
        C.BN_free(b1)
        C.BN_free(b2)
        if C.EC_POINT_is_at_infinity(grp, Q):
            return 4, bytes(64)
        buf = ctypes.create_string_buffer(65)
        assert C.EC_POINT_point2oct(grp, Q, 4, buf, 65, None) == 65
        return 0, buf.raw[1:]
    finally:
        C.EC_POINT_free(R)
        C.EC_POINT_free(Q)
        C.EC_KEY_free(key)


def point_mul2(curve: int, P: bytes, u1: int, u2: int):
    """u1 G + u2 P for P = X||Y with OpenSSL's EC_POINT_mul; None for infinity"""
    C = crypto()
    key = C.EC_KEY_new_by_curve_name(NID[curve])
    grp = C.EC_KEY_get0_group(key)
    pt = C.EC_POINT_new(grp)
    res = C.EC_POINT_new(grp)
    try:
        assert C.EC_POINT_oct2point(grp, pt, b"\x04" + P, 65, None) == 1
        b1, b2 = _bn(u1), _bn(u2)
        assert C.EC_POINT_mul(grp, res, b1, pt, b2, None) == 1
        C.BN_free(b1)
        C.BN_free(b2)
        if C.EC_POINT_is_at_infinity(grp, res):
            return None
        buf = ctypes.create_string_buffer(65)
        assert C.EC_POINT_point2oct(grp, res, 4, buf, 65, None) == 65
        return buf.raw[1:]
    finally:
        C.EC_POINT_free(pt)
        C.EC_POINT_free(res)
        C.EC_KEY_free(key)


# ---- CPU baseline timing (oracle/libossl_baseline.so, C over libcrypto, threaded) ----
_base = None


def baseline_lib():
    """oracle/libossl_baseline.so (built by oracle/Makefile); None when it is not built"""
    global _base
    if _base is None:
        import os
        so = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "libossl_baseline.so")
        if not os.path.exists(so):
            return None
        L = ctypes.CDLL(so)
        vp = ctypes.c_void_p
        L.ossl_ecdsa_verify_time.restype = ctypes.c_double
        L.ossl_ecdsa_verify_time.argtypes = [ctypes.c_int, vp, vp, vp, ctypes.c_size_t, ctypes.c_int, vp]
        L.ossl_ed25519_verify_time.restype = ctypes.c_double
        L.ossl_ed25519_verify_time.argtypes = [vp, vp, vp, vp, ctypes.c_size_t, ctypes.c_int, vp]
        _base = L
    return _base


def time_ecdsa_verify(curve, sigs, digs, pubs, nthreads):
    """(seconds, valid flags) for OpenSSL ECDSA verification of fuel-format signatures (numpy
    uint8 arrays n x 64, n x 32, n x 64 X||Y) on nthreads threads"""
    import numpy as np
    sigs, digs, pubs = (np.ascontiguousarray(a, dtype=np.uint8) for a in (sigs, digs, pubs))
    n = sigs.shape[0]
    valid = np.zeros(n, np.uint8)
    t = baseline_lib().ossl_ecdsa_verify_time(curve, sigs.ctypes.data, digs.ctypes.data, pubs.ctypes.data, n,
                                              nthreads, valid.ctypes.data)
    return t, valid


def time_ed25519_verify(sigs, pubs, msgs, off, nthreads):
    """(seconds, valid flags) for OpenSSL Ed25519 verification on nthreads threads"""
    import numpy as np
    sigs, pubs, msgs = (np.ascontiguousarray(a, dtype=np.uint8) for a in (sigs, pubs, msgs))
    off = np.ascontiguousarray(off, dtype=np.uint64)
    n = sigs.shape[0]
    valid = np.zeros(n, np.uint8)
    mp = msgs.ctypes.data if msgs.size else np.zeros(1, np.uint8).ctypes.data
    t = baseline_lib().ossl_ed25519_verify_time(sigs.ctypes.data, pubs.ctypes.data, mp, off.ctypes.data, n, nthreads,
                                                valid.ctypes.data)
    return t, valid
