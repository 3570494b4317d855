"""Independent pure-Python closed forms used to pin the oracle (small inputs only).

Affine short-Weierstrass arithmetic with Python integers and pow(), and Ed25519 in the style of
RFC 8032 section 6 (affine Edwards addition, recover_x via the (p+3)/8 exponent). These use
different formulas from the oracle's Montgomery/Jacobian/extended-coordinate code, so a
dropped term, a sign or an index error in either shows up as a mismatch.
"""
import hashlib

CURVES = {
    0: dict(
        p=2**256 - 2**32 - 977,
        n=0xFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFEBAAEDCE6AF48A03BBFD25E8CD0364141,
        a=0,
        b=7,
        G=(0x79BE667EF9DCBBAC55A06295CE870B07029BFCDB2DCE28D959F2815B16F81798,
           0x483ADA7726A3C4655DA4FBFC0E1108A8FD17B448A68554199C47D08FFB10D4B8),
    ),
    1: dict(
        p=2**256 - 2**224 + 2**192 + 2**96 - 1,
        n=0xFFFFFFFF00000000FFFFFFFFFFFFFFFFBCE6FAADA7179E84F3B9CAC2FC632551,
        a=-3,
        b=0x5AC635D8AA3A93E7B3EBBD55769886BC651D06B0CC53B0F63BCE3C3E27D2604B,
        G=(0x6B17D1F2E12C4247F8BCE6E563A440F277037D812DEB33A0F4A13945D898C296,
           0x4FE342E2FE1A7F9B8EE7EB4A7C0F9E162BCE33576B315ECECBB6406837BF51F5),
    ),
}


def w_add(c, P, Q):
    p, a = c["p"], c["a"]
    if P is None:
        return Q
    if Q is None:
        return P
    if P[0] == Q[0]:
        if (P[1] + Q[1]) % p == 0:
            return None
        lam = (3 * P[0] * P[0] + a) * pow(2 * P[1], -1, p) % p
    else:
        lam = (Q[1] - P[1]) * pow(Q[0] - P[0], -1, p) % p
    x = (lam * lam - P[0] - Q[0]) % p
    return (x, (lam * (P[0] - x) - P[1]) % p)


def w_mul(c, P, k):
    R = None
    for bit in bin(k)[2:] if k > 0 else "":
        R = w_add(c, R, R)
        if bit == "1":
            R = w_add(c, R, P)
    return R


def secp_recover(curve, sig: bytes, digest: bytes):
    """fuel-crypto / libsecp256k1 / ecdsa-0.16 recovery rules; returns (status, 64 bytes)."""
    c = CURVES[curve]
    p, n = c["p"], c["n"]
    r = int.from_bytes(sig[:32], "big")
    sb = bytearray(sig[32:])
    odd = sb[0] >> 7
    sb[0] &= 0x7F
    s = int.from_bytes(bytes(sb), "big")
    z = int.from_bytes(digest, "big") % n
    if not 0 < r < n:
        return 1, bytes(64)
    if not 0 < s < n:
        return 2, bytes(64)
    rhs = (r**3 + c["a"] * r + c["b"]) % p
    y = pow(rhs, (p + 1) // 4, p)
    if y * y % p != rhs:
        return 3, bytes(64)
    if y % 2 != odd:
        y = p - y
    rinv = pow(r, -1, n)
    u1 = (-z * rinv) % n
    u2 = (s * rinv) % n
    Q = w_add(c, w_mul(c, c["G"], u1), w_mul(c, (r, y), u2))
    if Q is None:
        return 4, bytes(64)
    return 0, Q[0].to_bytes(32, "big") + Q[1].to_bytes(32, "big")


def secp_verify(curve, sig: bytes, digest: bytes, pub: bytes) -> int:
    """fuel Signature::verify(pub, digest): libsecp256k1 rules for secp256k1 (lower-S only),
    p256 / ecdsa 0.16 rules for secp256r1; returns 0 or the SIGOPS_ST_* rejection code."""
    c = CURVES[curve]
    p, n = c["p"], c["n"]
    r = int.from_bytes(sig[:32], "big")
    sb = bytearray(sig[32:])
    sb[0] &= 0x7F
    s = int.from_bytes(bytes(sb), "big")
    if not 0 < r < n:
        return 1
    if not 0 < s < n:
        return 2
    if curve == 0 and s > (n - 1) // 2:
        return 5
    x, y = int.from_bytes(pub[:32], "big"), int.from_bytes(pub[32:], "big")
    if x >= p or y >= p or (y * y - (x**3 + c["a"] * x + c["b"])) % p != 0:
        return 6
    w = pow(s, -1, n)
    z = int.from_bytes(digest, "big") % n
    P = w_add(c, w_mul(c, c["G"], z * w % n), w_mul(c, (x, y), r * w % n))
    if P is None:
        return 4
    return 0 if P[0] % n == r else 7


# ---- Ed25519, RFC 8032 section 6 style (affine) ----
EP = 2**255 - 19
EL = 2**252 + 27742317777372353535851937790883648493
ED = -121665 * pow(121666, -1, EP) % EP
SQRTM1 = pow(2, (EP - 1) // 4, EP)


def e_add(P, Q):
    x1, y1 = P
    x2, y2 = Q
    t = ED * x1 * x2 * y1 * y2 % EP
    x3 = (x1 * y2 + x2 * y1) * pow(1 + t, -1, EP) % EP
    y3 = (y1 * y2 + x1 * x2) * pow(1 - t, -1, EP) % EP
    return (x3, y3)


def e_mul(P, k):
    R = (0, 1)
    for bit in bin(k)[2:] if k > 0 else "":
        R = e_add(R, R)
        if bit == "1":
            R = e_add(R, P)
    return R


def recover_x(y, sign):
    """x with sign bit `sign` for y (y < p), or None (RFC 8032 5.1.3 steps 2-4, no x=0/sign reject:
    dalek keeps the point and the negation of 0 is 0)."""
    x2 = (y * y - 1) * pow(ED * y * y + 1, -1, EP) % EP
    if x2 == 0:
        return 0
    x = pow(x2, (EP + 3) // 8, EP)
    if (x * x - x2) % EP != 0:
        x = x * SQRTM1 % EP
    if (x * x - x2) % EP != 0:
        return None
    if x & 1 != sign:
        x = EP - x
    return x


def e_decompress(b: bytes):
    y = int.from_bytes(b, "little") & ((1 << 255) - 1)
    sign = b[31] >> 7
    y %= EP  # dalek: non-canonical y is accepted and taken mod p
    x = recover_x(y, sign)
    if x is None:
        return None
    return (x, y)


def e_compress(P):
    x, y = P
    return (y | ((x & 1) << 255)).to_bytes(32, "little")


EB = (recover_x(4 * pow(5, -1, EP) % EP, 0), 4 * pow(5, -1, EP) % EP)


def ed25519_verify(sig: bytes, pk: bytes, msg: bytes) -> bool:
    s = int.from_bytes(sig[32:], "little")
    if s >= EL:
        return False
    A = e_decompress(pk)
    if A is None:
        return False
    k = int.from_bytes(hashlib.sha512(sig[:32] + pk + msg).digest(), "little") % EL
    negA = ((EP - A[0]) % EP, A[1])
    P = e_add(e_mul(EB, s), e_mul(negA, k))
    return e_compress(P) == sig[:32]


def e_small_order(P) -> bool:
    """order divides 8: three doublings reach the identity (0, 1)"""
    for _ in range(3):
        P = e_add(P, P)
    return P == (0, 1)


def ed25519_verify_strict(sig: bytes, pk: bytes, msg: bytes) -> bool:
    """dalek 2.1 VerifyingKey::verify_strict: verify, plus R must decompress and neither R nor A
    may have small order; the final comparison is on the compressed bytes of R"""
    R = e_decompress(sig[:32])
    A = e_decompress(pk)
    if R is None or A is None or e_small_order(R) or e_small_order(A):
        return False
    return ed25519_verify(sig, pk, msg)


def ed25519_torsion():
    """the eight points of order dividing 8, as [l]P for a point P with a torsion component"""
    pts = set()
    y = 2
    while len(pts) < 8:
        x = recover_x(y, 0)
        if x is not None:
            pts.add(e_mul((x, y), EL))
        y += 1
    return sorted(pts)


def ed25519_sign(seed: bytes, msg: bytes) -> bytes:
    h = hashlib.sha512(seed).digest()
    a = int.from_bytes(h[:32], "little")
    a &= (1 << 254) - 8
    a |= 1 << 254
    A = e_compress(e_mul(EB, a))
    r = int.from_bytes(hashlib.sha512(h[32:] + msg).digest(), "little") % EL
    R = e_compress(e_mul(EB, r))
    k = int.from_bytes(hashlib.sha512(R + A + msg).digest(), "little") % EL
    S = (r + k * a) % EL
    return R + S.to_bytes(32, "little")
