#!/usr/bin/env python3
"""Generates csrc/consts.hpp: field and curve constants for the device code, and the fixed-base
generator tables that the library uploads to HBM once per context.

Representation (shared with field.hpp): 9 limbs of 29 bits, Montgomery form with R = 2^261 for
every modulus (secp256k1 p and n, P-256 p and n, 2^255-19, l).

Generator tables (the reference's precompute.rs:12-69 builds [1..16]G at w = 4; these are wider):
  - secp256k1 / P-256: odd multiples (2j+1)G, j < 2^(W-1), affine (x, y), Montgomery form.
  - Ed25519: odd multiples (2j+1)B in "affine Niels" form (y+x, y-x, 2d*x*y), Montgomery form.
Run: python3 gen_consts.py > consts.hpp
"""
import sys

LIMB_BITS = 29
NLIMBS = 9
RBITS = LIMB_BITS * NLIMBS  # 261
MASK = (1 << LIMB_BITS) - 1
WG = 8  # fixed-base window (odd multiples 1..2^W-1)

P_K1 = 2**256 - 2**32 - 977
N_K1 = 0xFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFEBAAEDCE6AF48A03BBFD25E8CD0364141
P_R1 = 2**256 - 2**224 + 2**192 + 2**96 - 1
N_R1 = 0xFFFFFFFF00000000FFFFFFFFFFFFFFFFBCE6FAADA7179E84F3B9CAC2FC632551
B_R1 = 0x5AC635D8AA3A93E7B3EBBD55769886BC651D06B0CC53B0F63BCE3C3E27D2604B
P_ED = 2**255 - 19
L_ED = 2**252 + 27742317777372353535851937790883648493
G_K1 = (0x79BE667EF9DCBBAC55A06295CE870B07029BFCDB2DCE28D959F2815B16F81798,
        0x483ADA7726A3C4655DA4FBFC0E1108A8FD17B448A68554199C47D08FFB10D4B8)
G_R1 = (0x6B17D1F2E12C4247F8BCE6E563A440F277037D812DEB33A0F4A13945D898C296,
        0x4FE342E2FE1A7F9B8EE7EB4A7C0F9E162BCE33576B315ECECBB6406837BF51F5)
D_ED = -121665 * pow(121666, -1, P_ED) % P_ED
SQRTM1 = pow(2, (P_ED - 1) // 4, P_ED)


def limbs(x):
    assert 0 <= x < 2**RBITS
    return [(x >> (LIMB_BITS * i)) & MASK for i in range(NLIMBS)]


def mont(x, m):
    return x * 2**RBITS % m


def arr(name, vals, ctype="uint32_t"):
    body = ", ".join(("0x%016xull" if ctype == "uint64_t" else "0x%08xu") % v for v in vals)
    return f"static constexpr {ctype} {name}[{len(vals)}] = {{{body}}};\n"


def w_add(P, Q, p, a):
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


def e_add(P, Q):
    x1, y1 = P
    x2, y2 = Q
    t = D_ED * x1 * x2 * y1 * y2 % P_ED
    return ((x1 * y2 + x2 * y1) * pow(1 + t, -1, P_ED) % P_ED, (y1 * y2 + x1 * x2) * pow(1 - t, -1, P_ED) % P_ED)


def odd_multiples_w(G, p, a, count):
    out = [G]
    G2 = w_add(G, G, p, a)
    for _ in range(count - 1):
        out.append(w_add(out[-1], G2, p, a))
    return out


def ed_base():
    y = 4 * pow(5, -1, P_ED) % P_ED
    x2 = (y * y - 1) * pow(D_ED * y * y + 1, -1, P_ED) % P_ED
    x = pow(x2, (P_ED + 3) // 8, P_ED)
    if (x * x - x2) % P_ED:
        x = x * SQRTM1 % P_ED
    if x & 1:
        x = P_ED - x
    return (x, y)


# Sparse signed forms m = sum coef * 2^(29 * off) of the three field primes: a Montgomery step then
# adds coef * m_i at column i + off for each term (2 to 5 terms instead of 9 limb products).
SPARSE = {
    P_K1: [(0, -977), (1, -8), (8, 2**24)],                                # 2^256 - 2^32 - 977
    # 2^256 - 2^224 + 2^192 + 2^96 - 1, with -2^224 = (2^29 - 2^21) 2^203 - 2^232 so that every term
    # but the column-0 -1 (which needs no instruction) is non-negative
    P_R1: [(0, -1), (3, 2**9), (6, 2**18), (7, 2**29 - 2**21), (8, 2**24 - 1)],
    P_ED: [(0, -19), (8, 2**23)],                                          # 2^255 - 19
}


def sparse_block(m):
    terms = SPARSE.get(m)
    if terms is None:
        return "  static constexpr bool SPARSE = false;\n  static constexpr int SP_N = 1;\n" \
               "  static constexpr int SP_OFF[1] = {0};\n  static constexpr int32_t SP_COEF[1] = {0};\n"
    assert sum(c * 2**(LIMB_BITS * o) for o, c in terms) == m
    assert terms[0][0] == 0 and all(o > 0 for o, _ in terms[1:])
    assert all(abs(c) < 2**31 for _, c in terms)
    # the Montgomery digit only sees the column-0 coefficient: m = coef0 (mod 2^29)
    assert (m - terms[0][1]) % 2**LIMB_BITS == 0
    offs = ", ".join(str(o) for o, _ in terms)
    coefs = ", ".join(str(c) for _, c in terms)
    return (f"  static constexpr bool SPARSE = true;\n  static constexpr int SP_N = {len(terms)};\n"
            f"  static constexpr int SP_OFF[{len(terms)}] = {{{offs}}};\n"
            f"  static constexpr int32_t SP_COEF[{len(terms)}] = {{{coefs}}};\n")


def modulus_block(tag, m):
    out = f"// ---- {tag}: m = 0x{m:064x}\n"
    out += f"struct {tag} {{\n"
    out += sparse_block(m)
    out += "  " + arr("M", limbs(m))
    out += "  " + arr("M2", limbs(2 * m))
    out += f"  static constexpr uint32_t N0 = 0x{(-pow(m, -1, 2**LIMB_BITS)) % 2**LIMB_BITS:08x}u;\n"
    out += "  " + arr("ONE", limbs(mont(1, m)))
    out += "  " + arr("R2", limbs(2**(2 * RBITS) % m))
    out += "  " + arr("EXP_INV", [(m - 2) >> (32 * i) & 0xFFFFFFFF for i in range(8)])
    out += "};\n"
    return out


def _iroot(x, k):
    """floor(x ** (1/k)) for integers"""
    lo, hi = 0, 1 << (x.bit_length() // k + 1)
    while hi - lo > 1:
        mid = (lo + hi) // 2
        if mid**k <= x:
            lo = mid
        else:
            hi = mid
    return lo


def sha512_constants():
    """FIPS 180-4: first 64 bits of the fractional parts of the cube roots of the first 80 primes
    (K) and of the square roots of the first 8 primes (IV)."""
    primes = []
    x = 2
    while len(primes) < 80:
        if all(x % q for q in primes if q * q <= x):
            primes.append(x)
        x += 1
    K = [_iroot(q << 192, 3) & (2**64 - 1) for q in primes]
    IV = [_iroot(q << 128, 2) & (2**64 - 1) for q in primes[:8]]
    return K, IV


def main():
    o = sys.stdout
    o.write("// Generated by gen_consts.py -- do not edit.\n#pragma once\n#include <cstdint>\n\n")
    o.write(f"#define SIGOPS_LIMB_BITS {LIMB_BITS}\n#define SIGOPS_NLIMBS {NLIMBS}\n#define SIGOPS_WG {WG}\n\n")
    o.write("namespace sigops {\n")
    for tag, m in [("ModPk1", P_K1), ("ModNk1", N_K1), ("ModPr1", P_R1), ("ModNr1", N_R1), ("ModPed", P_ED),
                   ("ModL", L_ED)]:
        o.write(modulus_block(tag, m))
    o.write("// curve constants (Montgomery form)\n")
    o.write(arr("K1_B", limbs(mont(7, P_K1))))
    o.write(arr("K1_SQRT_EXP", [((P_K1 + 1) // 4) >> (32 * i) & 0xFFFFFFFF for i in range(8)]))
    o.write(arr("R1_B", limbs(mont(B_R1, P_R1))))
    o.write(arr("R1_SQRT_EXP", [((P_R1 + 1) // 4) >> (32 * i) & 0xFFFFFFFF for i in range(8)]))
    o.write(arr("ED_D", limbs(mont(D_ED, P_ED))))
    o.write(arr("ED_D2", limbs(mont(2 * D_ED % P_ED, P_ED))))
    o.write(arr("ED_SQRTM1", limbs(mont(SQRTM1, P_ED))))
    o.write(arr("ED_P58_EXP", [((P_ED - 5) // 8) >> (32 * i) & 0xFFFFFFFF for i in range(8)]))
    # n and l as plain 32-bit words (for range checks and scalar negation)
    o.write(arr("K1_N32", [N_K1 >> (32 * i) & 0xFFFFFFFF for i in range(8)]))
    o.write(arr("R1_N32", [N_R1 >> (32 * i) & 0xFFFFFFFF for i in range(8)]))
    o.write(arr("ED_L32", [L_ED >> (32 * i) & 0xFFFFFFFF for i in range(8)]))
    o.write(arr("ED_P32", [P_ED >> (32 * i) & 0xFFFFFFFF for i in range(8)]))
    # 2^512 mod l and 2^256 mod l in Montgomery-of-l form for the wide hash reduction:
    #   h = hi*2^256 + lo ->  mont(hi, 2^(2*261+256-261) ...) : we need x*2^256 mod l = mont_mul(x, 2^(256+261) mod l)
    o.write(arr("L_C256", limbs(2**(256 + RBITS) % L_ED)))
    o.write(arr("L_C0", limbs(2**RBITS % L_ED)))
    K, IV = sha512_constants()
    o.write(arr("SHA512_K", K, "uint64_t"))
    o.write(arr("SHA512_IV", IV, "uint64_t"))
    o.write("\n// host-side generator tables: entries of 20 u32 (x[9], y[9], pad[2]) for secp,\n")
    o.write("// 28 u32 (ypx[9], ymx[9], t2d[9], pad) for ed25519; Montgomery form\n")
    ntab = 1 << (WG - 1)
    for tag, G, p, a in [("K1", G_K1, P_K1, 0), ("R1", G_R1, P_R1, -3)]:
        pts = odd_multiples_w(G, p, a, ntab)
        vals = []
        for (x, y) in pts:
            vals += limbs(mont(x, p)) + limbs(mont(y, p)) + [0, 0]
        o.write(arr(f"{tag}_GTABLE", vals))
    B = ed_base()
    B2 = e_add(B, B)
    pts = [B]
    for _ in range(ntab - 1):
        pts.append(e_add(pts[-1], B2))
    vals = []
    for (x, y) in pts:
        vals += limbs(mont((y + x) % P_ED, P_ED)) + limbs(mont((y - x) % P_ED, P_ED))
        vals += limbs(mont(2 * D_ED * x * y % P_ED, P_ED)) + [0]
    o.write(arr("ED_BTABLE", vals))
    o.write("}  // namespace sigops\n")


if __name__ == "__main__":
    main()
