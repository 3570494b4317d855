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
ED_LAT_BITS = 136  # Ed25519 lattice-reduced check: bound on its short scalars (34 windows of 4 bits)

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


def limbs30(x):
    assert 0 <= x < 2**270
    return [(x >> (30 * i)) & (2**30 - 1) for i in range(9)]


def mont(x, m):
    return x * 2**RBITS % m


def arr(name, vals, ctype="uint32_t"):
    fmt = {"uint64_t": "0x%016xull", "int32_t": "%d"}.get(ctype, "0x%08xu")
    body = ", ".join(fmt % v for v in vals)
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
    # binary-GCD inversion (inv.hpp): m and the seeds R mod m, R^2 mod m in signed 30-bit limbs
    # (all limbs below 2^30 here), and m^-1 mod 2^30
    out += "  " + arr("M30", limbs30(m), "int32_t")
    out += "  " + arr("RMOD30", limbs30(2**RBITS % m), "int32_t")
    out += "  " + arr("R2MOD30", limbs30(2**(2 * RBITS) % m), "int32_t")
    out += f"  static constexpr uint32_t MINV30 = 0x{pow(m, -1, 2**30):08x}u;\n"
    out += "};\n"
    return out


def w_mul(k, P, p, a):
    R, Q = None, P
    while k:
        if k & 1:
            R = w_add(R, Q, p, a)
        Q = w_add(Q, Q, p, a)
        k >>= 1
    return R


def glv_k1():
    """GLV endomorphism of secp256k1: phi(x, y) = (beta x, y) = [lambda](x, y), with a reduced basis
    (a1, b1), (a2, b2) of the lattice {(u, v): u + v lambda = 0 mod n} from the extended Euclidean
    algorithm on (n, lambda) (Gallant-Lambert-Vanstone 2001; Hankerson-Menezes-Vanstone, Guide to
    ECC, algorithms 3.74 and 3.73), and the fixed-point roundings g1 = round(2^384 b2 / n),
    g2 = round(2^384 |b1| / n) used by the split."""
    def cube_roots(m):
        for g in range(2, 100):
            r = pow(g, (m - 1) // 3, m)
            if r != 1:
                return sorted({r, r * r % m})
    pairs = [(bt, lm) for bt in cube_roots(P_K1) for lm in cube_roots(N_K1)
             if w_mul(lm, G_K1, P_K1, 0) == (bt * G_K1[0] % P_K1, G_K1[1])]
    beta, lam = pairs[0]
    r0, r1, t0, t1 = N_K1, lam, 0, 1
    seq = [(r0, t0), (r1, t1)]
    while r1:
        q = r0 // r1
        r0, r1 = r1, r0 - q * r1
        t0, t1 = t1, t0 - q * t1
        seq.append((r1, t1))
    sq = _iroot(N_K1, 2)
    li = max(i for i, (r, _) in enumerate(seq) if r >= sq)
    a1, b1 = seq[li + 1][0], -seq[li + 1][1]
    a2, b2 = min([(seq[li][0], -seq[li][1]), (seq[li + 2][0], -seq[li + 2][1])], key=lambda v: v[0] ** 2 + v[1] ** 2)
    assert (a1 + b1 * lam) % N_K1 == 0 and (a2 + b2 * lam) % N_K1 == 0
    assert a1 * b2 - a2 * b1 == N_K1 and b1 < 0 < a1 and a2 > 0 and b2 > 0
    # parities used by the device split to make both halves odd
    assert a1 & 1 and b1 & 1 and not a2 & 1 and b2 & 1
    g1 = (2**384 * b2 + N_K1 // 2) // N_K1
    g2 = (2**384 * -b1 + N_K1 // 2) // N_K1
    return beta, lam, a1, b1, a2, b2, g1, g2


def glv_split_ref(k, c):
    """the device split (secp.hpp glv_split) restated on integers"""
    beta, lam, a1, b1, a2, b2, g1, g2 = c
    def signed(x):
        x %= 2**256
        return x - 2**256 if x >> 255 else x
    c1 = (k * g1 + 2**383) >> 384
    c2 = (k * g2 + 2**383) >> 384
    k1 = signed(k - c1 * a1 - c2 * a2)
    k2 = signed(-c1 * b1 - c2 * b2)
    e1, e2 = k1 % 2 == 0, k2 % 2 == 0
    if e1 and e2:
        k1, k2 = k1 + a1, k2 + b1
    elif e2:
        k1, k2 = k1 + a2, k2 + b2
    elif e1:
        k1, k2 = k1 + a1 + a2, k2 + b1 + b2
    return k1, k2


def exp_chain(e, w=4):
    """Addition chain for x -> x^e as steps t[dst] = t[src]^(2^nsq) * t[mul] (mul = -1: none), slot 0
    holding x on entry. The exponent is scanned from the top: a run of >= w ones becomes one factor
    x^(2^L - 1), built by the binary method on L (x^(2^2k - 1) = (x^(2^k - 1))^(2^k) x^(2^k - 1),
    x^(2^(k+1) - 1) = (x^(2^k - 1))^2 x); anything else a sliding window with an odd digit < 2^w,
    taken from the precomputed odd powers x, x^3, ..., x^(2^w - 1). Slots are reused once dead.
    Returns (steps, nslots, result_slot, squarings, multiplications)."""
    bits = bin(e)[2:]
    nb = len(bits)
    terms = []  # (kind, value, position of the lowest bit)
    top_run = len(bits) - len(bits.lstrip("1"))
    i = 0
    while i < nb:
        if bits[i] == "0":
            i += 1
            continue
        j = i
        while j < nb and bits[j] == "1":
            j += 1
        # a run costs no squarings of its own when it is no longer than the top run (its length is
        # on the way to the top run's); a longer run lower down would, so it takes windows instead
        if j - i >= w and j - i <= max(top_run, w):
            terms.append(("run", j - i, nb - j))
            i = j
        else:
            k = min(i + w, nb)
            while bits[k - 1] == "0":
                k -= 1
            terms.append(("odd", int(bits[i:k], 2), nb - k))
            i = k
    # symbolic program over named values; names: ("run", L) or ("odd", d); exponent of each
    prog = []  # (dst_name, src_name, nsq, mul_name)
    have = {("run", 1): 1, ("odd", 1): 1}

    def alias(name):
        return ("x",) if name in (("run", 1), ("odd", 1)) else name

    # one chain of run lengths through every needed length: x^(2^(m+c) - 1) = (x^(2^m - 1))^(2^c)
    # x^(2^c - 1) costs c squarings, so reaching the longest run L costs L - 1 squarings in all
    lengths = [1]
    for L in sorted({v for k, v, _ in terms if k == "run"}):
        while lengths[-1] < L:
            m = lengths[-1]
            c = max(x for x in lengths if m + x <= L)
            prog.append((("run", m + c), ("run", m), c, ("run", c)))
            have[("run", m + c)] = 2**(m + c) - 1
            lengths.append(m + c)

    odds = sorted({v for k, v, _ in terms if k == "odd"})
    if odds and odds[-1] > 1:
        prog.append((("sq",), ("odd", 1), 1, None))
        for d in range(3, odds[-1] + 1, 2):
            prog.append((("odd", d), ("odd", d - 2), 0, ("sq",)))
    acc = ("acc",)
    first = terms[0]
    prev_pos = first[2]
    src0 = ("run", first[1]) if first[0] == "run" else ("odd", first[1])
    prog.append((acc, src0, 0, None))
    for kind, v, pos in terms[1:]:
        prog.append((acc, acc, prev_pos - pos, ("run", v) if kind == "run" else ("odd", v)))
        prev_pos = pos
    if prev_pos:
        prog.append((acc, acc, prev_pos, None))
    # drop pure copies (nsq 0, no multiply) by renaming
    ren = {}
    prog2 = []
    for dst, src, nsq, mul in prog:
        src, mul = ren.get(src, src), ren.get(mul, mul)
        if nsq == 0 and mul is None:
            ren[dst] = src
            continue
        ren.pop(dst, None)
        prog2.append((dst, src, nsq, mul))
    # check the exponent symbolically
    val = {alias(("run", 1)): 1}
    for dst, src, nsq, mul in prog2:
        x = val[alias(src)] << nsq
        if mul is not None:
            x += val[alias(mul)]
        val[alias(dst)] = x
    res = alias(ren.get(acc, acc))
    assert val[res] == e, "chain does not compute the exponent"
    # slot allocation by liveness
    last_use = {}
    for idx, (dst, src, nsq, mul) in enumerate(prog2):
        for nm in (src, mul):
            if nm is not None:
                last_use[alias(nm)] = idx
    last_use[res] = len(prog2)
    slot = {("x",): 0}
    free, nslots, steps = [], 1, []
    for idx, (dst, src, nsq, mul) in enumerate(prog2):
        s_src = slot[alias(src)]
        s_mul = slot[alias(mul)] if mul is not None else -1
        for nm in list(slot):
            if last_use.get(nm, -1) <= idx and nm != alias(dst):
                free.append(slot.pop(nm))
        d = alias(dst)
        if d in slot:
            s_dst = slot[d]
        elif free:
            s_dst = free.pop()
        else:
            s_dst = nslots
            nslots += 1
        slot[d] = s_dst
        steps.append((s_dst, s_src, nsq, s_mul))
    # replay on slots to verify the allocation
    reg = {0: 1}
    for s_dst, s_src, nsq, s_mul in steps:
        x = reg[s_src] << nsq
        if s_mul >= 0:
            x += reg[s_mul]
        reg[s_dst] = x
    assert reg[slot[res]] == e, "slot allocation broke the chain"
    nsqr = sum(st[2] for st in steps)
    nmul = sum(1 for st in steps if st[3] >= 0)
    return steps, nslots, slot[res], nsqr, nmul


def chain_block(tag, e):
    steps, nslots, res, nsqr, nmul = exp_chain(e)
    n = len(steps)
    out = f"// x^e, e = 0x{e:x}: {nsqr} squarings + {nmul} multiplications\n"
    out += f"struct {tag} {{\n  static constexpr int N = {n}, SLOTS = {nslots}, RESULT = {res};\n"
    for nm, col in (("DST", 0), ("SRC", 1), ("NSQ", 2), ("MUL", 3)):
        out += f"  static constexpr int16_t {nm}[{n}] = {{{', '.join(str(st[col]) for st in steps)}}};\n"
    out += "};\n"
    return out


def words8(x):
    return [(x % 2**256) >> (32 * i) & 0xFFFFFFFF for i in range(8)]


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


def sha256_constants():
    """FIPS 180-4: first 32 bits of the fractional parts of the cube roots of the first 64 primes
    (K) and of the square roots of the first 8 primes (IV)."""
    primes = []
    x = 2
    while len(primes) < 64:
        if all(x % q for q in primes if q * q <= x):
            primes.append(x)
        x += 1
    K = [_iroot(q << 96, 3) & (2**32 - 1) for q in primes]
    IV = [_iroot(q << 64, 2) & (2**32 - 1) for q in primes[:8]]
    return K, IV


def sha256_pad_kw(bitlen):
    """K[t] + W[t] (mod 2^32) for the padding-only block of a message whose length is a multiple
    of 64 bytes (0x80, zeros, 64-bit bit length): its message schedule is a constant."""
    K, _ = sha256_constants()
    M32 = 2**32 - 1
    rotr = lambda x, n: ((x >> n) | (x << (32 - n))) & M32
    W = [0x80000000] + [0] * 13 + [bitlen >> 32, bitlen & M32]
    for t in range(16, 64):
        s0 = rotr(W[t - 15], 7) ^ rotr(W[t - 15], 18) ^ (W[t - 15] >> 3)
        s1 = rotr(W[t - 2], 17) ^ rotr(W[t - 2], 19) ^ (W[t - 2] >> 10)
        W.append((W[t - 16] + s0 + W[t - 7] + s1) & M32)
    return [(k + w) & M32 for k, w in zip(K, W)]


def main():
    o = sys.stdout
    o.write("// Generated by gen_consts.py -- do not edit.\n#pragma once\n#include <cstdint>\n\n")
    o.write(f"#define SIGOPS_LIMB_BITS {LIMB_BITS}\n#define SIGOPS_NLIMBS {NLIMBS}\n#define SIGOPS_WG {WG}\n"
            f"#define SIGOPS_ED_LAT_BITS {ED_LAT_BITS}\n\n")
    o.write("namespace sigops {\n")
    for tag, m in [("ModPk1", P_K1), ("ModNk1", N_K1), ("ModPr1", P_R1), ("ModNr1", N_R1), ("ModPed", P_ED),
                   ("ModL", L_ED)]:
        o.write(modulus_block(tag, m))
    o.write("// curve constants (Montgomery form)\n")
    o.write(arr("K1_B", limbs(mont(7, P_K1))))
    o.write(arr("R1_B", limbs(mont(B_R1, P_R1))))
    o.write(arr("ED_D", limbs(mont(D_ED, P_ED))))
    o.write(arr("ED_D2", limbs(mont(2 * D_ED % P_ED, P_ED))))
    o.write(arr("ED_SQRTM1", limbs(mont(SQRTM1, P_ED))))
    # n and l as plain 32-bit words (for range checks and scalar negation)
    o.write(arr("K1_N32", [N_K1 >> (32 * i) & 0xFFFFFFFF for i in range(8)]))
    o.write(arr("R1_N32", [N_R1 >> (32 * i) & 0xFFFFFFFF for i in range(8)]))
    o.write(arr("ED_L32", [L_ED >> (32 * i) & 0xFFFFFFFF for i in range(8)]))
    o.write(arr("ED_P32", [P_ED >> (32 * i) & 0xFFFFFFFF for i in range(8)]))
    # verification with a known key: p (key range), p - n (x(P) in [n, p) compares as x - n) and
    # (n - 1) / 2 (secp256k1 accepts lower-S signatures only)
    w8 = lambda v: [v >> (32 * i) & 0xFFFFFFFF for i in range(8)]
    o.write(arr("K1_P32", w8(P_K1)))
    o.write(arr("R1_P32", w8(P_R1)))
    o.write(arr("K1_PMN32", w8(P_K1 - N_K1)))
    o.write(arr("R1_PMN32", w8(P_R1 - N_R1)))
    o.write(arr("K1_NHALF32", w8((N_K1 - 1) // 2)))
    # 2^512 mod l and 2^256 mod l in Montgomery-of-l form for the wide hash reduction:
    #   h = hi*2^256 + lo ->  mont(hi, 2^(2*261+256-261) ...) : we need x*2^256 mod l = mont_mul(x, 2^(256+261) mod l)
    o.write(arr("L_C256", limbs(2**(256 + RBITS) % L_ED)))
    o.write(arr("L_C0", limbs(2**RBITS % L_ED)))
    # GLV for secp256k1: beta (Montgomery form), split constants as 8 little-endian u32 words
    # (negative ones in two's complement mod 2^256)
    glv = glv_k1()
    beta, lam, a1, b1, a2, b2, g1, g2 = glv
    import random
    rnd = random.Random(1)
    for k in [0, 1, 2, N_K1 - 1, N_K1 - 2, lam, N_K1 - lam] + [rnd.randrange(N_K1) for _ in range(20000)]:
        h1, h2 = glv_split_ref(k, glv)
        assert (h1 + h2 * lam - k) % N_K1 == 0 and h1 & 1 and h2 & 1
        assert abs(h1) < 2**129 and abs(h2) < 2**129
    o.write(arr("K1_BETA", limbs(mont(beta, P_K1))))
    o.write(arr("GLV_G1", words8(g1)))
    o.write(arr("GLV_G2", words8(g2)))
    o.write(arr("GLV_A1", words8(a1)))
    o.write(arr("GLV_A2", words8(a2)))
    o.write(arr("GLV_B1N", words8(-b1)))  # |b1|
    o.write(arr("GLV_B2", words8(b2)))
    # parity corrections (k1, k2) += v1, v2 or v1 + v2
    o.write(arr("GLV_FIX_K1", words8(a1) + words8(a2) + words8(a1 + a2)))
    o.write(arr("GLV_FIX_K2", words8(b1) + words8(b2) + words8(b1 + b2)))
    # exponentiation chains: inverses (Fermat), square roots ((p+1)/4) and dalek's pow_p58
    for tag, e in [("CH_K1_INV", P_K1 - 2), ("CH_K1_SQRT", (P_K1 + 1) // 4), ("CH_R1_INV", P_R1 - 2),
                   ("CH_R1_SQRT", (P_R1 + 1) // 4), ("CH_ED_INV", P_ED - 2), ("CH_ED_P58", (P_ED - 5) // 8),
                   ("CH_NK1_INV", N_K1 - 2), ("CH_NR1_INV", N_R1 - 2), ("CH_L_INV", L_ED - 2)]:
        o.write(chain_block(tag, e))
    K, IV = sha512_constants()
    o.write(arr("SHA512_K", K, "uint64_t"))
    o.write(arr("SHA512_IV", IV, "uint64_t"))
    K2, IV2 = sha256_constants()
    o.write(arr("SHA256_K", K2))
    o.write(arr("SHA256_IV", IV2))
    o.write(arr("SHA256_PAD64_KW", sha256_pad_kw(512)))  # second block of every 64-byte message
    # caller-supplied tables in the reference's layout hold x * 2^(nl*l) mod p in nl = 256 // l + 1
    # limbs of l = 11..15 bits (src/tests/mod.rs:136-149); one Montgomery product with
    # C_l = 2^(-nl*l) * 2^522 mod p turns that value into x * 2^261 mod p (this engine's form)
    for tag, p in [("K1", P_K1), ("R1", P_R1), ("ED", P_ED)]:
        vals = []
        for lb in range(11, 16):
            nl = 256 // lb + 1
            vals += limbs(pow(2, -nl * lb, p) * 2**522 % p)
        o.write(arr(f"REFCONV_{tag}", vals))
    o.write("\n// host-side generator tables: entries of 20 u32 (x[9], y[9], pad[2]) for secp (odd multiples 1..511),\n")
    o.write("// 28 u32 (ypx[9], ymx[9], t2d[9], pad) for ed25519 (odd multiples of B, then of [2^136] B);\n")
    o.write("// Montgomery form\n")
    ntab = 1 << (WG - 1)
    # secp: odd multiples 1..511 of G, for fixed-base windows of 8 bits (the first 128) or 9
    for tag, G, p, a in [("K1", G_K1, P_K1, 0), ("R1", G_R1, P_R1, -3)]:
        pts = odd_multiples_w(G, p, a, 2 * ntab)
        vals = []
        for (x, y) in pts:
            vals += limbs(mont(x, p)) + limbs(mont(y, p)) + [0, 0]
        o.write(arr(f"{tag}_GTABLE", vals))
    # Ed25519: odd multiples of B, then odd multiples of [2^ED_LAT_BITS] B (the high half of the
    # lattice-reduced check's fixed-base scalar)
    B = ed_base()
    B_hi = (0, 1)
    for bit in bin(1 << ED_LAT_BITS)[2:]:
        B_hi = e_add(B_hi, B_hi)
        if bit == "1":
            B_hi = e_add(B_hi, B)
    pts = []
    for base in (B, B_hi):
        base2 = e_add(base, base)
        pts.append(base)
        for _ in range(ntab - 1):
            pts.append(e_add(pts[-1], base2))
    vals = []
    for (x, y) in pts:
        vals += limbs(mont((y + x) % P_ED, P_ED)) + limbs(mont((y - x) % P_ED, P_ED))
        vals += limbs(mont(2 * D_ED * x * y % P_ED, P_ED)) + [0]
    o.write(arr("ED_BTABLE", vals))
    o.write("}  // namespace sigops\n")


if __name__ == "__main__":
    main()
