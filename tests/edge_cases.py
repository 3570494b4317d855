"""Hand-built edge cases for the signature path, with expected results from the independent
pure-Python closed forms in pyref.py (not from the oracle or the GPU).

secp: r / s range violations, r that is not an x-coordinate, digests >= n and == 0, recovered key
at infinity, and signatures crafted so that R is a small multiple of G and (u1, u2) are chosen
values, which makes the windowed GPU ladders hit doubling / inverse-point additions.
ed25519: s >= l, small-order and mixed-order public keys, non-canonical key encodings, the
x = 0 / sign-bit key, undecodable keys, SHA-512 padding-boundary message lengths, encodings of R
(non-canonical y, sign bit on x = 0, order-4 points), and challenges that the lattice-reduced
check hands to its full-length check (tests/golden/ed25519_long_lattice.json).
"""
import hashlib
import json
import os
import random

import pyref

ZERO64 = bytes(64)


def _enc_sig(curve, r, s, odd):
    """fuel format; if s needs bit 255, use (n - s) with the other parity (same recovered key)."""
    n = pyref.CURVES[curve]["n"]
    if s >= 2**255:
        s = n - s
        odd ^= 1
    sig = bytearray(r.to_bytes(32, "big") + s.to_bytes(32, "big"))
    sig[32] |= odd << 7
    return bytes(sig)


def secp_cases(curve, seed=7):
    """list of (name, sig64, digest32); expected via pyref.secp_recover"""
    c = pyref.CURVES[curve]
    p, n, G = c["p"], c["n"], c["G"]
    rnd = random.Random(seed + 1000 * curve)
    out = []
    # a valid signature to mutate
    sk = rnd.randrange(1, n)
    d = rnd.randrange(0, 2**256).to_bytes(32, "big")
    k = rnd.randrange(1, n)
    R = pyref.w_mul(c, G, k)
    r = R[0] % n
    s = pow(k, -1, n) * (int.from_bytes(d, "big") % n + r * sk) % n
    odd = R[1] & 1
    base = _enc_sig(curve, r, s, odd)
    out.append(("valid", base, d))
    for name, rr in [("r_zero", 0), ("r_eq_n", n), ("r_n_plus_1", n + 1), ("r_max", 2**256 - 1)]:
        out.append((name, rr.to_bytes(32, "big") + base[32:], d))
    out.append(("s_zero", base[:32] + bytes(32), d))
    out.append(("s_zero_recid", base[:32] + b"\x80" + bytes(31), d))
    # r that is not an x-coordinate
    x = 1
    while True:
        rhs = (x**3 + c["a"] * x + c["b"]) % p
        if pow(rhs, (p - 1) // 2, p) != 1:
            break
        x += 1
    out.append(("r_not_on_curve", x.to_bytes(32, "big") + base[32:], d))
    # digest special values
    for name, z in [("z_zero", 0), ("z_eq_n", n), ("z_n_plus_5", n + 5), ("z_max", 2**256 - 1), ("z_one", 1)]:
        out.append((name, base, z.to_bytes(32, "big")))
    # parity flipped (other root)
    flip = bytearray(base)
    flip[32] ^= 0x80
    out.append(("parity_flip", bytes(flip), d))

    # crafted (R = m*G, u1, u2): s = u2*r, z = -u1*r (mod n)
    def crafted(name, m, u1, u2):
        Rm = pyref.w_mul(c, G, m % n) if m % n else None
        if Rm is None or Rm[0] >= n:
            return
        rr = Rm[0]
        ss = u2 * rr % n
        zz = (-u1 * rr) % n
        if ss == 0:
            return
        out.append((name, _enc_sig(curve, rr, ss, Rm[1] & 1), zz.to_bytes(32, "big")))

    for m in [1, n - 1, 2, 3, 15, 16, 17, 31, 255, 256, 2**128 + 1]:
        for (u1, u2) in [(1, 1), (1, n - 1), (n - 1, 1), (2, 3), (0, 1), (5, n - 5), (3, 5),
                         (2**128, 2**128), (n - 1, n - 1), (2**252, 2**4), (17, 1), (1, 17)]:
            crafted(f"crafted_m{m % 1000}_u{u1 % 1000}_{u2 % 1000}", m, u1, u2)
            # make u1*G + u2*R collide: u1 = -u2*m (Q = infinity)
        crafted(f"crafted_inf_m{m % 1000}", m, (-7 * m) % n, 7)
    # R = G exactly, u2*R = -(u1)G at the last step
    crafted("crafted_inf_G", 1, n - 9, 9)
    if curve == 0:
        # the reference's Strauss-Shamir vector (x G + y B, tests/golden/secp256k1_strauss.json)
        # as a recovery: R = B (B.x < n), u1 = x, u2 = y
        v = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                        "secp256k1_strauss.json")))
        bx, by = int(v["bx_decimal"]), int(v["by_decimal"])
        u1, u2 = int(v["x"], 16) % n, int(v["y"], 16) % n
        assert bx < n and (bx**3 + 7 - by * by) % p == 0
        out.append(("ref_strauss_vector", _enc_sig(curve, bx, u2 * bx % n, by & 1),
                    ((-u1 * bx) % n).to_bytes(32, "big")))
    return out


def secp_expected(curve, cases):
    return [pyref.secp_recover(curve, sig, dig) for (_, sig, dig) in cases]


def secp_verify_cases(curve, seed=8):
    """list of (name, sig64, digest32, pub64) for verification with a known key; expected via
    pyref.secp_verify. The key is solved from a chosen signature, Q = r^-1 (s P - z G), which gives
    full control over r, s, z and the point P the verifier reaches."""
    c = pyref.CURVES[curve]
    p, n, G = c["p"], c["n"], c["G"]
    rnd = random.Random(seed + 1000 * curve)
    half = (n - 1) // 2
    out = []

    def enc_pub(Q):
        return Q[0].to_bytes(32, "big") + Q[1].to_bytes(32, "big")

    def neg(P):
        return (P[0], (p - P[1]) % p)

    def solve_key(P, r, s, z):
        """Q with (z/s) G + (r/s) Q = P"""
        T = pyref.w_add(c, pyref.w_mul(c, P, s), neg(pyref.w_mul(c, G, z % n)) if z % n else None)
        return pyref.w_mul(c, T, pow(r, -1, n)) if T is not None else None

    def sig(r, s, odd=0):
        b = bytearray(r.to_bytes(32, "big") + s.to_bytes(32, "big"))
        b[32] |= odd << 7
        return bytes(b)

    # an honest low-S signature
    sk = rnd.randrange(1, n)
    Q = pyref.w_mul(c, G, sk)
    z = rnd.randrange(0, 2**256)
    k = rnd.randrange(1, n)
    R = pyref.w_mul(c, G, k)
    r = R[0] % n
    s = pow(k, -1, n) * (z % n + r * sk) % n
    if s > half:
        s = n - s
    d = z.to_bytes(32, "big")
    pub = enc_pub(Q)
    out.append(("valid", sig(r, s), d, pub))
    out.append(("valid_recid_bit_ignored", sig(r, s, 1), d, pub))
    out.append(("wrong_digest", sig(r, s), ((z + 1) % 2**256).to_bytes(32, "big"), pub))
    out.append(("wrong_key", sig(r, s), d, enc_pub(pyref.w_mul(c, G, sk + 1))))
    out.append(("key_is_G", sig(r, s), d, enc_pub(G)))
    for name, rr in [("r_zero", 0), ("r_eq_n", n), ("r_max", 2**256 - 1)]:
        out.append((name, sig(rr, s), d, pub))
    # fuel's format keeps s below 2^255 (bit 255 is the recid) and n > 2^255, so s >= n cannot occur
    out.append(("s_zero", sig(r, 0), d, pub))
    # digest >= n is reduced
    for name, zz in [("z_plus_n", z % n + n), ("z_max", 2**256 - 1), ("z_zero", 0)]:
        if zz < 2**256:
            P = pyref.w_mul(c, G, rnd.randrange(1, n))
            rr = P[0] % n
            ss = rnd.randrange(1, half)
            Qz = solve_key(P, rr, ss, zz)
            if Qz is not None:
                out.append((name, sig(rr, ss), zz.to_bytes(32, "big"), enc_pub(Qz)))
    # low-S boundary, with otherwise valid equations: s = (n-1)/2, (n+1)/2 and the largest s the
    # format holds (secp256k1 rejects the last two, P-256 accepts them)
    for name, ss in [("s_half", half), ("s_half_plus_1", half + 1), ("s_max", 2**255 - 1)]:
        P = pyref.w_mul(c, G, rnd.randrange(1, n))
        rr = P[0] % n
        zz = rnd.randrange(0, n)
        out.append((name, sig(rr, ss), zz.to_bytes(32, "big"), enc_pub(solve_key(P, rr, ss, zz))))
    # bad keys
    bad_y = bytearray(pub)
    bad_y[63] ^= 1
    out.append(("key_not_on_curve", sig(r, s), d, bytes(bad_y)))
    out.append(("key_zero", sig(r, s), d, bytes(64)))
    out.append(("key_x_ge_p", sig(r, s), d, (Q[0] + p).to_bytes(32, "big") + pub[32:] if Q[0] + p < 2**256
                else b"\xff" * 32 + pub[32:]))
    out.append(("key_y_plus_p", sig(r, s), d, pub[:32] + (Q[1] + p).to_bytes(32, "big") if Q[1] + p < 2**256
                else pub[:32] + b"\xff" * 32))
    # the verifier's point at infinity: Q = -(z/r) G
    rr, ss, zz = rnd.randrange(1, n), rnd.randrange(1, half), rnd.randrange(1, n)
    Qi = pyref.w_mul(c, G, (-zz * pow(rr, -1, n)) % n)
    out.append(("infinity", sig(rr, ss), zz.to_bytes(32, "big"), enc_pub(Qi)))
    # x(P) in [n, p): accepted through x mod n == r (the r + n comparison)
    x = n + 1
    while x < p:
        rhs = (x**3 + c["a"] * x + c["b"]) % p
        y = pow(rhs, (p + 1) // 4, p)
        if y * y % p == rhs:
            break
        x += 1
    if x < p:
        P = (x, y)
        rr = x - n
        ss, zz = rnd.randrange(1, half), rnd.randrange(0, n)
        out.append(("x_ge_n", sig(rr, ss), zz.to_bytes(32, "big"), enc_pub(solve_key(P, rr, ss, zz))))
        out.append(("x_ge_n_r_is_x", sig(x, ss), zz.to_bytes(32, "big"), enc_pub(solve_key(P, rr, ss, zz))))
    # exceptional additions in windowed ladders: Q = m G with chosen (u1, u2); P = (u1 + u2 m) G
    for m in [1, n - 1, 2, 3, 15, 16, 17, 255, 2**128 + 1]:
        for (u1, u2) in [(1, 1), (1, n - 1), (n - 1, 1), (2, 3), (0, 1), (5, n - 5), (2**128, 2**128),
                         (17, 1), (1, 17), ((-7 * m) % n, 7)]:
            Pm = pyref.w_mul(c, G, (u1 + u2 * m) % n) if (u1 + u2 * m) % n else None
            if Pm is None:
                continue
            rr = Pm[0] % n
            if rr == 0:
                continue
            ss = rr * pow(u2, -1, n) % n
            zz = u1 * ss % n
            if ss > half:  # (r, n - s) negates u1, u2 and P (same x) and fits below 2^255
                ss = n - ss
            out.append((f"crafted_m{m % 1000}_u{u1 % 1000}_{u2 % 1000}", sig(rr, ss), zz.to_bytes(32, "big"),
                        enc_pub(pyref.w_mul(c, G, m))))
    return out


def secp_verify_expected(curve, cases):
    return [pyref.secp_verify(curve, sg, dg, pb) for (_, sg, dg, pb) in cases]


# ---------------- ed25519 ----------------
def _torsion8():
    """a point of exact order 8 (in the torsion subgroup), found as l * P for a curve point P"""
    y = 2
    while True:
        x = pyref.recover_x(y, 0)
        if x is not None:
            T = pyref.e_mul((x, y), pyref.EL)
            if pyref.e_mul(T, 4) != (0, 1):
                return T
        y += 1


def ed_cases(seed=11):
    """list of (name, sig64, pk32, msg)"""
    rnd = random.Random(seed)
    out = []
    EL, EP = pyref.EL, pyref.EP
    seedk = bytes(rnd.randrange(256) for _ in range(32))
    msg = bytes(rnd.randrange(256) for _ in range(40))
    sig = pyref.ed25519_sign(seedk, msg)
    h = hashlib.sha512(seedk).digest()
    a = (int.from_bytes(h[:32], "little") & ((1 << 254) - 8)) | (1 << 254)
    pk = pyref.e_compress(pyref.e_mul(pyref.EB, a))
    out.append(("valid", sig, pk, msg))
    s = int.from_bytes(sig[32:], "little")
    out.append(("s_plus_l", sig[:32] + (s + EL).to_bytes(32, "little"), pk, msg))
    out.append(("s_eq_l", sig[:32] + EL.to_bytes(32, "little"), pk, msg))
    out.append(("s_high_bit", sig[:32] + (s | (1 << 255)).to_bytes(32, "little"), pk, msg))
    out.append(("s_max", sig[:32] + b"\xff" * 32, pk, msg))
    out.append(("s_l_minus_1", sig[:32] + (EL - 1).to_bytes(32, "little"), pk, msg))
    out.append(("s_zero", sig[:32] + bytes(32), pk, msg))
    # key bit flipped -> usually undecodable or a different key
    for bit in (0, 100, 254, 255):
        pk2 = bytearray(pk)
        pk2[bit // 8] ^= 1 << (bit % 8)
        out.append((f"pk_flip{bit}", sig, bytes(pk2), msg))
    # R sign bit flipped
    r2 = bytearray(sig[:32])
    r2[31] ^= 0x80
    out.append(("R_sign_flip", bytes(r2) + sig[32:], pk, msg))

    def forge(name, A_bytes, m):
        """R = [t]B, s = t for a random t: the cofactorless equation [s]B - [k]A == R then holds
        exactly when [k]A is the identity (A small-order with k a multiple of its order), so these
        cases accept or reject depending on k = H(R || A || M); pyref decides which."""
        t = rnd.randrange(1, EL)
        R = pyref.e_compress(pyref.e_mul(pyref.EB, t))
        out.append((name, R + t.to_bytes(32, "little"), A_bytes, m))

    ident = pyref.e_compress((0, 1))
    forge("A_identity", ident, msg)
    ident_nc = (1 + EP).to_bytes(32, "little")  # non-canonical encoding of y = 1
    forge("A_identity_noncanonical", ident_nc, msg)
    ident_sign = bytearray(ident)
    ident_sign[31] |= 0x80  # x = 0 with the sign bit set: dalek accepts
    forge("A_identity_signbit", bytes(ident_sign), msg)
    order2 = pyref.e_compress((0, EP - 1))
    for i in range(4):
        forge(f"A_order2_{i}", order2, msg + bytes([i]))
    T8 = _torsion8()
    for j in range(1, 8):
        Tj = pyref.e_mul(T8, j)
        for i in range(2):
            forge(f"A_torsion{j}_{i}", pyref.e_compress(Tj), msg + bytes([j, i]))
    # mixed-order key A = aB + T: honest signatures verify only when [k]T is the identity
    for j in (1, 2, 4):
        Aj = pyref.e_add(pyref.e_mul(pyref.EB, a), pyref.e_mul(T8, j))
        Ab = pyref.e_compress(Aj)
        for i in range(4):
            m = msg + bytes([0xA0 + j, i])
            r = int.from_bytes(hashlib.sha512(h[32:] + m).digest(), "little") % EL
            R = pyref.e_compress(pyref.e_mul(pyref.EB, r))
            k = int.from_bytes(hashlib.sha512(R + Ab + m).digest(), "little") % EL
            S = (r + k * a) % EL
            out.append((f"A_mixed{j}_{i}", R + S.to_bytes(32, "little"), Ab, m))
    # non-canonical y for small-y points that decompress
    for y in range(0, 19):
        for sgn in (0, 1):
            b = bytearray((y + EP).to_bytes(32, "little"))
            b[31] |= sgn << 7
            if pyref.e_decompress(bytes(b)) is not None:
                forge(f"A_noncanon_y{y}_s{sgn}", bytes(b), msg)
    # undecodable keys: y with no x
    y = 2
    bad = []
    while len(bad) < 3:
        if pyref.recover_x(y, 0) is None:
            bad.append(y)
        y += 1
    for y in bad:
        out.append((f"A_undecodable_y{y}", sig, y.to_bytes(32, "little"), msg))
    out.append(("A_all_ff", sig, b"\xff" * 32, msg))
    # SHA-512 padding boundaries: preimage = 64 + L bytes
    for L in (0, 1, 47, 48, 63, 64, 111, 112, 127, 128, 175, 176, 239, 240, 1000, 1024, 2047):
        m = bytes(rnd.randrange(256) for _ in range(L))
        out.append((f"len{L}", pyref.ed25519_sign(seedk, m), pk, m))
    # encodings of R: with A the identity and s = 0, [s]B - [k]A is the identity for every k, so only
    # the bytes of R decide: compress() yields y = 1 with the sign bit clear and nothing else
    out.append(("R_identity", ident + bytes(32), ident, msg))
    out.append(("R_identity_noncanonical", ident_nc + bytes(32), ident, msg))
    out.append(("R_identity_signbit", bytes(ident_sign) + bytes(32), ident, msg))
    # A of order 4 (y = 0) and s = 0: [s]B - [k]A is A, the identity or -A by k mod 4; R spelled with
    # y = 0 or the non-canonical y = p, and either sign bit
    T4 = pyref.e_mul(T8, 2)
    A4 = pyref.e_compress(T4)
    for i in range(4):
        for yb, name in ((0, "y0"), (EP, "yp")):
            for sgn in (0, 1):
                Rb = bytearray(yb.to_bytes(32, "little"))
                Rb[31] |= sgn << 7
                out.append((f"R_order4_{name}_s{sgn}_{i}", bytes(Rb) + bytes(32), A4, msg + bytes([0xB0 + i])))
    # challenges without a short lattice pair: the Ed25519-only kernels' lattice-reduced check hands
    # them to its full-length check (tests/golden/make_ed25519_long_lattice.py found them)
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ed25519_long_lattice.json")) as f:
        for j, it in enumerate(json.load(f)["items"]):
            out.append((f"long_lattice_{j}_{it['name']}", bytes.fromhex(it["sig"]), bytes.fromhex(it["pk"]),
                        bytes.fromhex(it["msg"])))
    return out


def ed_expected(cases):
    return [pyref.ed25519_verify(sig, pk, m) for (_, sig, pk, m) in cases]


def ed_strict_cases(seed=12):
    """ed_cases plus cases aimed at verify_strict: R of small order (with a small-order key and
    s = 0, so that the cofactorless equation holds and plain verify accepts), and R encodings that
    do not decompress or are non-canonical"""
    rnd = random.Random(seed)
    out = list(ed_cases())
    EL, EP = pyref.EL, pyref.EP
    tors = pyref.ed25519_torsion()
    T8 = _torsion8()
    A = pyref.e_compress(T8)
    hits = 0
    for i in range(400):
        Rj = tors[i % 8]
        Rb = pyref.e_compress(Rj)
        m = bytes([i & 255, i >> 8]) + b"strict"
        k = int.from_bytes(hashlib.sha512(Rb + A + m).digest(), "little") % EL
        negkA = pyref.e_mul(((EP - T8[0]) % EP, T8[1]), k)
        if pyref.e_compress(negkA) == Rb:  # [0]B - [k]A == R: verify accepts, verify_strict must not
            out.append((f"R_small_order_{i}", Rb + bytes(32), A, m))
            hits += 1
            if hits == 6:
                break
    seedk = bytes(rnd.randrange(256) for _ in range(32))
    msg = b"verify_strict"
    sig = pyref.ed25519_sign(seedk, msg)
    h = hashlib.sha512(seedk).digest()
    a = (int.from_bytes(h[:32], "little") & ((1 << 254) - 8)) | (1 << 254)
    pk = pyref.e_compress(pyref.e_mul(pyref.EB, a))
    out.append(("strict_valid", sig, pk, msg))
    y = 2
    while pyref.recover_x(y, 0) is not None:
        y += 1
    out.append(("R_undecodable", y.to_bytes(32, "little") + sig[32:], pk, msg))
    for yy in range(0, 19):
        b = (yy + EP).to_bytes(32, "little")
        if pyref.e_decompress(b) is not None:
            out.append((f"R_noncanon_y{yy}", b + sig[32:], pk, msg))
    return out


def ed_strict_expected(cases):
    return [pyref.ed25519_verify_strict(sig, pk, m) for (_, sig, pk, m) in cases]
