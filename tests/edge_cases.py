"""Hand-built edge cases for the signature path, with expected results from the independent
pure-Python closed forms in pyref.py (not from the oracle or the GPU).

secp: r / s range violations, r that is not an x-coordinate, digests >= n and == 0, recovered key
at infinity, and signatures crafted so that R is a small multiple of G and (u1, u2) are chosen
values, which makes the windowed GPU ladders hit doubling / inverse-point additions.
ed25519: s >= l, small-order and mixed-order public keys, non-canonical key encodings, the
x = 0 / sign-bit key, undecodable keys, and SHA-512 padding-boundary message lengths.
"""
import hashlib
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
    return out


def secp_expected(curve, cases):
    return [pyref.secp_recover(curve, sig, dig) for (_, sig, dig) in cases]


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
    return out


def ed_expected(cases):
    return [pyref.ed25519_verify(sig, pk, m) for (_, sig, pk, m) in cases]
