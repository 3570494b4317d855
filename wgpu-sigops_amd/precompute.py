"""Generator tables in the reference's layout (src/precompute.rs:12-69).

The reference uploads these tables with every call (``table_limbs`` argument of ``ecrecover`` /
``ecverify``) and uses them as the fixed base. The mirrored functions pass a caller's table to the
library, which converts it on the device into its own wider odd-multiple table (the resident
tables of G and B are what the reference's own tables convert to) and refuses a table that is not
[P, 2P, ..., 16P] for a curve point P. ``secp_table_for_point`` / ``ed25519_table_for_point``
build the layout for any point. Layout: [G, 2G, ..., 16G] (precompute_table,
src/curve_algos/precompute.rs:3-17), affine, each coordinate times R = 2^(num_limbs*log_limb_size)
mod p, split into little-endian limbs of log_limb_size bits (src/tests/mod.rs:94-149); secp tables
hold (x, y), the Ed25519 table (x, y, x*y).
"""
WINDOW_SIZE = 4

_P_K1 = 2**256 - 2**32 - 977
_P_R1 = 2**256 - 2**224 + 2**192 + 2**96 - 1
_P_ED = 2**255 - 19
_G_K1 = (0x79BE667EF9DCBBAC55A06295CE870B07029BFCDB2DCE28D959F2815B16F81798,
         0x483ADA7726A3C4655DA4FBFC0E1108A8FD17B448A68554199C47D08FFB10D4B8)
_G_R1 = (0x6B17D1F2E12C4247F8BCE6E563A440F277037D812DEB33A0F4A13945D898C296,
         0x4FE342E2FE1A7F9B8EE7EB4A7C0F9E162BCE33576B315ECECBB6406837BF51F5)
_D_ED = -121665 * pow(121666, -1, _P_ED) % _P_ED


def calc_num_limbs(log_limb_size, bits=256):
    """smallest limb count whose capacity exceeds `bits` (multiprecision::utils::calc_num_limbs)"""
    return bits // log_limb_size + 1


def check_log_limb_size(log_limb_size):
    # the reference's Montgomery multiplication supports 11..15-bit limbs (src/wgsl/mont.wgsl:12,37)
    if not 11 <= int(log_limb_size) <= 15:
        raise ValueError(f"log_limb_size must be in 11..15, got {log_limb_size}")


def _limbs(x, n, l):
    return [(x >> (l * i)) & ((1 << l) - 1) for i in range(n)]


def _w_add(P, Q, p, a):
    if P[0] == Q[0]:
        lam = (3 * P[0] * P[0] + a) * pow(2 * P[1], -1, p) % p
    else:
        lam = (Q[1] - P[1]) * pow(Q[0] - P[0], -1, p) % p
    x = (lam * lam - P[0] - Q[0]) % p
    return (x, (lam * (P[0] - x) - P[1]) % p)


def _secp_table(G, p, a, log_limb_size):
    check_log_limb_size(log_limb_size)
    n = calc_num_limbs(log_limb_size)
    r = 2 ** (n * log_limb_size)
    out, cur = [], G
    for i in range(1 << WINDOW_SIZE):
        if i:
            cur = _w_add(cur, G, p, a)
        out += _limbs(cur[0] * r % p, n, log_limb_size) + _limbs(cur[1] * r % p, n, log_limb_size)
    return out


def secp256k1_bases(log_limb_size):
    return _secp_table(_G_K1, _P_K1, 0, log_limb_size)


def secp256r1_bases(log_limb_size):
    return _secp_table(_G_R1, _P_R1, -3, log_limb_size)


def _ed_base():
    y = 4 * pow(5, -1, _P_ED) % _P_ED
    x2 = (y * y - 1) * pow(_D_ED * y * y + 1, -1, _P_ED) % _P_ED
    x = pow(x2, (_P_ED + 3) // 8, _P_ED)
    if (x * x - x2) % _P_ED:
        x = x * pow(2, (_P_ED - 1) // 4, _P_ED) % _P_ED
    return (_P_ED - x if x & 1 else x, y)


def secp_table_for_point(curve, P, log_limb_size):
    """the reference's table layout for the multiples of an arbitrary curve point P = (x, y)"""
    return _secp_table(P, (_P_K1, _P_R1)[curve], (0, -3)[curve], log_limb_size)


def ed25519_bases(log_limb_size):
    return ed25519_table_for_point(_ed_base(), log_limb_size)


def ed25519_table_for_point(B, log_limb_size):
    """the reference's Ed25519 table layout (x, y, x*y per entry) for the multiples of B = (x, y)"""
    check_log_limb_size(log_limb_size)
    n = calc_num_limbs(log_limb_size)
    r = 2 ** (n * log_limb_size)
    p = _P_ED
    out, cur = [], B
    for i in range(1 << WINDOW_SIZE):
        if i:
            x1, y1 = cur
            x2, y2 = B
            t = _D_ED * x1 * x2 * y1 * y2 % p
            cur = ((x1 * y2 + x2 * y1) * pow(1 + t, -1, p) % p, (y1 * y2 + x1 * x2) * pow(1 - t, -1, p) % p)
        x, y = cur
        out += _limbs(x * r % p, n, log_limb_size) + _limbs(y * r % p, n, log_limb_size)
        out += _limbs(x * y % p * r % p, n, log_limb_size)
    return out


def check_table(table_limbs, coords, log_limb_size):
    check_log_limb_size(log_limb_size)
    want = (1 << WINDOW_SIZE) * coords * calc_num_limbs(log_limb_size)
    if len(table_limbs) != want:
        raise ValueError(f"table_limbs has {len(table_limbs)} limbs, expected {want}")
