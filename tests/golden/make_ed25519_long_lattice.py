"""Writes ed25519_long_lattice.json: Ed25519 items whose challenge k = SHA-512(R || A || M) mod l has
no short lattice pair, so the engine's lattice-reduced check gives way to its full-length check.

The reduction looks for c0 = c1 k (mod 8l) with c1 odd and |c0|, c1 < 2^136 along Euclid's remainder
sequence on (8l, k) (quotients below 2^32); about 3 in 10^6 random k have none. The messages are
found by search: a fixed key and a fixed nonce R, message counter i, until the model below rejects k.
Expected results come from tests/pyref.py (dalek `verify` rules). Run: python tests/golden/make_ed25519_long_lattice.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import pyref  # noqa: E402

EL, EB = pyref.EL, pyref.EB
N8 = 8 * EL


def short_pair(k):
    """the engine's reduction on exact integers: (c0, c1), or None when it finds no pair"""
    a, ta, b, tb = N8, 0, k, 1
    while b >> 128:
        q = a // b
        if q >> 32:
            return None
        a, ta, b, tb = b, tb, a - q * b, ta - q * tb
    if tb & 1:
        c0, c1 = b, tb
    else:
        m = max(0, (a - abs(ta)) // (b + abs(tb)))
        c0, c1 = a - m * b, ta - m * tb
    if c1 < 0:
        c0, c1 = -c0, -c1
    return (c0, c1) if c1 >> 136 == 0 and abs(c0) >> 136 == 0 else None


def search(A, Ab, a, r, tag, count):
    """messages tag || i whose challenge has no short pair, with honest signatures (s = r + k a)"""
    R = pyref.e_compress(pyref.e_mul(EB, r))
    out = []
    i = 0
    while len(out) < count:
        m = tag + i.to_bytes(8, "little")
        k = int.from_bytes(hashlib.sha512(R + Ab + m).digest(), "little") % EL
        if short_pair(k) is None:
            s = (r + k * a) % EL
            out.append((R + s.to_bytes(32, "little"), m, k))
        i += 1
    return out


def main():
    items = []
    a = int.from_bytes(hashlib.sha256(b"long lattice key").digest(), "little") % EL
    A = pyref.e_mul(EB, a)
    Ab = pyref.e_compress(A)
    r = int.from_bytes(hashlib.sha256(b"long lattice nonce").digest(), "little") % EL
    for sig, m, k in search(A, Ab, a, r, b"long-lattice ", 3):
        items.append(("valid", sig, Ab, m))
        bad = bytearray(sig)
        bad[40] ^= 1  # another s: the same k, a wrong equation
        items.append(("s_changed", bytes(bad), Ab, m))
    # a key with a component of order 8: the honest signature verifies only when [k]T is the identity
    T8 = next(P for P in pyref.ed25519_torsion() if pyref.e_mul(P, 4) != (0, 1))  # order 8
    Am = pyref.e_add(A, T8)
    Amb = pyref.e_compress(Am)
    for sig, m, k in search(Am, Amb, a, r, b"long-lattice torsion ", 2):
        items.append(("torsion_key", sig, Amb, m))
    rows = []
    for name, sig, pk, m in items:
        rows.append({"name": name, "sig": sig.hex(), "pk": pk.hex(), "msg": m.hex(),
                     "valid": pyref.ed25519_verify(sig, pk, m)})
    with open(os.path.join(HERE, "ed25519_long_lattice.json"), "w") as f:
        json.dump({"source": "tests/golden/make_ed25519_long_lattice.py; expected from tests/pyref.py", "items": rows}, f,
                  indent=1)


if __name__ == "__main__":
    main()
