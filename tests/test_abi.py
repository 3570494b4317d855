"""CPU checks of the C ABI library: it exists, loads, and exports every function declared in
include/sigops.h (no compute calls: there is no GPU here)."""
import ctypes
import os
import re

from sigops_pkg import sigops

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for hdr in os.listdir(os.path.join(REPO, "include")):
        if hdr.endswith(".h"):
            txt = open(os.path.join(REPO, "include", hdr)).read()
            names |= set(re.findall(r"\b(sigops_[a-z0-9_]+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(sigops.LIB_PATH), "libsigops.so not built"
    lib = ctypes.CDLL(sigops.LIB_PATH)
    names = declared_functions()
    assert {"sigops_create", "sigops_secp256k1_recover", "sigops_ed25519_verify"} <= names
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_reference_layout_tables():
    # sizes of the reference tables: 16 points x {x,y} x 20 limbs; Ed25519 16 x {x,y,t} x 20
    # (src/precompute.rs:12-69, 13-bit limbs -> 20 limbs, src/shader.rs:51)
    assert len(sigops.precompute.secp256k1_bases(13)) == 640
    assert len(sigops.precompute.secp256r1_bases(13)) == 640
    assert len(sigops.precompute.ed25519_bases(13)) == 960
    t = sigops.precompute.secp256k1_bases(13)
    # entry 0 is G in Montgomery form with R = 2^260
    p = 2**256 - 2**32 - 977
    gx = sum(v << (13 * i) for i, v in enumerate(t[:20]))
    assert gx == 0x79BE667EF9DCBBAC55A06295CE870B07029BFCDB2DCE28D959F2815B16F81798 * 2**260 % p


def test_mirror_argument_errors_without_gpu():
    import pytest
    with pytest.raises(ValueError):
        sigops.secp256k1_ecdsa.ecrecover([b"\0" * 64], [], sigops.precompute.secp256k1_bases(13), 13)
    with pytest.raises(ValueError):
        sigops.secp256k1_ecdsa.ecrecover_single_shader([], [], 9)
    assert sigops.secp256k1_ecdsa.ecrecover([], [], sigops.precompute.secp256k1_bases(13), 13) == []
    assert sigops.ed25519_eddsa.ecverify_single([], [], [], 13) == []
