"""MI355X batch signature engine: Python binding of libsigops.so (C ABI in include/sigops.h).

Two layers:

* array API on a ``Context`` (numpy uint8 arrays in, numpy arrays out), used by the tests and the
  benchmark: ``secp256k1_recover``, ``secp256r1_recover``, ``ed25519_verify``, plus ``*_dev``
  variants over device pointers (inputs already resident in HBM);
* a mirror of the reference's public functions with the same names, argument meaning and error
  behaviour -- ``secp256k1_ecdsa.ecrecover`` / ``ecrecover_single_shader`` (src/secp256k1_ecdsa.rs:61,215),
  ``secp256r1_ecdsa.*`` (src/secp256r1_ecdsa.rs:62,216), ``ed25519_eddsa.ecverify`` /
  ``ecverify_single`` (src/ed25519_eddsa.rs:67,259), ``precompute.*_bases`` (src/precompute.rs:36-69)
  and ``ShaderFailureError`` (src/lib.rs:12-14).

There is no CPU fallback: if the library or a GPU is missing, every call raises.
"""
import ctypes
import os
import subprocess
import threading

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libsigops.so")

OK, E_ARG, E_DEVICE, E_NOMEM, E_CANARY = 0, -1, -2, -3, -4
ST_OK, ST_BAD_R, ST_BAD_S, ST_NO_POINT, ST_INFINITY = 0, 1, 2, 3, 4
ED_OK, ED_BAD_S, ED_BAD_KEY, ED_MISMATCH = 0, 1, 2, 3
SCHEME_SECP256K1, SCHEME_SECP256R1, SCHEME_ED25519 = 0, 1, 2


class ShaderFailureError(RuntimeError):
    """Raised when a kernel did not produce every result (the reference's success-canary error)."""


class SigopsError(RuntimeError):
    """Any other failure of the native library (bad argument, HIP error, allocation failure)."""


class _Config(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("blocks_per_cu", ctypes.c_uint32), ("max_batch", ctypes.c_uint32),
                ("tile_rounds", ctypes.c_uint32), ("no_graph", ctypes.c_uint32), ("waves_per_simd", ctypes.c_uint32),
                ("bucket_order", ctypes.c_uint32), ("g_window", ctypes.c_uint32)]


class _ServerConfig(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("slots", ctypes.c_uint32), ("hbm_budget", ctypes.c_uint64),
                ("max_batch", ctypes.c_uint32), ("blocks_per_cu", ctypes.c_uint32), ("mode", ctypes.c_uint32),
                ("ring_tiles", ctypes.c_uint32), ("ring_msg_bytes", ctypes.c_uint64),
                ("idle_close_us", ctypes.c_uint32)]


SERVE_SLOTS = 0
SERVE_RING = 1


_lib = None
_lib_lock = threading.Lock()


def build(force=False):
    """Compile libsigops.so for gfx950 in-tree (make -C wgpu-sigops_amd)."""
    cmd = ["make", "-C", PKG_DIR, "-j4"]
    if force:
        cmd.insert(1, "-B")
    subprocess.run(cmd, check=True)
    return LIB_PATH


def lib():
    """Load libsigops.so (raises if it is missing: there is no fallback path)."""
    global _lib
    with _lib_lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise SigopsError(f"{LIB_PATH} is not built; run wgpu-sigops_amd build (make -C wgpu-sigops_amd)")
            L = ctypes.CDLL(LIB_PATH)
            vp, sz, c_p = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p
            L.sigops_create.restype = vp
            L.sigops_create.argtypes = [ctypes.POINTER(_Config)]
            L.sigops_destroy.argtypes = [vp]
            L.sigops_last_error.restype = ctypes.c_char_p
            L.sigops_last_error.argtypes = [vp]
            for name in ("sigops_secp256k1_recover", "sigops_secp256r1_recover"):
                getattr(L, name).argtypes = [vp, c_p, c_p, sz, c_p, c_p]
            L.sigops_ed25519_verify.argtypes = [vp, c_p, c_p, c_p, c_p, sz, c_p, c_p]
            for name in ("sigops_secp256k1_recover_table", "sigops_secp256r1_recover_table"):
                getattr(L, name).argtypes = [vp, c_p, c_p, sz, vp, sz, ctypes.c_uint32, c_p, c_p]
            L.sigops_ed25519_verify_table.argtypes = [vp, c_p, c_p, c_p, c_p, sz, vp, sz, ctypes.c_uint32, c_p, c_p]
            for name in ("sigops_secp256k1_verify", "sigops_secp256r1_verify"):
                getattr(L, name).argtypes = [vp, c_p, c_p, c_p, sz, c_p, c_p]
            for name in ("sigops_secp256k1_verify_dev", "sigops_secp256r1_verify_dev"):
                getattr(L, name).argtypes = [vp, vp, vp, vp, sz, vp, vp]
            for name in ("sigops_secp256k1_recover_dev", "sigops_secp256r1_recover_dev"):
                getattr(L, name).argtypes = [vp, vp, vp, sz, vp, vp, vp]
            L.sigops_ed25519_verify_dev.argtypes = [vp, vp, vp, vp, vp, sz, vp, vp]
            L.sigops_ed25519_verify_strict.argtypes = [vp, c_p, c_p, c_p, c_p, sz, c_p, c_p]
            L.sigops_ed25519_verify_strict_dev.argtypes = [vp, vp, vp, vp, vp, sz, vp, vp]
            L.sigops_batch.argtypes = [vp, c_p, c_p, c_p, c_p, c_p, sz, c_p, c_p]
            L.sigops_batch_dev.argtypes = [vp, vp, vp, vp, vp, vp, sz, vp, vp, vp]
            L.sigops_message_new.argtypes = [vp, c_p, c_p, sz, c_p]
            L.sigops_message_new_dev.argtypes = [vp, vp, vp, sz, vp, vp]
            L.sigops_message_new_noalloc.argtypes = [vp, vp, vp, sz, vp, vp]
            L.sigops_address_from_pubkeys_dev.argtypes = [vp, vp, vp, sz, vp, vp]
            for name in ("sigops_secp256k1_recover_address", "sigops_secp256r1_recover_address"):
                getattr(L, name).argtypes = [vp, c_p, c_p, sz, c_p, c_p]
            L.sigops_set_stage_timing.argtypes = [vp, ctypes.c_int]
            L.sigops_set_fault.argtypes = [vp, ctypes.c_uint32]
            L.sigops_set_streaming.argtypes = [vp, ctypes.c_int]
            L.sigops_batch_host_ms.argtypes = [vp, ctypes.POINTER(ctypes.c_float)]
            L.sigops_batch_stage_ms.argtypes = [vp, ctypes.POINTER(ctypes.c_float)]
            L.sigops_gen_secp_dev.argtypes = [vp, ctypes.c_int, ctypes.c_uint64, sz, ctypes.c_uint32, vp, vp, vp, vp,
                                              vp]
            L.sigops_gen_ed25519_dev.argtypes = [vp, ctypes.c_uint64, sz, ctypes.c_uint32, vp, vp, vp, vp, vp, vp]
            L.sigops_probe_valu.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                            ctypes.POINTER(ctypes.c_double)]
            L.sigops_server_create.restype = vp
            L.sigops_server_create.argtypes = [ctypes.POINTER(_ServerConfig)]
            L.sigops_server_destroy.argtypes = [vp]
            L.sigops_server_submit.restype = ctypes.c_int64
            L.sigops_server_submit.argtypes = [vp, c_p, c_p, c_p, c_p, c_p, sz]
            L.sigops_server_wait.argtypes = [vp, ctypes.c_int64, c_p, c_p, ctypes.c_int,
                                             ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]
            L.sigops_server_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]
            L.sigops_server_replay.argtypes = [vp, c_p, c_p, c_p, c_p, c_p, c_p, c_p, sz, c_p, c_p, c_p]
            _lib = L
        return _lib


def source_digest():
    """SHA-256 over the sources the mixed-batch kernel's code object is built from (csrc/kern_batch.hip,
    the headers it includes, directly or not, and the Makefile): ties a committed counter profile of
    that kernel to the build it measured, while host-only sources may change"""
    import hashlib
    import re
    csrc = os.path.join(PKG_DIR, "csrc")
    names, todo = set(), ["kern_batch.hip"]
    while todo:
        name = todo.pop()
        path = os.path.join(csrc, name)
        if name in names or not os.path.isfile(path):
            continue
        names.add(name)
        with open(path, encoding="utf-8", errors="replace") as f:
            todo.extend(re.findall(r'^\s*#\s*include\s+"([^"/]+)"', f.read(), re.M))
    h = hashlib.sha256()
    for name in sorted(names) + ["../Makefile"]:
        path = os.path.join(csrc, name)
        if os.path.isfile(path):
            h.update(name.encode())
            with open(path, "rb") as f:
                h.update(f.read())
    return h.hexdigest()


def probe_valu(device=0):
    """(v_mad_u64_u32 ops/s, v_fma_f64 ops/s) measured over the whole GPU"""
    mad, fma = ctypes.c_double(), ctypes.c_double()
    rc = lib().sigops_probe_valu(device, ctypes.byref(mad), ctypes.byref(fma))
    if rc:
        raise SigopsError(f"sigops_probe_valu failed ({rc})")
    return mad.value, fma.value


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None and a.size else None


def _u8(a, shape_tail, name):
    a = np.ascontiguousarray(a, dtype=np.uint8)
    if a.ndim == 1 and shape_tail:
        a = a.reshape(-1, shape_tail)
    if shape_tail and (a.ndim != 2 or a.shape[1] != shape_tail):
        raise ValueError(f"{name} must be (n, {shape_tail}) bytes")
    return a


def _tuned(device):
    try:
        from . import autotune
        return autotune.load(device)
    except Exception:
        return None


FAULT_SKIP_COMPUTE = 1
NOALLOC_MAX_MESSAGES = 1 << 20
DEFAULT_MAX_BATCH = 1 << 22


class Context:
    """One GPU context: resident generator tables, per-lane scratch and reusable I/O buffers.

    The native context serves one caller at a time (its staging buffers, stream and scratch are
    shared), so every native call on a Context holds the Context's lock: threads that share one
    Context (the mirrors share the default context) are serialized, as the reference's per-call
    devices are independent. The host-array methods split inputs larger than max_batch into
    consecutive calls of at most max_batch items."""

    def __init__(self, device=0, blocks_per_cu=0, max_batch=0, tile_rounds=0, use_graph=True, tuned=True, waves=0,
                 bucket_order=0, g_window=0):
        self._lock = threading.RLock()
        self.max_batch = max_batch or DEFAULT_MAX_BATCH
        if tuned and not (blocks_per_cu or tile_rounds or waves or bucket_order or g_window):
            # the variant the autotuner measured fastest on this device (autotune.py), if any
            t = _tuned(device)
            if t:
                waves, tile_rounds, bucket_order = t.get("waves", 0), t.get("tile_rounds", 0), t.get("bucket_order", 0)
                g_window = t.get("g_window", 0)
        self.waves = waves or 2
        self.blocks_per_cu = blocks_per_cu or self.waves
        self.tile_rounds = tile_rounds or 1
        self.bucket_order = bucket_order
        self.g_window = g_window or (8 if self.waves == 1 else 9)
        cfg = _Config(device, blocks_per_cu, max_batch, tile_rounds, 0 if use_graph else 1, waves, bucket_order,
                      g_window)
        self._h = lib().sigops_create(ctypes.byref(cfg))
        if not self._h:
            raise SigopsError("sigops_create failed (no usable GPU?)")

    def close(self):
        if getattr(self, "_h", None):
            lib().sigops_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    def _call(self, fn, *args):
        """fn(handle, *args) under the context's lock; raises on a failure status"""
        with self._lock:
            rc = fn(self._h, *args)
            msg = lib().sigops_last_error(self._h).decode(errors="replace") if rc != OK else ""
        if rc == OK:
            return
        if rc == E_CANARY:
            raise ShaderFailureError(msg)
        raise SigopsError(f"libsigops error {rc}: {msg}")

    def _chunks(self, n):
        """[a, b) ranges of at most max_batch items covering n"""
        step = max(1, self.max_batch)
        return [(a, min(n, a + step)) for a in range(0, n, step)] or [(0, 0)]

    def set_streaming(self, on):
        """host batches streamed through pinned memory while the kernel runs (default), or copied
        to the device first (sigops_set_streaming); the results are the same"""
        self._call(lib().sigops_set_streaming, 1 if on else 0)

    HOST_PHASES = ("head_packed_launched", "all_packed", "all_unpacked", "launch_done", "call_done", "launch_gpu_ms")

    def batch_host_ms(self):
        """phases of the last streamed host batch (ms from the call's start; the last is the launch's
        own GPU time), keyed by HOST_PHASES"""
        ms = (ctypes.c_float * len(self.HOST_PHASES))()
        self._call(lib().sigops_batch_host_ms, ms)
        return dict(zip(self.HOST_PHASES, (round(float(x), 3) for x in ms)))

    def set_fault(self, flags):
        """test switch (sigops_set_fault): FAULT_SKIP_COMPUTE makes every recovery, verification
        and batch call skip its compute kernels, so the host entry points report the unwritten
        status canaries as ShaderFailureError; 0 turns it off"""
        self._call(lib().sigops_set_fault, int(flags))

    def _secp(self, fn, sigs, digests):
        sigs = _u8(sigs, 64, "sigs")
        digests = _u8(digests, 32, "digests")
        n = sigs.shape[0]
        if digests.shape[0] != n:
            raise ValueError("sigs and digests differ in length")
        out = np.zeros((n, 64), np.uint8)
        st = np.zeros(n, np.uint8)
        for a, b in self._chunks(n):
            self._call(fn, _ptr(sigs[a:b]), _ptr(digests[a:b]), b - a, _ptr(out[a:b]), _ptr(st[a:b]))
        return out, st

    def secp256k1_recover(self, sigs, digests):
        """(n,64) fuel signatures, (n,32) digests -> ((n,64) X||Y, (n,) status)"""
        return self._secp(lib().sigops_secp256k1_recover, sigs, digests)

    def secp256r1_recover(self, sigs, digests):
        return self._secp(lib().sigops_secp256r1_recover, sigs, digests)

    def secp_recover_table(self, curve, sigs, digests, table_limbs, log_limb_size):
        """recovery with the caller's fixed-base table in the reference's layout: u1 * P + u2 * R for
        the table's point P (raises SigopsError when the table is not [P, ..., 16P])"""
        tab = np.ascontiguousarray(table_limbs, dtype=np.uint32).reshape(-1)
        fn = lib().sigops_secp256k1_recover_table if curve == 0 else lib().sigops_secp256r1_recover_table

        def call(h, ps, pd, n, po, pst):
            return fn(h, ps, pd, n, tab.ctypes.data, tab.size, int(log_limb_size), po, pst)
        return self._secp(call, sigs, digests)

    def ed25519_verify_table(self, sigs, pubkeys, msgs, msg_off, table_limbs, log_limb_size):
        """verification with the caller's fixed-base table (reference layout, x, y, x*y per entry)"""
        sigs = _u8(sigs, 64, "sigs")
        pubkeys = _u8(pubkeys, 32, "pubkeys")
        n = sigs.shape[0]
        msg_off = np.ascontiguousarray(msg_off, dtype=np.uint64)
        if pubkeys.shape[0] != n or msg_off.shape[0] != n + 1:
            raise ValueError("sigs, pubkeys and msg_off (n+1) differ in length")
        msgs = np.ascontiguousarray(msgs, dtype=np.uint8).reshape(-1)
        tab = np.ascontiguousarray(table_limbs, dtype=np.uint32).reshape(-1)
        valid = np.zeros(n, np.uint8)
        st = np.zeros(n, np.uint8)
        for a, b in self._chunks(n):
            self._call(lib().sigops_ed25519_verify_table, _ptr(sigs[a:b]), _ptr(pubkeys[a:b]),
                       _ptr(msgs) if msgs.size else None, _ptr(msg_off[a:b + 1]), b - a, tab.ctypes.data, tab.size,
                       int(log_limb_size), _ptr(valid[a:b]), _ptr(st[a:b]))
        return valid.astype(bool), st

    def _secp_verify(self, fn, sigs, digests, pubkeys):
        sigs = _u8(sigs, 64, "sigs")
        digests = _u8(digests, 32, "digests")
        pubkeys = _u8(pubkeys, 64, "pubkeys")
        n = sigs.shape[0]
        if digests.shape[0] != n or pubkeys.shape[0] != n:
            raise ValueError("sigs, digests and pubkeys differ in length")
        valid = np.zeros(n, np.uint8)
        st = np.zeros(n, np.uint8)
        for a, b in self._chunks(n):
            self._call(fn, _ptr(sigs[a:b]), _ptr(digests[a:b]), _ptr(pubkeys[a:b]), b - a, _ptr(valid[a:b]),
                       _ptr(st[a:b]))
        return valid.astype(bool), st

    def secp256k1_verify(self, sigs, digests, pubkeys):
        """fuel Signature::verify: (n,64) signatures, (n,32) digests, (n,64) X||Y keys
        -> ((n,) bool, (n,) status)"""
        return self._secp_verify(lib().sigops_secp256k1_verify, sigs, digests, pubkeys)

    def secp256r1_verify(self, sigs, digests, pubkeys):
        return self._secp_verify(lib().sigops_secp256r1_verify, sigs, digests, pubkeys)

    def secp256k1_verify_dev(self, d_sigs, d_digests, d_pubkeys, n, d_status, stream=0):
        self._call(lib().sigops_secp256k1_verify_dev, d_sigs, d_digests, d_pubkeys, n, d_status, stream or None)

    def secp256r1_verify_dev(self, d_sigs, d_digests, d_pubkeys, n, d_status, stream=0):
        self._call(lib().sigops_secp256r1_verify_dev, d_sigs, d_digests, d_pubkeys, n, d_status, stream or None)

    def ed25519_verify(self, sigs, pubkeys, msgs, msg_off, strict=False):
        """(n,64) R||s, (n,32) keys, concatenated messages + (n+1) offsets -> ((n,) bool, (n,) status);
        strict: ed25519-dalek verify_strict rules"""
        sigs = _u8(sigs, 64, "sigs")
        pubkeys = _u8(pubkeys, 32, "pubkeys")
        n = sigs.shape[0]
        msg_off = np.ascontiguousarray(msg_off, dtype=np.uint64)
        if pubkeys.shape[0] != n or msg_off.shape[0] != n + 1:
            raise ValueError("sigs, pubkeys and msg_off (n+1) differ in length")
        msgs = np.ascontiguousarray(msgs, dtype=np.uint8).reshape(-1)
        if n and int(msg_off[-1]) > msgs.size:
            raise ValueError("msg_off points past the message buffer")
        valid = np.zeros(n, np.uint8)
        st = np.zeros(n, np.uint8)
        fn = lib().sigops_ed25519_verify_strict if strict else lib().sigops_ed25519_verify
        for a, b in self._chunks(n):
            self._call(fn, _ptr(sigs[a:b]), _ptr(pubkeys[a:b]), _ptr(msgs) if msgs.size else None,
                       _ptr(msg_off[a:b + 1]), b - a, _ptr(valid[a:b]), _ptr(st[a:b]))
        return valid.astype(bool), st

    def ed25519_verify_strict(self, sigs, pubkeys, msgs, msg_off):
        return self.ed25519_verify(sigs, pubkeys, msgs, msg_off, strict=True)

    def ed25519_verify_strict_dev(self, d_sigs, d_pubkeys, d_msgs, d_msg_off, n, d_status, stream=0):
        self._call(lib().sigops_ed25519_verify_strict_dev, d_sigs, d_pubkeys, d_msgs, d_msg_off, n, d_status,
                   stream or None)

    def batch(self, schemes, sigs, aux, msgs, msg_off):
        """mixed batch: (n,) scheme bytes, (n,64) sigs, (n,32) digest-or-key, messages + (n+1) offsets
        -> ((n,64) recovered keys (zeros for Ed25519 items), (n,) status)"""
        schemes = np.ascontiguousarray(schemes, dtype=np.uint8).reshape(-1)
        sigs = _u8(sigs, 64, "sigs")
        aux = _u8(aux, 32, "aux")
        n = schemes.shape[0]
        msg_off = np.ascontiguousarray(msg_off, dtype=np.uint64)
        if sigs.shape[0] != n or aux.shape[0] != n or msg_off.shape[0] != n + 1:
            raise ValueError("schemes, sigs, aux and msg_off (n+1) differ in length")
        msgs = np.ascontiguousarray(msgs, dtype=np.uint8).reshape(-1)
        out = np.zeros((n, 64), np.uint8)
        st = np.zeros(n, np.uint8)
        for a, b in self._chunks(n):
            self._call(lib().sigops_batch, _ptr(schemes[a:b]), _ptr(sigs[a:b]), _ptr(aux[a:b]),
                       _ptr(msgs) if msgs.size else None, _ptr(msg_off[a:b + 1]), b - a, _ptr(out[a:b]),
                       _ptr(st[a:b]))
        return out, st

    def batch_dev(self, d_schemes, d_sigs, d_aux, d_msgs, d_msg_off, n, d_out, d_status, stream=0):
        self._call(lib().sigops_batch_dev, d_schemes, d_sigs, d_aux, d_msgs, d_msg_off, n, d_out, d_status,
                   stream or None)

    # ---- Fuel hashing (SHA-256) either side of the path ----
    def message_new(self, data, off=None):
        """Fuel Message::new on the device: SHA-256 of each message. ``data`` is a list of bytes
        objects, or one uint8 array with ``off`` (n + 1 offsets). Returns an (n, 32) uint8 array."""
        if off is None:
            off = np.zeros(len(data) + 1, np.uint64)
            np.cumsum([len(m) for m in data], out=off[1:])
            data = np.frombuffer(b"".join(bytes(m) for m in data), np.uint8)
        data = np.ascontiguousarray(data, dtype=np.uint8)
        off = np.ascontiguousarray(off, dtype=np.uint64)
        n = len(off) - 1
        out = np.zeros((n, 32), np.uint8)
        for a, b in self._chunks(n):
            self._call(lib().sigops_message_new, _ptr(data) if data.size else None, _ptr(off[a:b + 1]), b - a,
                       _ptr(out[a:b]))
        return out

    def message_new_dev(self, d_data, d_off, n, d_digests, stream=0):
        self._call(lib().sigops_message_new_dev, d_data, d_off, n, d_digests, stream or None)

    def message_new_noalloc_dev(self, d_data, d_off, n, d_digests, stream=0):
        """message_new_dev with the kernels' scratch in the library's fixed device array: the call
        allocates nothing (at most NOALLOC_MAX_MESSAGES messages; calls on one GPU must not overlap)"""
        self._call(lib().sigops_message_new_noalloc, d_data, d_off, n, d_digests, stream or None)

    def address_from_pubkeys_dev(self, d_pubs, d_status, n, d_addr, stream=0):
        self._call(lib().sigops_address_from_pubkeys_dev, d_pubs, d_status, n, d_addr, stream or None)

    def recover_address(self, curve, sigs, digests):
        """recovery then Fuel address SHA-256(X || Y) on the device: (n, 32) addresses (zeros for
        failed items) and the status bytes"""
        sigs = _u8(sigs, 64, "sigs")
        digests = _u8(digests, 32, "digests")
        n = sigs.shape[0]
        if digests.shape[0] != n:
            raise ValueError("signatures and digests differ in length")
        out = np.zeros((n, 32), np.uint8)
        st = np.zeros(n, np.uint8)
        fn = lib().sigops_secp256k1_recover_address if curve == 0 else lib().sigops_secp256r1_recover_address
        for a, b in self._chunks(n):
            self._call(fn, _ptr(sigs[a:b]), _ptr(digests[a:b]), b - a, _ptr(out[a:b]), _ptr(st[a:b]))
        return out, st

    STAGES = ("bucketing", "ed25519_hash", "verify_kernel", "careful")

    def set_stage_timing(self, on):
        """Eager batch launches with HIP events between the stages (see batch_stage_ms)."""
        self._call(lib().sigops_set_stage_timing, 1 if on else 0)

    def batch_stage_ms(self):
        """Stage durations (ms) of the last batch launched with stage timing on, keyed by STAGES."""
        ms = (ctypes.c_float * len(self.STAGES))()
        self._call(lib().sigops_batch_stage_ms, ms)
        return dict(zip(self.STAGES, (float(x) for x in ms)))

    def gen_secp_dev(self, curve, seed, n, corrupt_ppm, d_sigs, d_digests, d_pubs, d_corrupted=None, stream=0):
        self._call(lib().sigops_gen_secp_dev, curve, seed, n, corrupt_ppm, d_sigs, d_digests, d_pubs, d_corrupted,
                   stream or None)

    def gen_ed25519_dev(self, seed, n, corrupt_ppm, d_msg_off, d_msgs, d_sigs, d_pubkeys, d_corrupted=None, stream=0):
        self._call(lib().sigops_gen_ed25519_dev, seed, n, corrupt_ppm, d_msg_off, d_msgs, d_sigs, d_pubkeys,
                   d_corrupted, stream or None)

    # device-resident variants (torch tensors or raw ints as device pointers)
    def secp_recover_dev(self, curve, d_sigs, d_digests, n, d_out, d_status, stream=0):
        fn = lib().sigops_secp256k1_recover_dev if curve == 0 else lib().sigops_secp256r1_recover_dev
        self._call(fn, d_sigs, d_digests, n, d_out, d_status, stream or None)

    def ed25519_verify_dev(self, d_sigs, d_pubkeys, d_msgs, d_msg_off, n, d_status, stream=0):
        self._call(lib().sigops_ed25519_verify_dev, d_sigs, d_pubkeys, d_msgs, d_msg_off, n, d_status, stream or None)


class Server:
    """Continuous-batching request server (sigops_server_*): submit() returns a ticket at once,
    wait() returns (out, status, t_submit_ns, t_done_ns). mode SERVE_SLOTS packs requests into
    batches of per-slot contexts; SERVE_RING feeds them, tile by tile, to one running persistent
    launch (ring_tiles positions, ring_msg_bytes of messages, ended after idle_close_us idle)."""

    def __init__(self, device=0, slots=0, hbm_budget=0, max_batch=0, blocks_per_cu=0, mode=SERVE_SLOTS,
                 ring_tiles=0, ring_msg_bytes=0, idle_close_us=0):
        cfg = _ServerConfig(device, slots, hbm_budget, max_batch, blocks_per_cu, mode, ring_tiles, ring_msg_bytes,
                            idle_close_us)
        self._h = lib().sigops_server_create(ctypes.byref(cfg))
        if not self._h:
            raise SigopsError("sigops_server_create failed")
        self._n = {}

    def close(self):
        if getattr(self, "_h", None):
            lib().sigops_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, schemes, sigs, aux, msgs, msg_off):
        schemes = np.ascontiguousarray(schemes, dtype=np.uint8).reshape(-1)
        n = schemes.shape[0]
        sigs = _u8(sigs, 64, "sigs")
        aux = _u8(aux, 32, "aux")
        msgs = np.ascontiguousarray(msgs, dtype=np.uint8).reshape(-1)
        msg_off = np.ascontiguousarray(msg_off, dtype=np.uint64)
        t = lib().sigops_server_submit(self._h, _ptr(schemes), _ptr(sigs), _ptr(aux), _ptr(msgs) if msgs.size else None,
                                       _ptr(msg_off), n)
        if t < 0:
            raise SigopsError(f"submit failed ({t})")
        self._n[t] = n
        return t

    def wait(self, ticket, timeout_ms=-1):
        n = self._n[ticket]
        out = np.zeros((n, 64), np.uint8)
        st = np.zeros(n, np.uint8)
        ts, td = ctypes.c_uint64(), ctypes.c_uint64()
        rc = lib().sigops_server_wait(self._h, ticket, _ptr(out), _ptr(st), timeout_ms, ctypes.byref(ts),
                                      ctypes.byref(td))
        if rc == 1:
            return None
        del self._n[ticket]
        if rc == E_CANARY:
            raise ShaderFailureError("request results were not written")
        if rc:
            raise SigopsError(f"request failed ({rc})")
        return out, st, ts.value, td.value

    def stats(self):
        b, i = ctypes.c_uint64(), ctypes.c_uint64()
        lib().sigops_server_stats(self._h, ctypes.byref(b), ctypes.byref(i))
        return b.value, i.value

    def replay(self, schemes, sigs, aux, msgs, msg_off, req_off, arrival_ns):
        """native client: request r = items req_off[r]..req_off[r+1], submitted at arrival_ns[r];
        returns (latency_ns per request, out, status)"""
        schemes = np.ascontiguousarray(schemes, dtype=np.uint8)
        sigs = _u8(sigs, 64, "sigs")
        aux = _u8(aux, 32, "aux")
        msgs = np.ascontiguousarray(msgs, dtype=np.uint8).reshape(-1)
        msg_off = np.ascontiguousarray(msg_off, dtype=np.uint64)
        req_off = np.ascontiguousarray(req_off, dtype=np.uint64)
        arrival_ns = np.ascontiguousarray(arrival_ns, dtype=np.uint64)
        nreq = arrival_ns.shape[0]
        lat = np.zeros(nreq, np.uint64)
        out = np.zeros((schemes.shape[0], 64), np.uint8)
        st = np.zeros(schemes.shape[0], np.uint8)
        rc = lib().sigops_server_replay(self._h, _ptr(schemes), _ptr(sigs), _ptr(aux), _ptr(msgs) if msgs.size else None,
                                        _ptr(msg_off), _ptr(req_off), _ptr(arrival_ns), nreq, _ptr(lat), _ptr(out),
                                        _ptr(st))
        if rc:
            raise SigopsError(f"replay failed ({rc})")
        return lat, out, st


_default_ctx = None
_default_lock = threading.Lock()


def default_context():
    """the Context the reference-API mirrors use when none is given (created once, thread-safe)"""
    global _default_ctx
    with _default_lock:
        if _default_ctx is None:
            _default_ctx = Context()
        return _default_ctx


from . import ed25519_eddsa, fuel, precompute, secp256k1_ecdsa, secp256r1_ecdsa  # noqa: E402,F401
