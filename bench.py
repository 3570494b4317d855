#!/usr/bin/env python3
"""Benchmark of the MI355X signature engine (one GPU).

Workload (the headline configuration): a mixed batch of 2^20 signatures -- 50 % secp256k1
recoveries, 30 % secp256r1 recoveries, 20 % Ed25519 verifications with message lengths
log-uniform on 0..4096 bytes, schemes randomly interleaved, 5 % corrupted -- generated on the GPU
by the engine's seeded signer and resident in HBM when the timed region starts.

One step = one batched, graph-captured pass over the whole batch (device-side bucketing + one
persistent kernel + careful pass). `value` is signatures per second over the K timed steps, inputs
resident in HBM. Also reported, each checked for exactness:
  * incl_transfers: the batch through the host API (host arrays in, host arrays out), the
    reference's benchmark method (src/benchmarks/secp256k1_ecdsa.rs:126-138);
  * plain: the three per-scheme kernels on the same items (device-resident, and through the host
    API with transfers: three calls against the batch's one), and the batched/plain ratios;
  * configs: the plain per-scheme configurations -- secp256k1 recovery at N = 1024, secp256r1
    recovery at N = 65536, Ed25519 verification at N = 65536 (messages 0..1024 bytes, 10 %
    corrupted) -- through the host API with transfers and device-resident, each with its CPU
    baselines on the same items;
  * roofline: the persistent kernel's VALU use per launch, from the committed counter profile of
    this build at this launch shape (busy SIMD cycles, instructions), over its duration measured
    here with HIP events, against MI355X's VALU issue peak; the canonical textbook work per
    signature over the same peak is `vs_textbook`;
  * cpu_baseline: the CPU oracle (a plain C restatement) on a bounded sample of the same items on
    this host's cores, with OpenSSL libcrypto (a tuned library; ECDSA verification stands in for
    recovery, which it does not offer) timed on the same sample beside it;
  * served: Poisson arrivals of 1..4096-signature requests at three offered loads.
Prints ONE JSON line.
"""
import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)
# the CPU baseline leg times the oracle and OpenSSL through the tests' bindings
sys.path.insert(1, os.path.join(REPO, "tests"))

# canonical multiply-accumulates per signature (u32 x u32 -> u64) of a textbook schedule: fixed-window
# Fermat inversions, 252 doublings, dense 2 x 8^2 MACs per 256-bit product
MAC_PER_SIG = {"secp256k1": 595072, "secp256r1": 627456, "ed25519_per_block": 24 * 128, "ed25519_base": 3850 * 128}
# MI355X VALU issue peak: 256 CUs x 4 SIMDs, each issuing one wave64 VALU instruction every 4 cycles
# (16 lanes per cycle) at the 2.4 GHz peak engine clock: 39.3 T lane-operation slots per second
CUS, SIMDS_PER_CU, CLOCK_HZ = 256, 4, 2.4e9
VALU_PEAK = CUS * SIMDS_PER_CU * 16 * CLOCK_HZ
COUNTER_PROFILE = os.path.join(REPO, "profiles", "counters_k_batch_run_mixed_2p20.md")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def fuel_hashing(ctx, n, mx, out, st, stream, s_ptr, steps, dev):
    """Fuel hashing either side of the path, each form timed with events on the launch stream and
    checked against hashlib on a sample: Message::new over 2^20 64-byte messages (the reference's
    benchmarks sign Message::new(hex of a 32-byte value), src/benchmarks/secp256k1_ecdsa.rs:162),
    Message::new over the mixed batch's own layout (secp items hash empty messages, Ed25519 items
    their log-uniform 0..4096-byte messages), and the addresses of the 2^20 recovered keys."""
    import hashlib
    import torch
    hexchars = torch.tensor(list(b"0123456789abcdef"), dtype=torch.uint8, device=dev)
    g = torch.Generator(device=dev).manual_seed(11)
    hex_msgs = torch.zeros(n * 64 + 16, dtype=torch.uint8, device=dev)
    hex_msgs[: n * 64] = hexchars[torch.randint(0, 16, (n * 64,), device=dev, generator=g)]
    hex_off = torch.arange(0, n + 1, dtype=torch.int64, device=dev) * 64
    forms = {"hex64": (hex_msgs, hex_off, hex_off.cpu().numpy()), "mixed": (mx.msgs, mx.msg_off, mx.msg_off_host)}
    dig = torch.zeros((n, 32), dtype=torch.uint8, device=dev)
    addr = torch.zeros((n, 32), dtype=torch.uint8, device=dev)
    rng = np.random.default_rng(5)
    res = {"messages": n}
    ok = True
    for name, (msgs, off, off_h) in forms.items():
        ctx.message_new_dev(msgs.data_ptr(), off.data_ptr(), n, dig.data_ptr(), s_ptr)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            ctx.message_new_dev(msgs.data_ptr(), off.data_ptr(), n, dig.data_ptr(), s_ptr)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        h_dig, h_msgs = dig.cpu().numpy(), msgs.cpu().numpy()
        for i in rng.choice(n, 2000, replace=False):
            ok &= h_dig[i].tobytes() == hashlib.sha256(h_msgs[off_h[i]:off_h[i + 1]].tobytes()).digest()
        nbytes = int(off_h[-1])
        blocks = int(((np.diff(off_h) + 72) >> 6).sum())
        res[f"message_new_{name}"] = {"ms": round(ms, 3), "messages_per_s": round(n / (ms / 1e3), 1),
                                      "GBps": round(nbytes / (ms / 1e3) / 1e9, 2), "bytes": nbytes,
                                      "sha256_blocks": blocks, "blocks_per_s": round(blocks / (ms / 1e3), 1)}
        log(f"fuel Message::new ({name}): {ms:.3f} ms, {n / ms / 1e3:.1f} M msgs/s, "
            f"{nbytes / (ms / 1e3) / 1e9:.1f} GB/s, {blocks / ms / 1e3:.1f} M blocks/s")
    ctx.address_from_pubkeys_dev(out.data_ptr(), st.data_ptr(), n, addr.data_ptr(), s_ptr)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        ctx.address_from_pubkeys_dev(out.data_ptr(), st.data_ptr(), n, addr.data_ptr(), s_ptr)
    e1.record(stream)
    torch.cuda.synchronize()
    adr_ms = e0.elapsed_time(e1) / steps
    h_addr, h_out, h_st = addr.cpu().numpy(), out.cpu().numpy(), st.cpu().numpy()
    for i in rng.choice(n, 2000, replace=False):
        want = hashlib.sha256(h_out[i].tobytes()).digest() if h_st[i] == 0 else bytes(32)
        ok &= h_addr[i].tobytes() == want
    res["address"] = {"ms": round(adr_ms, 3), "addresses_per_s": round(n / (adr_ms / 1e3), 1)}
    res["sample_matches_hashlib"] = bool(ok)
    log(f"fuel addresses: {adr_ms:.3f} ms, {n / adr_ms / 1e3:.1f} M/s; sample exact: {ok}")
    return res


def predicates(ctx, mx, stream, s_ptr, steps, dev):
    """The predicates next to the path on the batch's own items, each scheme timed alone with events
    on the launch stream: secp verification with the generator's public keys (fuel
    Signature::verify) and Ed25519 verify_strict, with plain Ed25519 verification of the same items
    beside it. Every uncorrupted item must come out valid."""
    import torch
    k1, r1, ed = mx.parts
    res = {}
    for name, it in (("secp256k1_verify", k1), ("secp256r1_verify", r1), ("ed25519_verify", ed),
                     ("ed25519_verify_strict", ed)):
        st = torch.empty(it.n, dtype=torch.uint8, device=dev)

        def run():
            if it is k1 or it is r1:
                fn = ctx.secp256k1_verify_dev if it is k1 else ctx.secp256r1_verify_dev
                fn(it.sigs.data_ptr(), it.aux.data_ptr(), it.pubs.data_ptr(), it.n, st.data_ptr(), s_ptr)
            else:
                fn = ctx.ed25519_verify_strict_dev if name.endswith("strict") else ctx.ed25519_verify_dev
                fn(it.sigs.data_ptr(), it.aux.data_ptr(), it.msgs.data_ptr(), it.msg_off.data_ptr(), it.n,
                   st.data_ptr(), s_ptr)

        run()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            run()
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        st_h, cor = st.cpu().numpy(), it.corrupted.cpu().numpy()
        ok = bool((st_h[cor == 0] == 0).all())
        res[name] = {"n": it.n, "ms": round(ms, 3), "sigs_per_s": round(it.n / (ms / 1e3), 1),
                     "uncorrupted_all_valid": ok}
        log(f"{name}: n={it.n} {ms:.3f} ms, {it.n / ms / 1e3:.2f} M sigs/s, uncorrupted all valid: {ok}")
    return res


def mixed_mac_per_sig(schemes, lens):
    """average canonical MACs per signature of this exact batch"""
    k1 = (schemes == 0).sum()
    r1 = (schemes == 1).sum()
    ed = schemes == 2
    blocks = ((64 + lens[ed] + 17 + 127) // 128).sum()
    total = k1 * MAC_PER_SIG["secp256k1"] + r1 * MAC_PER_SIG["secp256r1"] + \
        ed.sum() * MAC_PER_SIG["ed25519_base"] + blocks * MAC_PER_SIG["ed25519_per_block"]
    return float(total) / len(schemes)


def read_counter_profile(digest, shape):
    """Counters of one launch of the persistent kernel from the committed counter profile
    (rocprofv3 --pmc passes of this benchmark): used only when the profile was taken on this build
    (same source digest) at this launch shape; otherwise the status says why not."""
    out = {"source": os.path.relpath(COUNTER_PROFILE, REPO), "status": None, "values": {}}
    try:
        kv = {}
        for line in open(COUNTER_PROFILE):
            if ":" in line and not line.startswith(("#", "|", " ", "-", "`")):
                k, v = line.split(":", 1)
                kv[k.strip()] = v.strip()
    except OSError:
        out["status"] = "missing"
        return out
    if kv.get("sources_sha256") != digest:
        out["status"] = "stale: the library sources changed after the profile was taken"
    elif kv.get("launch_shape") != shape:
        out["status"] = f"taken at another launch shape ({kv.get('launch_shape')})"
    else:
        out["status"] = "matches this build and launch shape"
        for k in ("valu_insts_per_launch", "valu_busy_quads_per_launch", "valu_int64_per_launch",
                  "valu_int32_per_launch", "traffic_bytes_per_launch", "kernel_ms_profiled", "valu_busy"):
            if k in kv:
                out["values"][k] = float(kv[k])
    return out


def roofline(n, mac_sig, kern_ms, stage_ms, prof):
    """VALU roofline of the persistent kernel: busy SIMD quad-cycles per launch (each one wave64
    VALU instruction, 64 lane-operations) over the launch's duration, against the issue peak.
    frac is at most 1 by construction: a SIMD is busy at most every cycle."""
    t = kern_ms / 1e3
    v = prof["values"]
    r = {"bound": "valu", "achieved": None, "peak": round(VALU_PEAK / 1e12, 3),
         "unit": "T VALU lane-op slots/s (busy SIMD quad-cycles x 64 lanes)", "frac": None, "traffic": None,
         "peak_basis": "256 CU x 4 SIMD x one wave64 VALU instruction per 4 cycles at 2.4 GHz",
         "kernel": "k_batch_run", "kernel_ms": round(kern_ms, 3), "stage_ms": {k: round(x, 3) for k, x in stage_ms.items()},
         "vs_textbook": round(n * mac_sig / t / VALU_PEAK, 4), "textbook_mac_per_sig": round(mac_sig),
         "vs_textbook_note": "canonical textbook MACs (fixed-window Fermat, 252 doublings, dense products) per "
                             "second over the peak; this kernel executes far fewer (GLV, sparse reduction, chains, "
                             "divsteps), so this ratio is not a utilisation",
         "counter_profile": {"source": prof["source"], "status": prof["status"]}}
    if v:
        busy = v.get("valu_busy_quads_per_launch")
        if busy:
            r["achieved"] = round(busy * 64 / t / 1e12, 3)
            r["frac"] = round(busy * 64 / t / VALU_PEAK, 4)
        if v.get("valu_insts_per_launch"):
            r["valu_insts_per_launch"] = v["valu_insts_per_launch"]
            r["issued_T_lane_ops_per_s"] = round(v["valu_insts_per_launch"] * 64 / t / 1e12, 3)
        if v.get("valu_int64_per_launch"):
            r["int64_insts_per_launch"] = v["valu_int64_per_launch"]
            r["int64_frac_of_peak"] = round(v["valu_int64_per_launch"] * 64 / t / VALU_PEAK, 4)
        if v.get("traffic_bytes_per_launch"):
            r["traffic"] = int(v["traffic_bytes_per_launch"])
            r["traffic_unit"] = "HBM bytes per launch (FETCH_SIZE + WRITE_SIZE, copy-calibrated)"
            r["hbm_GBps"] = round(r["traffic"] / t / 1e9, 1)
            r["hbm_frac"] = round(r["traffic"] / t / 8e12, 4)
        if v.get("kernel_ms_profiled"):
            r["counter_profile"]["kernel_ms_profiled"] = v["kernel_ms_profiled"]
        if v.get("valu_busy"):
            r["counter_profile"]["valu_busy_of_profiled_run"] = v["valu_busy"]
    return r


def cpu_time(fn, *args, **kw):
    t0 = time.perf_counter()
    res = fn(*args, **kw)
    return time.perf_counter() - t0, res


def cpu_secp_sample(curve, sigs, digs, pubs, threads):
    """oracle recovery (threads, and one core) and OpenSSL verification (threads) of a sample"""
    import oracle_lib as O
    import openssl_lib as S
    t, (o, st) = cpu_time(O.recover_batch, curve, sigs, digs, nthreads=threads)
    m1 = min(len(sigs), 1024)
    t1, _ = cpu_time(O.recover_batch, curve, sigs[:m1], digs[:m1], nthreads=1)
    res = {"oracle": {"sigs_per_s": round(len(sigs) / t, 1), "threads": threads,
                      "serial_1_core": round(m1 / t1, 1)}, "outputs": (o, st)}
    if S.available() and S.baseline_lib() is not None:
        to, valid = S.time_ecdsa_verify(curve, sigs, digs, pubs, threads)
        t1o, _ = S.time_ecdsa_verify(curve, sigs[:m1], digs[:m1], pubs[:m1], 1)
        res["openssl"] = {"sigs_per_s": round(len(sigs) / to, 1), "threads": threads,
                          "serial_1_core": round(m1 / t1o, 1), "op": "ECDSA_do_verify", "valid": valid}
    return res


def cpu_ed_sample(sigs, pks, msgs, off, threads):
    import oracle_lib as O
    import openssl_lib as S
    t, v = cpu_time(O.verify_batch, sigs, pks, msgs, off, nthreads=threads)
    m1 = min(len(sigs), 1024)
    t1, _ = cpu_time(O.verify_batch, sigs[:m1], pks[:m1], msgs, off[:m1 + 1], nthreads=1)
    res = {"oracle": {"sigs_per_s": round(len(sigs) / t, 1), "threads": threads,
                      "serial_1_core": round(m1 / t1, 1)}, "outputs": v}
    if S.available() and S.baseline_lib() is not None:
        to, valid = S.time_ed25519_verify(sigs, pks, msgs, off, threads)
        t1o, _ = S.time_ed25519_verify(sigs[:m1], pks[:m1], msgs, off[:m1 + 1], 1)
        res["openssl"] = {"sigs_per_s": round(len(sigs) / to, 1), "threads": threads,
                          "serial_1_core": round(m1 / t1o, 1), "op": "EVP_DigestVerify (Ed25519)", "valid": valid}
    return res


def plain_configs(ctx, W, dev, stream, steps, threads, with_cpu, cpu_sample):
    """The plain per-scheme configurations through the host API (host arrays in and out, transfers
    included, the reference's method) and device-resident, each checked on every item and, on a
    bounded sample, against the oracle, with the CPU baselines of that sample."""
    import torch
    res = []
    specs = (("config 1: secp256k1 recovery, N=1024, all valid", "secp", 0, 1024),
             ("config 2: secp256r1 recovery, N=65536, all valid", "secp", 1, 65536),
             ("config 3: ed25519 verification, N=65536, messages 0..1024 B, 10% corrupted", "ed", 2, 65536))
    for name, kind, curve, n in specs:
        if kind == "secp":
            it = W.secp(ctx, curve, n, seed=41 + curve, corrupt_ppm=0)
            torch.cuda.synchronize()
            h_sigs, h_dig, h_pub = it.sigs.cpu().numpy(), it.aux.cpu().numpy(), it.pubs.cpu().numpy()
            host_fn = ctx.secp256k1_recover if curve == 0 else ctx.secp256r1_recover
            out = torch.zeros((n, 64), dtype=torch.uint8, device=dev)
            st = torch.zeros(n, dtype=torch.uint8, device=dev)

            def host_call():
                return host_fn(h_sigs, h_dig)

            def dev_call():
                ctx.secp_recover_dev(curve, it.sigs.data_ptr(), it.aux.data_ptr(), n, out.data_ptr(), st.data_ptr(),
                                     stream.cuda_stream)

            def exact(r):
                return bool(np.array_equal(r[0], h_pub) and not r[1].any())
        else:
            lens = W.uniform_lengths(n, 1024, 43)
            it = W.ed25519(ctx, n, lens, seed=44, corrupt_ppm=100000)
            torch.cuda.synchronize()
            h_sigs, h_pks, h_off = it.sigs.cpu().numpy(), it.aux.cpu().numpy(), it.msg_off_host.astype(np.uint64)
            h_msgs = it.msgs.cpu().numpy()[: int(h_off[-1])]
            cor = it.corrupted.cpu().numpy()
            st = torch.zeros(n, dtype=torch.uint8, device=dev)

            def host_call():
                return ctx.ed25519_verify(h_sigs, h_pks, h_msgs, h_off)

            def dev_call():
                ctx.ed25519_verify_dev(it.sigs.data_ptr(), it.aux.data_ptr(), it.msgs.data_ptr(),
                                       it.msg_off.data_ptr(), n, st.data_ptr(), stream.cuda_stream)

            def exact(r):
                return bool(np.array_equal(r[0], cor == 0))
        first = host_call()
        reps = max(3, steps)
        t0 = time.perf_counter()
        for _ in range(reps):
            r = host_call()
        incl_ms = (time.perf_counter() - t0) / reps * 1e3
        dev_call()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            dev_call()
        e1.record(stream)
        torch.cuda.synchronize()
        dev_ms = e0.elapsed_time(e1) / reps
        entry = {"config": name, "n": n,
                 "incl_transfers": {"ms_per_call": round(incl_ms, 3), "sigs_per_s": round(n / incl_ms * 1e3, 1)},
                 "device_resident": {"ms_per_call": round(dev_ms, 3), "sigs_per_s": round(n / dev_ms * 1e3, 1)},
                 "all_items_as_generated": exact(first) and exact(r)}
        log(f"{name}: {incl_ms:.3f} ms incl. transfers ({n / incl_ms / 1e3:.2f} M sigs/s), {dev_ms:.3f} ms "
            f"device-resident; every item as generated: {entry['all_items_as_generated']}")
        if with_cpu:
            m = min(n, cpu_sample)
            if kind == "secp":
                c = cpu_secp_sample(curve, h_sigs[:m], h_dig[:m], h_pub[:m], threads)
                o, t = c.pop("outputs")
                entry["sample_bit_exact_vs_oracle"] = bool(np.array_equal(o, r[0][:m]) and np.array_equal(t, r[1][:m]))
            else:
                off_m = h_off[: m + 1]
                c = cpu_ed_sample(h_sigs[:m], h_pks[:m], h_msgs, off_m, threads)
                v = c.pop("outputs")
                entry["sample_bit_exact_vs_oracle"] = bool(np.array_equal(v.astype(bool), r[0][:m]))
            if "openssl" in c:
                valid = c["openssl"].pop("valid")
                want = np.ones(m, bool) if kind == "secp" else (cor[:m] == 0)
                c["openssl"]["agrees_with_gpu"] = bool(np.array_equal(valid.astype(bool), want))
            entry["cpu_baseline"] = dict(c, sample=f"the first {m} items", unit="sigs/s")
            log(f"  cpu: oracle {c['oracle']['sigs_per_s']:.0f} sigs/s on {threads} threads"
                + (f", OpenSSL {c['openssl']['sigs_per_s']:.0f}" if "openssl" in c else ""))
        res.append(entry)
        del it
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--cpu-sample", type=int, default=32768)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the Fuel hashing and verification-predicate timings (shorter profiled runs)")
    ap.add_argument("--blocks-per-cu", type=int, default=0)
    ap.add_argument("--tile-rounds", type=int, default=0)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-host", action="store_true",
                    help="skip the host-API runs (transfers included, serving): the persistent kernel is then "
                         "launched only as the timed steps launch it, for profiling")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the plain per-scheme configurations (secp256k1 1024, secp256r1 / Ed25519 65536)")
    ap.add_argument("--serve-seconds", type=float, default=1.0)
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="run a plain device copy of this size first (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); every signature is independent, "
              "so more GPUs would shard the batch by item ranges without any collective.", file=sys.stderr)
        sys.exit(2)

    import torch

    sigops = importlib.import_module("wgpu-sigops_amd")
    W = importlib.import_module("wgpu-sigops_amd.workload")
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    n = args.n

    if args.calib_copy_mib:
        # a plain copy kernel of known size, read by the FETCH_SIZE / WRITE_SIZE counter passes
        src = torch.empty(args.calib_copy_mib << 20, dtype=torch.uint8, device=dev).fill_(1)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    mad_per_s, fma_per_s = sigops.probe_valu(0)
    log(f"probe: v_mad_u64_u32 {mad_per_s / 1e12:.2f} T/s, v_fma_f64 {fma_per_s / 1e12:.2f} T/s")
    tuned_shape = not args.blocks_per_cu and not args.tile_rounds and sigops._tuned(0) is not None
    t_setup = time.perf_counter()
    ctx = sigops.Context(blocks_per_cu=args.blocks_per_cu, tile_rounds=args.tile_rounds, max_batch=max(n, 1 << 22))
    torch.cuda.synchronize()
    setup_ms = (time.perf_counter() - t_setup) * 1e3  # context, table upload, buffers (one time)
    t0 = time.time()
    mx = W.mixed(ctx, n, seed=args.seed)
    torch.cuda.synchronize()
    log(f"workload: {n} mixed items generated on the GPU in {time.time() - t0:.2f} s "
        f"({int(mx.msg_off_host[-1])} message bytes)")
    stream = torch.cuda.current_stream()
    s_ptr = stream.cuda_stream
    out = torch.zeros((n, 64), dtype=torch.uint8, device=dev)
    st = torch.zeros(n, dtype=torch.uint8, device=dev)

    def batched_step():
        ctx.batch_dev(mx.schemes.data_ptr(), mx.sigs.data_ptr(), mx.aux.data_ptr(), mx.msgs.data_ptr(),
                      mx.msg_off.data_ptr(), n, out.data_ptr(), st.data_ptr(), s_ptr)

    for _ in range(args.warmup):
        batched_step()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    ev[0].record(stream)
    for k in range(args.steps):
        batched_step()
        ev[k + 1].record(stream)
    torch.cuda.synchronize()
    step_ms = [ev[k].elapsed_time(ev[k + 1]) for k in range(args.steps)]
    total_ms = ev[0].elapsed_time(ev[-1])
    ms_per_step = total_ms / args.steps
    value = n / (ms_per_step / 1e3)
    log(f"batched: {ms_per_step:.2f} ms/step, {value / 1e6:.2f} M sigs/s (steps {min(step_ms):.2f}..{max(step_ms):.2f})")

    # the dominant kernel alone: eager launches with HIP events recorded by the library on the
    # launch stream around each stage; averaged over the same number of steps
    ctx.set_stage_timing(True)
    batched_step()
    stage_runs = []
    for _ in range(args.steps):
        batched_step()
        stage_runs.append(ctx.batch_stage_ms())
    ctx.set_stage_timing(False)
    stage_ms = {k: sum(r[k] for r in stage_runs) / len(stage_runs) for k in stage_runs[0]}
    log("stages (ms): " + ", ".join(f"{k} {v:.3f}" for k, v in stage_ms.items()))

    fuel = preds = None
    if not args.no_extras:
        fuel = fuel_hashing(ctx, n, mx, out, st, stream, s_ptr, args.steps, dev)
        preds = predicates(ctx, mx, stream, s_ptr, args.steps, dev)

    # correctness of the timed output by the round-trip property (all 2^20 items)
    out_h, st_h, cor = out.cpu().numpy(), st.cpu().numpy(), mx.corrupted.cpu().numpy()
    sch = mx.schemes_host
    roundtrip_ok = True
    for s in (0, 1):
        ii = mx.idx[s]
        pubs = mx.parts[s].pubs.cpu().numpy()
        ok = cor[ii] == 0
        roundtrip_ok &= bool(np.array_equal(out_h[ii][ok], pubs[ok]) and not st_h[ii][ok].any())
    roundtrip_ok &= bool(np.array_equal(st_h[mx.idx[2]] == 0, cor[mx.idx[2]] == 0))
    # lanes that skip the scalar multiplication: secp items the prologue rejects (r or s out of
    # range, r not an x-coordinate) and Ed25519 items with s >= l or an undecodable key. This is
    # the most a wavefront-ballot compaction of rejected items could give back.
    sec = sch < 2
    early = int(np.isin(st_h[sec], (1, 2, 3)).sum()) + int(np.isin(st_h[~sec], (1, 2)).sum())
    rejected = {"early_rejected_items": early, "fraction_of_lanes": round(early / n, 5),
                "corrupted_items": int(cor.sum())}

    # plain per-scheme kernels on the same mixed workload
    plain = None
    if not args.no_plain:
        d_idx = mx.d_idx
        parts = []
        for s, di in enumerate(d_idx):
            m = len(di)
            if s < 2:
                parts.append(dict(sigs=mx.sigs[di].contiguous(), aux=mx.aux[di].contiguous(),
                                  out=torch.zeros((m, 64), dtype=torch.uint8, device=dev),
                                  st=torch.zeros(m, dtype=torch.uint8, device=dev)))
            else:
                ed = mx.parts[2]
                parts.append(dict(sigs=mx.sigs[di].contiguous(), aux=mx.aux[di].contiguous(), msgs=ed.msgs,
                                  off=ed.msg_off, st=torch.zeros(m, dtype=torch.uint8, device=dev)))

        def plain_step(gather):
            if gather:  # the mixed layout must be split by scheme first, and merged back after
                for s, di in enumerate(d_idx):
                    parts[s]["sigs"].copy_(mx.sigs[di])
                    parts[s]["aux"].copy_(mx.aux[di])
            for s in (0, 1):
                p = parts[s]
                ctx.secp_recover_dev(s, p["sigs"].data_ptr(), p["aux"].data_ptr(), len(d_idx[s]), p["out"].data_ptr(),
                                     p["st"].data_ptr(), s_ptr)
            p = parts[2]
            ctx.ed25519_verify_dev(p["sigs"].data_ptr(), p["aux"].data_ptr(), p["msgs"].data_ptr(), p["off"].data_ptr(),
                                   len(d_idx[2]), p["st"].data_ptr(), s_ptr)
            if gather:
                for s in (0, 1):
                    out[d_idx[s]] = parts[s]["out"]
                    st[d_idx[s]] = parts[s]["st"]
                st[d_idx[2]] = parts[2]["st"]

        res = {}
        for gather in (False, True):
            for _ in range(max(1, args.warmup)):
                plain_step(gather)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(args.steps):
                plain_step(gather)
            e1.record(stream)
            torch.cuda.synchronize()
            res[gather] = e0.elapsed_time(e1) / args.steps
        plain = dict(ms_per_step_kernels_only=res[False], ms_per_step_with_split=res[True],
                     sigs_per_s_kernels_only=n / (res[False] / 1e3), sigs_per_s_with_split=n / (res[True] / 1e3),
                     batched_over_plain_kernels_only=res[False] / ms_per_step,
                     batched_over_plain_with_split=res[True] / ms_per_step)
        log(f"plain: {res[False]:.2f} ms kernels only, {res[True]:.2f} ms with split/merge; "
            f"batched/plain = {res[False] / ms_per_step:.2f}x / {res[True] / ms_per_step:.2f}x")

    # transfers included: host arrays in, host arrays out (pageable memory, as a caller has it)
    h_sch, h_sigs, h_aux = mx.schemes_host, mx.sigs.cpu().numpy(), mx.aux.cpu().numpy()
    h_msgs = mx.msgs.cpu().numpy()[: int(mx.msg_off_host[-1])]
    h_off = mx.msg_off_host.astype(np.uint64)
    reps = max(3, min(10, args.steps))

    def host_batch(streamed):
        ctx.set_streaming(streamed)
        ctx.batch(h_sch, h_sigs, h_aux, h_msgs, h_off)
        t0 = time.perf_counter()
        for _ in range(reps):
            o, t = ctx.batch(h_sch, h_sigs, h_aux, h_msgs, h_off)
        ms = (time.perf_counter() - t0) / reps * 1e3
        return ms, bool(np.array_equal(o, out_h) and np.array_equal(t, st_h))

    # copy-then-launch (transfers, then the graph, then transfers back), then streamed (the
    # default: the host packs chunks into pinned memory while the persistent kernel drains them)
    incl = None
    if not args.no_host:
        copy_ms, copy_ok = host_batch(False)
        incl_ms, incl_ok = host_batch(True)
        phases = ctx.batch_host_ms()
        log("streamed batch phases (ms): " + ", ".join(f"{k} {v}" for k, v in phases.items()))
        log(f"with transfers: streamed {incl_ms:.2f} ms per batch ({n / incl_ms * 1e3 / 1e6:.2f} M sigs/s), "
            f"copy-then-launch {copy_ms:.2f} ms; identical to the device path: {incl_ok}, {copy_ok}")
        incl = {"ms_per_batch": round(incl_ms, 3), "sigs_per_s": round(n / incl_ms * 1e3, 1),
                "identical_to_device_path": incl_ok, "method": "streamed host batch (sigops_batch)",
                "phases_ms_last_call": phases,
                "copy_then_launch": {"ms_per_batch": round(copy_ms, 3), "sigs_per_s": round(n / copy_ms * 1e3, 1),
                                     "identical_to_device_path": copy_ok}}

    # the same items through the three plain host entry points (one call per scheme, inputs split
    # by scheme beforehand): what a caller of the per-scheme API pays, transfers included
    plain_incl = None
    if not args.no_plain and not args.no_host:
        ii = [np.nonzero(h_sch == s)[0] for s in (0, 1, 2)]
        ed_off = np.concatenate([h_off[ii[2]], h_off[ii[2][-1] + 1: ii[2][-1] + 2]]) if len(ii[2]) else \
            np.zeros(1, np.uint64)
        split = [(h_sigs[ii[0]], h_aux[ii[0]]), (h_sigs[ii[1]], h_aux[ii[1]]),
                 (h_sigs[ii[2]], h_aux[ii[2]], h_msgs, ed_off)]

        def plain_host():
            a = ctx.secp256k1_recover(*split[0])
            b = ctx.secp256r1_recover(*split[1])
            c = ctx.ed25519_verify(*split[2])
            return a, b, c

        plain_host()
        t0 = time.perf_counter()
        for _ in range(reps):
            pa, pb, pc = plain_host()
        p_ms = (time.perf_counter() - t0) / reps * 1e3
        p_ok = bool(np.array_equal(pa[0], out_h[ii[0]]) and np.array_equal(pa[1], st_h[ii[0]]) and
                    np.array_equal(pb[0], out_h[ii[1]]) and np.array_equal(pb[1], st_h[ii[1]]) and
                    np.array_equal(pc[1], st_h[ii[2]]))
        plain_incl = {"ms_per_batch": round(p_ms, 3), "sigs_per_s": round(n / p_ms * 1e3, 1),
                      "batched_over_plain": round(p_ms / incl_ms, 4), "identical_to_batch": p_ok,
                      "calls": "secp256k1_recover, secp256r1_recover, ed25519_verify on the items of each scheme"}
        log(f"plain with transfers: {p_ms:.2f} ms (3 host calls); batched/plain with transfers = {p_ms / incl_ms:.2f}x")

    configs = None
    if not args.no_configs:
        configs = plain_configs(ctx, W, dev, stream, args.steps, min(16, os.cpu_count() or 1), not args.no_cpu,
                                args.cpu_sample)

    # served: Poisson arrivals of mixed requests (1..4096 items, log-uniform) at three offered loads,
    # replayed by the library's native client against both servers: slot batches (requests packed
    # into per-slot batches) and the device-drained ring (requests packed into tiles of one running
    # persistent launch)
    served = None
    if not args.no_serve and not args.no_host:
        served = []
        rng = np.random.default_rng(7)
        # serving moves every request over PCIe, so its capacity is the transfer-inclusive batch rate
        capacity = n / (incl_ms / 1e3)
        mean_req = (4096 - 1) / np.log(4096)
        for server, mk in (("ring", lambda: sigops.Server(mode=sigops.SERVE_RING)),
                           ("slots", lambda: sigops.Server(slots=3))):
            with mk() as srv:
                warm = True
                # the first pass warms up and is dropped: at the highest load, so that every buffer
                # reaches its largest size before anything is measured; the last pass submits every
                # request at once, so its sustained rate is the server's capacity
                for load in (0.8, 0.25, 0.5, 0.8, "saturated"):
                    rate_items = capacity if load == "saturated" else load * capacity
                    nreq = int(max(2000, (0.5 if warm else args.serve_seconds) * rate_items / mean_req))
                    sizes = np.floor(np.exp(rng.random(nreq) * np.log(4096))).astype(np.int64).clip(1, 4096)
                    starts = rng.integers(0, n - 4096, nreq)
                    # requests are contiguous slices of the pool; the replay takes one offsets array, so
                    # lay the requests out back to back in a gathered copy
                    idx = np.concatenate([np.arange(a, a + k) for a, k in zip(starts, sizes)])
                    req_off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
                    lens = (h_off[idx + 1] - h_off[idx]).astype(np.int64)
                    g_off = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
                    g_msgs = np.concatenate([h_msgs[h_off[i]:h_off[i + 1]] for i in idx[h_sch[idx] == 2]]) \
                        if (h_sch[idx] == 2).any() else np.zeros(0, np.uint8)
                    # messages of non-Ed25519 items are empty, so the Ed25519 messages concatenate in order
                    arrivals = np.cumsum(rng.exponential(1e9 * mean_req / rate_items, nreq)).astype(np.uint64)
                    if load == "saturated":
                        arrivals[:] = 0
                    r_sch, r_sigs, r_aux = h_sch[idx], h_sigs[idx], h_aux[idx]
                    lat, s_out, s_st = srv.replay(r_sch, r_sigs, r_aux, g_msgs, g_off, req_off, arrivals)
                    # on the replay's own clock: from the first arrival to the last completion
                    span_s = float((arrivals.astype(np.float64) + lat.astype(np.float64)).max()) / 1e9
                    same = bool(np.array_equal(s_out, out_h[idx]) and np.array_equal(s_st, st_h[idx]))
                    lat_ms = lat.astype(np.float64) / 1e6
                    if warm:
                        warm = False
                        continue
                    served.append({"server": server, "load": load, "capacity_basis": "batched incl. transfers",
                                   "offered_sigs_per_s": None if load == "saturated" else round(rate_items, 1),
                                   "latency_from": "scheduled arrival", "requests": nreq,
                                   "items": int(req_off[-1]), "p50_ms": round(float(np.percentile(lat_ms, 50)), 3),
                                   "p99_ms": round(float(np.percentile(lat_ms, 99)), 3),
                                   "sustained_sigs_per_s": round(float(req_off[-1]) / span_s, 1),
                                   "results_identical_to_batch": same})
                    log(f"served ({server}) load {load}: {nreq} requests, p50 {served[-1]['p50_ms']} ms, "
                        f"p99 {served[-1]['p99_ms']} ms, {served[-1]['sustained_sigs_per_s'] / 1e6:.2f} M sigs/s "
                        f"sustained, identical: {same}")

    # the served form's capacity (everything submitted at once) over the three plain per-scheme host
    # calls on the same items, transfers included on both sides
    served_vs_plain = None
    if served and plain_incl:
        served_vs_plain = {srv_name: round(e["sustained_sigs_per_s"] / plain_incl["sigs_per_s"], 3)
                           for e in served if e["load"] == "saturated" for srv_name in [e["server"]]}

    # roofline of the dominant kernel (the persistent verification kernel): its VALU use per launch
    # from the counter profile of this build and launch shape, over its duration measured here
    mac_sig = mixed_mac_per_sig(sch, mx.lens)
    kern_ms = stage_ms["verify_kernel"]
    shape = (f"waves={ctx.waves} blocks_per_cu={ctx.blocks_per_cu} tile_rounds={ctx.tile_rounds} "
             f"bucket_order={ctx.bucket_order} g_window={ctx.g_window} n={n} seed={args.seed}")
    prof = read_counter_profile(sigops.source_digest(), shape)
    roof = roofline(n, mac_sig, kern_ms, stage_ms, prof)
    log(f"roofline: frac {roof['frac']}, achieved {roof['achieved']} of {roof['peak']} T lane-op slots/s; "
        f"counter profile: {prof['status']}; vs_textbook {roof['vs_textbook']}")

    # CPU baseline: the oracle on a bounded sample of the same items (rank 0, N=1)
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(REPO, "tests"))
        import oracle_lib as O

        threads = min(16, os.cpu_count() or 1)
        rng = np.random.default_rng(1)
        sample = np.sort(rng.choice(n, min(args.cpu_sample, n), replace=False))
        t0 = time.perf_counter()
        exact = True
        for s in (0, 1):
            ss = sample[sch[sample] == s]
            o, t = O.recover_batch(s, h_sigs[ss], h_aux[ss], nthreads=threads)
            exact &= bool(np.array_equal(o, out_h[ss]) and np.array_equal(t, st_h[ss]))
        ss = sample[sch[sample] == 2]
        lens = h_off[ss + 1] - h_off[ss]
        sub_off = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
        sub_msgs = np.concatenate([h_msgs[h_off[i]:h_off[i + 1]] for i in ss]) if len(ss) else np.zeros(0, np.uint8)
        v = O.verify_batch(h_sigs[ss], h_aux[ss], sub_msgs, sub_off, nthreads=threads)
        exact &= bool(np.array_equal(v.astype(bool), st_h[ss] == 0))
        cpu_s = time.perf_counter() - t0
        # OpenSSL on the same sample: ECDSA verification under the generator's keys for the secp
        # items, Ed25519 verification for the rest
        import openssl_lib as S
        ossl = None
        if S.available() and S.baseline_lib() is not None:
            h_pubs = np.zeros((n, 64), np.uint8)
            for s_ in (0, 1):
                h_pubs[mx.idx[s_]] = mx.parts[s_].pubs.cpu().numpy()
            t_o, agree, recid_flips = 0.0, True, 0
            for s_ in (0, 1):
                ss_ = sample[sch[sample] == s_]
                t_, valid = S.time_ecdsa_verify(s_, h_sigs[ss_], h_aux[ss_], h_pubs[ss_], threads)
                t_o += t_
                # a corrupted secp item usually still recovers, to another key: OpenSSL rejects it
                # under the signer's key exactly when the GPU did not recover the signer's key;
                # except when the flipped bit is the recovery bit (bit 7 of s's first byte), which
                # verification ignores: then recovery with that bit restored must give the signer
                signer = (st_h[ss_] == 0) & (out_h[ss_] == h_pubs[ss_]).all(axis=1)
                odd = np.nonzero(valid.astype(bool) != signer)[0]
                if len(odd):
                    fixed = h_sigs[ss_][odd].copy()
                    fixed[:, 32] ^= 0x80
                    o_, t_ = O.recover_batch(s_, fixed, h_aux[ss_][odd])
                    explained = valid.astype(bool)[odd] & (t_ == 0) & (o_ == h_pubs[ss_][odd]).all(axis=1)
                    agree &= bool(explained.all())
                    recid_flips += int(explained.sum())
            t_, valid = S.time_ed25519_verify(h_sigs[ss], h_aux[ss], sub_msgs, sub_off, threads)
            t_o += t_
            agree &= bool(np.array_equal(valid.astype(bool), st_h[ss] == 0))
            ossl = {"value": round(len(sample) / t_o, 1), "unit": "sigs/s", "cores": threads,
                    "kind": "tuned library (OpenSSL libcrypto 3, C, threaded)",
                    "op": "ECDSA_do_verify under the signer's key (secp items; OpenSSL has no recovery) and "
                          "EVP_DigestVerify (Ed25519 items)", "agrees_with_gpu_status": agree,
                    "recovery_bit_flips_explained": recid_flips}
        # serial, one core, on the first 2048 items of the sample (the reference's CPU method)
        s1 = sample[:2048]
        t1 = time.perf_counter()
        for s in (0, 1):
            ss = s1[sch[s1] == s]
            O.recover_batch(s, h_sigs[ss], h_aux[ss], nthreads=1)
        ss = s1[sch[s1] == 2]
        lens1 = h_off[ss + 1] - h_off[ss]
        off1 = np.concatenate([[0], np.cumsum(lens1)]).astype(np.uint64)
        msgs1 = np.concatenate([h_msgs[h_off[i]:h_off[i + 1]] for i in ss]) if len(ss) else np.zeros(0, np.uint8)
        O.verify_batch(h_sigs[ss], h_aux[ss], msgs1, off1, nthreads=1)
        serial = len(s1) / (time.perf_counter() - t1)
        cpu = {"value": round(len(sample) / cpu_s, 1), "unit": "sigs/s", "cores": threads, "kind": "port",
               "sample": f"{len(sample)} items drawn uniformly from the same 2^20 mixed batch, C oracle "
                         f"(oracle/), {threads} threads; outputs also compared with the GPU's",
               "bit_exact_vs_gpu": exact, "serial_1_core": round(serial, 1),
               "host_cpus": os.cpu_count(), "openssl_same_sample": ossl}
        log(f"cpu oracle: {cpu['value']:.0f} sigs/s on {threads} threads, {serial:.0f} sigs/s on 1 core; "
            f"GPU sample bit-exact: {exact}")

    line = {
        "metric": "sigs/sec at 2^20 (batched vs plain, bit-exact) + p50/p99 latency served",
        "value": round(value, 1),
        "unit": "sigs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (seeded, signed on the GPU)",
        "config": {"workload": "mixed 2^20: 50% secp256k1 recover, 30% secp256r1 recover, 20% ed25519 verify "
                               "(messages log-uniform 0..4096 B), interleaved, 5% corrupted; one graph-captured "
                               "persistent launch per step",
                   "n": n, "seed": args.seed, "waves_per_simd": ctx.waves, "blocks_per_cu": ctx.blocks_per_cu,
                   "tile_rounds": ctx.tile_rounds, "bucket_order": ctx.bucket_order, "g_window": ctx.g_window,
                   "launch_shape_from": "autotune (tuned.json)" if tuned_shape else "arguments or library defaults"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "plain": plain,
        "incl_transfers": incl,
        "plain_incl_transfers": plain_incl,
        "configs": configs,
        "roundtrip_all_items_ok": roundtrip_ok,
        "rejected_lanes": rejected,
        "served": served,
        "served_over_plain_incl_transfers": served_vs_plain,
        "fuel_hashing": fuel,
        "predicates": preds,
        "probe": {"mad_per_s": mad_per_s, "fma64_per_s": fma_per_s},
        "setup_ms": round(setup_ms, 1),
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
