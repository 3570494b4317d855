#!/usr/bin/env python3
"""Benchmark of the MI355X signature engine (one GPU).

Workload (the headline configuration): a mixed batch of 2^20 signatures -- 50 % secp256k1
recoveries, 30 % secp256r1 recoveries, 20 % Ed25519 verifications with message lengths
log-uniform on 0..4096 bytes, schemes randomly interleaved, 5 % corrupted -- generated on the GPU
by the engine's seeded signer and resident in HBM when the timed region starts.

One step = one batched, graph-captured pass over the whole batch (device-side bucketing + one
persistent kernel + careful pass). `value` is signatures per second over the K timed steps.
Also reported, each on the same mixed workload:
  * plain: the three per-scheme kernels (gather by scheme, one launch per scheme, scatter back),
    and the batched/plain ratio;
  * transfers: the batched path through the host API (H2D + graph + D2H), the reference's
    benchmark method (src/benchmarks/secp256k1_ecdsa.rs:126-138);
  * roofline: algorithmic u32 multiply-accumulates per signature (the canonical textbook schedule:
    fixed-window Fermat inversions, 252 doublings, dense 2 x 8^2 MACs per 256-bit product) divided
    by the launch time, against the measured v_mad_u64_u32 issue rate;
  * cpu_baseline: the CPU oracle timed on a bounded sample of the same items on this host's cores,
    also used to check the GPU outputs of that sample bit for bit.
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

# canonical multiply-accumulates per signature (u32 x u32 -> u64) of the textbook schedule above
MAC_PER_SIG = {"secp256k1": 595072, "secp256r1": 627456, "ed25519_per_block": 24 * 128, "ed25519_base": 3850 * 128}
LIMBS_MAC = 128  # MACs per canonical 256-bit modular multiplication


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


TRAFFIC_PROFILE = os.path.join(REPO, "profiles", "hbm_traffic_and_valu_k_batch_run_mixed_2p20.md")


def read_profile():
    """From the committed summary of the counter passes over this workload: HBM bytes per launch of
    the persistent kernel (FETCH_SIZE / WRITE_SIZE, each scaled by a copy of known size) and the
    fraction of its active cycles that issue VALU instructions (SQ_INSTS_VALU x 4 / GRBM_GUI_ACTIVE)."""
    out = {"traffic": None, "valu_issue_frac": None, "source": None}
    try:
        for line in open(TRAFFIC_PROFILE):
            if line.startswith("traffic_bytes_per_launch:"):
                out["traffic"] = int(line.split(":")[1])
            if line.startswith("valu_issue_frac:"):
                out["valu_issue_frac"] = float(line.split(":")[1])
        out["source"] = os.path.relpath(TRAFFIC_PROFILE, REPO)
    except OSError:
        pass
    return out


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
    ctx.batch(h_sch, h_sigs, h_aux, h_msgs, h_off)
    t0 = time.perf_counter()
    reps = max(1, min(3, args.steps))
    for _ in range(reps):
        h_out, h_st = ctx.batch(h_sch, h_sigs, h_aux, h_msgs, h_off)
    incl_ms = (time.perf_counter() - t0) / reps * 1e3
    incl_ok = bool(np.array_equal(h_out, out_h) and np.array_equal(h_st, st_h))
    log(f"with transfers: {incl_ms:.2f} ms per batch, {n / incl_ms * 1e3 / 1e6:.2f} M sigs/s")

    # served: Poisson arrivals of mixed requests (1..4096 items, log-uniform) at three offered loads,
    # replayed by the library's native client against the continuous-batching server
    served = None
    if not args.no_serve:
        served = []
        rng = np.random.default_rng(7)
        # serving moves every request over PCIe, so its capacity is the transfer-inclusive batch rate
        capacity = n / (incl_ms / 1e3)
        with sigops.Server(slots=3) as srv:
            warm = True
            # the first pass warms up and is dropped: at the highest load, so that every slot's buffers
            # reach their largest size (and their graphs are captured) before anything is measured
            # the last pass submits every request at once: its sustained rate is the server's capacity
            for load in (0.8, 0.25, 0.5, 0.8, "saturated"):
                rate_items = capacity if load == "saturated" else load * capacity
                mean_req = (4096 - 1) / np.log(4096)
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
                served.append({"load": load, "capacity_basis": "batched incl. transfers",
                               "offered_sigs_per_s": None if load == "saturated" else round(rate_items, 1),
                               "latency_from": "scheduled arrival", "requests": nreq,
                               "items": int(req_off[-1]), "p50_ms": round(float(np.percentile(lat_ms, 50)), 3),
                               "p99_ms": round(float(np.percentile(lat_ms, 99)), 3),
                               "sustained_sigs_per_s": round(float(req_off[-1]) / span_s, 1),
                               "results_identical_to_batch": same})
                log(f"served load {load}: {nreq} requests, p50 {served[-1]['p50_ms']} ms, p99 {served[-1]['p99_ms']} ms, "
                    f"{served[-1]['sustained_sigs_per_s'] / 1e6:.2f} M sigs/s sustained, identical: {same}")
            served_batches = srv.stats()

    # roofline of the dominant kernel (the persistent verification kernel): algorithmic MACs of the
    # launch over its own event-timed duration; traffic from the calibrated counter passes
    mac_sig = mixed_mac_per_sig(sch, mx.lens)
    kern_ms = stage_ms["verify_kernel"]
    achieved = n * mac_sig / (kern_ms / 1e3) / 1e12
    peak = mad_per_s / 1e12
    prof = read_profile()
    roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": round(peak, 3),
                "unit": "TMAC/s (u32 x u32 -> u64 multiply-accumulate)", "frac": round(achieved / peak, 4),
                "work": "canonical textbook schedule (fixed-window Fermat, 252 doublings, dense Montgomery); "
                        "the kernel executes fewer MACs than that, so frac can exceed 1",
                "valu_issue_frac": prof["valu_issue_frac"],
                "traffic": prof["traffic"], "traffic_unit": "B per launch", "traffic_source": prof["source"],
                # the same traffic over this run's kernel time, against MI355X's 8 TB/s HBM peak
                "hbm_GBps": round(prof["traffic"] / (kern_ms / 1e3) / 1e9, 1) if prof["traffic"] else None,
                "hbm_frac": round(prof["traffic"] / (kern_ms / 1e3) / 8e12, 4) if prof["traffic"] else None,
                "kernel": "k_batch_run", "kernel_ms": round(kern_ms, 3), "stage_ms": {k: round(v, 3) for k, v in stage_ms.items()},
                "mac_per_sig": round(mac_sig), "peak_source": "sigops_probe_valu (v_mad_u64_u32)"}

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
               "host_cpus": os.cpu_count()}
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
                   "n": n, "seed": args.seed, "blocks_per_cu": ctx.blocks_per_cu or 2,
                   "tile_rounds": ctx.tile_rounds or 1,
                   "launch_shape_from": "autotune (tuned.json)" if tuned_shape else "arguments or library defaults"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "plain": plain,
        "incl_transfers": {"ms_per_batch": round(incl_ms, 3), "sigs_per_s": round(n / incl_ms * 1e3, 1),
                           "identical_to_device_path": incl_ok},
        "roundtrip_all_items_ok": roundtrip_ok,
        "served": served,
        "fuel_hashing": fuel,
        "predicates": preds,
        "probe": {"mad_per_s": mad_per_s, "fma64_per_s": fma_per_s},
        "setup_ms": round(setup_ms, 1),
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
