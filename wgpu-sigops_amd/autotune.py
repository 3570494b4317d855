"""Autotuner of the mixed batch: built-in kernel variants and launch knobs, measured on the real workload.

Candidates (each a Context setting; the library holds every variant's code):
  * waves:        the persistent kernel built for 2 waves per SIMD (256 registers per lane, some
                  values spilled to scratch memory) or for 1 (512 registers, accumulation registers
                  included, no scratch spills, half the waves to hide latencies);
  * tile_rounds:  64-item rounds per tile popped from the work counter (larger tiles mean fewer
                  atomics but a coarser tail at the end of the batch);
  * bucket_order: the processing order of the buckets (0: secp256r1, secp256k1, Ed25519 by
                  decreasing block count; 1: secp256r1, Ed25519, secp256k1; 2: secp256k1 first).
  * g_window:     the fixed-base window of the secp streams, 8 bits (128 odd multiples of G in LDS)
                  or 9 (256, 1.75x the LDS; 4 fewer G additions per secp256k1 item and 3 per P-256
                  item, on a ladder that takes one doubling per loop iteration).
blocks_per_cu follows waves (one 256-thread workgroup per SIMD wave): more workgroups than fit
only queue behind the resident ones.

Measurement: the seeded mixed 2^20 batch, signed on the GPU, in HBM. The candidates are timed in
turn, `rounds` times over (so a slow spell of the box hits every candidate), each over `steps`
graph launches bracketed by HIP events on the launch stream; a candidate's time is its median
over the rounds, and its noise the spread of its rounds. Every candidate's outputs must equal the
first one's. tuned.json (beside this file, keyed by the device) records the winner only when it
beats the default by more than the larger of 0.5 % and both candidates' spread; otherwise it
records the default.

Fixed by design rather than tuned: 29-bit limbs (one v_mad_u64_u32 per partial product with an
even-aligned 64-bit accumulator), windows of 4 bits for per-lane tables, 8 bits for Ed25519's LDS
table.

Run on a GPU:  python wgpu-sigops_amd/autotune.py [--n 1048576] [--steps 5] [--rounds 3]
"""
import argparse
import importlib
import json
import os
import statistics
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
TUNED_PATH = os.path.join(HERE, "tuned.json")
# the library's defaults (a Context with no settings)
DEFAULT = {"waves": 2, "tile_rounds": 1, "bucket_order": 0, "g_window": 9}
# each knob changed alone, and the 8-bit windows with each of them (the one-wave kernel is built
# with 8-bit windows only)
_W8 = dict(DEFAULT, g_window=8)
CANDIDATES = (DEFAULT, dict(DEFAULT, waves=1, g_window=8), dict(DEFAULT, tile_rounds=2), dict(DEFAULT, bucket_order=1),
              dict(DEFAULT, bucket_order=2), _W8, dict(_W8, tile_rounds=2), dict(_W8, bucket_order=1),
              dict(_W8, bucket_order=2))


def device_key(device=0):
    """the ISA target (with its feature flags) and the CU count; not the marketing name, which the
    runtime reads from a table that some hosts lack"""
    import torch

    p = torch.cuda.get_device_properties(device)
    return f"{getattr(p, 'gcnArchName', '?')}|{p.multi_processor_count}"


def load(device=0, path=TUNED_PATH):
    """the tuned variant for this device, or None"""
    try:
        with open(path) as f:
            data = json.load(f)
        return data.get(device_key(device))
    except (OSError, ValueError, KeyError):
        return None


def _name(c):
    return (f"waves {c['waves']} tile_rounds {c['tile_rounds']} bucket_order {c['bucket_order']} "
            f"g_window {c.get('g_window', 8)}")


def tune(n=1 << 20, seed=2024, steps=5, rounds=3, candidates=CANDIDATES, write=True, path=TUNED_PATH, log=print):
    import torch

    pkg = sys.modules.get("wgpu-sigops_amd") or importlib.import_module("wgpu-sigops_amd")
    W = importlib.import_module("wgpu-sigops_amd.workload")
    dev = torch.device("cuda:0")
    base = pkg.Context(tuned=False, max_batch=n)
    mx = W.mixed(base, n, seed=seed)
    out = torch.zeros((n, 64), dtype=torch.uint8, device=dev)
    st = torch.zeros(n, dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream()
    ctxs = [pkg.Context(tuned=False, max_batch=n, **c) for c in candidates]
    times = [[] for _ in candidates]
    ref = None
    for r in range(rounds):
        for k, (c, ctx) in enumerate(zip(candidates, ctxs)):
            def step():
                ctx.batch_dev(mx.schemes.data_ptr(), mx.sigs.data_ptr(), mx.aux.data_ptr(), mx.msgs.data_ptr(),
                              mx.msg_off.data_ptr(), n, out.data_ptr(), st.data_ptr(), stream.cuda_stream)

            step()
            torch.cuda.synchronize()
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            for _ in range(steps):
                step()
            b.record(stream)
            torch.cuda.synchronize()
            times[k].append(a.elapsed_time(b) / steps)
            got = (out.cpu(), st.cpu())
            if ref is None:
                ref = got
            elif not (torch.equal(got[0], ref[0]) and torch.equal(got[1], ref[1])):
                raise RuntimeError(f"variant {_name(c)} changed the results")
            log(f"autotune round {r}: {_name(c)}: {times[k][-1]:.3f} ms per {n}-item batch")
    for ctx in ctxs:
        ctx.close()
    base.close()
    results = [dict(c, ms_median=round(statistics.median(t), 3), ms_min=round(min(t), 3), ms_max=round(max(t), 3))
               for c, t in zip(candidates, times)]
    d = next(i for i, c in enumerate(candidates) if c == DEFAULT) if DEFAULT in candidates else 0
    best = min(range(len(results)), key=lambda i: results[i]["ms_median"])
    spread = max(results[best]["ms_max"] - results[best]["ms_min"], results[d]["ms_max"] - results[d]["ms_min"])
    margin = max(0.005 * results[d]["ms_median"], spread)
    win = best != d and results[d]["ms_median"] - results[best]["ms_median"] > margin
    choice = candidates[best] if win else candidates[d]
    entry = dict(choice, workload=f"mixed {n} items, seed {seed}, {rounds} alternated rounds of {steps} launches",
                 decided_by=("beats the default by more than the noise" if win else
                             "no variant beats the default by more than the noise"), results=results)
    log(f"autotune: {_name(choice)} ({entry['decided_by']})")
    if write:
        data = {}
        try:
            with open(path) as f:
                data = json.load(f)
        except (OSError, ValueError):
            pass
        data[device_key()] = entry
        with open(path, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
            f.write("\n")
    return entry


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--rounds", type=int, default=3)
    ap.add_argument("--out", default=TUNED_PATH)
    args = ap.parse_args()
    sys.path.insert(0, os.path.dirname(HERE))
    entry = tune(n=args.n, steps=args.steps, rounds=args.rounds, path=args.out)
    print(json.dumps({k: v for k, v in entry.items() if k != "results"}))


if __name__ == "__main__":
    main()
