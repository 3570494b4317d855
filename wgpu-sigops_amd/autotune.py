"""Autotuner for the mixed-batch launch shape.

Sweeps the runtime knobs of the persistent batch kernel on the real workload (the seeded mixed
2^20 batch, signed on the GPU) and keeps the fastest:
  * blocks_per_cu: workgroups of 256 threads per CU (the kernel is compiled for 2 waves per SIMD,
    so 2 fills the CU; fewer leaves it idle, more only queues workgroups behind resident ones),
  * tile_rounds:   64-item rounds per tile popped from the work counter (1..4: larger tiles mean
    fewer atomics but a coarser tail at the end of the batch).
Each candidate gets its own context and is timed over several graph launches with HIP events on
the launch stream. The result is written to ``tuned.json`` beside this file, keyed by the device
(architecture, name and CU count); ``Context()`` applies it when the caller does not choose.

Fixed by design rather than tuned: 29-bit limbs (one v_mad_u64_u32 per partial product with an
even-aligned 64-bit accumulator), windows of 4 bits for per-lane tables and 8 bits for the LDS
generator tables (the LDS tables of both curves and Ed25519 fit two workgroups per CU).

Run on a GPU:  python wgpu-sigops_amd/autotune.py [--n 1048576] [--steps 5]
"""
import argparse
import importlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
TUNED_PATH = os.path.join(HERE, "tuned.json")
CANDIDATES = ((1, 1), (2, 1), (2, 2), (2, 4), (3, 1), (4, 1))


def device_key(device=0):
    import torch

    p = torch.cuda.get_device_properties(device)
    return f"{getattr(p, 'gcnArchName', '?')}|{p.name}|{p.multi_processor_count}"


def load(device=0, path=TUNED_PATH):
    """the tuned launch shape for this device, or None"""
    try:
        with open(path) as f:
            data = json.load(f)
        return data.get(device_key(device))
    except (OSError, ValueError, KeyError):
        return None


def tune(n=1 << 20, seed=2024, steps=5, candidates=CANDIDATES, write=True, path=TUNED_PATH, log=print):
    import torch

    pkg = sys.modules.get("wgpu-sigops_amd") or importlib.import_module("wgpu-sigops_amd")
    W = importlib.import_module("wgpu-sigops_amd.workload")
    dev = torch.device("cuda:0")
    base = pkg.Context(blocks_per_cu=2, tile_rounds=1, max_batch=n)
    mx = W.mixed(base, n, seed=seed)
    out = torch.zeros((n, 64), dtype=torch.uint8, device=dev)
    st = torch.zeros(n, dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream()
    ref = None
    results = []
    for bpc, tr in candidates:
        ctx = pkg.Context(blocks_per_cu=bpc, tile_rounds=tr, max_batch=n)

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
        ms = a.elapsed_time(b) / steps
        got = (out.cpu(), st.cpu())
        same = ref is None or (torch.equal(got[0], ref[0]) and torch.equal(got[1], ref[1]))
        ref = ref or got
        if not same:
            raise RuntimeError(f"launch shape blocks_per_cu={bpc} tile_rounds={tr} changed the results")
        results.append({"blocks_per_cu": bpc, "tile_rounds": tr, "ms_per_batch": round(ms, 3)})
        log(f"autotune: blocks_per_cu {bpc} tile_rounds {tr}: {ms:.3f} ms per {n}-item batch")
        ctx.close()
    base.close()
    best = min(results, key=lambda r: r["ms_per_batch"])
    entry = {"blocks_per_cu": best["blocks_per_cu"], "tile_rounds": best["tile_rounds"],
             "workload": f"mixed {n} items, seed {seed}", "results": results}
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
    ap.add_argument("--out", default=TUNED_PATH)
    args = ap.parse_args()
    sys.path.insert(0, os.path.dirname(HERE))
    entry = tune(n=args.n, steps=args.steps, path=args.out)
    print(json.dumps({k: v for k, v in entry.items() if k != "results"}))


if __name__ == "__main__":
    main()
