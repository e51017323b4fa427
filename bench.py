"""Benchmark: MLUPS and HBM-roofline fraction of the fused D3Q19 step, 256^3.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload: lid-driven cavity, D3Q19, FP32 storage and arithmetic, 256^3 nodes,
Re 1000, lid speed 0.05, started from rest, with the function objects k, kMean,
S and SMean on.  One "step" is one full LBM update of every node together with
that step's function-object work (one launch of the fused step kernel, which
updates the FP64 time means and advances the device step counter they divide
by); the timed region holds K steps on the solver's
HIP stream, bracketed by hipEvents, with all state resident in HBM.  Prints ONE
JSON line.

Roofline basis: algorithmic bytes per node update = the 10 stored FP32 moments
read once and written once (80 B) plus the 7 FP64 time means (kMean, 6 SMean
components) read once and written once (112 B): 192 B.  The instantaneous k
and S depend only on the moments of the step they are read at, and are formed
from them when read.  Achieved GB/s = bytes per step / average step duration.
Variants report the plain step (80 B/node FP32, 40 B/node FP16), D3Q27 and the
batched multi-case sweep.  The CPU baseline is the FP64 oracle (oracle/, a
scalar C restatement of the reference step) timed on this host on a bounded
sample of the same case.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0  # MI355X published HBM3E bandwidth
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic.json")
# warm-start network: the 64^3-grid model where present (closer to the 128^3
# meshes of the warm-vs-cold leg), else the 32^3 one
MODEL_FILE = next((p for p in (os.path.join(ROOT, "cudalbm_amd", "models", f) for f in ("unet64.npz", "unet32.npz"))
                   if os.path.exists(p)), os.path.join(ROOT, "cudalbm_amd", "models", "unet32.npz"))


def measured_traffic(key):
    """HBM bytes per launch of a kernel/workload from the committed PMC summary
    (FETCH_SIZE and WRITE_SIZE runs of rocprofv3, calibrated on a device copy)."""
    try:
        with open(TRAFFIC_FILE) as f:
            return json.load(f)[key]["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def warm_vs_cold(n_draw=3, edge=128, max_steps=200000, re_steady=1500.0):
    """Held-out cavities at edge^3 (the first n_draw of the evaluation draw, Re
    log-uniform [400, 2500]) solved to residual 1e-6 from rest and from the
    trained surrogate's prediction.  Only the cases with Re <= re_steady are
    run: above it the cavity is not steady at this resolution (Re 1941 of the
    draw keeps r ~ 1e-3 through 200k steps from either start)."""
    import numpy as np

    from cudalbm_amd import warmstart as W

    net, meta = W.load_model(MODEL_FILE)
    with net:
        cases = [c for c in W.draw_cases(n_draw, W.EVAL_SEED, (edge,), re_range=(400.0, 2500.0)) if c[1] <= re_steady]
        res = W.evaluate(net, cases, max_steps=max_steps, log=lambda m: print(m, file=sys.stderr, flush=True))
    return {"cases": [{"N": r["N"], "Re": round(r["Re"], 1), "U": round(r["U"], 4),
                       "cold_steps": r["cold"]["steps"], "cold_ms": round(r["cold"]["ms"], 1),
                       "cold_residual": r["cold"]["residual"], "warm_steps": r["warm"]["steps"],
                       "warm_ms": round(r["warm"]["ms"], 1), "warm_residual": r["warm"]["residual"],
                       "warm_kept": bool(r["warm"]["warm"]), "time_ratio": round(r["time_ratio"], 4),
                       "final_velocity_rel_l2": r["final_rel_l2"]} for r in res],
            "median_time_ratio": float(np.median([r["time_ratio"] for r in res])),
            "max_final_rel_l2": float(max(r["final_rel_l2"] for r in res)),
            "model": {"file": os.path.basename(MODEL_FILE),
                      **{k: meta[k] for k in ("S", "cases", "epochs", "final_loss") if k in meta}}}


def measure(n, Q, storage, steps, warmup, fo_mask=0, Re=1000.0, u=0.05, tile=(0, 0, 0)):
    from cudalbm_amd import Solver

    with Solver(n, n, n, Re, u, velocity_set=Q, storage=storage, tile=tile) as s:
        if fo_mask:
            s.fo_enable(fo_mask)
        if warmup:
            s.step(warmup)
        s.synchronize()
        ms = s.step_timed(steps)
        tiling = s.tiling
        mom = s.get_moments() if steps + warmup >= 1 else None
    nodes = n ** 3
    mlups = nodes * steps / (ms * 1e-3) / 1e6
    bpn = (40 if storage == "fp16" else 80) + (112 if fo_mask & 10 else 0)
    gbs = nodes * bpn / (ms * 1e-3 / steps) / 1e9
    finite = bool(mom is not None and all(abs(float(m.max())) < 1e3 for m in mom))
    return {"mlups": mlups, "ms_per_step": ms / steps, "achieved_gbs": gbs, "bytes_per_node": bpn,
            "tiling": list(tiling), "finite": finite}


def batch_sweep(steps, n_cases=64):
    """Aggregate MLUPS of the seeded uneven sweep (D3Q19 FP32 cavities, N in
    {48..160}, Re log-uniform [100, 3200], lid in [0.02, 0.08]): every case in
    one launch per step (autotuned tiling) versus the same cases one after
    another as single-case solvers with their default tiling."""
    from cudalbm_amd import Batch, Solver
    from cudalbm_amd.sweep import nodes, sweep_cases

    cases = sweep_cases(n_cases)
    total = nodes(cases)
    t0 = time.perf_counter()
    with Batch(cases) as b:
        tune_s = time.perf_counter() - t0
        b.step(4)
        b.synchronize()
        ms_b = b.step_timed(steps)
        tiling = b.tiling
        items = b.work_items
    ms_s = 0.0
    for nx, ny, nz, Re, u in cases:
        with Solver(nx, ny, nz, Re, u) as s:
            s.step(4)
            s.synchronize()
            ms_s += s.step_timed(steps)
    mb = total * steps / (ms_b * 1e-3) / 1e6
    ms = total * steps / (ms_s * 1e-3) / 1e6
    return {"cases": n_cases, "nodes": total, "steps": steps, "mlups_batched": round(mb, 1),
            "mlups_sequential": round(ms, 1), "speedup": round(mb / ms, 3),
            "ms_per_step_batched": round(ms_b / steps, 4), "ms_per_step_sequential": round(ms_s / steps, 4),
            "tiling": list(tiling[:3]), "persistent": bool(tiling[3]), "work_items": items,
            "autotune_s": round(tune_s, 2),
            "roofline_frac_batched": round(total * 80 / (ms_b * 1e-3 / steps) / 1e9 / PEAK_HBM_GBS, 4)}


def cpu_baseline(target_s=25.0):
    """FP64 oracle, D3Q19 96^3 cavity (Re 1000, lid 0.05) from rest, on this
    host's cores; the step count is sized to about target_s seconds.  The
    oracle step carries no function-object work, so the CPU figure is an upper
    bound for the CPU on the headline workload."""
    from oracle import oracle as O

    n = 96
    threads = max(1, min(os.cpu_count() or 1, 16))
    om = O.omega(0.05, 1000.0, n)
    mom = O.rest_state(n, n, n)
    O.run(19, n, n, n, om, 0.05, mom, 2, threads)  # first touch and thread start-up, untimed
    t0 = time.perf_counter()
    O.run(19, n, n, n, om, 0.05, mom, 10, threads)
    per = (time.perf_counter() - t0) / 10
    steps = max(2, int(target_s / max(per, 1e-6)))
    t0 = time.perf_counter()
    O.run(19, n, n, n, om, 0.05, mom, steps, threads)
    dt = time.perf_counter() - t0
    # one thread, a few steps, for the single-core figure
    t1 = time.perf_counter()
    O.run(19, n, n, n, om, 0.05, mom, 3, 1)
    single = n ** 3 * 3 / (time.perf_counter() - t1) / 1e6
    model = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as f:
            model = next(l.split(":", 1)[1].strip() for l in f if l.startswith("model name"))
    except (OSError, StopIteration):
        pass
    return {"value": n ** 3 * steps / dt / 1e6, "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": f"FP64 oracle, D3Q19 {n}^3 lid cavity Re 1000 u 0.05 from rest, {steps} steps, "
                      f"{dt:.1f} s on {threads} threads of {os.cpu_count()} ({model}); "
                      f"one thread: {single:.1f} MLUPS"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-variants", action="store_true", help="only the headline D3Q19 FP32 step")
    ap.add_argument("--no-warm", action="store_true", help="skip the warm-vs-cold leg")
    ap.add_argument("--only", default="",
                    help="profile one variant: q19fo|q19fp32|q19fp16|q19fp16fo|q27fp32|q27fp16|batch|warm")
    ap.add_argument("--tile", default="0,0,0")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run only a plain 256 MiB device-to-device copy (x4), to calibrate HBM byte counters")
    args = ap.parse_args()
    if args.copy_probe:
        import torch

        src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(4):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read_per_copy": src.numel() * 4, "copies": 4}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: the solver has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)
    tile = tuple(int(v) for v in args.tile.split(","))
    FO_ALL = 15  # k | kMean | S | SMean

    if args.only:
        if args.only == "batch":
            print(json.dumps({"variant": "batch", **batch_sweep(args.steps)}), flush=True)
            return
        if args.only == "warm":
            print(json.dumps({"variant": "warm", **warm_vs_cold()}), flush=True)
            return
        spec = {"q19fo": (19, "fp32", FO_ALL), "q19fp32": (19, "fp32", 0), "q19fp16": (19, "fp16", 0),
                "q19fp16fo": (19, "fp16", FO_ALL),
                "q27fp32": (27, "fp32", 0), "q27fp16": (27, "fp16", 0)}[args.only]
        r = measure(args.n, spec[0], spec[1], args.steps, args.warmup, fo_mask=spec[2], tile=tile)
        print(json.dumps({"variant": args.only, **r}), flush=True)
        return

    head = measure(args.n, 19, "fp32", args.steps, args.warmup, fo_mask=FO_ALL, tile=tile)
    print(f"# D3Q19 fp32 {args.n}^3 + k/kMean/S/SMean: {head['mlups']:.0f} MLUPS", file=sys.stderr, flush=True)
    variants = {}
    if not args.no_variants:
        k = max(50, args.steps // 4)
        n = args.n
        for name, (nv, Q, st, fo) in {"d3q19_fp32_step_only": (n, 19, "fp32", 0),
                                      "d3q19_fp16_step_only": (n, 19, "fp16", 0),
                                      "d3q19_fp16_fo": (n, 19, "fp16", FO_ALL),
                                      "d3q27_fp32_step_only": (n, 27, "fp32", 0),
                                      "d3q27_fp16_step_only": (n, 27, "fp16", 0),
                                      # the larger D3Q27 mesh (2.3 GB of moments)
                                      "d3q27_fp32_step_only_384": (384, 27, "fp32", 0),
                                      "d3q27_fp16_step_only_384": (384, 27, "fp16", 0)}.items():
            r = measure(nv, Q, st, k, 10, fo_mask=fo)
            variants[name] = {"n": nv, "mlups": round(r["mlups"], 1), "ms_per_step": round(r["ms_per_step"], 4),
                              "roofline_frac": round(r["achieved_gbs"] / PEAK_HBM_GBS, 4),
                              "bytes_per_node": r["bytes_per_node"], "tiling": r["tiling"]}
            print(f"# {name}: {r['mlups']:.0f} MLUPS", file=sys.stderr, flush=True)
        variants["batch_sweep64_d3q19_fp32"] = batch_sweep(100)
        print(f"# batch sweep: {variants['batch_sweep64_d3q19_fp32']}", file=sys.stderr, flush=True)
        if os.path.exists(MODEL_FILE) and not args.no_warm:
            variants["warm_vs_cold_time_to_residual_1e-6"] = warm_vs_cold()
    cpu = None if args.no_cpu else cpu_baseline()
    out = {
        "metric": "MLUPS (D3Q19 256^3 lid-driven cavity, FP32, k/kMean/S/SMean on)",
        "value": round(head["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: lid-driven cavity from rest (no dataset)",
        "config": {"workload": f"lid-driven cavity D3Q19 {args.n}^3 FP32, Re 1000, lid 0.05, rest start, "
                               "function objects k, kMean, S, SMean on",
                   "velocity_set": "D3Q19", "nodes": args.n ** 3, "tiling": head["tiling"]},
        "roofline": {"bound": "hbm", "achieved": round(head["achieved_gbs"], 1), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(head["achieved_gbs"] / PEAK_HBM_GBS, 4),
                     "traffic": measured_traffic(f"step_kernel<19,32,16,float,fo>@{args.n}^3"),
                     "bytes_per_node": head["bytes_per_node"]},
        "cpu_baseline": cpu,
        "variants": variants,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
