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
import subprocess
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


def warm_vs_cold(n_cases=3, n_cases_192=1, max_steps=200000, max_re=1000.0):
    """The first n_cases 128^3 cavities and the first n_cases_192 192^3 cavities
    with Re <= max_re of the held-out evaluation draw (64 cavities, a seed no
    training set uses, edges {128, 192}, Re log-uniform [400, 2500]) solved to
    residual 1e-6 from rest and from the trained surrogate's prediction.  The
    Re filter is fixed in advance: over the whole draw no 128^3 cavity above Re
    1117 reached a steady state in 200k steps, so higher-Re cases only time two
    capped runs.  Every selected case is reported; one whose cold run does not
    reach 1e-6 in max_steps is flagged and left out of the medians.  The whole
    64-case evaluation is `python -m cudalbm_amd.warmstart eval` (a run of many
    minutes)."""
    from cudalbm_amd import warmstart as W

    net, meta = W.load_model(MODEL_FILE)
    with net:
        draw = W.draw_cases(64, W.HELDOUT_SEED, (128, 192), re_range=(400.0, 2500.0))
        cases = ([c for c in draw if c[0] == 128 and c[1] <= max_re][:n_cases] +
                 [c for c in draw if c[0] == 192 and c[1] <= max_re][:n_cases_192])
        res = W.evaluate(net, cases, max_steps=max_steps, log=lambda m: print(m, file=sys.stderr, flush=True))
    by_edge = {f"summary_{e}": W.summarise([r for r in res if r["N"] == e]) for e in sorted({r["N"] for r in res})}
    return {"cases": [{"N": r["N"], "Re": round(r["Re"], 1), "U": round(r["U"], 4), "steady": r["steady"],
                       "cold_steps": r["cold"]["steps"], "cold_ms": round(r["cold"]["ms"], 1),
                       "cold_residual": r["cold"]["residual"], "warm_steps": r["warm"]["steps"],
                       "warm_ms": round(r["warm"]["ms"], 1), "warm_residual": r["warm"]["residual"],
                       "warm_kept": bool(r["warm"]["warm"]), "time_ratio": round(r["time_ratio"], 4),
                       "final_velocity_rel_l2": r["final_rel_l2"]} for r in res],
            **W.summarise(res), **by_edge,
            "model": {"file": os.path.basename(MODEL_FILE),
                      **{k: meta[k] for k in ("S", "cases", "epochs", "final_loss", "edges") if k in meta}}}


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


def _cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        with open("/proc/cpuinfo") as f:
            return next(l.split(":", 1)[1].strip() for l in f if l.startswith("model name"))
    except (OSError, StopIteration):
        return "unknown CPU"


def _oracle_mlups(n, steps, threads, Re=1000.0, u=0.05):
    """FP64 oracle on an n^3 D3Q19 cavity from rest: (MLUPS, seconds) of `steps`
    steps after one untimed step (first touch, thread start-up)."""
    from oracle import oracle as O

    om = O.omega(u, Re, n)
    mom = O.rest_state(n, n, n)
    O.run(19, n, n, n, om, u, mom, 1, threads)
    t0 = time.perf_counter()
    O.run(19, n, n, n, om, u, mom, steps, threads)
    dt = time.perf_counter() - t0
    return n ** 3 * steps / dt / 1e6, dt


def cpu_baseline(full_single=False):
    """The FP64 oracle (the reference has no CPU path; this is its restatement)
    on this host: D3Q19 cavities from rest at 64^3 x 1000 steps (Re 100, the
    oracle's reference run) and at the benchmark size 256^3 x 20 steps (Re
    1000), on all cores (up to 16 threads) and on one core (64^3 x 100 and
    256^3 x 2 steps, or the full step counts with full_single).  The oracle
    step carries no function-object work, so the figures bound the CPU from
    above on the headline workload.  value = the all-core 256^3 MLUPS."""
    threads = max(1, min(os.cpu_count() or 1, 16))
    m64, t64 = _oracle_mlups(64, 1000, threads, Re=100.0)
    m256, t256 = _oracle_mlups(256, 20, threads)
    s64, ts64 = _oracle_mlups(64, 1000 if full_single else 100, 1, Re=100.0)
    s256, ts256 = _oracle_mlups(256, 20 if full_single else 2, 1)
    ns = (1000, 20) if full_single else (100, 2)
    return {"value": round(m256, 2), "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": f"FP64 oracle, D3Q19 lid cavity from rest: 256^3 x 20 steps (Re 1000) {m256:.1f} MLUPS "
                      f"({t256:.1f} s) and 64^3 x 1000 steps (Re 100) {m64:.1f} MLUPS ({t64:.1f} s) on {threads} "
                      f"threads; one thread: 256^3 x {ns[1]} {s256:.2f} MLUPS ({ts256:.1f} s), 64^3 x {ns[0]} "
                      f"{s64:.2f} MLUPS ({ts64:.1f} s); host nproc {os.cpu_count()}, {_cpu_model()}",
            "all_cores": {"d3q19_256_x20": round(m256, 2), "d3q19_64_x1000": round(m64, 2)},
            "one_core": {f"d3q19_256_x{ns[1]}": round(s256, 3), f"d3q19_64_x{ns[0]}": round(s64, 3)}}


def sweep_full(n_cases=64, tol=1e-6, check=100, max_steps=200000, segment=20000):
    """Config 4 at its workload: the seeded 64-case sweep run by lbm_batch_run
    until each case's FP64 residual is below tol or it took max_steps steps.
    The run is split into segments of `segment` steps (cases stopped by a
    segment's cap rejoin the next one; converged cases stay out), and prints
    a progress line per segment.  Aggregate MLUPS = node updates of all cases /
    wall time of all segments."""
    import numpy as np

    from cudalbm_amd import Batch
    from cudalbm_amd.sweep import sweep_cases

    cases = sweep_cases(n_cases)
    nodes = np.array([c[0] * c[1] * c[2] for c in cases], dtype=np.int64)
    res = np.full(n_cases, np.inf)
    total_ms = 0.0
    with Batch(cases) as b:
        tiling = b.tiling
        cap = 0
        while cap < max_steps:
            cap = min(cap + segment, max_steps)
            steps, r, ms = b.run(tol, check, cap)
            total_ms += ms
            ran = np.isfinite(r)
            res[ran] = r[ran]
            again = (res >= tol) & (steps >= cap)
            print(f"# sweep segment to {cap} steps: {ms / 1e3:.1f} s, {int((res < tol).sum())} converged, "
                  f"{int(again.sum())} continue", file=sys.stderr, flush=True)
            if not again.any():
                break
            b.set_active(again.astype(np.uint8))
        steps = np.asarray(b.steps)
    upd = int((steps * nodes).sum())
    return {"cases": n_cases, "tol": tol, "check": check, "max_steps": max_steps, "tiling": list(tiling[:3]),
            "persistent": bool(tiling[3]), "wall_s": round(total_ms / 1e3, 2), "node_updates": upd,
            "mlups_batched": round(upd / (total_ms * 1e-3) / 1e6, 1),
            "converged": int((res < tol).sum()), "steps": steps.tolist(),
            "residual": [float(f"{v:.3e}") for v in res], "seed_cases": [list(c) for c in cases]}


def sweep_sequential(steps_per_case, tol=1e-6, check=100):
    """The same cases as one-after-another single-case solves to the same step
    counts (the batched and single-case steps are bit-identical, so each case
    stops at the same step): aggregate MLUPS over the whole sequence."""
    from cudalbm_amd import Solver
    from cudalbm_amd import native as NT
    from cudalbm_amd.sweep import sweep_cases

    cases = sweep_cases(len(steps_per_case))
    total_ms, upd = 0.0, 0
    for i, ((nx, ny, nz, Re, u), k) in enumerate(zip(cases, steps_per_case)):
        with Solver(nx, ny, nz, Re, u) as s:
            t0 = time.perf_counter()
            info = NT.solve_to_residual(s, Re, u, tol, check, int(k))
            dt = (time.perf_counter() - t0) * 1e3
        total_ms += dt
        upd += nx * ny * nz * int(info["steps"])
        if i % 8 == 7:
            print(f"# sequential: {i + 1} cases, {total_ms / 1e3:.1f} s", file=sys.stderr, flush=True)
    return {"wall_s": round(total_ms / 1e3, 2), "node_updates": upd,
            "mlups_sequential": round(upd / (total_ms * 1e-3) / 1e6, 1)}


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
                    help="one leg: q19fo|q19fp32|q19fp16|q19fp16fo|q27fp32|q27fp16|batch|warm|sweep-full|sweep-seq")
    ap.add_argument("--tile", default="0,0,0")
    ap.add_argument("--steps-json", default="", help="sweep-seq: JSON line of a sweep-full run (its per-case steps)")
    ap.add_argument("--cpu-full", action="store_true",
                    help="CPU baseline only, with the one-core figures at the full step counts")
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

    if args.cpu_full:
        print(json.dumps({"cpu_baseline": cpu_baseline(full_single=True)}), flush=True)
        return
    if args.only:
        if args.only == "sweep-full":
            r = sweep_full()
            print(f"# batched: {r['mlups_batched']} MLUPS over {r['wall_s']} s", file=sys.stderr, flush=True)
            print(json.dumps({"variant": "sweep-full", **r}), flush=True)
            return
        if args.only == "sweep-seq":
            with open(args.steps_json) as f:
                steps = json.load(f)["steps"]
            print(json.dumps({"variant": "sweep-seq", **sweep_sequential(steps)}), flush=True)
            return
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
                              "bytes_per_node": r["bytes_per_node"], "tiling": r["tiling"],
                              "traffic": measured_traffic(
                                  f"step_kernel<{Q},{r['tiling'][0]},{r['tiling'][1]},"
                                  f"{'_Float16' if st == 'fp16' else 'float'}{',fo' if fo else ''}>@{nv}^3")}
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
                     "traffic": measured_traffic(f"step_kernel<19,{head['tiling'][0]},{head['tiling'][1]},float,fo>"
                                                 f"@{args.n}^3"),
                     "bytes_per_node": head["bytes_per_node"]},
        "cpu_baseline": cpu,
        "variants": variants,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
