"""Learned warm start: data generation, training and evaluation.

    python -m cudalbm_amd.warmstart train --out cudalbm_amd/models/unet32.npz
    python -m cudalbm_amd.warmstart gen --set hi --out hi.npz      (data only)
    python -m cudalbm_amd.warmstart train --data base.npz hi.npz --out model.npz
    python -m cudalbm_amd.warmstart eval  --model cudalbm_amd/models/unet32.npz
    python -m cudalbm_amd.warmstart bound --indices 6,25   (perfect-network bound)

Data: lid-driven cavities (D3Q19 FP32) drawn like the batched sweep (Re
log-uniform, lid speed uniform) on edges N in {32, 48, 64, 96, 128} (and 128/192
for the "x192" set), solved together by
the batched solver (lbm_batch_run) until the residual
||u_t - u_{t-100}|| / ||u_t|| < 1e-6 or a step cap, then resampled onto the
network's 32^3 grid (network node i at solver coordinate i (N-1)/31) and
normalised: u/U, v/U, w/U and p'/U^2 with p' = (rho - 1)/3 minus its mean.

Training: the U-Net of include/lbm_surrogate.h on the GPU (BF16 MFMA forward
and backward, Adam).  Evaluation: held-out cavities at N = 128 (and 192),
solved to residual 1e-6 from rest (cold) and from the network's prediction
(warm, lbm_solve_to_residual with its fallback rule); reports steps, wall time
and the rel-L2 difference of the two final velocity fields.

All compute runs in the HIP library; this module only orchestrates the C ABI.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import native as N

TRAIN_SEED = 20261203
EVAL_SEED = 20261204
# held-out evaluation draws: a seed no training set uses (the training sets use
# TRAIN_SEED, TRAIN_SEED + 1 -- which equals EVAL_SEED -- and X192_SEED)
HELDOUT_SEED = 20261207
S_NET = 32
# training edges: the sweep's range up to 128, weighted towards the cheap ones
TRAIN_EDGES = (32, 48, 64, 96, 128)
TRAIN_EDGE_P = (0.25, 0.25, 0.25, 0.15, 0.10)
# data sets: "base" as above; "hi" adds cavities at the large edges, where a
# 32^3 prediction is interpolated furthest, with a higher step cap
# "x192": cavities at the evaluation edges 128 and 192 (192 lies beyond the
# other sets' edges) in the Re range where those meshes reach a steady state,
# drawn with a seed that no other set and no evaluation draw uses
X192_SEED = 20261211
DATA_SETS = {
    "base": {"cases": 512, "seed": TRAIN_SEED, "edges": TRAIN_EDGES, "edge_p": TRAIN_EDGE_P, "max_steps": 80000},
    "hi": {"cases": 160, "seed": TRAIN_SEED + 1, "edges": (96, 128), "edge_p": (0.4, 0.6), "max_steps": 150000},
    "x192": {"cases": 96, "seed": X192_SEED, "edges": (128, 192), "edge_p": (0.5, 0.5), "max_steps": 150000,
             "re_range": (200.0, 1400.0)},
}


def draw_cases(n, seed, edges, re_range=(100.0, 3200.0), u_range=(0.02, 0.08), edge_p=None):
    rng = np.random.default_rng(seed)
    E = rng.choice(np.asarray(edges), size=n, p=edge_p)
    Re = np.exp(rng.uniform(np.log(re_range[0]), np.log(re_range[1]), size=n))
    U = rng.uniform(u_range[0], u_range[1], size=n)
    return [(int(e), float(r), float(u)) for e, r, u in zip(E, Re, U)]


def resample(field, N, S=S_NET):
    """Trilinear samples of an N^3 field (x fastest) at network nodes i (N-1)/(S-1): -> S^3."""
    f = np.asarray(field, dtype=np.float64).reshape(N, N, N)
    if N == S:
        return f.reshape(-1)
    sc = np.arange(S) * (N - 1) / (S - 1)
    i0 = np.minimum(N - 2, np.floor(sc).astype(int))
    fr = sc - i0
    out = np.zeros((S, S, S))
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                w = ((fr if dz else 1 - fr)[:, None, None] * (fr if dy else 1 - fr)[None, :, None]
                     * (fr if dx else 1 - fr)[None, None, :])
                out += w * f[np.ix_(i0 + dz, i0 + dy, i0 + dx)]
    return out.reshape(-1)


P_CLIP = 4.0


def normalised_fields(mom, N, U, S=S_NET):
    """Stored moments of an N^3 solution -> [S^3][4] (u/U, v/U, w/U, p'/U^2).

    p' is the pressure p = (rho - 1)/3 minus its domain mean: the mean density of
    a cavity drifts with the elapsed steps (the lid closure does not conserve
    mass), so it is history, not a property of the steady flow.  p'/U^2 is
    clipped to +-P_CLIP at the singular lid corners."""
    u, v, w = (np.asarray(mom[k], dtype=np.float64) / 3.0 for k in (1, 2, 3))
    p = np.asarray(mom[0], dtype=np.float64) / 3.0
    pn = np.clip((p - p.mean()) / (U * U), -P_CLIP, P_CLIP)
    cols = [resample(u / U, N, S), resample(v / U, N, S), resample(w / U, N, S), resample(pn, N, S)]
    return np.stack(cols, -1).astype(np.float32)


def generate(cases, tol=1e-6, check=100, max_steps=60000, log=print, S=S_NET):
    """Solve cases [(N, Re, U)] with one batch per edge; returns (inputs, targets, info)
    with inputs and targets on the network's S^3 grid."""
    inputs, targets, info = [], [], []
    by_edge = {}
    for i, c in enumerate(cases):
        by_edge.setdefault(c[0], []).append(i)
    results = {}
    for edge, idx in sorted(by_edge.items()):
        t0 = time.perf_counter()
        with N.Batch([(edge, edge, edge, cases[i][1], cases[i][2]) for i in idx], tile=(16, 16, 16),
                     persistent=0) as b:
            # in chunks of 10k steps, so that progress shows; unconverged cases rejoin
            active = np.ones(len(idx), bool)
            res = np.full(len(idx), np.inf)
            cap = 0
            while cap < max_steps and active.any():
                cap = min(max_steps, cap + 10000)
                b.set_active(active)
                steps, r, ms = b.run(tol, check, cap)
                res[active] = r[active]  # cases that sat out report no residual
                active = np.isfinite(res) & (res >= tol) & (steps < max_steps)
                log(f"#   edge {edge}: {cap} steps, {int(active.sum())} still running, {time.perf_counter() - t0:.0f} s")
            for j, i in enumerate(idx):
                results[i] = (b.get_moments(j), int(steps[j]), float(res[j]))
        log(f"# edge {edge}: {len(idx)} cases, steps {int(np.min(steps))}..{int(np.max(steps))}, "
            f"converged {int(np.sum(res < tol))}/{len(idx)}, {time.perf_counter() - t0:.1f} s")
    dropped = 0
    for i, (n, Re, U) in enumerate(cases):
        mom, steps, res = results[i]
        if not (np.isfinite(res) and all(np.all(np.isfinite(m)) for m in mom)):
            dropped += 1  # a case the scheme cannot hold at this resolution
            continue
        inputs.append(N.encode_inputs(S, Re, n))
        targets.append(normalised_fields(mom, n, U, S))
        info.append({"N": n, "Re": Re, "U": U, "steps": steps, "residual": res})
    if dropped:
        log(f"# dropped {dropped} non-finite cases")
    return np.stack(inputs), np.stack(targets), info


def train(inputs, targets, epochs=150, batch=8, lr=1e-3, seed=TRAIN_SEED, log=print, S=S_NET, init=None):
    """Adam in three learning-rate stages; `init`: parameters to start from (the
    network is fully convolutional, so weights trained on one grid size are a
    starting point on another)."""
    net = N.Surrogate(S=S, seed=seed)
    if init is not None:
        net.set_params(init)
    net.load_data(inputs, targets)
    t0 = time.perf_counter()
    losses = []
    # step decay of the learning rate in three stages
    for stage, frac in enumerate((0.6, 0.25, 0.15)):
        e = max(1, int(round(epochs * frac)))
        done = 0
        while done < e:
            k = min(10, e - done)
            losses += list(net.train(k, batch=batch, lr=lr * (0.2 ** stage), seed=seed + 100 * stage + done))
            done += k
            log(f"# stage {stage} (lr {lr * 0.2 ** stage:.1e}): epoch {len(losses)}, loss {losses[-1]:.3e}, "
                f"{time.perf_counter() - t0:.0f} s")
    return net, np.asarray(losses), time.perf_counter() - t0


def velocity_rel_l2(a, b):
    va = np.stack([np.asarray(a[k], np.float64) for k in (1, 2, 3)])
    vb = np.stack([np.asarray(b[k], np.float64) for k in (1, 2, 3)])
    return float(np.linalg.norm(va - vb) / max(np.linalg.norm(vb), 1e-300))


def evaluate(net, cases, tol=1e-6, check=100, max_steps=200000, log=print, converge_tol=None, converge_steps=0):
    """Cold and warm solves of each (N, Re, U) to residual tol; returns a list of
    per-case dicts.  With converge_tol, the cold run then continues to residual
    converge_tol (at most converge_steps more steps) and both stopped fields are
    also compared with that converged one.  A case whose cold run does not reach
    tol within max_steps is marked unsteady."""
    out = []
    for n, Re, U in cases:
        conv = None
        with N.Solver(n, n, n, Re, U) as s:
            cold = N.solve_to_residual(s, Re, U, tol, check, max_steps)
            mc = s.get_moments()
            if converge_tol is not None and cold["residual"] < tol:
                conv = N.solve_to_residual(s, Re, U, converge_tol, check, converge_steps)
                mconv = s.get_moments()
        with N.Solver(n, n, n, Re, U) as s:
            warm = N.solve_to_residual(s, Re, U, tol, check, max_steps, surrogate=net)
            mw = s.get_moments()
        r = {"N": n, "Re": Re, "U": U, "cold": cold, "warm": warm, "final_rel_l2": velocity_rel_l2(mw, mc),
             "time_ratio": warm["ms"] / cold["ms"], "step_ratio": warm["steps"] / max(cold["steps"], 1),
             "steady": bool(cold["residual"] < tol)}
        if conv is not None:
            r["converged"] = {"steps": int(cold["steps"] + conv["steps"]), "residual": conv["residual"]}
            r["warm_vs_converged_rel_l2"] = velocity_rel_l2(mw, mconv)
            r["cold_vs_converged_rel_l2"] = velocity_rel_l2(mc, mconv)
        log(f"# N {n} Re {Re:.0f} U {U:.3f}: cold {cold['steps']} steps {cold['ms']:.0f} ms (r {cold['residual']:.1e}), "
            f"warm {warm['steps']} steps {warm['ms']:.0f} ms (r {warm['residual']:.1e}, first {warm['first_residual']:.2e}, "
            f"kept {warm['warm']}), time ratio {r['time_ratio']:.3f}, final rel-L2 {r['final_rel_l2']:.2e}"
            + (f", vs converged (r {conv['residual']:.1e}): warm {r['warm_vs_converged_rel_l2']:.2e} "
               f"cold {r['cold_vs_converged_rel_l2']:.2e}" if conv is not None else "")
            + ("" if r["steady"] else ", UNSTEADY (cold run did not reach the tolerance)"))
        out.append(r)
    return out


def upsample(f4, S, n):
    """[S^3][4] network-grid fields -> n^3 x 4 (trilinear, network node i at
    solver coordinate i (n-1)/(S-1)); the host twin of the warm start's
    device interpolation."""
    g = np.asarray(f4, np.float64).reshape(S, S, S, 4)
    sc = np.arange(n) * (S - 1) / (n - 1)
    k = np.minimum(S - 2, np.floor(sc).astype(int))
    fr = sc - k
    out = np.zeros((n, n, n, 4))
    for dz in (0, 1):
        wz = (fr if dz else 1 - fr)[:, None, None, None]
        for dy in (0, 1):
            wy = (fr if dy else 1 - fr)[None, :, None, None]
            for dx in (0, 1):
                wx = (fr if dx else 1 - fr)[None, None, :, None]
                out += wz * wy * wx * g[np.ix_(k + dz, k + dy, k + dx)]
    return out


def equilibrium_moments(f, n, U):
    """n^3 x 4 normalised fields -> the ten stored moments of the equilibrium
    state (as lbm_sur_warm_start writes them: walls at rest, the lid row moving
    at U along +x)."""
    u, v, w, rm1 = U * f[..., 0], U * f[..., 1], U * f[..., 2], 3 * U * U * f[..., 3]
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    wall = (x == 0) | (x == n - 1) | (y == 0) | (z == 0) | (z == n - 1)
    lid = y == n - 1
    u = np.where(lid, U, np.where(wall, 0.0, u))
    v = np.where(lid | wall, 0.0, v)
    w = np.where(lid | wall, 0.0, w)
    m = [rm1, 3 * u, 3 * v, 3 * w, 4.5 * u * u, 9 * u * v, 9 * u * w, 4.5 * v * v, 9 * v * w, 4.5 * w * w]
    return [a.astype(np.float32).reshape(-1) for a in m]


def perfect_bound(cases, grids, tol=1e-6, check=100, max_steps=200000, log=print):
    """What a perfect network on an S^3 grid could gain: each case solved cold to
    tol, its solution sampled on the S^3 grid and interpolated back exactly as a
    prediction is, and solved again to tol from there."""
    out = []
    for n, Re, U in cases:
        with N.Solver(n, n, n, Re, U) as s:
            cold = N.solve_to_residual(s, Re, U, tol, check, max_steps)
            mc = s.get_moments()
        row = {"N": n, "Re": Re, "U": U, "cold_steps": int(cold["steps"]), "steady": bool(cold["residual"] < tol)}
        for S in grids:
            with N.Solver(n, n, n, Re, U) as s:
                s.set_moments(equilibrium_moments(upsample(normalised_fields(mc, n, U, S), S, n), n, U))
                warm = N.solve_to_residual(s, Re, U, tol, check, max_steps)
            row[f"S{S}"] = {"steps": int(warm["steps"]), "step_ratio": warm["steps"] / max(cold["steps"], 1),
                            "first_residual": warm["first_residual"]}
            log(f"# N {n} Re {Re:.0f}: cold {cold['steps']} steps; perfect S={S}: {warm['steps']} steps "
                f"(ratio {warm['steps'] / max(cold['steps'], 1):.3f}, first r {warm['first_residual']:.3g})")
        out.append(row)
    return out


def summarise(res):
    """Medians over the steady cases; unsteady ones counted separately."""
    st = [r for r in res if r["steady"]]
    out = {"cases": len(res), "steady": len(st), "unsteady": len(res) - len(st)}
    if st:
        out["median_time_ratio"] = float(np.median([r["time_ratio"] for r in st]))
        out["max_time_ratio"] = float(max(r["time_ratio"] for r in st))
        out["median_step_ratio"] = float(np.median([r["step_ratio"] for r in st]))
        out["max_final_rel_l2"] = float(max(r["final_rel_l2"] for r in st))
        conv = [r for r in st if "warm_vs_converged_rel_l2" in r]
        if conv:
            out["max_warm_vs_converged_rel_l2"] = float(max(r["warm_vs_converged_rel_l2"] for r in conv))
            out["max_cold_vs_converged_rel_l2"] = float(max(r["cold_vs_converged_rel_l2"] for r in conv))
    out["fell_back"] = int(sum(r["warm"]["fell_back"] for r in res))
    return out


def save_data(path, targets, info, S=S_NET):
    """Targets (FP16: their values are O(1) and the fit error is ~1e-2) and case list."""
    np.savez_compressed(path, targets=np.asarray(targets, np.float16), info=json.dumps(info), S=S)


def load_data(path):
    z = np.load(path, allow_pickle=False)
    info = json.loads(str(z["info"]))
    S = int(z["S"]) if "S" in z else S_NET
    inputs = np.stack([N.encode_inputs(S, c["Re"], c["N"]) for c in info])
    return inputs, z["targets"].astype(np.float32), info, S


def save_model(path, net, meta):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    np.savez_compressed(path, params=net.get_params(), meta=json.dumps(meta))


def load_model(path):
    z = np.load(path, allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    net = N.Surrogate(S=int(meta.get("S", S_NET)), seed=0)
    net.set_params(z["params"])
    return net, meta


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    g = sub.add_parser("gen")
    g.add_argument("--set", choices=sorted(DATA_SETS), default="base")
    g.add_argument("--min-edge", type=int, default=0, help="keep only the set's cases with N >= this")
    g.add_argument("--max-edge", type=int, default=1 << 30, help="keep only the set's cases with N <= this")
    g.add_argument("--S", type=int, default=S_NET, help="network grid edge of the targets")
    g.add_argument("--first", type=int, default=0, help="solve cases first .. first+count-1 of the (filtered) set")
    g.add_argument("--count", type=int, default=0, help="number of cases of the slice (0: all)")
    g.add_argument("--out", required=True)
    t = sub.add_parser("train")
    t.add_argument("--out", required=True)
    t.add_argument("--data", nargs="*", default=[], help="data files from `gen` (default: generate the base set)")
    t.add_argument("--S", type=int, default=S_NET, help="network grid edge (must match the data files)")
    t.add_argument("--init", default="", help="model file whose parameters start the training")
    t.add_argument("--cases", type=int, default=512)
    t.add_argument("--max-steps", type=int, default=80000)
    t.add_argument("--epochs", type=int, default=150)
    t.add_argument("--batch", type=int, default=8)
    t.add_argument("--lr", type=float, default=1e-3)
    e = sub.add_parser("eval")
    e.add_argument("--model", required=True)
    e.add_argument("--edges", default="128")
    e.add_argument("--cases", type=int, default=4)
    e.add_argument("--max-steps", type=int, default=200000)
    e.add_argument("--seed", type=int, default=HELDOUT_SEED)
    e.add_argument("--first", type=int, default=0, help="evaluate cases first .. first+count-1 of the draw")
    e.add_argument("--count", type=int, default=0, help="number of cases of the slice (0: all)")
    e.add_argument("--indices", default="", help="comma-separated case indices of the draw (instead of a slice)")
    e.add_argument("--converge-tol", type=float, default=0.0,
                   help="continue each cold run to this residual and compare both fields with it (0: off)")
    e.add_argument("--converge-steps", type=int, default=200000)
    e.add_argument("--re-range", default="400,2500", help="Re interval of the log-uniform draw")
    e.add_argument("--json", default="")
    bd = sub.add_parser("bound", help="perfect-network bound: converged fields interpolated from S^3 grids")
    bd.add_argument("--edges", default="128,192")
    bd.add_argument("--cases", type=int, default=64)
    bd.add_argument("--seed", type=int, default=HELDOUT_SEED)
    bd.add_argument("--re-range", default="400,2500")
    bd.add_argument("--indices", required=True, help="comma-separated case indices of the draw")
    bd.add_argument("--grids", default="64,96,128")
    a = ap.parse_args(argv)
    log = lambda s: print(s, file=sys.stderr, flush=True)  # noqa: E731
    if a.cmd == "gen":
        d = DATA_SETS[a.set]
        cases = [c for c in draw_cases(d["cases"], d["seed"], d["edges"], edge_p=d["edge_p"],
                                       re_range=d.get("re_range", (100.0, 3200.0)))
                 if a.min_edge <= c[0] <= a.max_edge]
        cases = cases[a.first:a.first + a.count] if a.count else cases[a.first:]
        t0 = time.perf_counter()
        _, y, info = generate(cases, max_steps=d["max_steps"], log=log, S=a.S)
        save_data(a.out, y, info, S=a.S)
        print(json.dumps({"set": a.set, "S": a.S, "cases": len(info), "generate_s": time.perf_counter() - t0,
                          "converged": int(sum(i["residual"] < 1e-6 for i in info))}), flush=True)
    elif a.cmd == "train":
        t0 = time.perf_counter()
        if a.data:
            parts = [load_data(p) for p in a.data]
            if any(p[3] != a.S for p in parts):
                raise SystemExit(f"data files were generated for S = {[p[3] for p in parts]}, not --S {a.S}")
            x = np.concatenate([p[0] for p in parts])
            y = np.concatenate([p[1] for p in parts])
            info = [c for p in parts for c in p[2]]
            sets = [os.path.basename(p) for p in a.data]
        else:
            cases = draw_cases(a.cases, TRAIN_SEED, TRAIN_EDGES, edge_p=TRAIN_EDGE_P)
            x, y, info = generate(cases, max_steps=a.max_steps, log=log, S=a.S)
            sets = ["base"]
        gen_s = time.perf_counter() - t0
        init = None
        if a.init:
            init_net, init_meta = load_model(a.init)
            with init_net:
                init = init_net.get_params()
        net, losses, train_s = train(x, y, epochs=a.epochs, batch=a.batch, lr=a.lr, log=log, S=a.S, init=init)
        edges = sorted({c["N"] for c in info})
        meta = {"S": a.S, "cases": len(info), "data": sets, "edges": edges, "init": os.path.basename(a.init),
                "per_edge": {str(e): sum(c["N"] == e for c in info) for e in edges}, "generate_s": gen_s,
                "train_s": train_s, "final_loss": float(losses[-1]),
                "converged": int(sum(i["residual"] < 1e-6 for i in info)), "epochs": len(losses), "batch": a.batch,
                "lr": a.lr}
        save_model(a.out, net, meta)
        print(json.dumps(meta), flush=True)
    elif a.cmd == "bound":
        edges = tuple(int(v) for v in a.edges.split(","))
        lo, hi = (float(v) for v in a.re_range.split(","))
        draw = draw_cases(a.cases, a.seed, edges, re_range=(lo, hi))
        idx = [int(v) for v in a.indices.split(",")]
        rows = perfect_bound([draw[i] for i in idx], [int(v) for v in a.grids.split(",")], log=log)
        for i, r in zip(idx, rows):
            r["index"] = i
        print(json.dumps(rows), flush=True)
    else:
        net, meta = load_model(a.model)
        edges = tuple(int(v) for v in a.edges.split(","))
        lo, hi = (float(v) for v in a.re_range.split(","))
        cases = draw_cases(a.cases, a.seed, edges, re_range=(lo, hi))
        if a.indices:
            idx = [int(v) for v in a.indices.split(",")]
            cases = [cases[i] for i in idx]
        else:
            idx = list(range(len(cases)))[a.first:a.first + a.count] if a.count else list(range(len(cases)))[a.first:]
            cases = cases[a.first:a.first + a.count] if a.count else cases[a.first:]
        res = evaluate(net, cases, max_steps=a.max_steps, log=log,
                       converge_tol=a.converge_tol or None, converge_steps=a.converge_steps)
        for i, r in zip(idx, res):
            r["index"] = i
        summary = {**summarise(res), "seed": a.seed, "first": a.first, "edges": list(edges),
                   "re_range": [lo, hi], "model": os.path.basename(a.model), "per_case": res}
        if a.json:
            with open(a.json, "w") as f:
                json.dump(summary, f)
        print(json.dumps({k: v for k, v in summary.items() if k != "per_case"}), flush=True)


if __name__ == "__main__":
    main()
