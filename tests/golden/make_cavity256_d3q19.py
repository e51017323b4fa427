"""Writes tests/golden/cavity256_d3q19_re1000_t1000.npz: the FP64 oracle's
D3Q19 lid-driven cavity at 256^3 (Re 1000, lid 0.05 along +x on the y-max
face, rest start) after 1000 steps with the function objects k, kMean, S and
SMean on, sampled for the GPU parity test (tests/test_gpu_cavity256.py).

    python tests/golden/make_cavity256_d3q19.py [threads]

Contents (float32 samples of FP64 values; node ids are x + nx*(y + ny*z)):
  nodes            sampled node ids: 8192 drawn uniformly (seed 20261205) and
                   the 8^3 block at the lid corner x, y = 255, z = 0 plus the
                   lid row y = 255, z = 128
  moments          (10, nodes) stored moments at step 1000
  kMean, SMean     (nodes), (6, nodes) FP64 running means over steps 0..999
  norms_*          full-field L2 norms (FP64) of every moment, kMean, SMean
                   component, k and S component
  kinetic_energy   sum over nodes of 0.5 rho |u|^2 (physical u = m/3)
The run takes about 10 minutes on 8 cores.
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

N, RE, U, STEPS, SEED = 256, 1000.0, 0.05, 1000, 20261205
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cavity256_d3q19_re1000_t1000.npz")


def sample_nodes():
    rng = np.random.default_rng(SEED)
    rand = rng.choice(N ** 3, size=8192, replace=False)
    z, y, x = np.meshgrid(np.arange(0, 8), np.arange(N - 8, N), np.arange(N - 8, N), indexing="ij")
    corner = (x + N * (y + N * z)).ravel()
    row = np.arange(N) + N * ((N - 1) + N * 128)
    return np.unique(np.concatenate([rand, corner, row])).astype(np.int64)


def main():
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else 0
    om = O.omega(U, RE, N)
    tau = 1.0 / om
    mom = O.rest_state(N, N, N)
    t0 = time.time()
    means = {"kMean": np.zeros(N ** 3), "SMean": np.zeros((6, N ** 3))}
    O.run_means(19, N, N, N, om, U, mom, STEPS, tau, means=means, nthreads=threads, fields=("kMean", "SMean"))
    print(f"oracle {N}^3 x {STEPS}: {time.time() - t0:.0f} s", flush=True)
    nodes = sample_nodes()
    k = O.k_field(mom)
    S = O.S_field(mom, tau)
    out = {
        "nodes": nodes,
        "moments": np.stack([m[nodes] for m in mom]).astype(np.float32),
        "kMean": means["kMean"][nodes].astype(np.float32),
        "SMean": means["SMean"][:, nodes].astype(np.float32),
        "norms_moments": np.array([np.linalg.norm(m) for m in mom]),
        "norms_kMean": np.array([np.linalg.norm(means["kMean"])]),
        "norms_SMean": np.array([np.linalg.norm(a) for a in means["SMean"]]),
        "norms_k": np.array([np.linalg.norm(k)]),
        "norms_S": np.array([np.linalg.norm(a) for a in S]),
        "kinetic_energy": np.array([0.5 * np.sum((1.0 + mom[0]) * (mom[1] ** 2 + mom[2] ** 2 + mom[3] ** 2) / 9.0)]),
        "case": np.array([N, RE, U, STEPS], dtype=np.float64),
    }
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({os.path.getsize(OUT) / 1e6:.2f} MB, {nodes.size} nodes)", flush=True)


if __name__ == "__main__":
    main()
