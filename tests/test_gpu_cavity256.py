"""The headline workload checked against the FP64 oracle at its own size: D3Q19
lid-driven cavity 256^3, FP32 storage and arithmetic, Re 1000, lid 0.05, rest
start, function objects k, kMean, S and SMean on, 1000 steps.

The oracle's run (about 10 minutes on 8 cores) is precomputed by
tests/golden/make_cavity256_d3q19.py into a fixture of sampled nodes and
full-field norms.  Tolerances (the FP32 step's parity bar, 1e-5 rel-L2):
  * rho-1, u, v, w and the velocity vector: rel-L2 over the sampled nodes
    <= 1e-5 (8192 uniform samples estimate the full-field rel-L2; plus the
    nodes around the lid corner and one lid row);
  * Pi moments, kMean, SMean, k, S: rel-L2 over the samples <= 1e-5;
  * every field's full-field L2 norm within 1e-6 relative, and the kinetic
    energy (FP64 reduction over all 16.8 M nodes) within 1e-6 relative;
  * the fixture is float32: it adds at most 6e-8 relative per value.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cavity256_d3q19_re1000_t1000.npz")


def _rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_cavity256_d3q19_fp32_against_oracle(native_lib):
    from cudalbm_amd import Solver
    from cudalbm_amd import native as N

    fx = np.load(FIXTURE)
    n, Re, u, steps = (int(fx["case"][0]), float(fx["case"][1]), float(fx["case"][2]), int(fx["case"][3]))
    assert (n, Re, u, steps) == (256, 1000.0, 0.05, 1000)
    nodes = fx["nodes"]
    print(f"cavity D3Q19 {n}^3 Re {Re} lid {u}, {steps} steps, FO k/kMean/S/SMean; {nodes.size} sampled nodes",
          flush=True)
    with Solver(n, n, n, Re, u, velocity_set=19, storage="fp32") as s:
        s.fo_enable(N.LBM_FO_K | N.LBM_FO_KMEAN | N.LBM_FO_S | N.LBM_FO_SMEAN)
        s.step(steps)
        tau = s.tau
        mom = s.get_moments()
        ke, _ = s.reduce()
        fields = {"kMean": s.fo_get(N.LBM_FIELD_KMEAN), "SMean": s.fo_get(N.LBM_FIELD_SMEAN),
                  "k": s.fo_get(N.LBM_FIELD_K), "S": s.fo_get(N.LBM_FIELD_S)}
    errs, norm_errs = {}, {}
    names = ["rho-1", "u", "v", "w", "mxx", "mxy", "mxz", "myy", "myz", "mzz"]
    ref_m = fx["moments"].astype(np.float64)
    for a in range(10):
        errs[names[a]] = _rel(mom[a][nodes], ref_m[a])
        norm_errs[names[a]] = abs(np.linalg.norm(mom[a].astype(np.float64)) / fx["norms_moments"][a] - 1.0)
    errs["velocity"] = _rel(np.stack([mom[a][nodes] for a in (1, 2, 3)]), ref_m[1:4])
    # instantaneous k and S of the oracle at the sampled nodes, from its moments
    ref_k = O.k_field(list(ref_m))
    ref_S = O.S_field(list(ref_m), tau)
    errs["kMean"] = _rel(fields["kMean"][nodes], fx["kMean"])
    errs["k"] = _rel(fields["k"][nodes], ref_k)
    norm_errs["kMean"] = abs(np.linalg.norm(fields["kMean"]) / fx["norms_kMean"][0] - 1.0)
    norm_errs["k"] = abs(np.linalg.norm(fields["k"]) / fx["norms_k"][0] - 1.0)
    for a, c in enumerate(["xx", "xy", "xz", "yy", "yz", "zz"]):
        errs[f"SMean_{c}"] = _rel(fields["SMean"][a][nodes], fx["SMean"][a])
        errs[f"S_{c}"] = _rel(fields["S"][a][nodes], ref_S[a])
        norm_errs[f"SMean_{c}"] = abs(np.linalg.norm(fields["SMean"][a]) / fx["norms_SMean"][a] - 1.0)
        norm_errs[f"S_{c}"] = abs(np.linalg.norm(fields["S"][a]) / fx["norms_S"][a] - 1.0)
    ke_err = abs(ke / fx["kinetic_energy"][0] - 1.0)
    print("sampled rel-L2: " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
    print("full-field norm rel. diff: " + ", ".join(f"{k} {v:.1e}" for k, v in norm_errs.items()), flush=True)
    print(f"kinetic energy rel. diff {ke_err:.2e}", flush=True)
    for key in ("rho-1", "u", "v", "w", "velocity"):
        assert errs[key] <= 1e-5, (key, errs[key])
    assert max(errs.values()) <= 1e-5, errs
    assert max(norm_errs.values()) <= 1e-6, norm_errs
    assert ke_err <= 1e-6
