"""Writes the golden fixtures of tests/test_golden.py (small .npz files).

For three small problems (1-D Poisson 64 and 2-D Poisson 20x20 with
aggregation + Jacobi; 3-D heterogeneous 8^3 with the hypre-default HMIS /
ext+i hierarchy, seeded) it stores the matrix, every level's P, transpose map
(Pt_off / Pt_rows / Pt_pnz), Galerkin triple map and coarse pattern, the
coarsest operator, a fixed x0 / b, the oracle's V-cycle output from x0, the
independent dense restatement's V-cycle output, and the first five PCG
iterates of the oracle.  Inputs come from the portable SplitMix64 generator.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "draugr.jl_amd"), os.path.dirname(HERE)):
    if p not in sys.path:
        sys.path.insert(0, p)

SEED = 20261202


def cases():
    import draugr_hip as dh
    P = dh.problems
    agg = dh.AMGConfig(coarsening=dh.AggregationCoarsening(0.25), smoother=dh.JacobiSmootherType(),
                       max_coarse_size=10)
    hyp = dh.hypre_default_config(smoother=dh.L1JacobiSmootherType(), max_coarse_size=20)
    return {
        "poisson1d_64_agg": (P.poisson1d(64), agg, "jacobi"),
        "poisson2d_20_agg": (P.poisson2d(20), agg, "jacobi"),
        "hetero3d_8_hypre": (P.heterogeneous3d(8, P.log_permeability(8, 1.5, SEED + 8)), hyp, "l1_jacobi"),
    }


def build(name, A, cfg, smoother):
    """All arrays of one fixture, from the current library and oracle."""
    import draugr_hip as dh
    from oracle import oracle as orc
    from dense_ref import DenseAMG
    h = dh.amg_setup(A, cfg, upload=False)
    levels = [h.level(l) for l in range(h.nlevels)]
    c = h.coarsest()
    out = {"A_rp": A.rowptr, "A_cv": A.colval, "A_nz": A.nzval, "nlevels": np.array([h.nlevels]),
           "coarse_rp": c["rp"], "coarse_cv": c["cv"], "coarse_nz": c["nz"]}
    for l, lv in enumerate(levels):
        for k in ("A_rp", "A_cv", "A_nz", "P_rp", "P_cv", "P_nz", "Pt_off", "Pt_rows", "Pt_pnz",
                  "R_off", "R_pi", "R_ai", "R_pj"):
            if lv.get(k) is not None:
                out["L%d_%s" % (l, k)] = lv[k]
    n = A.n
    x0 = dh.problems.splitmix_normal(SEED + 900, n)
    b = dh.problems.splitmix_normal(SEED + 901, n)
    out["x0"], out["b"] = x0, b
    oh = orc.OracleHierarchy(levels, c["rp"], c["cv"], c["nz"], smoother=smoother)
    out["cycle_oracle"] = oh.cycle(x0, b)
    out["cycle_dense"] = DenseAMG(levels, c, smoother).cycle(x0, b)
    out["pcg_iterates"] = np.stack([oh.pcg(np.zeros(n), b, 0.0, k)[0] for k in range(1, 6)])
    return out


def main():
    for name, (A, cfg, sm) in cases().items():
        arrs = build(name, A, cfg, sm)
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **arrs)
        print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
