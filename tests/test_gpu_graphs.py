"""The preconditioner cycle replayed from a captured hipGraph gives exactly
the results of the same cycle launched eagerly: every PCG iteration replays
the graph once, so bit-identical solutions and iteration counts check every
replay.  Replays are interleaved with eager work (solves without the graph,
a resetup that rewrites the values the graph reads, host synchronizations)."""
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 20261202


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16"])
def test_graph_replays_match_eager(dh, torch_cuda, prec):
    torch = torch_cuda
    P = dh.problems
    m = 20
    logk = P.log_permeability(m, 1.5, SEED + 80)
    A = P.heterogeneous3d(m, logk)
    cfg = dh.hypre_default_config(smoother=dh.L1JacobiSmootherType(), precision=prec)
    cfg_g = dataclasses.replace(cfg, graphs=True)
    cfg_e = dataclasses.replace(cfg, graphs=False)
    h = dh.amg_setup(A, cfg)
    for rep in range(4):
        if rep == 2:  # new values under the same graph
            A2 = P.heterogeneous3d(m, logk + 0.3 * P.splitmix_normal(SEED + 81, m ** 3))
            dh.amg_resetup_(h, A2, cfg)
        b = torch.tensor(P.splitmix_normal(SEED + 82 + rep, A.n), device="cuda")
        xg = torch.zeros_like(b)
        _, itg, relg = dh.pcg_solve_(xg, b, h, cfg_g, tol=1e-10, maxiter=300)
        torch.cuda.synchronize()
        xe = torch.zeros_like(b)
        _, ite, rele = dh.pcg_solve_(xe, b, h, cfg_e, tol=1e-10, maxiter=300)
        torch.cuda.synchronize()
        assert itg == ite and relg == rele, (rep, itg, ite)
        assert torch.equal(xg, xe), rep
        # the mixed / FP64 single cycle (zero guess) through the same graph
        zg = torch.zeros_like(b)
        ze = torch.zeros_like(b)
        dh.amg_cycle_(zg, b, h, cfg_g, zero_guess=True)
        dh.amg_cycle_(ze, b, h, cfg_e, zero_guess=True)
        torch.cuda.synchronize()
        assert torch.equal(zg, ze), rep
