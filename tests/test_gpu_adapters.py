"""Parity of the preconditioner adapters (the roles of
ext/DraugrLinearSolveExt.jl:18-62 and ext/DraugrJutulExt.jl:146-185 over
src/preconditioner.jl:29-102): DraugrPreconditioner.ldiv_ / apply_ is
z = 0 followed by one cycle, so on host and on device vectors it equals the
oracle's cycle from a zero guess bit for bit (FP64), before and after
update_ (a partial resetup with new values on the same pattern), and the
scipy LinearOperator built from it applies the same map."""
import numpy as np
import pytest

from conftest import oracle_of

pytestmark = pytest.mark.gpu

SEED = 20261202


@pytest.mark.parametrize("smoother", ["l1_jacobi", "jacobi"])
def test_ldiv_equals_oracle_zero_guess_cycle(dh, orc, torch_cuda, smoother):
    torch = torch_cuda
    P = dh.problems
    m = 14
    logk = P.log_permeability(m, 1.0, SEED + 1000)
    A = P.heterogeneous3d(m, logk)
    sm = dh.L1JacobiSmootherType() if smoother == "l1_jacobi" else dh.JacobiSmootherType()
    cfg = dh.hypre_default_config(smoother=sm)
    prec = dh.DraugrPreconditioner(cfg)
    prec.update_(A)
    oh = oracle_of(prec.hierarchy, orc, smoother)
    for stage in range(2):
        if stage == 1:  # new values, same pattern: update_ resets up
            A2 = P.heterogeneous3d(m, logk + 0.3 * P.splitmix_normal(SEED + 1001, m ** 3))
            prec.update_(A2)
            oh.resetup(A2.nzval)
        r = P.splitmix_normal(SEED + 1002 + stage, A.n)
        ref = oh.cycle(np.zeros(A.n), r)
        z = np.full(A.n, 7.0)  # ldiv_ zeroes z first
        prec.ldiv_(z, r)
        assert np.array_equal(z, ref), stage
        zd = torch.full((A.n,), 7.0, dtype=torch.float64, device="cuda")
        prec.ldiv_(zd, torch.tensor(r, device="cuda"))
        torch.cuda.synchronize()
        assert np.array_equal(zd.cpu().numpy(), ref), stage
        op = dh.aslinearoperator(prec)
        assert np.array_equal(op.matvec(r), ref), stage
