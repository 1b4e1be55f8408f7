"""GPU parity of the adjoint (reverse-mode rule of x = A(k)^-1 b).

* Device kernels (assembly of the 7-point operator, dL/dA, dL/dlog k) are
  bit-identical to the oracle (oracle/adjoint.py) and to the host generator.
* The torch.autograd rule agrees with the oracle's FP64 CPU adjoint (same
  hierarchy, both solves to 1e-12): 1e-7 relative.
* At 96^3 (sigma = 1.5, L = 1/2 |x - x*|^2) the gradients w.r.t. log k and b
  agree with central finite differences of the GPU forward (h = 1e-6, 20
  random unit directions) to 1e-6 relative, and with the CPU adjoint.
"""
import numpy as np
import pytest

from conftest import oracle_of, rel_err

pytestmark = pytest.mark.gpu

SEED = 20261202
H_FD = 1e-6


@pytest.fixture(scope="module")
def adj():
    from oracle import adjoint
    return adjoint


def _cfg(dh, **kw):
    return dh.hypre_default_config(smoother=dh.L1JacobiSmootherType(), **kw)


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def _host(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("m", [1, 2, 5, 16])
def test_assemble_bit_exact(dh, adj, torch_cuda, m):
    from draugr_hip import adjoint as dadj
    P = dh.problems
    logk = P.log_permeability(m, 1.5, SEED + m)
    A = P.heterogeneous3d(m, logk)
    h = dh.amg_setup(A, _cfg(dh, max_coarse_size=20))
    k = np.exp(logk)
    nz = dadj.assemble_7pt_(h, m, _dev(torch_cuda, k))
    torch_cuda.cuda.synchronize()
    assert np.array_equal(_host(nz), A.nzval)
    assert np.array_equal(_host(nz), adj.assemble_7pt(m, k)[2])


def test_assemble_rejects_wrong_grid(dh, torch_cuda):
    from draugr_hip import adjoint as dadj
    h = dh.amg_setup(dh.problems.poisson2d(16), dh.AMGConfig())
    with pytest.raises(dh.DraugrError, match="7-point"):
        dadj.assemble_7pt_(h, 4, _dev(torch_cuda, np.ones(64)))


def test_gradient_kernels_bit_exact(dh, adj, torch_cuda):
    from draugr_hip import adjoint as dadj
    P = dh.problems
    m = 16
    n = m ** 3
    logk = P.log_permeability(m, 1.5, SEED + 40)
    A = P.heterogeneous3d(m, logk)
    h = dh.amg_setup(A, _cfg(dh))
    k = np.exp(logk)
    x, lam = P.splitmix_normal(SEED + 41, n), P.splitmix_normal(SEED + 42, n)
    gnz = dadj.matrix_grad(h, _dev(torch_cuda, x), _dev(torch_cuda, lam))
    gk = dadj.grad_logk_7pt(m, _dev(torch_cuda, k), _dev(torch_cuda, x), _dev(torch_cuda, lam))
    torch_cuda.cuda.synchronize()
    assert np.array_equal(_host(gnz), adj.matrix_grad(A.rowptr, A.colval, x, lam))
    assert np.array_equal(_host(gk), adj.grad_logk_7pt(m, k, x, lam))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_rrule_matches_cpu_adjoint(dh, orc, adj, torch_cuda, prec):
    from draugr_hip import adjoint as dadj
    P = dh.problems
    m = 20
    n = m ** 3
    logk0 = P.log_permeability(m, 1.5, SEED + 50)
    cfg = _cfg(dh, precision=prec)
    h = dh.amg_setup(P.heterogeneous3d(m, logk0), cfg)
    oh = oracle_of(h, orc, "l1_jacobi")
    logk = logk0 + 0.3 * P.splitmix_normal(SEED + 51, n)  # new values on the same hierarchy
    b, xs = P.splitmix_normal(SEED + 52, n), P.splitmix_normal(SEED + 53, n)
    lk = _dev(torch_cuda, logk).requires_grad_(True)
    bt = _dev(torch_cuda, b).requires_grad_(True)
    x = dadj.solve_7pt(lk, bt, h, m, tol=1e-12)
    loss = 0.5 * ((x - _dev(torch_cuda, xs)) ** 2).sum()
    loss.backward()
    xo, gko, gbo, _ = adj.adjoint(oh, m, np.exp(logk), b, lambda v: v - xs, tol=1e-12)
    assert rel_err(_host(x), xo) < 1e-9
    assert rel_err(_host(bt.grad), gbo) < 1e-7
    assert rel_err(_host(lk.grad), gko) < 1e-7


def test_sparse_solve_matrix_gradient(dh, orc, adj, torch_cuda):
    """dL/dA values of the generic rule against the oracle's FP64 adjoint."""
    from draugr_hip import adjoint as dadj
    P = dh.problems
    A = P.poisson2d(40)
    cfg = dh.AMGConfig(coarsening=dh.AggregationCoarsening(0.25), smoother=dh.JacobiSmootherType())
    h = dh.amg_setup(A, cfg)
    oh = oracle_of(h, orc, "jacobi")
    b, xs = P.splitmix_normal(SEED + 60, A.n), P.splitmix_normal(SEED + 61, A.n)
    nz = _dev(torch_cuda, A.nzval).requires_grad_(True)
    x = dadj.sparse_solve(nz, _dev(torch_cuda, b), h, tol=1e-12)
    (0.5 * ((x - _dev(torch_cuda, xs)) ** 2).sum()).backward()
    xo, _, _ = oh.pcg(np.zeros(A.n), b, tol=1e-12, maxiter=500)
    lo, _, _ = oh.pcg(np.zeros(A.n), xo - xs, tol=1e-12, maxiter=500)
    assert rel_err(_host(nz.grad), adj.matrix_grad(A.rowptr, A.colval, xo, lo)) < 1e-7


def test_backward_after_another_forward(dh, torch_cuda):
    """A backward after a second forward on the same hierarchy resets up the
    operator of its own forward first."""
    from draugr_hip import adjoint as dadj
    P = dh.problems
    m = 12
    n = m ** 3
    logk = P.log_permeability(m, 1.0, SEED + 70)
    h = dh.amg_setup(P.heterogeneous3d(m, logk), _cfg(dh))
    b = _dev(torch_cuda, P.splitmix_normal(SEED + 71, n))
    l1 = _dev(torch_cuda, logk).requires_grad_(True)
    x1 = dadj.solve_7pt(l1, b, h, m, tol=1e-12)
    x1.sum().backward()
    g_ref = l1.grad.clone()
    l1.grad = None
    x1 = dadj.solve_7pt(l1, b, h, m, tol=1e-12)
    l2 = _dev(torch_cuda, logk + 1.0).requires_grad_(True)
    dadj.solve_7pt(l2, b, h, m, tol=1e-12).sum().backward()
    x1.sum().backward()
    assert rel_err(_host(l1.grad), _host(g_ref)) < 1e-9


def test_c5_adjoint_96(dh, orc, adj, torch_cuda):
    """96^3, sigma = 1.5, L = 1/2 |x - x*|^2: gradients w.r.t. log k and b
    against central differences of the GPU forward (h = 1e-6, 20 random unit
    directions) and against the FP64 CPU adjoint.

    A unit direction spreads over n = 884,736 cells, so h d_i is about 1e-9
    and theta +- h d is rounded by ~1e-7 of the step: the adjoint is applied
    to the step actually taken, ((theta + h d) - (theta - h d)) / 2h (the
    difference is exact), and the differenced solves run to 1e-14 so that
    their own error stays well below the signal x(theta + h d) - x(theta - h d)."""
    import torch
    from draugr_hip import adjoint as dadj
    P = dh.problems
    m = 96
    n = m ** 3
    tol, tol_fd = 1e-12, 1e-14
    logk = P.log_permeability(m, 1.5, SEED + 5)
    h = dh.amg_setup(P.heterogeneous3d(m, logk), _cfg(dh))
    b = P.splitmix_normal(SEED + 105, n)
    xs = P.splitmix_normal(SEED + 205, n)
    lk = _dev(torch, logk).requires_grad_(True)
    bt = _dev(torch, b).requires_grad_(True)
    xst = _dev(torch, xs)
    x = dadj.solve_7pt(lk, bt, h, m, tol=tol)
    (0.5 * ((x - xst) ** 2).sum()).backward()
    gk, gb = lk.grad.detach(), bt.grad.detach()

    dirs = P.splitmix_normal(SEED + 305, 20 * n).reshape(20, n)
    fd, ad, fdb, adb = [], [], [], []
    with torch.no_grad():
        for t, d in enumerate(dirs):
            dt = _dev(torch, d / np.linalg.norm(d))
            lp, lm = lk.detach() + H_FD * dt, lk.detach() - H_FD * dt
            xp = dadj.solve_7pt(lp, bt.detach(), h, m, tol=tol_fd, maxiter=2000)
            xm = dadj.solve_7pt(lm, bt.detach(), h, m, tol=tol_fd, maxiter=2000)
            fd.append(float(torch.dot(xp - xm, xp + xm - 2 * xst)) / 2 / (2 * H_FD))
            ad.append(float(torch.dot(gk, (lp - lm) / (2 * H_FD))))
            # dL/db along the same 20 directions
            bp, bm = bt.detach() + H_FD * dt, bt.detach() - H_FD * dt
            xp = dadj.solve_7pt(lk.detach(), bp, h, m, tol=tol_fd, maxiter=2000)
            xm = dadj.solve_7pt(lk.detach(), bm, h, m, tol=tol_fd, maxiter=2000)
            fdb.append(float(torch.dot(xp - xm, xp + xm - 2 * xst)) / 2 / (2 * H_FD))
            adb.append(float(torch.dot(gb, (bp - bm) / (2 * H_FD))))
    e_k, e_b = rel_err(fd, ad), rel_err(fdb, adb)
    print("C5 96^3: FD vs adjoint rel (log k) %.3e, (b) %.3e" % (e_k, e_b), flush=True)
    assert e_k < 1e-6 and e_b < 1e-6

    oh = oracle_of(h, orc, "l1_jacobi")
    _, gko, gbo, _ = adj.adjoint(oh, m, np.exp(logk), b, lambda v: v - xs, tol=tol)
    e_cpu_k, e_cpu_b = rel_err(_host(gk), gko), rel_err(_host(gb), gbo)
    print("C5 96^3: GPU vs CPU adjoint rel (log k) %.3e, (b) %.3e" % (e_cpu_k, e_cpu_b), flush=True)
    assert e_cpu_k < 1e-7 and e_cpu_b < 1e-7
