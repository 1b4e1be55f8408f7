"""Pins the oracle's adjoint (oracle/adjoint.py) on the CPU: the operator's
values against the independent host generator (bit-exact), the gradients
against central finite differences of direct (scipy) solves, and the
oracle-PCG adjoint against the direct-solve adjoint.  No GPU needed."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from conftest import oracle_of, rel_err

SEED = 20261202
H_FD = 1e-6


@pytest.fixture(scope="module")
def adj():
    from oracle import adjoint
    return adjoint


@pytest.mark.parametrize("m", [1, 2, 3, 7])
def test_assemble_matches_generator(dh, adj, m):
    logk = dh.problems.log_permeability(m, 1.5, SEED + m)
    A = dh.problems.heterogeneous3d(m, logk)
    rp, cv, nz = adj.assemble_7pt(m, np.exp(logk))
    assert np.array_equal(rp, A.rowptr) and np.array_equal(cv, A.colval)
    assert np.array_equal(nz, A.nzval)


def _direct(m, k, b):
    rp, cv, nz = __import__("oracle.adjoint", fromlist=["x"]).assemble_7pt(m, k)
    A = sp.csr_matrix((nz, cv, rp), shape=(m ** 3, m ** 3)).tocsc()
    return spla.splu(A)


def _loss_diff(xp, xm, xs):
    """L(x+) - L(x-) for L = 1/2 |x - x*|^2, in the cancellation-free form
    (x+ - x-) . (x+ + x- - 2 x*) / 2."""
    return float(np.dot(xp - xm, xp + xm - 2.0 * xs) / 2.0)


def test_gradients_vs_central_fd(dh, adj):
    """dL/dlog k, dL/db and dL/dA of L = 1/2 |A(k)^-1 b - x*|^2 against central
    differences (h = 1e-6) along 20 random unit directions, direct solves."""
    m = 5
    n = m ** 3
    P = dh.problems
    logk = P.log_permeability(m, 1.0, SEED + 11)
    k = np.exp(logk)
    b = P.splitmix_normal(SEED + 12, n)
    xs = P.splitmix_normal(SEED + 13, n)
    lu = _direct(m, k, b)
    x = lu.solve(b)
    lam = lu.solve(x - xs)  # A symmetric
    rp, cv, nz = adj.assemble_7pt(m, k)
    g_logk = adj.grad_logk_7pt(m, k, x, lam)
    g_nz = adj.matrix_grad(rp, cv, x, lam)
    dirs = P.splitmix_normal(SEED + 14, 20 * n).reshape(20, n)
    fd, ad = [], []
    for d in dirs:
        d = d / np.linalg.norm(d)
        xp = _direct(m, np.exp(logk + H_FD * d), b).solve(b)
        xm = _direct(m, np.exp(logk - H_FD * d), b).solve(b)
        fd.append(_loss_diff(xp, xm, xs) / (2 * H_FD))
        ad.append(float(np.dot(g_logk, d)))
    fd, ad = np.array(fd), np.array(ad)
    assert rel_err(fd, ad) < 1e-6, (fd, ad)
    # b: L is quadratic in b, so central differences are exact up to rounding
    for d in dirs[:5]:
        d = d / np.linalg.norm(d)
        fdb = _loss_diff(lu.solve(b + H_FD * d), lu.solve(b - H_FD * d), xs) / (2 * H_FD)
        assert abs(fdb - np.dot(lam, d)) <= 1e-6 * max(abs(fdb), 1e-3)
    # matrix values: perturb symmetric pairs so that A stays symmetric
    A0 = sp.csr_matrix((nz, cv, rp), shape=(n, n))
    for s in range(3):
        E = sp.random(n, n, density=0.02, random_state=SEED + s, format="csr")
        E = (E + E.T).multiply(A0 != 0).tocsr()
        E.sort_indices()
        D = np.asarray(sp.csr_matrix((np.ones_like(nz), cv, rp), shape=(n, n)).multiply(E)[
            np.repeat(np.arange(n), np.diff(rp)), cv]).ravel()
        D /= np.linalg.norm(D)
        Ap = sp.csr_matrix((nz + H_FD * D, cv, rp), shape=(n, n)).tocsc()
        Am = sp.csr_matrix((nz - H_FD * D, cv, rp), shape=(n, n)).tocsc()
        fda = _loss_diff(spla.spsolve(Ap, b), spla.spsolve(Am, b), xs) / (2 * H_FD)
        assert abs(fda - np.dot(g_nz, D)) <= 1e-6 * abs(fda)


def test_grad_logk_is_matrix_grad_chained(dh, adj):
    """dL/dlog k equals the matrix gradient pushed through dA/dlog k, which is
    formed here column by column by complex-step differentiation of the
    assembly (exact to rounding)."""
    m = 3
    n = m ** 3
    P = dh.problems
    k = np.exp(P.log_permeability(m, 1.0, SEED + 21))
    x, lam = P.splitmix_normal(SEED + 22, n), P.splitmix_normal(SEED + 23, n)
    rp, cv, _ = adj.assemble_7pt(m, k)
    g_nz = adj.matrix_grad(rp, cv, x, lam)
    ref = np.zeros(n)
    for c in range(n):
        lk = np.log(k).astype(complex)
        lk[c] += 1e-30j
        _, _, nzc = adj.assemble_7pt(m, np.exp(lk))
        ref[c] = np.dot(g_nz, nzc.imag / 1e-30)
    assert rel_err(adj.grad_logk_7pt(m, k, x, lam), ref) < 1e-13


def test_oracle_pcg_adjoint_matches_direct(dh, orc, adj):
    """The oracle's PCG adjoint on an AMG hierarchy agrees with the direct-solve
    adjoint to the solve tolerance."""
    m = 10
    n = m ** 3
    P = dh.problems
    logk = P.log_permeability(m, 1.5, SEED + 31)
    A = P.heterogeneous3d(m, logk)
    h = dh.amg_setup(A, dh.hypre_default_config(smoother=dh.L1JacobiSmootherType(), max_coarse_size=20),
                     upload=False)
    oh = oracle_of(h, orc, "l1_jacobi")
    b = P.splitmix_normal(SEED + 32, n)
    xs = P.splitmix_normal(SEED + 33, n)
    k2 = np.exp(logk + 0.3 * P.splitmix_normal(SEED + 34, n))  # resetup to new values
    x, g_logk, g_b, g_nz = adj.adjoint(oh, m, k2, b, lambda x: x - xs, tol=1e-13)
    lu = _direct(m, k2, b)
    xd = lu.solve(b)
    lamd = lu.solve(xd - xs)
    assert rel_err(x, xd) < 1e-9
    assert rel_err(g_b, lamd) < 1e-9
    assert rel_err(g_logk, adj.grad_logk_7pt(m, k2, xd, lamd)) < 1e-9
