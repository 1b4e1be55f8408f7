"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker:
known answers from the reference's own tests (test/runtests.jl), scipy /
numpy results, an independent dense restatement of the cycle
(tests/dense_ref.py), and invariants the mathematics fixes.  No GPU needed."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from conftest import oracle_of, rel_err
from dense_ref import DenseAMG

SEED = 20261202


def _sp(A):
    return sp.csr_matrix((A.nzval, A.colval, A.rowptr), shape=(A.n, A.n))


def _hier(dh, A, cfg):
    return dh.amg_setup(A, cfg, upload=False)


# ── known answers of the reference tests ────────────────────────────────────
def test_spmv_known_answer(dh, orc):
    """A*1 for 1-D Poisson is [1, 0, ..., 0, 1] (test/runtests.jl:52-94)."""
    A = dh.problems.poisson1d(10)
    y = orc.spmv(A.rowptr, A.colval, A.nzval, np.ones(10))
    assert np.array_equal(y, np.r_[1.0, np.zeros(8), 1.0])


def test_invdiag_known_answers(dh, orc):
    A = dh.problems.poisson1d(10)
    # Jacobi: 1/2 everywhere (test/runtests.jl:272-284)
    assert np.array_equal(orc.invdiag(A.rowptr, A.colval, A.nzval, False), np.full(10, 0.5))
    # l1: interior 1/4, boundary 1/3 (test/runtests.jl:696-706, 1857-1865)
    d = orc.invdiag(A.rowptr, A.colval, A.nzval, True)
    assert d[4] == 0.25 and d[0] == 1.0 / 3.0 and d[9] == 1.0 / 3.0


def test_invdiag_safe_for_tiny_diagonal(dh, orc):
    """Row with a 1e-20 diagonal gets 0, never Inf (test/runtests.jl:1784-1796)."""
    rp = np.array([0, 2, 5, 7], np.int32)
    cv = np.array([0, 1, 0, 1, 2, 1, 2], np.int32)
    nz = np.array([2.0, -1.0, -1.0, 1e-20, -1.0, -1.0, 2.0])
    d = orc.invdiag(rp, cv, nz, False)
    assert np.all(np.isfinite(d)) and d[1] == 0.0 and d[0] == 0.5


def test_trivial_system_known_answer(dh, orc):
    """5 x = 3 -> 0.6 (test/runtests.jl:1798-1815)."""
    A = dh.problems.CSR(np.array([0, 1], np.int32), np.array([0], np.int32), np.array([5.0]))
    h = _hier(dh, A, dh.AMGConfig(max_coarse_size=10))
    oh = oracle_of(h, orc, "jacobi")
    x, it, _ = oh.solve(np.zeros(1), np.array([3.0]), 1e-10)
    assert abs(x[0] - 0.6) < 1e-15 and it == 1


# ── single kernels against scipy / numpy ────────────────────────────────────
@pytest.fixture(scope="module")
def agg_case(dh):
    A = dh.problems.poisson2d(12)
    cfg = dh.AMGConfig(coarsening=dh.AggregationCoarsening(0.25), smoother=dh.JacobiSmootherType())
    return A, cfg, _hier(dh, A, cfg)


@pytest.fixture(scope="module")
def hypre_case(dh):
    P = dh.problems
    A = P.heterogeneous3d(8, P.log_permeability(8, 1.5, SEED))
    cfg = dh.hypre_default_config(smoother=dh.L1JacobiSmootherType())
    return A, cfg, _hier(dh, A, cfg)


def test_kernels_vs_scipy(dh, orc, agg_case, hypre_case):
    rng = np.random.default_rng(SEED)
    for A, cfg, h in (agg_case, hypre_case):
        for l in range(h.nlevels):
            lv = h.level(l)
            As = sp.csr_matrix((lv["A_nz"], lv["A_cv"], lv["A_rp"]))
            Ps = sp.csr_matrix((lv["P_nz"], lv["P_cv"], lv["P_rp"]), shape=(lv["n"], lv["n_coarse"]))
            x, b, xc = rng.standard_normal(lv["n"]), rng.standard_normal(lv["n"]), rng.standard_normal(lv["n_coarse"])
            assert rel_err(orc.spmv(lv["A_rp"], lv["A_cv"], lv["A_nz"], x), As @ x) < 1e-14
            assert rel_err(orc.residual(lv["A_rp"], lv["A_cv"], lv["A_nz"], x, b), b - As @ x) < 1e-14
            d = orc.invdiag(lv["A_rp"], lv["A_cv"], lv["A_nz"], True)
            assert rel_err(d, 1.0 / abs(As).sum(axis=1).A1) < 1e-15
            w = 2.0 / 3.0
            assert rel_err(orc.jacobi_step(lv["A_rp"], lv["A_cv"], lv["A_nz"], x, b, d, w),
                           x + w * d * (b - As @ x)) < 1e-14
            assert rel_err(orc.restrict(lv["Pt_off"], lv["Pt_rows"], lv["Pt_pnz"], lv["P_nz"], x), Ps.T @ x) < 1e-14
            assert rel_err(orc.prolong(lv["P_rp"], lv["P_cv"], lv["P_nz"], xc, x), x + Ps @ xc) < 1e-14


def test_galerkin_vs_scipy_and_pattern(dh, orc, agg_case, hypre_case):
    """Numeric RAP vs explicit P^T A P (test/runtests.jl:181-270); the coarse
    pattern is the structural pattern of |P|^T |A| |P| (structural zeros kept)."""
    for A, cfg, h in (agg_case, hypre_case):
        for l in range(h.nlevels):
            lv = h.level(l)
            As = sp.csr_matrix((lv["A_nz"], lv["A_cv"], lv["A_rp"]))
            Ps = sp.csr_matrix((lv["P_nz"], lv["P_cv"], lv["P_rp"]), shape=(lv["n"], lv["n_coarse"]))
            vals = orc.galerkin(lv["R_off"], lv["R_pi"], lv["R_ai"], lv["R_pj"], lv["P_nz"], lv["A_nz"])
            if l + 1 < h.nlevels:
                nxt = h.level(l + 1)
                rp, cv = nxt["A_rp"], nxt["A_cv"]
            else:
                c = h.coarsest()
                rp, cv = c["rp"], c["cv"]
            Ac = sp.csr_matrix((vals, cv, rp), shape=(lv["n_coarse"], lv["n_coarse"]))
            ref = (Ps.T @ As @ Ps).toarray()
            assert np.abs(Ac.toarray() - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())
            B = abs(Ps).T @ abs(As) @ abs(Ps)  # positive terms: no cancellation
            B.sort_indices()
            assert np.array_equal(B.indptr, rp) and np.array_equal(B.indices, cv)


def test_lu_solve_vs_numpy(orc):
    rng = np.random.default_rng(SEED)
    for n in (1, 2, 7, 50):
        M = rng.standard_normal((n, n)) + n * np.eye(n)
        b = rng.standard_normal(n)
        assert rel_err(orc.lu_solve_dense(M, b), np.linalg.solve(M, b)) < 1e-13
    with pytest.raises(ZeroDivisionError):
        orc.lu_solve_dense(np.zeros((3, 3)), np.ones(3))


# ── the cycle against the independent dense restatement ──────────────────────
@pytest.mark.parametrize("smoother,cycle,pre,post", [
    ("jacobi", "v", 1, 1), ("jacobi", "w", 2, 1), ("l1_jacobi", "v", 1, 1), ("l1_jacobi", "v", 0, 2),
    ("colored_gs", "v", 1, 1), ("l1_colored_gs", "w", 1, 2)])
def test_cycle_vs_dense(dh, orc, smoother, cycle, pre, post):
    P = dh.problems
    A = P.heterogeneous3d(7, P.log_permeability(7, 1.0, SEED + 4))
    sm = {"jacobi": dh.JacobiSmootherType(), "l1_jacobi": dh.L1JacobiSmootherType(),
          "colored_gs": dh.ColoredGaussSeidelType(), "l1_colored_gs": dh.L1ColoredGaussSeidelType()}[smoother]
    cfg = dh.AMGConfig(coarsening=dh.HMISCoarsening(0.25, dh.ExtendedIInterpolation(0.3)), smoother=sm,
                       max_coarse_size=20)
    h = _hier(dh, A, cfg)
    assert h.nlevels >= 2
    oh = oracle_of(h, orc, smoother, cycle=cycle, pre=pre, post=post)
    levels = [h.level(l) for l in range(h.nlevels)]
    ref = DenseAMG(levels, h.coarsest(), smoother, cycle=cycle, pre=pre, post=post)
    rng = np.random.default_rng(SEED)
    x0, b = rng.standard_normal(A.n), rng.standard_normal(A.n)
    assert rel_err(oh.cycle(x0, b), ref.cycle(x0, b)) < 1e-12
    assert rel_err(oh.cycle(np.zeros(A.n), b), ref.cycle(np.zeros(A.n), b)) < 1e-12


def test_dense_check_catches_planted_mistakes(dh, orc, hypre_case):
    """The comparison above is sensitive: a restatement with a wrong sign, a
    dropped smoothing step, the wrong diagonal, no coarse correction or a
    transfer operator off by 1e-6 is told apart from the oracle."""
    A, cfg, h = hypre_case
    oh = oracle_of(h, orc, "l1_jacobi")
    levels = [h.level(l) for l in range(h.nlevels)]
    rng = np.random.default_rng(SEED + 1)
    x0, b = rng.standard_normal(A.n), rng.standard_normal(A.n)
    good = oh.cycle(x0, b)
    assert rel_err(good, DenseAMG(levels, h.coarsest(), "l1_jacobi").cycle(x0, b)) < 1e-12
    bads = [DenseAMG(levels, h.coarsest(), "l1_jacobi", omega=-2.0 / 3.0),   # wrong sign
            DenseAMG(levels, h.coarsest(), "l1_jacobi", post=0),              # dropped step
            DenseAMG(levels, h.coarsest(), "jacobi")]                         # wrong diagonal
    no_coarse = DenseAMG(levels, h.coarsest(), "l1_jacobi")
    no_coarse.Ac = 1e30 * no_coarse.Ac                                          # dropped coarse correction
    bads.append(no_coarse)
    perturbed = DenseAMG(levels, h.coarsest(), "l1_jacobi")
    perturbed.P[0] = perturbed.P[0] * (1.0 + 1e-6)                              # slightly wrong transfer
    bads.append(perturbed)
    for bad in bads:
        assert rel_err(good, bad.cycle(x0, b)) > 1e-9


def test_vcycle_jacobi_is_symmetric(dh, orc, agg_case):
    """V(1,1) with Jacobi is a symmetric operator: <u, M v> = <M u, v>."""
    A, cfg, h = agg_case
    oh = oracle_of(h, orc, "jacobi")
    rng = np.random.default_rng(SEED)
    u, v = rng.standard_normal(A.n), rng.standard_normal(A.n)
    Mu, Mv = oh.cycle(np.zeros(A.n), u), oh.cycle(np.zeros(A.n), v)
    assert abs(u @ Mv - Mu @ v) < 1e-12 * abs(u @ Mv)


@pytest.mark.parametrize("smoother", ["colored_gs", "l1_colored_gs"])
def test_symmetric_gs_cycle(dh, orc, smoother):
    """The reverse-colour post-sweep (gs_symmetric; no reference counterpart)
    pinned twice: against the dense restatement, and by the symmetry it is
    for -- <u, M v> = <M u, v> with it, but not with the reference's
    same-order post-sweep (the planted asymmetry is detected)."""
    P = dh.problems
    A = P.heterogeneous3d(7, P.log_permeability(7, 1.0, SEED + 4))
    sm = {"colored_gs": dh.ColoredGaussSeidelType(), "l1_colored_gs": dh.L1ColoredGaussSeidelType()}[smoother]
    cfg = dh.AMGConfig(coarsening=dh.HMISCoarsening(0.25, dh.ExtendedIInterpolation(0.3)), smoother=sm,
                       max_coarse_size=20)
    h = _hier(dh, A, cfg)
    levels = [h.level(l) for l in range(h.nlevels)]
    rng = np.random.default_rng(SEED + 5)
    u, v = rng.standard_normal(A.n), rng.standard_normal(A.n)
    for cycle in ("v", "w"):
        sym = oracle_of(h, orc, smoother, cycle=cycle, gs_symmetric=True)
        ref = DenseAMG(levels, h.coarsest(), smoother, cycle=cycle, gs_symmetric=True)
        assert rel_err(sym.cycle(np.zeros(A.n), u), ref.cycle(np.zeros(A.n), u)) < 1e-12
        Mu, Mv = sym.cycle(np.zeros(A.n), u), sym.cycle(np.zeros(A.n), v)
        assert abs(u @ Mv - Mu @ v) < 1e-12 * abs(u @ Mv)
        fwd = oracle_of(h, orc, smoother, cycle=cycle)
        Mu, Mv = fwd.cycle(np.zeros(A.n), u), fwd.cycle(np.zeros(A.n), v)
        assert abs(u @ Mv - Mu @ v) > 1e-6 * abs(u @ Mv)


# ── solvers: reference invariants and scipy ──────────────────────────────────
def test_stationary_solve_reference_thresholds(dh, orc):
    """Relative residual < 1e-8 in < 200 cycles on 2-D Poisson 10x10
    (test/runtests.jl:332-344); W-cycle needs no more cycles than V
    (test/runtests.jl:1828-1851)."""
    A = dh.problems.poisson2d(10)
    b = np.random.default_rng(SEED).random(A.n)
    its = {}
    for cyc in ("v", "w"):
        cfg = dh.AMGConfig(cycle_type=cyc.upper(), pre_smoothing_steps=2, post_smoothing_steps=2)
        h = _hier(dh, A, cfg)
        oh = oracle_of(h, orc, "l1_colored_gs", cycle=cyc, pre=2, post=2)
        x, it, rel = oh.solve(np.zeros(A.n), b, 1e-8, 200)
        assert rel < 1e-8 and it < 200
        assert np.linalg.norm(b - _sp(A) @ x) / np.linalg.norm(b) < 1e-8
        its[cyc] = it
    assert its["w"] <= its["v"]


def test_pcg_matches_scipy_cg(dh, orc, hypre_case):
    """The oracle PCG is textbook preconditioned CG: same iterations (+-1) and
    solution as scipy's CG with the oracle cycle as preconditioner."""
    A, cfg, h = hypre_case
    oh = oracle_of(h, orc, "l1_jacobi")
    b = dh.problems.splitmix_normal(SEED, A.n)
    x, it, rel = oh.pcg(np.zeros(A.n), b, 1e-8, 200)
    M = spla.LinearOperator((A.n, A.n), matvec=lambda r: oh.cycle(np.zeros(A.n), r))
    count = [0]
    xs, info = spla.cg(_sp(A), b, rtol=1e-8, atol=0.0, maxiter=200, M=M,
                       callback=lambda xk: count.__setitem__(0, count[0] + 1))
    assert info == 0 and abs(count[0] - it) <= 1
    assert rel_err(x, xs) < 1e-7
    assert np.linalg.norm(b - _sp(A) @ x) / np.linalg.norm(b) < 1.01e-8


def test_resetup_vs_scipy(dh, orc, hypre_case):
    """Partial resetup keeps P and recomputes every coarse operator as P^T A P
    of the new values (src/resetup.jl:27-45)."""
    A, cfg, h = hypre_case
    oh = oracle_of(h, orc, "l1_jacobi")
    rng = np.random.default_rng(SEED)
    s = np.exp(0.5 * rng.standard_normal(A.n))
    rows = np.repeat(np.arange(A.n), np.diff(A.rowptr))
    new = A.nzval * s[rows] * s[A.colval]
    oh.resetup(new)
    Al = sp.csr_matrix((new, A.colval, A.rowptr), shape=(A.n, A.n))
    for l in range(h.nlevels):
        lv = h.level(l)
        Ps = sp.csr_matrix((lv["P_nz"], lv["P_cv"], lv["P_rp"]), shape=(lv["n"], lv["n_coarse"]))
        Ac = (Ps.T @ Al @ Ps).tocsr()
        Ac.sort_indices()
        got = oh.level_values(l + 1)
        if l + 1 < h.nlevels:
            pat = h.level(l + 1)
            rp, cv = pat["A_rp"], pat["A_cv"]
        else:
            c = h.coarsest()
            rp, cv = c["rp"], c["cv"]
        G = sp.csr_matrix((got, cv, rp), shape=Ac.shape).toarray()
        assert np.abs(G - Ac.toarray()).max() < 1e-12 * np.abs(Ac).max()
        Al = sp.csr_matrix((got, cv, rp), shape=Ac.shape)
        assert rel_err(oh.dinv(l), 1.0 / abs(sp.csr_matrix((oh.level_values(l), lv["A_cv"], lv["A_rp"]))).sum(axis=1).A1) < 1e-15


# ── SPAI(0) and Chebyshev smoothers ──────────────────────────────────────────
def _cheb_bounds(h, orc, seed=SEED + 77):
    """lambda bounds per level from the numpy power iteration on D^-1 A."""
    out = []
    for l in range(h.nlevels):
        lv = h.level(l, with_rmap=False)
        d = orc.invdiag(lv["A_rp"], lv["A_cv"], lv["A_nz"], 0)
        v0 = np.random.default_rng(seed + l).standard_normal(len(d))
        v0 /= np.linalg.norm(v0)
        rho = orc.power_rho(lv["A_rp"], lv["A_cv"], lv["A_nz"], d, v0)
        out.append((1.1 * rho / 30.0, 1.1 * rho))
    return out


def test_spai0_known_answers(dh, orc):
    """Interior 1/3 and boundary 2/5 of 1-D Poisson (test/runtests.jl:754-762);
    ten SPAI(0) steps from zero reduce the residual (:764-773)."""
    A = dh.problems.poisson1d(10)
    m = orc.invdiag(A.rowptr, A.colval, A.nzval, 2)
    assert m[4] == 2.0 / 6.0 and m[0] == 2.0 / 5.0
    x = np.zeros(10)
    for _ in range(10):
        x = orc.jacobi_step(A.rowptr, A.colval, A.nzval, x, np.ones(10), m, 1.0)
    assert np.linalg.norm(np.ones(10) - A.matvec(x)) < np.linalg.norm(np.ones(10))


def test_power_iteration_bounds(dh, orc):
    """The estimate never exceeds the true spectral radius of D^-1 A (1-D
    Poisson with Jacobi: 1 + cos(pi / (n + 1))) and approaches it."""
    n = 64
    A = dh.problems.poisson1d(n)
    d = orc.invdiag(A.rowptr, A.colval, A.nzval, 0)
    v0 = np.random.default_rng(1).standard_normal(n)
    v0 /= np.linalg.norm(v0)
    rho = orc.power_rho(A.rowptr, A.colval, A.nzval, d, v0)
    exact = 1.0 + np.cos(np.pi / (n + 1))
    assert 0.9 * exact < rho <= exact * (1 + 1e-12)
    more = orc.power_rho(A.rowptr, A.colval, A.nzval, d, v0, niter=200)
    assert rho <= more <= exact * (1 + 1e-12)


def test_chebyshev_smoothing_known_answer(dh, orc):
    """Five Chebyshev steps from zero reduce the residual (test/runtests.jl:1924-1933);
    the oracle's smoothing equals the dense restatement's."""
    A = dh.problems.poisson1d(10)
    d = orc.invdiag(A.rowptr, A.colval, A.nzval, 0)
    v0 = np.random.default_rng(3).standard_normal(10)
    v0 /= np.linalg.norm(v0)
    lmax = 1.1 * orc.power_rho(A.rowptr, A.colval, A.nzval, d, v0)
    x = orc.cheb_smooth(A.rowptr, A.colval, A.nzval, np.zeros(10), np.ones(10), d, lmax / 30, lmax, steps=5)
    assert np.linalg.norm(np.ones(10) - A.matvec(x)) < np.linalg.norm(np.ones(10))
    from dense_ref import dense
    Ad = dense(A.rowptr, A.colval, A.nzval)
    th, de = (lmax + lmax / 30) / 2, (lmax - lmax / 30) / 2
    xd = np.zeros(10)
    for _ in range(5):
        dd = d * (1 - Ad @ xd) / th
        xd = xd + dd
        so = th / de
        for _k in range(2):
            sn = 1 / (2 * th / de - so)
            dd = 2 * sn / de * d * (1 - Ad @ xd) + sn * so * dd
            xd = xd + dd
            so = sn
    assert rel_err(x, xd) < 1e-13


@pytest.mark.parametrize("smoother", ["spai0", "chebyshev"])
def test_cycle_vs_dense_new_smoothers(dh, orc, smoother):
    P = dh.problems
    A = P.heterogeneous3d(8, P.log_permeability(8, 1.0, SEED + 5))
    st = dh.SPAI0SmootherType() if smoother == "spai0" else dh.ChebyshevSmootherType()
    cfg = dh.AMGConfig(coarsening=dh.HMISCoarsening(0.5, dh.ExtendedIInterpolation(0.3)), smoother=st,
                       max_coarse_size=10)
    h = _hier(dh, A, cfg)
    levels = [h.level(l) for l in range(h.nlevels)]
    c = h.coarsest()
    cheb = _cheb_bounds(h, orc) if smoother == "chebyshev" else None
    oh = orc.OracleHierarchy(levels, c["rp"], c["cv"], c["nz"], smoother=smoother, cheb=cheb, pre=2, post=2)
    dn = DenseAMG(levels, c, smoother, cheb=cheb, pre=2, post=2)
    b = P.splitmix_normal(SEED + 6, A.n)
    x0 = P.splitmix_normal(SEED + 7, A.n)
    assert rel_err(oh.cycle(x0, b), dn.cycle(x0, b)) < 1e-12
    x, it, rel = oh.solve(np.zeros(A.n), b, 1e-8, 200)
    assert rel < 1e-8 and it < 200


# ── ILU(0) ───────────────────────────────────────────────────────────────────
def test_ilu0_tridiagonal_is_exact(dh, orc):
    """On a tridiagonal matrix ILU(0) is the exact LU: one step solves."""
    A = dh.problems.poisson1d(50)
    b = dh.problems.splitmix_normal(SEED + 90, 50)
    x = orc.ilu0_smooth(A.rowptr, A.colval, A.nzval, np.zeros(50), b)
    assert np.linalg.norm(b - A.matvec(x)) < 1e-12 * np.linalg.norm(b)


@pytest.mark.parametrize("name", ["poisson2d", "hetero3d"])
def test_ilu0_factors_vs_dense_and_host(dh, orc, name):
    """The oracle's factors equal the dense restatement's and the library's
    host factorization (the one the ilu0 smoother uploads) bit for bit."""
    import ctypes
    from dense_ref import dense, ilu0
    P = dh.problems
    A = P.poisson2d(9) if name == "poisson2d" else P.heterogeneous3d(5, P.log_permeability(5, 1.5, SEED + 91))
    diag, L, U = orc.ilu0_factor(A.rowptr, A.colval, A.nzval)
    Ld, Ud = ilu0(dense(A.rowptr, A.colval, A.nzval))
    rows = np.repeat(np.arange(A.n), np.diff(A.rowptr))
    low = A.colval < rows
    assert np.array_equal(L[low], Ld[rows[low], A.colval[low]]) and not L[~low].any()
    up = A.colval >= rows
    assert np.array_equal(U[up], Ud[rows[up], A.colval[up]]) and not U[~up].any()
    d2, L2, U2 = np.zeros(A.n, np.int32), np.zeros(A.nnz), np.zeros(A.nnz)
    vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    dh._lib.check(dh.lib().draugr_hip_ilu0_factor(A.n, vp(A.rowptr), vp(A.colval), vp(A.nzval), vp(d2), vp(L2),
                                                  vp(U2)))
    assert np.array_equal(d2, diag) and np.array_equal(L2, L) and np.array_equal(U2, U)


def test_cycle_vs_dense_ilu0(dh, orc):
    P = dh.problems
    A = P.heterogeneous3d(8, P.log_permeability(8, 1.0, SEED + 92))
    cfg = dh.AMGConfig(coarsening=dh.HMISCoarsening(0.5, dh.ExtendedIInterpolation(0.3)),
                       smoother=dh.ILU0SmootherType(), max_coarse_size=10)
    h = _hier(dh, A, cfg)
    levels = [h.level(l) for l in range(h.nlevels)]
    c = h.coarsest()
    oh = orc.OracleHierarchy(levels, c["rp"], c["cv"], c["nz"], smoother="ilu0")
    dn = DenseAMG(levels, c, "ilu0")
    b = P.splitmix_normal(SEED + 93, A.n)
    x0 = P.splitmix_normal(SEED + 94, A.n)
    assert rel_err(oh.cycle(x0, b), dn.cycle(x0, b)) < 1e-11
    x, it, rel = oh.solve(np.zeros(A.n), b, 1e-8, 200)
    assert rel < 1e-8 and it < 50
