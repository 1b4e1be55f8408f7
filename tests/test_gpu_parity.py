"""GPU parity: every hot-path kernel, the V/W cycle, the stationary solve, the
FP64 / mixed-precision PCG and the partial resetup against the FP64 CPU
oracle, on the same hierarchy and the same seeded inputs.

Bar: FP64 kernels, cycles and resetup values are bit-identical to the oracle
(both sum every row in the reference's sequential order, no FMA contraction).
Solvers whose dot products are reduced in a different order (PCG,
stationary norms) agree to 1e-10 relative with iteration counts within 1.
FP32 kernels: 1e-5 relative 2-norm; BF16-value kernels: 1e-2 (8-bit
mantissa storage).
"""
import numpy as np
import pytest

from conftest import oracle_of, rel_err

pytestmark = pytest.mark.gpu

SEED = 20261202


def _case(dh, name):
    P = dh.problems
    if name == "agg2d":
        A = P.poisson2d(48)
        cfg = dh.AMGConfig(coarsening=dh.AggregationCoarsening(0.25), smoother=dh.JacobiSmootherType())
        sm = "jacobi"
    elif name == "hypre3d":
        A = P.heterogeneous3d(20, P.log_permeability(20, 1.5, SEED + 1))
        cfg = dh.hypre_default_config(smoother=dh.L1JacobiSmootherType())
        sm = "l1_jacobi"
    elif name == "gs2d":
        A = P.poisson2d(40)
        cfg = dh.AMGConfig(coarsening=dh.PMISCoarsening(0.25), smoother=dh.ColoredGaussSeidelType())
        sm = "colored_gs"
    elif name == "l1gs3d":
        A = P.heterogeneous3d(16, P.log_permeability(16, 1.0, SEED + 3))
        cfg = dh.AMGConfig(coarsening=dh.HMISCoarsening(0.5, dh.ExtendedIInterpolation(0.3)),
                           smoother=dh.L1ColoredGaussSeidelType())
        sm = "l1_colored_gs"
    elif name == "wide2d":
        # 2-D Poisson (73,984 rows) plus a coupling of the first and last rows:
        # their row blocks span more than 65,535 columns, so the 16-bit column
        # form reads them through the 32-bit columns
        import scipy.sparse as sp
        B = P.poisson2d(272).to_scipy().tolil()
        n = B.shape[0]
        B[0, n - 1] = B[n - 1, 0] = -0.5
        B[0, 0] += 0.5
        B[n - 1, n - 1] += 0.5
        B = B.tocsr()
        B.sort_indices()
        A = dh.as_csr(B)
        cfg = dh.AMGConfig(coarsening=dh.AggregationCoarsening(0.25), smoother=dh.JacobiSmootherType())
        sm = "jacobi"
    elif name == "hub2d":
        # 2-D Poisson (4,096 rows) plus one hub row coupled to 2,999 others: the
        # hub row (3,000 nonzeros) is longer than a row block's LDS stage, so
        # the stream kernels sum it chunk by chunk
        A = dh.as_csr(hub_matrix(P, 64, 3000))
        cfg = dh.AMGConfig(coarsening=dh.AggregationCoarsening(0.25), smoother=dh.JacobiSmootherType())
        sm = "jacobi"
    else:
        raise KeyError(name)
    return A, cfg, sm


def hub_matrix(P, m, hub):
    """2-D Poisson m x m with row/column 0 coupled to rows 1..hub-1 by -0.01
    (diagonals raised to keep the matrix SPD and diagonally dominant)."""
    import scipy.sparse as sp
    B = P.poisson2d(m).to_scipy().tolil()
    for j in range(1, hub):
        B[0, j] = B[0, j] - 0.01
        B[j, 0] = B[j, 0] - 0.01
        B[j, j] = B[j, j] + 0.01
    B[0, 0] = B[0, 0] + 0.01 * (hub - 1)
    B = B.tocsr()
    B.sort_indices()
    return B


CASES = ["agg2d", "hypre3d", "gs2d", "l1gs3d"]
_cache = {}


def _built(dh, orc, name):
    if name not in _cache:
        A, cfg, sm = _case(dh, name)
        h = dh.amg_setup(A, cfg)
        oh = oracle_of(h, orc, sm)
        _cache[name] = (A, cfg, sm, h, oh)
    return _cache[name]


TOL = {"fp32": 1e-5, "bf16": 1e-2}


def same(got, ref, prec="fp64"):
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    if prec == "fp64":
        bad = np.flatnonzero(got != ref)
        assert bad.size == 0, "%d of %d entries differ, first at %d: %r vs %r (rel err %.3e)" % (
            bad.size, ref.size, bad[0], got[bad[0]], ref[bad[0]], rel_err(got, ref))
    else:
        assert rel_err(got, ref) < TOL[prec], rel_err(got, ref)


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16"])
@pytest.mark.parametrize("name", CASES)
def test_level_kernels(dh, orc, torch_cuda, name, prec):
    torch = torch_cuda
    A, cfg, sm, h, oh = _built(dh, orc, name)
    tdt = torch.float64 if prec == "fp64" else torch.float32
    w = cfg.jacobi_omega
    rng = np.random.default_rng(SEED)
    for l in range(h.nlevels):
        lv = oh.levels[l]
        n, nc = len(lv["A_rp"]) - 1, len(lv["Pt_off"]) - 1
        x, b, xc = rng.standard_normal(n), rng.standard_normal(n), rng.standard_normal(nc)
        dev = lambda a: torch.tensor(a, dtype=tdt, device="cuda")  # noqa: E731
        out, out2 = torch.zeros(n, dtype=tdt, device="cuda"), torch.zeros(n, dtype=tdt, device="cuda")
        d = oh.dinv(l)
        rp, cv, nz = lv["A_rp"], lv["A_cv"], lv["A_nz"]
        print("case", name, prec, "level", l, "n", n, flush=True)
        h.level_op(l, 0, prec, dev(x), None, out)
        same(out.cpu().numpy(), orc.spmv(rp, cv, nz, x), prec)
        h.level_op(l, 1, prec, dev(x), dev(b), out)
        same(out.cpu().numpy(), orc.residual(rp, cv, nz, x, b), prec)
        if sm in ("jacobi", "l1_jacobi"):
            h.level_op(l, 2, prec, dev(x), dev(b), out)
            same(out.cpu().numpy(), orc.jacobi_step(rp, cv, nz, x, b, d, w), prec)
            h.level_op(l, 3, prec, None, dev(b), out, out2)
            x1 = orc.jacobi_step(rp, cv, nz, np.zeros(n), b, d, w)
            same(out.cpu().numpy(), x1, prec)
            same(out2.cpu().numpy(), orc.residual(rp, cv, nz, x1, b), prec)
        else:
            xg = dev(x)
            h.level_op(l, 6, prec, None, dev(b), xg)
            ref = orc.gs_sweep(rp, cv, nz, x, b, d, lv["c_off"], lv["c_order"])
            same(xg.cpu().numpy(), ref, prec)
        bc = torch.zeros(nc, dtype=tdt, device="cuda")
        h.level_op(l, 4, prec, dev(x), None, bc)
        same(bc.cpu().numpy(), orc.restrict(lv["Pt_off"], lv["Pt_rows"], lv["Pt_pnz"], lv["P_nz"], x), prec)
        xo = dev(x)
        h.level_op(l, 5, prec, dev(xc), None, xo)
        same(xo.cpu().numpy(), orc.prolong(lv["P_rp"], lv["P_cv"], lv["P_nz"], xc, x), prec)
    if prec == "fp64":
        for l in range(h.nlevels):
            n = oh.levels[l]["A_rp"].shape[0] - 1
            dd = torch.zeros(n, dtype=torch.float64, device="cuda")
            h.level_op(l, 9, 0, None, None, dd)
            same(dd.cpu().numpy(), oh.dinv(l))
    # coarse direct solve
    c = h.coarsest()
    bc = rng.standard_normal(c["n"])
    out = torch.zeros(c["n"], dtype=tdt, device="cuda")
    h.level_op(0, 7, prec, torch.tensor(bc, dtype=tdt, device="cuda"), None, out)
    # pinned twice: the oracle's LAPACK-order LU solve (bit for bit in FP64) and
    # numpy's LAPACK solve (to rounding)
    assert rel_err(out.cpu().numpy(), np.linalg.solve(c["dense"], bc)) < (1e-11 if prec == "fp64" else 1e-3)
    if prec == "fp64":
        same(out.cpu().numpy(), orc.lu_solve_dense(c["dense"], bc))


@pytest.mark.parametrize("name", CASES)
def test_galerkin_matches_oracle(dh, orc, torch_cuda, name):
    A, cfg, sm, h, oh = _built(dh, orc, name)
    for l in range(h.nlevels):
        lv = oh.levels[l]
        ref = orc.galerkin(lv["R_off"], lv["R_pi"], lv["R_ai"], lv["R_pj"], lv["P_nz"], lv["A_nz"])
        nxt = h.level_info(l + 1)["nnz"] if l + 1 < h.nlevels else h.coarsest()["nnz"]
        h.level_op(l, 8, 0)
        got = h.device_values(l + 1, 0, nxt) if l + 1 < h.nlevels else h.device_values(0, 3, nxt)
        same(got, ref)
        # and the setup-time values are the same sums in the same order
        same(oh.level_values(l + 1), ref)


@pytest.mark.parametrize("cycle", ["v", "w"])
@pytest.mark.parametrize("name", CASES)
def test_cycle_fp64(dh, orc, name, cycle):
    A, cfg, sm, h, oh = _built(dh, orc, name)
    oh.set_options(cycle=cycle, pre=1, post=1)
    c2 = dh.AMGConfig(**{**cfg.__dict__, "cycle_type": cycle.upper()})
    rng = np.random.default_rng(SEED + 7)
    x0, b = rng.standard_normal(A.n), rng.standard_normal(A.n)
    x = x0.copy()
    dh.amg_cycle_(x, b, h, c2)
    same(x, oh.cycle(x0, b))
    oh.set_options(cycle="v")


@pytest.mark.parametrize("steps", [(2, 1), (0, 2), (3, 3)])
def test_cycle_pre_post_steps(dh, orc, steps):
    A, cfg, sm, h, oh = _built(dh, orc, "hypre3d")
    pre, post = steps
    oh.set_options(pre=pre, post=post)
    c2 = dh.AMGConfig(**{**cfg.__dict__, "pre_smoothing_steps": pre, "post_smoothing_steps": post})
    rng = np.random.default_rng(SEED + 11)
    x0, b = rng.standard_normal(A.n), rng.standard_normal(A.n)
    x = x0.copy()
    dh.amg_cycle_(x, b, h, c2)
    same(x, oh.cycle(x0, b))
    # preconditioner form (zero guess) on device vectors
    import torch
    xd = torch.zeros(A.n, dtype=torch.float64, device="cuda")
    dh.amg_cycle_(xd, torch.tensor(b, device="cuda"), h, c2, zero_guess=True)
    same(xd.cpu().numpy(), oh.cycle(np.zeros(A.n), b))
    oh.set_options(pre=1, post=1)


@pytest.mark.parametrize("name", ["agg2d", "hypre3d", "l1gs3d"])
def test_stationary_solve(dh, orc, name):
    A, cfg, sm, h, oh = _built(dh, orc, name)
    b = 2.0 * dh.problems.splitmix_uniform(SEED, A.n) - 1.0
    x = np.zeros(A.n)
    x, it = dh.amg_solve_(x, b, h, cfg, tol=1e-8, maxiter=300)
    xr, itr, rel = oh.solve(np.zeros(A.n), b, 1e-8, 300)
    assert abs(it - itr) <= 1
    assert rel_err(x, xr) < 1e-9
    assert np.linalg.norm(b - A.matvec(x)) / np.linalg.norm(b) < 1e-8 or it == 300


@pytest.mark.parametrize("name", ["agg2d", "hypre3d", "l1gs3d", "gs2d"])
def test_pcg_fp64(dh, orc, torch_cuda, name):
    """FP64 PCG vs the oracle's.  Coloured GS cycles are symmetric only with
    the reverse-colour post-sweep (gs_symmetric): without it PCG refuses the
    preconditioner; with it the cycle is the oracle's bit for bit and PCG
    matches the oracle's PCG."""
    import dataclasses
    torch = torch_cuda
    A, cfg, sm, h, oh = _built(dh, orc, name)
    b = dh.problems.splitmix_normal(SEED + 5, A.n)
    if "gs" in sm:
        xd = torch.zeros(A.n, dtype=torch.float64, device="cuda")
        with pytest.raises(dh.DraugrError, match="symmetric preconditioner"):
            dh.pcg_solve_(xd, torch.tensor(b, device="cuda"), h, cfg, tol=1e-8, maxiter=500)
        cfg = dataclasses.replace(cfg, gs_symmetric=True)
        oh = oracle_of(h, orc, sm, gs_symmetric=True)
        z = torch.zeros(A.n, dtype=torch.float64, device="cuda")
        dh.amg_cycle_(z, torch.tensor(b, device="cuda"), h, cfg, zero_guess=True)
        same(z.cpu().numpy(), oh.cycle(np.zeros(A.n), b))
        for c in (cfg, dataclasses.replace(cfg, pre_smoothing_steps=1, post_smoothing_steps=2)):
            if c.pre_smoothing_steps != c.post_smoothing_steps:
                with pytest.raises(dh.DraugrError, match="must equal"):
                    dh.pcg_solve_(xd, torch.tensor(b, device="cuda"), h, c, tol=1e-8, maxiter=500)
    xd = torch.zeros(A.n, dtype=torch.float64, device="cuda")
    _, it, rel = dh.pcg_solve_(xd, torch.tensor(b, device="cuda"), h, cfg, tol=1e-8, maxiter=500)
    torch.cuda.synchronize()
    x = xd.cpu().numpy()
    xr, itr, relr = oh.pcg(np.zeros(A.n), b, 1e-8, 500)
    assert abs(it - itr) <= 1, (it, itr)
    assert rel_err(x, xr) < 1e-10
    assert np.linalg.norm(b - A.matvec(x)) / np.linalg.norm(b) < 1.01e-8
    # host-vector entry point gives the same
    xh = np.zeros(A.n)
    _, it2, _ = dh.pcg_solve_(xh, b, h, cfg, tol=1e-8, maxiter=500)
    assert it2 == it and rel_err(xh, x) < 1e-12


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
@pytest.mark.parametrize("name", ["agg2d", "hypre3d"])
def test_pcg_mixed(dh, orc, torch_cuda, name, prec):
    """FP32 / BF16-storage cycle inside FP64 PCG reaches the FP64 tolerance."""
    torch = torch_cuda
    A, cfg, sm, h, oh = _built(dh, orc, name)
    b = dh.problems.splitmix_normal(SEED + 6, A.n)
    c2 = dh.AMGConfig(**{**cfg.__dict__, "precision": prec})
    xd = torch.zeros(A.n, dtype=torch.float64, device="cuda")
    _, it, rel = dh.pcg_solve_(xd, torch.tensor(b, device="cuda"), h, c2, tol=1e-8, maxiter=500)
    x = xd.cpu().numpy()
    xr, itr, _ = oh.pcg(np.zeros(A.n), b, 1e-8, 500)
    true_rel = np.linalg.norm(b - A.matvec(x)) / np.linalg.norm(b)
    print(name, prec, "iters", it, "fp64 oracle iters", itr, "true relres", true_rel, "err", rel_err(x, xr))
    assert true_rel < 1.01e-8
    assert it <= itr + (2 if prec == "fp32" else 8)
    assert rel_err(x, xr) < 1e-6


@pytest.mark.parametrize("name", CASES)
def test_resetup_matches_oracle(dh, orc, torch_cuda, name):
    torch = torch_cuda
    A, cfg, sm, _, _ = _built(dh, orc, name)
    h = dh.amg_setup(A, cfg)
    oh = oracle_of(h, orc, sm)
    rng = np.random.default_rng(SEED + 9)
    # same pattern, new values: symmetric positive rescaling keeps an M-matrix
    s = np.exp(0.3 * rng.standard_normal(A.n))
    rows = np.repeat(np.arange(A.n), np.diff(A.rowptr))
    new = A.nzval * s[rows] * s[A.colval]
    dh.amg_resetup_(h, dh.problems.CSR(A.rowptr, A.colval, new), cfg)
    oh.resetup(new)
    for l in range(1, h.nlevels):
        same(h.device_values(l, 0, h.level_info(l)["nnz"]), oh.level_values(l))
    same(h.device_values(0, 3, h.coarsest()["nnz"]), oh.coarse["nz"])
    for l in range(h.nlevels):
        same(h.device_values(l, 1, h.level_info(l)["n"]), oh.dinv(l))
    b = rng.standard_normal(A.n)
    x = np.zeros(A.n)
    dh.amg_cycle_(x, b, h, cfg)
    same(x, oh.cycle(np.zeros(A.n), b))
    # resetup from device values
    new2 = A.nzval * 2.0
    dh.amg_resetup_(h, torch.tensor(new2, device="cuda"), cfg)
    oh.resetup(new2)
    x = np.zeros(A.n)
    dh.amg_cycle_(x, b, h, cfg)
    same(x, oh.cycle(np.zeros(A.n), b))


def test_trivial_systems(dh):
    """1x1 and 2x2 systems (test/runtests.jl:1798-1815): direct solve only."""
    A1 = dh.problems.CSR(np.array([0, 1], np.int32), np.array([0], np.int32), np.array([5.0]))
    cfg = dh.AMGConfig(max_coarse_size=10)
    h = dh.amg_setup(A1, cfg)
    x = np.zeros(1)
    x, it = dh.amg_solve_(x, np.array([3.0]), h, cfg, tol=1e-10)
    assert abs(x[0] - 0.6) < 1e-8 and it == 1
    A2 = dh.problems.CSR(np.array([0, 2, 4], np.int32), np.array([0, 1, 0, 1], np.int32),
                         np.array([4.0, -1.0, -1.0, 4.0]))
    h2 = dh.amg_setup(A2, cfg)
    b = np.array([1.0, 1.0])
    x = np.zeros(2)
    x, it = dh.amg_solve_(x, b, h2, cfg, tol=1e-10)
    assert np.linalg.norm(b - A2.matvec(x)) / np.linalg.norm(b) < 1e-10
    x, it = dh.amg_solve_(x, np.zeros(2), h2, cfg)
    assert it == 0 and not x.any()


def test_gpu_error_paths(dh):
    A = dh.problems.poisson2d(8)
    cfg = dh.AMGConfig(coarsening=dh.AggregationCoarsening(), smoother=dh.JacobiSmootherType())
    h = dh.amg_setup(A, cfg)
    import ctypes
    L = dh.lib()
    c = dh.core._Config(cfg)
    x = np.zeros(A.n - 1)
    assert L.draugr_amg_cycle(h.handle, A.n - 1, x.ctypes.data_as(ctypes.c_void_p),
                              x.ctypes.data_as(ctypes.c_void_p), c.handle) == -1
    assert "DimensionMismatch" in dh.last_error()
    with pytest.raises(dh.DraugrError, match="DimensionMismatch"):
        bad = dh.problems.poisson2d(9)
        dh.amg_resetup_(h, bad, cfg)
    h_np = dh.amg_setup(A, cfg, allow_partial_resetup=False)
    with pytest.raises(dh.DraugrError, match="restriction maps"):
        dh.amg_resetup_(h_np, A, cfg)
    # full resetup rebuilds
    dh.amg_resetup_(h_np, dh.problems.poisson2d(12), cfg, partial=False)
    assert h_np.n == 144 and h_np.info()["n_fine"] == 144


def test_preconditioner_adapter_with_scipy_cg(dh, torch_cuda):
    """DraugrPreconditioner (src/preconditioner.jl:29-102) as the M of scipy's
    CG: update_ sets up, then resets up; CG converges in about the PCG's
    iteration count."""
    import scipy.sparse.linalg as spla
    P = dh.problems
    A = P.heterogeneous3d(16, P.log_permeability(16, 1.0, SEED + 800))
    prec = dh.DraugrPreconditioner(dh.hypre_default_config(smoother=dh.L1JacobiSmootherType()))
    prec.update_(A)
    b = P.splitmix_normal(SEED + 801, A.n)
    its = []
    x, info = spla.cg(A.to_scipy(), b, rtol=1e-8, maxiter=200, M=dh.aslinearoperator(prec),
                      callback=lambda xk: its.append(1))
    assert info == 0 and np.linalg.norm(b - A.matvec(x)) / np.linalg.norm(b) < 1e-7
    assert len(its) < 40
    A2 = P.heterogeneous3d(16, P.log_permeability(16, 1.0, SEED + 802))
    prec.update_(A2)  # resetup path
    x2, info2 = spla.cg(A2.to_scipy(), b, rtol=1e-8, maxiter=200, M=dh.aslinearoperator(prec))
    assert info2 == 0


@pytest.mark.parametrize("kernel", ["stream", "rowseq", "pipe", "stream16", "sorted", "stream16pipe", "sortedpipe",
                                    "sortedquad", "pattern"])
@pytest.mark.parametrize("name", ["hypre3d", "agg2d", "wide2d", "hub2d"])
def test_every_kernel_choice_is_bit_exact(dh, orc, torch_cuda, kernel, name):
    """Each of the five sparse-pass kernels the autotuner chooses among,
    forced on every level, gives the oracle's FP64 cycle bit for bit, also
    after a resetup (new values: the reduced-precision and column-sorted
    copies follow), and the FP32-cycle PCG converges.  The 16-bit column form
    and its column-sorted copy are present on every level of these problems
    (wide blocks, if any, read the 32-bit columns in row order)."""
    import dataclasses
    torch = torch_cuda
    A, cfg, sm = _case(dh, name)
    c = dataclasses.replace(cfg, kernel=kernel)
    h = dh.amg_setup(A, c)
    if kernel in ("stream16", "stream16pipe") and name != "hub2d":
        assert all(h.level_info(l)["has16"] for l in range(h.nlevels))
    oh = oracle_of(h, orc, sm)
    b = dh.problems.splitmix_normal(SEED + 900, A.n)
    x0 = dh.problems.splitmix_normal(SEED + 901, A.n)
    x = torch.tensor(x0, device="cuda")
    dh.amg_cycle_(x, torch.tensor(b, device="cuda"), h, c)
    assert np.array_equal(x.cpu().numpy(), oh.cycle(x0, b))
    if kernel == "pattern":  # the grid operators have the form on their finest level; the hub row is too long
        assert ("level 0: A has the pattern form" in h.tune_report()) == (name != "hub2d")
    z = torch.zeros(A.n, dtype=torch.float64, device="cuda")
    c32 = dataclasses.replace(c, precision="fp32")
    _, it, rel = dh.pcg_solve_(z, torch.tensor(b, device="cuda"), h, c32, tol=1e-8)
    assert rel < 1e-8
    # new values, same pattern
    s = np.exp(0.3 * np.random.default_rng(SEED + 902).standard_normal(A.n))
    rows = np.repeat(np.arange(A.n), np.diff(A.rowptr))
    new = A.nzval * s[rows] * s[A.colval]
    dh.amg_resetup_(h, dh.problems.CSR(A.rowptr, A.colval, new), c)
    oh.resetup(new)
    x = torch.tensor(x0, device="cuda")
    dh.amg_cycle_(x, torch.tensor(b, device="cuda"), h, c)
    assert np.array_equal(x.cpu().numpy(), oh.cycle(x0, b))
    # the FP32 cycle after the resetup equals, bit for bit, the row-order
    # 32-bit-column kernel's on a second hierarchy with the same values
    cs = dataclasses.replace(cfg, kernel="stream")
    h2 = dh.amg_setup(A, cs)
    dh.amg_resetup_(h2, dh.problems.CSR(A.rowptr, A.colval, new), cs)
    y = [torch.tensor(x0, device="cuda") for _ in range(2)]
    dh.amg_cycle_(y[0], torch.tensor(b, device="cuda"), h, c32)
    dh.amg_cycle_(y[1], torch.tensor(b, device="cuda"), h2, dataclasses.replace(cs, precision="fp32"))
    assert np.array_equal(y[0].cpu().numpy(), y[1].cpu().numpy())
    # and so is the FP64 PCG (its q = A p pass reads the column-sorted form when that kernel is forced)
    out = []
    for hh, cc in ((h, c), (h2, cs)):
        zz = torch.zeros(A.n, dtype=torch.float64, device="cuda")
        _, it2, rel2 = dh.pcg_solve_(zz, torch.tensor(b, device="cuda"), hh, cc, tol=1e-8, maxiter=400)
        out.append((it2, rel2, zz.cpu().numpy()))
    assert out[0][0] == out[1][0] and out[0][1] == out[1][1] and np.array_equal(out[0][2], out[1][2])


def _row_patterns(A):
    """Distinct (length, col - row ...) tuples of a CSR matrix, counted
    independently of the library (numpy over the rows)."""
    rp, cv = np.asarray(A.rowptr), np.asarray(A.colval)
    return {tuple((cv[rp[i]:rp[i + 1]] - i).tolist()) for i in range(len(rp) - 1)}


def _six_point(dh, m):
    """2-D Poisson m x m plus an upwind coupling to row i - m - 1 (a
    nonsymmetric stencil of width 6, diagonally dominant)."""
    B = dh.problems.poisson2d(m).to_scipy().tolil()
    for i in range(m + 1, m * m):
        if i % m:
            B[i, i - m - 1] = -0.1
            B[i, i] = B[i, i] + 0.1
    B = B.tocsr()
    B.sort_indices()
    return dh.as_csr(B)


def _padded_chain(dh, n):
    """1-D Laplacian of n rows with row 0 coupled to rows 1..15: one row of 16
    entries among rows of 3, too much padding for the pattern form."""
    import scipy.sparse as sp
    B = sp.diags([-np.ones(n - 1), 2.0 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1]).tolil()
    for j in range(2, 16):
        B[0, j] = B[j, 0] = -0.01
        B[j, j] = B[j, j] + 0.01
    B[0, 0] = B[0, 0] + 0.01 * 14
    B = B.tocsr()
    B.sort_indices()
    return dh.as_csr(B)


@pytest.mark.parametrize("grid", ["7pt24", "5pt64", "6pt64", "padded"])
def test_pattern_form(dh, orc, torch_cuda, grid):
    """The pattern form (a 16-bit pattern index per row, values in 64-row
    slices, no column indices): present on the finest operator of a 3-D
    7-point grid (27 row patterns of width 7), of a 2-D 5-point grid (9 of
    width 5) and of a nonsymmetric 6-point stencil (an even width), the
    counts matching an independent count of the rows' col - row tuples;
    absent where the slices would pad too much (a chain with one 16-entry
    row, whose tridiagonal coarse levels have it again), the forced kind
    then falling back to the stream kernel.  Forced on every level, the FP64 cycle is the oracle's
    bit for bit and the FP32 and BF16 cycles are the stream kernel's bit for
    bit, before and after a resetup with new values; the FP64 PCG, whose
    q = A p pass then reads the pattern form, is the stream form's bit for bit."""
    import dataclasses
    torch = torch_cuda
    P = dh.problems
    agg = dh.AMGConfig(coarsening=dh.AggregationCoarsening(0.25), smoother=dh.JacobiSmootherType())
    if grid == "7pt24":
        A = P.heterogeneous3d(24, P.log_permeability(24, 1.5, SEED + 950))
        cfg = dh.hypre_default_config(smoother=dh.L1JacobiSmootherType())
        sm, want = "l1_jacobi", (27, 7)
    elif grid == "5pt64":
        A, cfg, sm, want = P.poisson2d(64), agg, "jacobi", (9, 5)
    elif grid == "6pt64":  # an even width: the BF16 pairs have no padding half
        A, cfg, sm = _six_point(dh, 64), agg, "jacobi"
        want = (len(_row_patterns(A)), 6)
    else:  # 17 patterns, but 64-row slices of width 16 would pad the values 5x: no pattern form
        A, cfg, sm, want = _padded_chain(dh, 4000), agg, "jacobi", None
    pats = _row_patterns(A)
    if want is not None:
        assert (len(pats), max(len(t) for t in pats)) == want and want[0] <= 64
    else:
        assert len(pats) <= 64 and max(len(t) for t in pats) == 16
    cp = dataclasses.replace(cfg, kernel="pattern")
    cs = dataclasses.replace(cfg, kernel="stream")
    hp, hs = dh.amg_setup(A, cp), dh.amg_setup(A, cs)
    oh = oracle_of(hp, orc, sm)
    b = torch.tensor(P.splitmix_normal(SEED + 951, A.n), device="cuda")
    x0 = P.splitmix_normal(SEED + 952, A.n)
    for stage in ("setup", "resetup"):
        if stage == "resetup":
            s = np.exp(0.3 * np.random.default_rng(SEED + 953).standard_normal(A.n))
            rows = np.repeat(np.arange(A.n), np.diff(A.rowptr))
            new = A.nzval * s[rows] * s[A.colval]
            for hh, cc in ((hp, cp), (hs, cs)):
                dh.amg_resetup_(hh, P.CSR(A.rowptr, A.colval, new), cc)
            oh.resetup(new)
        x = torch.tensor(x0, device="cuda")
        dh.amg_cycle_(x, b, hp, cp)
        assert np.array_equal(x.cpu().numpy(), oh.cycle(x0, b.cpu().numpy()))
        if stage == "setup":
            if want is not None:
                assert "level 0: A has the pattern form (%d row patterns, width %d)" % want in hp.tune_report()
            else:
                assert "level 0: A has the pattern form" not in hp.tune_report()
        for prec in ("fp32", "bf16"):
            y = []
            for hh, cc in ((hp, cp), (hs, cs)):
                yy = torch.tensor(x0, device="cuda")
                dh.amg_cycle_(yy, b, hh, dataclasses.replace(cc, precision=prec))
                y.append(yy.cpu().numpy())
            assert np.array_equal(y[0], y[1]), (stage, prec)
        if grid == "6pt64":  # nonsymmetric: no CG
            continue
        out = []
        for hh, cc in ((hp, cp), (hs, cs)):
            zz = torch.zeros(A.n, dtype=torch.float64, device="cuda")
            _, it, rel = dh.pcg_solve_(zz, b, hh, cc, tol=1e-8, maxiter=400)
            out.append((it, rel, zz.cpu().numpy()))
        assert out[0][0] == out[1][0] and out[0][1] == out[1][1] and np.array_equal(out[0][2], out[1][2])
        if want is not None:  # (aggregation AMG on the 1-D chain is a weak preconditioner)
            assert out[0][1] < 1e-8
    if grid != "6pt64":
        assert ("pcg q = A p: pattern" in hp.tune_report()) == (want is not None)


@pytest.mark.parametrize("name", ["hypre3d", "agg2d", "hub2d"])
def test_vector_kernel(dh, orc, torch_cuda, name):
    """The vector CSR kernel (lanes per row chosen by row length, partial sums
    combined by shuffles; FP32 / BF16 only): forced on every level, the FP64
    cycle keeps the stream kernel and is the oracle's bit for bit; the level
    kernels in FP32 are within 1e-5 of the oracle's; the FP32 and BF16 cycles
    are within 1e-5 / 1e-2 of the FP64 cycle; their PCG reaches 1e-8 with at
    most one iteration more than the FP64 cycle's."""
    import dataclasses
    torch = torch_cuda
    A, cfg, sm = _case(dh, name)
    c = dataclasses.replace(cfg, kernel="vector")
    h = dh.amg_setup(A, c)
    oh = oracle_of(h, orc, sm)
    b = dh.problems.splitmix_normal(SEED + 990, A.n)
    x0 = dh.problems.splitmix_normal(SEED + 991, A.n)
    x = torch.tensor(x0, device="cuda")
    dh.amg_cycle_(x, torch.tensor(b, device="cuda"), h, c)
    ref = oh.cycle(x0, b)
    assert np.array_equal(x.cpu().numpy(), ref)
    rp, cv, nz = A.rowptr, A.colval, A.nzval
    d = oh.dinv(0)
    f32 = lambda a: torch.tensor(a, dtype=torch.float32, device="cuda")  # noqa: E731
    out = torch.zeros(A.n, dtype=torch.float32, device="cuda")
    h.level_op(0, 1, "fp32", f32(x0), f32(b), out)
    same(out.cpu().numpy(), orc.residual(rp, cv, nz, x0, b), "fp32")
    h.level_op(0, 2, "fp32", f32(x0), f32(b), out)
    same(out.cpu().numpy(), orc.jacobi_step(rp, cv, nz, x0, b, d, cfg.jacobi_omega), "fp32")
    z64 = torch.zeros(A.n, dtype=torch.float64, device="cuda")
    _, it64, _ = dh.pcg_solve_(z64, torch.tensor(b, device="cuda"), h, c, tol=1e-8, maxiter=400)
    for prec, tol in (("fp32", 1e-5), ("bf16", 1e-2)):
        cp = dataclasses.replace(c, precision=prec)
        x = torch.tensor(x0, device="cuda")
        dh.amg_cycle_(x, torch.tensor(b, device="cuda"), h, cp)
        assert rel_err(x.cpu().numpy(), ref) < tol, prec
        z = torch.zeros(A.n, dtype=torch.float64, device="cuda")
        _, it, rel = dh.pcg_solve_(z, torch.tensor(b, device="cuda"), h, cp, tol=1e-8, maxiter=400)
        assert rel < 1e-8 and it <= it64 + 1 + (prec == "bf16") * 6, (prec, it, it64)


@pytest.mark.parametrize("name", ["hypre3d", "agg2d"])
@pytest.mark.parametrize("steps", [(1, 1), (2, 3), (1, 2)])
def test_one_pass_sweeps_are_bit_exact(dh, orc, torch_cuda, name, steps):
    """A level's first smoothing step from zero + residual + restriction in one
    pass, and its prolongation + first post-smoothing step in one pass
    (config fuse = "on", every level), give the oracle's FP64 cycle bit for
    bit and the separate passes' FP32 / BF16 cycles (and PCG, where the step
    counts are equal) bit for bit, for V- and W-cycles."""
    import dataclasses
    torch = torch_cuda
    A, cfg, sm = _case(dh, name)
    # one row-order kernel on both hierarchies, so that they differ only in the
    # fusions (each autotuner could pick the vector CSR kernel, whose FP32 sums
    # are in another order, on different levels)
    cfg = dataclasses.replace(cfg, pre_smoothing_steps=steps[0], post_smoothing_steps=steps[1], kernel="sorted")
    b = torch.tensor(dh.problems.splitmix_normal(SEED + 960, A.n), device="cuda")
    x0 = torch.tensor(dh.problems.splitmix_normal(SEED + 961, A.n), device="cuda")

    def run(c):
        h = dh.amg_setup(A, c)
        out = []
        for prec in ("fp64", "fp32", "bf16"):
            for cyc in ("V", "W"):
                cp = dataclasses.replace(c, precision=prec, cycle_type=cyc)
                x = x0.clone()
                dh.amg_cycle_(x, b, h, cp)
                z = torch.zeros_like(b)
                dh.amg_cycle_(z, b, h, cp, zero_guess=True)
                out.append((x.cpu().numpy(), z.cpu().numpy()))
            if steps[0] == steps[1]:  # PCG needs the symmetric cycle of equal step counts
                w = torch.zeros_like(b)
                _, it, rel = dh.pcg_solve_(w, b, h, dataclasses.replace(c, precision=prec), tol=1e-8)
                out.append((w.cpu().numpy(), np.array([it, rel])))
        return h, out

    h_on, on = run(dataclasses.replace(cfg, fuse="on"))
    assert "down=one-pass up=one-pass" in h_on.tune_report()
    _, off = run(dataclasses.replace(cfg, fuse="off"))
    for a, c in zip(on, off):
        for u, v in zip(a, c):
            assert u.tobytes() == v.tobytes(), (name, steps)
    # and the oracle's FP64 V-cycle from a nonzero guess and from zero
    oh = oracle_of(h_on, orc, sm, pre=steps[0], post=steps[1])
    assert np.array_equal(on[0][0], oh.cycle(x0.cpu().numpy(), b.cpu().numpy()))
    assert np.array_equal(on[0][1], oh.cycle(np.zeros(A.n), b.cpu().numpy()))
