"""GPU parity of the branches that the configuration problems do not reach:

* the coarsest direct solve (src/cycle.jl:88-93, src/setup.jl:177-209) for
  coarsest sizes above one wavefront: factors staged in LDS (n_c <= 87), the
  vector alone in LDS (n_c of a few hundred), the vector in a global
  workspace (any size), and the column-by-column factorisation used above 256
  rows -- each bit for bit against the oracle's LAPACK-order LU solve and,
  through a whole V-cycle, the oracle's cycle;
* rows longer than a row block's LDS stage (hub rows), summed chunk by chunk
  in row order: level kernels bit for bit against the oracle;
* row blocks whose four-entry quads exceed the LDS stage (a block within
  three nonzeros of the budget starting off a multiple of four): the
  four-entry kernel's entry-by-entry fallback, bit for bit against the oracle;
* row blocks spanning more than 65,536 columns (no 16-bit form) but fewer
  than 2^20: the column-sorted kernels read them through 20-bit offsets;
  level kernels bit for bit against the oracle;
* the stationary solve's non-convergence semantics (src/cycle.jl:135-159):
  a non-finite residual never passes the test, the iteration count returned
  is maxiter and no error is raised.
"""
import numpy as np
import pytest

from conftest import oracle_of, rel_err
from test_gpu_parity import hub_matrix, same

pytestmark = pytest.mark.gpu

SEED = 20261202


def _two_level(dh, m):
    """2-D Poisson m x m, aggregation, one AMG level: the coarsest operator is
    the aggregates' Galerkin product (69 rows at m = 20, 85 at 22, 288 at 40,
    655 at 60)."""
    A = dh.problems.poisson2d(m)
    cfg = dh.AMGConfig(coarsening=dh.AggregationCoarsening(0.25), smoother=dh.JacobiSmootherType(), max_levels=2)
    return A, cfg


@pytest.mark.parametrize("m,nc", [(20, 69), (22, 85), (40, 288), (60, 655)])
def test_coarse_solve_above_one_wavefront(dh, orc, torch_cuda, m, nc):
    torch = torch_cuda
    A, cfg = _two_level(dh, m)
    h = dh.amg_setup(A, cfg)
    c = h.coarsest()
    assert c["n"] == nc and h.nlevels == 1
    print("coarse solve n_c =", nc, flush=True)
    rng = np.random.default_rng(SEED + nc)
    bc = rng.standard_normal(nc)
    ref = orc.lu_solve_dense(c["dense"], bc)
    assert rel_err(ref, np.linalg.solve(c["dense"], bc)) < 1e-12
    outs = {}
    for op in (7, 12):  # 12: the vector in a global workspace
        out = torch.zeros(nc, dtype=torch.float64, device="cuda")
        h.level_op(0, op, "fp64", torch.tensor(bc, device="cuda"), None, out)
        outs[op] = out.cpu().numpy()
        same(outs[op], ref)
        o32 = torch.zeros(nc, dtype=torch.float32, device="cuda")
        h.level_op(0, op, "fp32", torch.tensor(bc, dtype=torch.float32, device="cuda"), None, o32)
        assert rel_err(o32.cpu().numpy(), ref) < 1e-4
    # the column-by-column factorisation gives the same factors, bit for bit
    f1 = h.device_values(0, 2, nc * nc)
    h.level_op(0, 13, "fp64")
    f2 = h.device_values(0, 2, nc * nc)
    assert np.array_equal(f1, f2)
    out = torch.zeros(nc, dtype=torch.float64, device="cuda")
    h.level_op(0, 7, "fp64", torch.tensor(bc, device="cuda"), None, out)
    same(out.cpu().numpy(), ref)
    # a whole V-cycle through this coarsest level, and a converged solve
    oh = oracle_of(h, orc, "jacobi")
    b = dh.problems.splitmix_normal(SEED + 970, A.n)
    x0 = dh.problems.splitmix_normal(SEED + 971, A.n)
    x = torch.tensor(x0, device="cuda")
    dh.amg_cycle_(x, torch.tensor(b, device="cuda"), h, cfg)
    same(x.cpu().numpy(), oh.cycle(x0, b))
    z = torch.zeros(A.n, dtype=torch.float64, device="cuda")
    _, it, rel = dh.pcg_solve_(z, torch.tensor(b, device="cuda"), h, cfg, tol=1e-10, maxiter=200)
    xo, ito, _ = oh.pcg(np.zeros(A.n), b, 1e-10, 200)
    assert abs(it - ito) <= 1 and rel_err(z.cpu().numpy(), xo) < 1e-10


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_long_row_level_kernels(dh, orc, torch_cuda, prec):
    """A 3,000-nonzero hub row on the finest level (longer than the 2,048
    products a row block stages): SpMV, residual, Jacobi and the fused
    zero-guess step are the oracle's bit for bit in FP64 (1e-5 in FP32),
    whichever sparse-pass kernel is forced."""
    import dataclasses
    torch = torch_cuda
    B = hub_matrix(dh.problems, 64, 3000)
    A = dh.as_csr(B)
    assert np.diff(A.rowptr).max() == 3000
    base = dh.AMGConfig(coarsening=dh.AggregationCoarsening(0.25), smoother=dh.JacobiSmootherType())
    tdt = torch.float64 if prec == "fp64" else torch.float32
    rng = np.random.default_rng(SEED + 975)
    x, b = rng.standard_normal(A.n), rng.standard_normal(A.n)
    rp, cv, nz = A.rowptr, A.colval, A.nzval
    for kernel in ("stream", "stream16", "sorted", "pipe", "rowseq", "stream16pipe", "sortedpipe", "sortedquad",
                   "vector"):
        cfg = dataclasses.replace(base, kernel=kernel)
        h = dh.amg_setup(A, cfg)
        oh = oracle_of(h, orc, "jacobi")
        d = oh.dinv(0)
        dev = lambda a: torch.tensor(a, dtype=tdt, device="cuda")  # noqa: E731
        out, out2 = torch.zeros(A.n, dtype=tdt, device="cuda"), torch.zeros(A.n, dtype=tdt, device="cuda")
        print("long row", kernel, prec, flush=True)
        h.level_op(0, 0, prec, dev(x), None, out)
        same(out.cpu().numpy(), orc.spmv(rp, cv, nz, x), prec)
        h.level_op(0, 1, prec, dev(x), dev(b), out)
        same(out.cpu().numpy(), orc.residual(rp, cv, nz, x, b), prec)
        h.level_op(0, 2, prec, dev(x), dev(b), out)
        same(out.cpu().numpy(), orc.jacobi_step(rp, cv, nz, x, b, d, cfg.jacobi_omega), prec)
        h.level_op(0, 3, prec, None, dev(b), out, out2)
        x1 = orc.jacobi_step(rp, cv, nz, np.zeros(A.n), b, d, cfg.jacobi_omega)
        same(out.cpu().numpy(), x1, prec)
        same(out2.cpu().numpy(), orc.residual(rp, cv, nz, x1, b), prec)


def test_stationary_solve_nonfinite_returns_maxiter(dh, torch_cuda):
    """A right-hand side with a NaN: ||b|| is NaN, no residual passes
    rel < tol, and amg_solve! returns (x, maxiter) without an error."""
    torch = torch_cuda
    P = dh.problems
    A = P.heterogeneous3d(10, P.log_permeability(10, 1.0, SEED + 980))
    cfg = dh.hypre_default_config(smoother=dh.L1JacobiSmootherType())
    h = dh.amg_setup(A, cfg)
    b = P.splitmix_normal(SEED + 981, A.n)
    b[17] = np.nan
    x = np.zeros(A.n)
    _, it = dh.amg_solve_(x, b, h, cfg, tol=1e-8, maxiter=4)
    assert it == 4
    xd = torch.zeros(A.n, dtype=torch.float64, device="cuda")
    _, it = dh.amg_solve_(xd, torch.tensor(b, device="cuda"), h, cfg, tol=1e-8, maxiter=3)
    assert it == 3
    # the same hierarchy still solves a finite system afterwards
    b[17] = 0.0
    x = np.zeros(A.n)
    _, it = dh.amg_solve_(x, b, h, cfg, tol=1e-8, maxiter=100)
    assert it < 100 and np.linalg.norm(b - A.matvec(x)) / np.linalg.norm(b) < 1e-8


def test_resetup_rejects_wrong_device_values(dh, torch_cuda):
    """amg_resetup_ with a tensor of values checks dtype, device and layout in
    Python before any device call."""
    torch = torch_cuda
    A = dh.problems.poisson2d(12)
    cfg = dh.AMGConfig(coarsening=dh.AggregationCoarsening(0.25), smoother=dh.JacobiSmootherType())
    h = dh.amg_setup(A, cfg)
    nz = torch.tensor(A.nzval, device="cuda")
    with pytest.raises(TypeError):
        dh.amg_resetup_(h, nz.float(), cfg)
    with pytest.raises(TypeError):
        dh.amg_resetup_(h, nz.cpu(), cfg)
    with pytest.raises(TypeError):
        dh.amg_resetup_(h, torch.stack([nz, nz], 1)[:, 0], cfg)  # strided view
    dh.amg_resetup_(h, nz, cfg)  # the valid form still works


@pytest.mark.parametrize("fuse", ["auto", "off", "on"])
@pytest.mark.parametrize("name", ["hypre3d", "agg2d", "wide2d", "hub2d"])
def test_tail_in_one_launch_is_bit_exact(dh, orc, torch_cuda, name, fuse):
    """The last AMG level and the coarse solve in one single-workgroup launch
    from LDS (config tail = "on") give the oracle's FP64 cycle bit for bit,
    V-cycles with several step counts (the W-cycle keeps the per-level
    launches), and the separate launches' FP32 / BF16 cycles bit for bit; PCG
    matches the oracle's; a resetup refreshes the LDS image.  With the
    one-pass sweeps off the parent level's restriction and prolongation ride
    in the tail's launch; with them on (where they apply) the parent's own
    one-pass launches do them."""
    import dataclasses
    from test_gpu_parity import _case
    torch = torch_cuda
    A, cfg, sm = _case(dh, name)
    # one row-order kernel on both hierarchies, so that the two differ only in
    # the tail (each autotuner could pick the vector CSR kernel, whose FP32
    # sums are in another order, on different levels)
    cfg = dataclasses.replace(cfg, kernel="sorted", fuse=fuse)
    on, off = dataclasses.replace(cfg, tail="on"), dataclasses.replace(cfg, tail="off")
    h_on, h_off = dh.amg_setup(A, on), dh.amg_setup(A, off)
    oh = oracle_of(h_on, orc, sm)
    b = torch.tensor(dh.problems.splitmix_normal(SEED + 990, A.n), device="cuda")
    x0 = torch.tensor(dh.problems.splitmix_normal(SEED + 991, A.n), device="cuda")
    for steps in ((1, 1), (2, 2), (0, 2), (2, 1)):
        for cyc in ("V", "W"):
            c = dataclasses.replace(on, pre_smoothing_steps=steps[0], post_smoothing_steps=steps[1], cycle_type=cyc)
            oh.set_options(cycle=cyc.lower(), pre=steps[0], post=steps[1])
            x = x0.clone()
            dh.amg_cycle_(x, b, h_on, c)
            z = torch.zeros_like(b)
            dh.amg_cycle_(z, b, h_on, c, zero_guess=True)
            torch.cuda.synchronize()
            print("tail", name, steps, cyc, flush=True)
            same(x.cpu().numpy(), oh.cycle(x0.cpu().numpy(), b.cpu().numpy()))
            same(z.cpu().numpy(), oh.cycle(np.zeros(A.n), b.cpu().numpy()))
            for prec in ("fp32", "bf16"):
                outs = []
                for h, base in ((h_on, on), (h_off, off)):
                    cp = dataclasses.replace(base, precision=prec, pre_smoothing_steps=steps[0],
                                             post_smoothing_steps=steps[1], cycle_type=cyc)
                    zz = torch.zeros_like(b)
                    dh.amg_cycle_(zz, b, h, cp, zero_guess=True)
                    outs.append(zz.cpu().numpy())
                assert np.array_equal(outs[0], outs[1]), (prec, steps, cyc)
    assert "tail: level" in h_on.tune_report() and "one launch from LDS" in h_on.tune_report()
    oh.set_options(cycle="v", pre=1, post=1)
    xd = torch.zeros_like(b)
    _, it, rel = dh.pcg_solve_(xd, b, h_on, on, tol=1e-10, maxiter=300)
    xo, ito, _ = oh.pcg(np.zeros(A.n), b.cpu().numpy(), 1e-10, 300)
    assert abs(it - ito) <= 1 and rel_err(xd.cpu().numpy(), xo) < 1e-10
    # new values on the same pattern: the image follows the resetup
    s = np.exp(0.3 * np.random.default_rng(SEED + 992).standard_normal(A.n))
    rows = np.repeat(np.arange(A.n), np.diff(A.rowptr))
    new = A.nzval * s[rows] * s[A.colval]
    dh.amg_resetup_(h_on, dh.problems.CSR(A.rowptr, A.colval, new), on)
    oh.resetup(new)
    for prec in ("fp64", "fp32"):
        z = torch.zeros_like(b)
        dh.amg_cycle_(z, b, h_on, dataclasses.replace(on, precision=prec), zero_guess=True)
        ref = oh.cycle(np.zeros(A.n), b.cpu().numpy())
        if prec == "fp64":
            same(z.cpu().numpy(), ref)
        else:
            assert rel_err(z.cpu().numpy(), ref) < 1e-5


def _wide_band(dh, n, w):
    """SPD, diagonally dominant: 4 on the diagonal, -1 at i +- 1 and i +- w.
    With w = 70,001 every row block spans about 2 w columns."""
    import scipy.sparse as sp
    off = [np.full(n - 1, -1.0), np.full(n - 1, -1.0), np.full(n - w, -1.0), np.full(n - w, -1.0)]
    B = sp.diags([np.full(n, 4.0)] + off, [0, 1, -1, w, -w], shape=(n, n), format="csr")
    B.sort_indices()
    return dh.as_csr(B)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_wide_blocks_column_sorted_kernels(dh, orc, torch_cuda, prec):
    """A 500,000-row band matrix whose row blocks span ~140,000 columns: no
    block has the 16-bit form, every one the column-sorted form (20-bit
    offsets).  SpMV, residual, Jacobi and the fused zero-guess step under the
    three column-sorted kernel kinds (budget 1024: the four-entry kernel runs)
    are the oracle's bit for bit in FP64 (1e-5 in FP32)."""
    import dataclasses
    torch = torch_cuda
    A = _wide_band(dh, 500_000, 70_001)
    base = dh.AMGConfig(coarsening=dh.AggregationCoarsening(0.25), smoother=dh.JacobiSmootherType())
    tdt = torch.float64 if prec == "fp64" else torch.float32
    rng = np.random.default_rng(SEED + 985)
    x, b = rng.standard_normal(A.n), rng.standard_normal(A.n)
    rp, cv, nz = A.rowptr, A.colval, A.nzval
    for kernel in ("sorted", "sortedquad", "sortedpipe"):
        cfg = dataclasses.replace(base, kernel=kernel)
        h = dh.amg_setup(A, cfg)
        li = h.level_info(0)
        assert li["has16"] == 0 and li["nnz_sorted"] == A.nnz, li
        oh = oracle_of(h, orc, "jacobi")
        d = oh.dinv(0)
        dev = lambda a: torch.tensor(a, dtype=tdt, device="cuda")  # noqa: E731
        out, out2 = torch.zeros(A.n, dtype=tdt, device="cuda"), torch.zeros(A.n, dtype=tdt, device="cuda")
        print("wide blocks", kernel, prec, flush=True)
        h.level_op(0, 0, prec, dev(x), None, out)
        same(out.cpu().numpy(), orc.spmv(rp, cv, nz, x), prec)
        h.level_op(0, 1, prec, dev(x), dev(b), out)
        same(out.cpu().numpy(), orc.residual(rp, cv, nz, x, b), prec)
        h.level_op(0, 2, prec, dev(x), dev(b), out)
        same(out.cpu().numpy(), orc.jacobi_step(rp, cv, nz, x, b, d, cfg.jacobi_omega), prec)
        h.level_op(0, 3, prec, None, dev(b), out, out2)
        x1 = orc.jacobi_step(rp, cv, nz, np.zeros(A.n), b, d, cfg.jacobi_omega)
        same(out.cpu().numpy(), x1, prec)
        same(out2.cpu().numpy(), orc.residual(rp, cv, nz, x1, b), prec)


def _misaligned_blocks(dh, n, w):
    """4 on the diagonal, -1 at i - 1, i + 1 and i + w: row 0 has three
    entries and most others four, so with the 1024-nonzero budget nearly
    every row block holds exactly 1024 entries starting 3 past a multiple of
    four, and its quads (257) exceed the stage (256 x 4)."""
    import scipy.sparse as sp
    idx = np.arange(n)
    rows, cols = [], []
    for off in (0, -1, 1, w):
        j = idx + off
        m = (j >= 0) & (j < n)
        rows.append(idx[m])
        cols.append(j[m])
    r, c = np.concatenate(rows), np.concatenate(cols)
    B = sp.csr_matrix((np.where(r == c, 4.0, -1.0), (r, c)), shape=(n, n))
    B.sort_indices()
    return dh.as_csr(B)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_quad_kernel_fallback_blocks(dh, orc, torch_cuda, prec):
    """Nearly every row block of this matrix takes the four-entry kernel's
    entry-by-entry path (checked on the host here): SpMV, residual, Jacobi
    and the fused zero-guess step with the sortedquad and sorted kinds are
    the oracle's bit for bit in FP64 (1e-5 in FP32)."""
    import dataclasses
    torch = torch_cuda
    A = _misaligned_blocks(dh, 600_000, 1000)
    rp, cv, nz = A.rowptr, A.colval, A.nzval
    # the host's view of the blocks: budget 1024, blocks of <= 256 rows
    lens = np.diff(rp)
    starts, acc, rows = [0], 0, 0
    for i, L in enumerate(lens):
        if rows > 0 and (acc + L > 1024 or rows == 256):
            starts.append(i)
            acc = rows = 0
        acc += L
        rows += 1
    z = rp[np.array(starts + [A.n])]
    nq = (z[1:] - (z[:-1] & ~3) + 3) >> 2
    assert np.mean(nq > 256) > 0.99
    base = dh.AMGConfig(coarsening=dh.AggregationCoarsening(0.25), smoother=dh.JacobiSmootherType())
    tdt = torch.float64 if prec == "fp64" else torch.float32
    rng = np.random.default_rng(SEED + 995)
    x, b = rng.standard_normal(A.n), rng.standard_normal(A.n)
    for kernel in ("sortedquad", "sorted"):
        cfg = dataclasses.replace(base, kernel=kernel)
        h = dh.amg_setup(A, cfg)
        assert h.level_info(0)["nblk"] == len(starts)
        oh = oracle_of(h, orc, "jacobi")
        d = oh.dinv(0)
        dev = lambda a: torch.tensor(a, dtype=tdt, device="cuda")  # noqa: E731
        out, out2 = torch.zeros(A.n, dtype=tdt, device="cuda"), torch.zeros(A.n, dtype=tdt, device="cuda")
        print("quad fallback", kernel, prec, flush=True)
        h.level_op(0, 0, prec, dev(x), None, out)
        same(out.cpu().numpy(), orc.spmv(rp, cv, nz, x), prec)
        h.level_op(0, 1, prec, dev(x), dev(b), out)
        same(out.cpu().numpy(), orc.residual(rp, cv, nz, x, b), prec)
        h.level_op(0, 2, prec, dev(x), dev(b), out)
        same(out.cpu().numpy(), orc.jacobi_step(rp, cv, nz, x, b, d, cfg.jacobi_omega), prec)
        h.level_op(0, 3, prec, None, dev(b), out, out2)
        x1 = orc.jacobi_step(rp, cv, nz, np.zeros(A.n), b, d, cfg.jacobi_omega)
        same(out.cpu().numpy(), x1, prec)
        same(out2.cpu().numpy(), orc.residual(rp, cv, nz, x1, b), prec)


def test_l1_inverse_diagonal_long_row(dh, orc, torch_cuda):
    """The l1 inverse diagonal over row blocks (|a| staged in LDS, row-order
    sums) with a 5,000-entry hub row, longer than the stage and summed by one
    thread from memory: every level's values are the oracle's bit for bit,
    also after a resetup with new values."""
    torch = torch_cuda
    B = hub_matrix(dh.problems, 100, 5000)
    A = dh.as_csr(B)
    assert np.diff(A.rowptr).max() == 5000
    cfg = dh.AMGConfig(coarsening=dh.AggregationCoarsening(0.25), smoother=dh.L1JacobiSmootherType())
    h = dh.amg_setup(A, cfg)
    oh = oracle_of(h, orc, "l1_jacobi")
    for stage in ("setup", "resetup"):
        if stage == "resetup":
            s = np.exp(0.3 * np.random.default_rng(SEED + 996).standard_normal(A.n))
            rows = np.repeat(np.arange(A.n), np.diff(A.rowptr))
            new = A.nzval * s[rows] * s[A.colval]
            dh.amg_resetup_(h, dh.problems.CSR(A.rowptr, A.colval, new), cfg)
            oh.resetup(new)
        for l in range(h.nlevels):
            n = h.level_info(l)["n"]
            dd = torch.zeros(n, dtype=torch.float64, device="cuda")
            h.level_op(l, 9, 0, None, None, dd)
            torch.cuda.synchronize()
            same(dd.cpu().numpy(), oh.dinv(l))
    # the oracle's l1 diagonal itself, on the fine level, from its definition
    d0 = oh.dinv(0)
    ref0 = 1.0 / np.add.reduceat(np.abs(new), A.rowptr[:-1])
    assert np.allclose(d0, ref0, rtol=1e-14, atol=0)
