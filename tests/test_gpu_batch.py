"""GPU parity of the batched path (many uneven systems, K right-hand sides).

* A batched V(1,1) cycle equals, for every system and right-hand side, the
  single-system cycle of that system on that right-hand side bit for bit,
  in FP64 and FP32 (same per-row sums, same epilogues, same coarse solve),
  and in FP64 it equals the oracle's cycle bit for bit.
* The batched FP64 PCG agrees with the oracle's PCG per (system, rhs) to
  1e-10 relative with iteration counts within 1; the FP32-cycle PCG reaches
  the tolerance (true residual) per pair.
* The graph-replayed PCG equals the eagerly launched one bit for bit.
"""
import dataclasses

import numpy as np
import pytest

from conftest import oracle_of, rel_err

pytestmark = pytest.mark.gpu

SEED = 20261202
EDGES = [6, 13, 9, 16, 11]  # uneven systems: different depths and coarsest sizes


def _systems(dh, edges, sigma=1.0):
    P = dh.problems
    return [P.heterogeneous3d(m, P.log_permeability(m, sigma, SEED + 400 + m)) for m in edges]


def _cfg(dh, **kw):
    kw.setdefault("max_coarse_size", 12)
    return dh.hypre_default_config(smoother=dh.L1JacobiSmootherType(), **kw)


def _rhs(dh, batch, seed):
    P = dh.problems
    return [P.splitmix_normal(seed + s, n * batch.nrhs).reshape(n, batch.nrhs) for s, n in enumerate(batch.sizes)]


_cache = {}


def _setup(dh, edges, K, **kw):
    key = (tuple(edges), K, tuple(sorted(kw.items())))
    if key not in _cache:
        mats = _systems(dh, edges)
        cfg = _cfg(dh, **kw)
        batch = dh.amg_batch_setup(mats, cfg, nrhs=K)
        # the single-system references sum every row in row order (the stream
        # kernel): the batched passes do, while the single-system autotuner may
        # pick the vector CSR kernel for FP32 passes, which sums in another order
        singles = [dh.amg_setup(A, dataclasses.replace(cfg, kernel="stream")) for A in mats]
        _cache[key] = (mats, cfg, batch, singles)
    return _cache[key]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("K", [1, 4])
def test_batched_cycle_equals_single_cycles(dh, torch_cuda, prec, K):
    torch = torch_cuda
    mats, cfg, batch, singles = _setup(dh, EDGES, K)
    assert len({h.nlevels for h in singles}) > 1  # uneven depths exercised
    c = dataclasses.replace(cfg, precision=prec)
    bs = _rhs(dh, batch, SEED + 500)
    b = batch.pack(bs)
    x = torch.zeros_like(b)
    batch.cycle_(x, b, c)
    torch.cuda.synchronize()
    got = batch.unpack(x)
    for s, h in enumerate(singles):
        for j in range(K):
            bj = torch.tensor(np.ascontiguousarray(bs[s][:, j]), device="cuda")
            xj = torch.zeros_like(bj)
            dh.amg_cycle_(xj, bj, h, c, zero_guess=True)
            torch.cuda.synchronize()
            assert np.array_equal(got[s][:, j], xj.cpu().numpy()), (prec, K, s, j)


def test_batched_cycle_equals_oracle(dh, orc, torch_cuda):
    torch = torch_cuda
    mats, cfg, batch, singles = _setup(dh, EDGES, 4)
    bs = _rhs(dh, batch, SEED + 510)
    b = batch.pack(bs)
    x = torch.zeros_like(b)
    batch.cycle_(x, b, dataclasses.replace(cfg, precision="fp64"))
    got = batch.unpack(x)
    for s, h in enumerate(singles):
        oh = oracle_of(h, orc, "l1_jacobi")
        for j in range(4):
            ref = oh.cycle(np.zeros(batch.sizes[s]), bs[s][:, j])
            assert np.array_equal(got[s][:, j], ref), (s, j)


def test_batched_cycle_large_coarsest(dh, torch_cuda):
    """Coarsest levels above 64 unknowns take the one-thread substitution."""
    torch = torch_cuda
    mats, cfg, batch, singles = _setup(dh, [12, 16, 14], 2, max_coarse_size=150)
    assert any(batch.system_info(s)["n_coarse"] > 64 for s in range(batch.nsys))
    bs = _rhs(dh, batch, SEED + 520)
    b = batch.pack(bs)
    x = torch.zeros_like(b)
    batch.cycle_(x, b, dataclasses.replace(cfg, precision="fp64"))
    got = batch.unpack(x)
    for s, h in enumerate(singles):
        for j in range(2):
            bj = torch.tensor(np.ascontiguousarray(bs[s][:, j]), device="cuda")
            xj = torch.zeros_like(bj)
            dh.amg_cycle_(xj, bj, h, dataclasses.replace(cfg, precision="fp64"), zero_guess=True)
            torch.cuda.synchronize()
            assert np.array_equal(got[s][:, j], xj.cpu().numpy()), (s, j)


def test_batched_pcg_fp64_matches_oracle(dh, orc, torch_cuda):
    torch = torch_cuda
    mats, cfg, batch, singles = _setup(dh, EDGES, 4)
    bs = _rhs(dh, batch, SEED + 530)
    b = batch.pack(bs)
    x = torch.zeros_like(b)
    _, its, rel = batch.pcg_solve_(x, b, dataclasses.replace(cfg, precision="fp64"), tol=1e-10, maxiter=300)
    got = batch.unpack(x)
    for s, h in enumerate(singles):
        oh = oracle_of(h, orc, "l1_jacobi")
        for j in range(4):
            xo, ito, _ = oh.pcg(np.zeros(batch.sizes[s]), bs[s][:, j], tol=1e-10, maxiter=300)
            assert abs(int(its[s, j]) - ito) <= 1, (s, j, its[s, j], ito)
            assert rel_err(got[s][:, j], xo) < 1e-10, (s, j)
            assert rel[s, j] < 1e-10


def test_batched_pcg_mixed_reaches_tolerance(dh, torch_cuda):
    torch = torch_cuda
    mats, cfg, batch, singles = _setup(dh, EDGES, 4)
    bs = _rhs(dh, batch, SEED + 540)
    b = batch.pack(bs)
    x64 = torch.zeros_like(b)
    _, its64, _ = batch.pcg_solve_(x64, b, dataclasses.replace(cfg, precision="fp64"), tol=1e-8)
    x32 = torch.zeros_like(b)
    _, its32, rel32 = batch.pcg_solve_(x32, b, dataclasses.replace(cfg, precision="fp32"), tol=1e-8)
    got = batch.unpack(x32)
    for s, A in enumerate(mats):
        for j in range(4):
            r = bs[s][:, j] - A.matvec(got[s][:, j])
            assert np.linalg.norm(r) / np.linalg.norm(bs[s][:, j]) < 1.5e-8, (s, j)
            assert its32[s, j] <= its64[s, j] + 1, (s, j, its32[s, j], its64[s, j])
    assert (rel32 < 1e-8).all()


def test_batched_pcg_graph_equals_eager(dh, torch_cuda):
    torch = torch_cuda
    mats, cfg, batch, singles = _setup(dh, EDGES, 4)
    bs = _rhs(dh, batch, SEED + 550)
    b = batch.pack(bs)
    for prec in ("fp64", "fp32"):
        for rep in range(2):
            xg = torch.zeros_like(b)
            _, itg, relg = batch.pcg_solve_(xg, b, dataclasses.replace(cfg, precision=prec, graphs=True), tol=1e-9)
            torch.cuda.synchronize()
            xe = torch.zeros_like(b)
            _, ite, rele = batch.pcg_solve_(xe, b, dataclasses.replace(cfg, precision=prec, graphs=False), tol=1e-9)
            torch.cuda.synchronize()
            assert np.array_equal(itg, ite) and np.array_equal(relg, rele), (prec, rep)
            assert torch.equal(xg, xe), (prec, rep)


def test_batched_pcg_zero_rhs_and_initial_guess(dh, torch_cuda):
    """A zero right-hand side gives x = 0 in 0 iterations; a converged initial
    guess stops at once."""
    torch = torch_cuda
    mats, cfg, batch, singles = _setup(dh, EDGES, 4)
    bs = _rhs(dh, batch, SEED + 560)
    bs[1][:, 2] = 0.0
    b = batch.pack(bs)
    x = torch.ones_like(b)
    _, its, rel = batch.pcg_solve_(x, b, dataclasses.replace(cfg, precision="fp64"), tol=1e-10, maxiter=300)
    got = batch.unpack(x)
    assert its[1, 2] == 0 and not got[1][:, 2].any()
    _, its2, _ = batch.pcg_solve_(x, b, dataclasses.replace(cfg, precision="fp64"), tol=1e-8, maxiter=300)
    assert (its2 == 0).all()


def test_batch_errors(dh, torch_cuda):
    mats = _systems(dh, [8, 10])
    with pytest.raises(dh.DraugrError, match="jacobi"):
        dh.amg_batch_setup(mats, dh.AMGConfig(smoother=dh.ColoredGaussSeidelType()), nrhs=2)
    with pytest.raises(dh.DraugrError, match="power of two"):
        dh.amg_batch_setup(mats, _cfg(dh), nrhs=3)
    with pytest.raises(dh.DraugrError, match="no AMG level"):
        dh.amg_batch_setup([dh.problems.poisson1d(5)] + mats, _cfg(dh), nrhs=1)
    batch = dh.amg_batch_setup(mats, _cfg(dh), nrhs=2)
    b = batch.pack(_rhs(dh, batch, SEED))
    with pytest.raises(dh.DraugrError, match="fp64 and fp32"):
        batch.cycle_(b.clone(), b, _cfg(dh, precision="bf16"))
    with pytest.raises(dh.DraugrError, match="V\\(1,1\\)"):
        batch.cycle_(b.clone(), b, _cfg(dh, cycle_type="w"))
    batch.free()
    with pytest.raises(dh.DraugrError, match="Invalid batch handle"):
        batch.handle = 99999
        batch.info()


# The C4 workload's shape at sizes the oracle finishes in seconds: 8 uneven
# systems (edges 9 to 30), 16 and 8 right-hand sides.  K = 16 runs the
# smoothing passes with four threads per row and four right-hand sides per
# thread; K = 8 with two threads per row.
C4_EDGES = [24, 9, 30, 13, 26, 16, 11, 28]


@pytest.mark.parametrize("K", [16, 8])
def test_c4_shape_cycle_equals_single_cycles(dh, torch_cuda, K):
    torch = torch_cuda
    mats, cfg, batch, singles = _setup(dh, C4_EDGES, K, max_coarse_size=50)
    assert len({h.nlevels for h in singles}) > 1
    for prec in ("fp64", "fp32"):
        c = dataclasses.replace(cfg, precision=prec)
        bs = _rhs(dh, batch, SEED + 600 + K)
        b = batch.pack(bs)
        x = torch.zeros_like(b)
        batch.cycle_(x, b, c)
        torch.cuda.synchronize()
        got = batch.unpack(x)
        print("C4-shape batched cycle", prec, "K", K, flush=True)
        for s, h in enumerate(singles):
            for j in range(K):
                bj = torch.tensor(np.ascontiguousarray(bs[s][:, j]), device="cuda")
                xj = torch.zeros_like(bj)
                dh.amg_cycle_(xj, bj, h, c, zero_guess=True)
                torch.cuda.synchronize()
                assert np.array_equal(got[s][:, j], xj.cpu().numpy()), (prec, K, s, j)


@pytest.mark.parametrize("K", [16, 8])
def test_c4_shape_pcg_fp64_matches_oracle(dh, orc, torch_cuda, K):
    torch = torch_cuda
    mats, cfg, batch, singles = _setup(dh, C4_EDGES, K, max_coarse_size=50)
    bs = _rhs(dh, batch, SEED + 620 + K)
    b = batch.pack(bs)
    x = torch.zeros_like(b)
    _, its, rel = batch.pcg_solve_(x, b, dataclasses.replace(cfg, precision="fp64"), tol=1e-10, maxiter=300)
    got = batch.unpack(x)
    for s, h in enumerate(singles):
        oh = oracle_of(h, orc, "l1_jacobi")
        for j in range(K):
            xo, ito, _ = oh.pcg(np.zeros(batch.sizes[s]), bs[s][:, j], tol=1e-10, maxiter=300)
            assert abs(int(its[s, j]) - ito) <= 1, (s, j, its[s, j], ito)
            assert rel_err(got[s][:, j], xo) < 1e-10, (s, j)
            assert rel[s, j] < 1e-10
    # the FP32-cycle batched PCG: every pair's true residual at the tolerance
    x32 = torch.zeros_like(b)
    _, its32, rel32 = batch.pcg_solve_(x32, b, dataclasses.replace(cfg, precision="fp32"), tol=1e-8)
    g32 = batch.unpack(x32)
    for s, A in enumerate(mats):
        for j in range(K):
            r = bs[s][:, j] - A.matvec(g32[s][:, j])
            assert np.linalg.norm(r) / np.linalg.norm(bs[s][:, j]) < 1.5e-8, (s, j)


def test_batch_setup_with_gpu_setup_hooks_concurrent(dh, torch_cuda):
    """gpu_setup_min_rows = 0: every level's interpolation, PMIS sweeps and
    Galerkin product run through the GPU setup hooks, from the batch setup's
    concurrent host threads; each system's hierarchy is still the one a lone
    setup builds (batched cycle = single cycles, bit for bit)."""
    torch = torch_cuda
    mats, cfg, batch, singles = _setup(dh, C4_EDGES, 4, max_coarse_size=50, gpu_setup_min_rows=0)
    c = dataclasses.replace(cfg, precision="fp64")
    bs = _rhs(dh, batch, SEED + 640)
    b = batch.pack(bs)
    x = torch.zeros_like(b)
    batch.cycle_(x, b, c)
    torch.cuda.synchronize()
    got = batch.unpack(x)
    # the lone setups of the host-only configuration give the same hierarchies
    host_cfg = dataclasses.replace(cfg, gpu_setup=False)
    for s, A in enumerate(mats):
        h = dh.amg_setup(A, host_cfg)
        for j in range(4):
            bj = torch.tensor(np.ascontiguousarray(bs[s][:, j]), device="cuda")
            xj = torch.zeros_like(bj)
            dh.amg_cycle_(xj, bj, h, dataclasses.replace(host_cfg, precision="fp64"), zero_guess=True)
            torch.cuda.synchronize()
            assert np.array_equal(got[s][:, j], xj.cpu().numpy()), (s, j)


@pytest.mark.parametrize("rpt", [1, 4, 8, 16, -1, -4])
def test_every_rhs_per_thread_mapping_is_bit_exact(dh, torch_cuda, rpt):
    """Each mapping of the K-wide passes the batch autotuner chooses among
    (1, 4, 8 or 16 right-hand sides per thread; -1 and -4 the LDS-staged
    kernels of 1 and 4), forced on every pass, gives the single-system cycles
    bit for bit at K = 16 (FP64 and FP32)."""
    torch = torch_cuda
    mats, cfg, batch, singles = _setup(dh, [9, 14, 11], 16, max_coarse_size=50, batch_rhs_per_thread=rpt)
    for prec in ("fp64", "fp32"):
        c = dataclasses.replace(cfg, precision=prec)
        bs = _rhs(dh, batch, SEED + 660)
        b = batch.pack(bs)
        x = torch.zeros_like(b)
        batch.cycle_(x, b, c)
        torch.cuda.synchronize()
        got = batch.unpack(x)
        for s, h in enumerate(singles):
            for j in range(16):
                bj = torch.tensor(np.ascontiguousarray(bs[s][:, j]), device="cuda")
                xj = torch.zeros_like(bj)
                dh.amg_cycle_(xj, bj, h, c, zero_guess=True)
                torch.cuda.synchronize()
                assert np.array_equal(got[s][:, j], xj.cpu().numpy()), (rpt, prec, s, j)
    tag = "%d%s" % (abs(rpt), "s" if rpt < 0 else "")
    assert ("%s, restrict %s" % (tag, tag)) in batch.tune_report()


def test_batched_jacobi_pass_equals_single_level_ops(dh, torch_cuda):
    """The measurement hook (one K-wide Jacobi step on packed level 0, tuned
    mapping) equals the single-system Jacobi step of every system and
    right-hand side bit for bit."""
    torch = torch_cuda
    mats, cfg, batch, singles = _setup(dh, C4_EDGES, 16, max_coarse_size=50)
    info = batch.level_info(0)
    assert info["n_smooth"] == info["n_vec"] == batch.rows
    for prec, tdt in (("fp64", torch.float64), ("fp32", torch.float32)):
        xs = _rhs(dh, batch, SEED + 670)
        bs = _rhs(dh, batch, SEED + 680)
        x = batch.pack(xs).to(tdt)
        b = batch.pack(bs).to(tdt)
        out = torch.zeros_like(x)
        batch.jacobi_pass(0, prec, x, b, out)
        torch.cuda.synchronize()
        got = out.cpu().numpy().reshape(batch.rows, 16)
        for s, h in enumerate(singles):
            o = batch.offsets[s]
            for j in (0, 7, 15):
                xj = torch.tensor(np.ascontiguousarray(xs[s][:, j]), dtype=tdt, device="cuda")
                bj = torch.tensor(np.ascontiguousarray(bs[s][:, j]), dtype=tdt, device="cuda")
                oj = torch.zeros_like(xj)
                h.level_op(0, 2, prec, xj, bj, oj)
                torch.cuda.synchronize()
                assert np.array_equal(got[o:o + batch.sizes[s], j], oj.cpu().numpy()), (prec, s, j)
