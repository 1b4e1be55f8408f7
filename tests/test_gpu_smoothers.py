"""GPU parity of the SPAI(0) and Chebyshev smoothers (src/smoothers.jl:493-558,
860-990) against the oracle.

* SPAI(0) inverse diagonal and the Chebyshev level step: bit-identical (FP64).
* The device power iteration (Chebyshev bounds) agrees with the oracle's
  numpy restatement from the same start vector to 1e-12; given the bounds,
  cycles are bit-identical to the oracle's.
* PCG with either smoother: FP64 within 1e-10 / +-1 iteration of the oracle;
  the FP32 cycle reaches the tolerance.  Resetup recomputes the bounds.
"""
import dataclasses

import numpy as np
import pytest

from conftest import oracle_of, rel_err

pytestmark = pytest.mark.gpu

SEED = 20261202
_cache = {}


def _built(dh, orc, smoother, steps=1):
    key = (smoother, steps)
    if key not in _cache:
        P = dh.problems
        A = P.heterogeneous3d(16, P.log_permeability(16, 1.0, SEED + 600))
        st = dh.SPAI0SmootherType() if smoother == "spai0" else dh.ChebyshevSmootherType()
        cfg = dh.AMGConfig(coarsening=dh.HMISCoarsening(0.5, dh.ExtendedIInterpolation(0.3)), smoother=st,
                           pre_smoothing_steps=steps, post_smoothing_steps=steps)
        h = dh.amg_setup(A, cfg)
        _cache[key] = (A, cfg, h)
    return _cache[key]


def _bounds(h):
    return [tuple(h.device_values(l, 4, 2)) for l in range(h.nlevels)]


def _oracle(h, orc, smoother, steps):
    oh = oracle_of(h, orc, smoother, pre=steps, post=steps)
    if smoother == "chebyshev":
        oh.set_cheb(_bounds(h))
    return oh


def test_spai0_inverse_diagonal(dh, orc, torch_cuda):
    torch = torch_cuda
    A, cfg, h = _built(dh, orc, "spai0")
    for l in range(h.nlevels):
        lv = h.level(l, with_rmap=False)
        out = torch.empty(lv["n"], dtype=torch.float64, device="cuda")
        h.level_op(l, 9, "fp64", out=out)
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy(), orc.invdiag(lv["A_rp"], lv["A_cv"], lv["A_nz"], 2)), l


def test_chebyshev_power_iteration(dh, orc, torch_cuda):
    A, cfg, h = _built(dh, orc, "chebyshev")
    for l, (lo, hi) in enumerate(_bounds(h)):
        lv = h.level(l, with_rmap=False)
        d = orc.invdiag(lv["A_rp"], lv["A_cv"], lv["A_nz"], 0)
        v0 = h.device_values(l, 5, lv["n"])
        rho = orc.power_rho(lv["A_rp"], lv["A_cv"], lv["A_nz"], d, v0)
        assert abs(hi - 1.1 * rho) <= 1e-12 * hi, l
        assert hi / 30.0 == lo


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_chebyshev_level_step(dh, orc, torch_cuda, prec):
    torch = torch_cuda
    A, cfg, h = _built(dh, orc, "chebyshev")
    tdt = torch.float64 if prec == "fp64" else torch.float32
    for l, (lo, hi) in enumerate(_bounds(h)):
        lv = h.level(l, with_rmap=False)
        n = lv["n"]
        x0 = dh.problems.splitmix_normal(SEED + 610 + l, n)
        b = dh.problems.splitmix_normal(SEED + 620 + l, n)
        out = torch.tensor(x0, dtype=tdt, device="cuda")
        h.level_op(l, 10, prec, in2=torch.tensor(b, dtype=tdt, device="cuda"), out=out)
        torch.cuda.synchronize()
        d = orc.invdiag(lv["A_rp"], lv["A_cv"], lv["A_nz"], 0)
        ref = orc.cheb_smooth(lv["A_rp"], lv["A_cv"], lv["A_nz"], x0, b, d, lo, hi)
        got = out.cpu().numpy().astype(np.float64)
        if prec == "fp64":
            assert np.array_equal(got, ref), l
        else:
            assert rel_err(got, ref) < 1e-4, l


@pytest.mark.parametrize("smoother,steps", [("spai0", 1), ("chebyshev", 1), ("chebyshev", 2)])
def test_cycles_match_oracle(dh, orc, torch_cuda, smoother, steps):
    torch = torch_cuda
    A, cfg, h = _built(dh, orc, smoother, steps)
    oh = _oracle(h, orc, smoother, steps)
    b = dh.problems.splitmix_normal(SEED + 630, A.n)
    x0 = dh.problems.splitmix_normal(SEED + 631, A.n)
    x = torch.tensor(x0, device="cuda")
    dh.amg_cycle_(x, torch.tensor(b, device="cuda"), h, cfg)
    assert np.array_equal(x.cpu().numpy(), oh.cycle(x0, b))
    z = torch.zeros(A.n, dtype=torch.float64, device="cuda")
    dh.amg_cycle_(z, torch.tensor(b, device="cuda"), h, cfg, zero_guess=True)
    assert np.array_equal(z.cpu().numpy(), oh.cycle(np.zeros(A.n), b))


@pytest.mark.parametrize("smoother", ["spai0", "chebyshev"])
def test_pcg_matches_oracle(dh, orc, torch_cuda, smoother):
    torch = torch_cuda
    A, cfg, h = _built(dh, orc, smoother)
    oh = _oracle(h, orc, smoother, 1)
    b = dh.problems.splitmix_normal(SEED + 640, A.n)
    x = torch.zeros(A.n, dtype=torch.float64, device="cuda")
    _, it, rel = dh.pcg_solve_(x, torch.tensor(b, device="cuda"), h, cfg, tol=1e-10, maxiter=300)
    xo, ito, _ = oh.pcg(np.zeros(A.n), b, 1e-10, 300)
    assert abs(it - ito) <= 1 and rel_err(x.cpu().numpy(), xo) < 1e-10
    x32 = torch.zeros_like(x)
    _, it32, rel32 = dh.pcg_solve_(x32, torch.tensor(b, device="cuda"), h, dataclasses.replace(cfg, precision="fp32"),
                                  tol=1e-8, maxiter=300)
    true = np.linalg.norm(b - A.matvec(x32.cpu().numpy())) / np.linalg.norm(b)
    assert rel32 < 1e-8 and true < 1.5e-8


def test_chebyshev_resetup_recomputes_bounds(dh, orc, torch_cuda):
    torch = torch_cuda
    P = dh.problems
    A = P.heterogeneous3d(12, P.log_permeability(12, 1.0, SEED + 650))
    cfg = dh.AMGConfig(coarsening=dh.HMISCoarsening(0.5, dh.ExtendedIInterpolation(0.3)),
                       smoother=dh.ChebyshevSmootherType(), pre_smoothing_steps=2, post_smoothing_steps=2)
    h = dh.amg_setup(A, cfg)
    before = _bounds(h)
    oh = oracle_of(h, orc, "chebyshev", pre=2, post=2)
    A2 = P.heterogeneous3d(12, P.log_permeability(12, 1.0, SEED + 650) + 0.5 * P.splitmix_normal(SEED + 651, A.n))
    dh.amg_resetup_(h, A2, cfg)
    after = _bounds(h)
    assert after != before
    oh.resetup(A2.nzval)
    oh.set_cheb(after)
    b = P.splitmix_normal(SEED + 652, A.n)
    z = torch.zeros(A.n, dtype=torch.float64, device="cuda")
    dh.amg_cycle_(z, torch.tensor(b, device="cuda"), h, cfg, zero_guess=True)
    assert np.array_equal(z.cpu().numpy(), oh.cycle(np.zeros(A.n), b))
    x = torch.zeros_like(z)
    _, it, rel = dh.pcg_solve_(x, torch.tensor(b, device="cuda"), h, cfg, tol=1e-8, maxiter=200)
    assert rel < 1e-8


# ── ILU(0), level-scheduled ─────────────────────────────────────────────────
def _ilu_built(dh):
    if "ilu" not in _cache:
        P = dh.problems
        A = P.heterogeneous3d(12, P.log_permeability(12, 1.0, SEED + 700))
        cfg = dh.AMGConfig(coarsening=dh.HMISCoarsening(0.5, dh.ExtendedIInterpolation(0.3)),
                           smoother=dh.ILU0SmootherType())
        _cache["ilu"] = (A, cfg, dh.amg_setup(A, cfg))
    return _cache["ilu"]


def test_ilu0_factors_and_step(dh, orc, torch_cuda):
    torch = torch_cuda
    A, cfg, h = _ilu_built(dh)
    for l in range(h.nlevels):
        lv = h.level(l, with_rmap=False)
        n, nnz = lv["n"], len(lv["A_cv"])
        _, L, U = orc.ilu0_factor(lv["A_rp"], lv["A_cv"], lv["A_nz"])
        assert np.array_equal(h.device_values(l, 6, nnz), L) and np.array_equal(h.device_values(l, 7, nnz), U), l
        x0 = dh.problems.splitmix_normal(SEED + 710 + l, n)
        b = dh.problems.splitmix_normal(SEED + 720 + l, n)
        out = torch.tensor(x0, device="cuda")
        h.level_op(l, 11, "fp64", in2=torch.tensor(b, device="cuda"), out=out)
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy(), orc.ilu0_smooth(lv["A_rp"], lv["A_cv"], lv["A_nz"], x0, b)), l


def test_ilu0_cycle_pcg_and_resetup(dh, orc, torch_cuda):
    torch = torch_cuda
    P = dh.problems
    A, cfg, h = _ilu_built(dh)
    oh = oracle_of(h, orc, "ilu0")
    b = P.splitmix_normal(SEED + 730, A.n)
    z = torch.zeros(A.n, dtype=torch.float64, device="cuda")
    dh.amg_cycle_(z, torch.tensor(b, device="cuda"), h, cfg, zero_guess=True)
    assert np.array_equal(z.cpu().numpy(), oh.cycle(np.zeros(A.n), b))
    # the ILU(0) cycle is not symmetric: PCG refuses it; the stationary solve
    # (the reference's own iteration) matches the oracle
    x = torch.zeros_like(z)
    with pytest.raises(dh.DraugrError, match="symmetric preconditioner"):
        dh.pcg_solve_(x, torch.tensor(b, device="cuda"), h, cfg, tol=1e-10, maxiter=200)
    _, it = dh.amg_solve_(x, torch.tensor(b, device="cuda"), h, cfg, tol=1e-10, maxiter=200)
    xo, ito, _ = oh.solve(np.zeros(A.n), b, 1e-10, 200)
    assert abs(it - ito) <= 1 and rel_err(x.cpu().numpy(), xo) < 1e-10
    # resetup refactorizes every level (on the host, as the reference)
    A2 = P.heterogeneous3d(12, P.log_permeability(12, 1.0, SEED + 700) + 0.4 * P.splitmix_normal(SEED + 731, A.n))
    dh.amg_resetup_(h, A2, cfg)
    oh.resetup(A2.nzval)
    z.zero_()
    dh.amg_cycle_(z, torch.tensor(b, device="cuda"), h, cfg, zero_guess=True)
    assert np.array_equal(z.cpu().numpy(), oh.cycle(np.zeros(A.n), b))
