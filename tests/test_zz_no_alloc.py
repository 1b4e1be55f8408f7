"""The allocation-free PCG entry point (draugr_hip_pcg_device_noalloc) gives
the results of draugr_hip_pcg_device bit for bit: same iterate, iteration
count and residual, on two calls in a row where the second hierarchy is
smaller than the first (the static scratch keeps the first call's contents),
in FP64 and with the FP32 cycle.  An unprepared hierarchy is refused with a
message before any kernel runs."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 20261202


def _pcg(L, fn, h, cfg_handle, x, b, tol, maxiter, torch):
    rel = ctypes.c_double(0.0)
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    it = fn(h.handle, h.n, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(b.data_ptr()), cfg_handle, tol, maxiter,
            st, ctypes.byref(rel))
    torch.cuda.synchronize()
    return it, rel.value


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_no_alloc_pcg_equals_pcg(dh, torch_cuda, prec):
    torch = torch_cuda
    from draugr_hip import _lib, core
    L = _lib.lib()
    P = dh.problems
    cfg = dh.hypre_default_config(smoother=dh.L1JacobiSmootherType(), precision=prec)
    ch = core._cached_config(cfg).handle
    hs = []
    for m, s in ((18, 1), (11, 2)):  # the second call's hierarchy is the smaller one
        A = P.heterogeneous3d(m, P.log_permeability(m, 1.5, SEED + 900 + s))
        hs.append((dh.amg_setup(A, cfg), torch.tensor(P.splitmix_normal(SEED + 910 + s, A.n), device="cuda")))
    print("no-alloc PCG, %s: hierarchies of %s rows" % (prec, [h.n for h, _ in hs]), flush=True)
    # prepare: one call of the allocating entry point per hierarchy
    for h, b in hs:
        x = torch.zeros_like(b)
        it, rel = _pcg(L, L.draugr_hip_pcg_device, h, ch, x, b, 1e-9, 300, torch)
        assert it > 0, _lib.last_error()
    for h, b in hs:
        xr = torch.zeros_like(b)
        itr, relr = _pcg(L, L.draugr_hip_pcg_device, h, ch, xr, b, 1e-9, 300, torch)
        xn = torch.zeros_like(b)
        itn, reln = _pcg(L, L.draugr_hip_pcg_device_noalloc, h, ch, xn, b, 1e-9, 300, torch)
        assert itn == itr > 0, (itn, itr, _lib.last_error())
        assert reln == relr and reln < 1e-9
        assert torch.equal(xn, xr)
        # again right away: the scratch holds this call's own leftovers now
        xn.zero_()
        itn2, reln2 = _pcg(L, L.draugr_hip_pcg_device_noalloc, h, ch, xn, b, 1e-9, 300, torch)
        assert itn2 == itr and reln2 == relr and torch.equal(xn, xr)


def test_no_alloc_pcg_refuses_unprepared_hierarchy(dh, torch_cuda):
    torch = torch_cuda
    from draugr_hip import _lib, core
    L = _lib.lib()
    P = dh.problems
    cfg = dh.hypre_default_config(smoother=dh.L1JacobiSmootherType(), precision="fp32")
    A = P.heterogeneous3d(8, P.log_permeability(8, 1.0, SEED + 930))
    h = dh.amg_setup(A, cfg)
    b = torch.tensor(P.splitmix_normal(SEED + 931, A.n), device="cuda")
    x = torch.zeros_like(b)
    it, _ = _pcg(L, L.draugr_hip_pcg_device_noalloc, h, core._cached_config(cfg).handle, x, b, 1e-8, 100, torch)
    assert it == -1
    assert "not prepared" in _lib.last_error()
    assert not bool(torch.any(x != 0))  # nothing ran
    np.testing.assert_equal(x.cpu().numpy(), np.zeros(A.n))
