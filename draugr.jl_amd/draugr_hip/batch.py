"""Batches of independent systems over the batched C-ABI (draugr_hip_batch_*).

Each system keeps its own hierarchy (the reference's setup, one per system);
the GPU runs every level's passes for all systems and all right-hand sides in
one launch (block-diagonal packed levels, K-wide vectors) and replays the PCG
iteration from captured hipGraphs.  No counterpart in the reference, which
solves one system with one right-hand side per call (src/cycle.jl:12, :110).

Packed vectors: a float64 CUDA tensor of ``rows * nrhs`` entries; system s
occupies rows ``[offset_s, offset_s + n_s)`` and row i holds its ``nrhs``
values contiguously (entry (i, j) at ``i * nrhs + j``).  ``pack`` / ``unpack``
convert from / to a list of ``(n_s, nrhs)`` arrays.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import DraugrError, check
from .config import AMGConfig
from .core import _Config, _torch_stream, as_csr


class AMGBatch:
    """Handle of a batch of hierarchies resident in HBM."""

    def __init__(self, handle: int, config: AMGConfig, nrhs: int):
        self.handle = handle
        self.config = config
        self.nrhs = nrhs
        info = self.info()
        self.nsys, self.rows = info["nsys"], info["rows"]
        self.sizes = [self.system_info(s)["n"] for s in range(self.nsys)]
        self.offsets = [self.system_info(s)["offset"] for s in range(self.nsys)]

    def __del__(self):
        try:
            if self.handle > 0:
                _lib.lib().draugr_hip_batch_free(self.handle)
        except Exception:
            pass

    def free(self):
        if self.handle > 0:
            check(_lib.lib().draugr_hip_batch_free(self.handle), "draugr_hip_batch_free")
            self.handle = -1

    def info(self) -> dict:
        out = (ctypes.c_int64 * 6)()
        m = check(_lib.lib().draugr_hip_batch_info(self.handle, out, 6), "draugr_hip_batch_info")
        keys = ["nsys", "nrhs", "rows", "max_levels", "setup_us", "resident_bytes"]
        return dict(zip(keys, [int(out[i]) for i in range(m)]))

    def system_info(self, s: int) -> dict:
        out = (ctypes.c_int64 * 4)()
        m = check(_lib.lib().draugr_hip_batch_system_info(self.handle, s, out, 4), "draugr_hip_batch_system_info")
        return dict(zip(["n", "offset", "levels", "n_coarse"], [int(out[i]) for i in range(m)]))

    def level_info(self, l: int) -> dict:
        out = (ctypes.c_int64 * 5)()
        m = check(_lib.lib().draugr_hip_batch_level_info(self.handle, l, out, 5), "draugr_hip_batch_level_info")
        return dict(zip(["n_smooth", "n_vec", "nnz", "nnz_P", "n_Pt"], [int(out[i]) for i in range(m)]))

    def tune_report(self) -> str:
        L = _lib.lib()
        n = check(L.draugr_hip_batch_tune_report(self.handle, None, 0), "draugr_hip_batch_tune_report")
        buf = ctypes.create_string_buffer(n + 1)
        L.draugr_hip_batch_tune_report(self.handle, buf, n + 1)
        return buf.value.decode()

    def jacobi_pass(self, l: int, prec: str, x, b, out):
        """One K-wide Jacobi step on packed level l (test / measurement hook):
        out = x + w d (b - A x) over the level's smoothing rows; tensors of
        the level's vector rows * nrhs entries, float64 (prec "fp64") or
        float32 ("fp32")."""
        p = {"fp64": 0, "fp32": 1}[prec]
        ptr = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
        check(_lib.lib().draugr_hip_batch_jacobi_pass(self.handle, l, p, ptr(x), ptr(b), ptr(out), _torch_stream()),
              "draugr_hip_batch_jacobi_pass")
        return out

    # ── packed vectors ──────────────────────────────────────────────────────
    def pack(self, vectors: Sequence, device="cuda"):
        """List of (n_s, nrhs) arrays (or (n_s,) when nrhs = 1) -> packed tensor."""
        import torch
        if len(vectors) != self.nsys:
            raise DraugrError("DimensionMismatch: %d vectors for %d systems" % (len(vectors), self.nsys))
        parts = []
        for s, v in enumerate(vectors):
            a = np.asarray(v, dtype=np.float64).reshape(self.sizes[s], self.nrhs)
            parts.append(a)
        return torch.from_numpy(np.ascontiguousarray(np.concatenate(parts, axis=0)).ravel()).to(device)

    def unpack(self, packed) -> List[np.ndarray]:
        a = packed.detach().cpu().numpy().reshape(self.rows, self.nrhs)
        return [a[o:o + n] for o, n in zip(self.offsets, self.sizes)]

    def _check_vec(self, t, what):
        import torch
        if not (type(t).__module__.startswith("torch") and t.is_cuda and t.dtype == torch.float64
                and t.is_contiguous()):
            raise TypeError("%s must be a contiguous float64 CUDA tensor" % what)
        if t.numel() != self.rows * self.nrhs:
            raise DraugrError("DimensionMismatch: %s has %d entries, expected rows * nrhs = %d"
                              % (what, t.numel(), self.rows * self.nrhs))
        return ctypes.c_void_p(t.data_ptr())

    def cycle_(self, x, b, config: Optional[AMGConfig] = None):
        """x = V(0; b) for every system and right-hand side (in place on x)."""
        c = _Config(config or self.config)
        check(_lib.lib().draugr_hip_batch_cycle_device(self.handle, self._check_vec(x, "x"), self._check_vec(b, "b"),
                                                       c.handle, _torch_stream()), "draugr_hip_batch_cycle_device")
        return x

    def pcg_solve_(self, x, b, config: Optional[AMGConfig] = None, tol: float = 1e-8, maxiter: int = 200):
        """FP64 PCG for every (system, rhs); x is the initial guess and the result.
        Returns (x, iterations (nsys, nrhs), relres (nsys, nrhs))."""
        c = _Config(config or self.config)
        its = np.zeros(self.nsys * self.nrhs, dtype=np.int32)
        rel = np.zeros(self.nsys * self.nrhs, dtype=np.float64)
        check(_lib.lib().draugr_hip_batch_pcg_device(self.handle, self._check_vec(x, "x"), self._check_vec(b, "b"),
                                                     c.handle, tol, maxiter, _torch_stream(),
                                                     its.ctypes.data_as(ctypes.c_void_p),
                                                     rel.ctypes.data_as(ctypes.c_void_p)),
              "draugr_hip_batch_pcg_device")
        return x, its.reshape(self.nsys, self.nrhs), rel.reshape(self.nsys, self.nrhs)


def amg_batch_setup(matrices: Sequence, config: Optional[AMGConfig] = None, nrhs: int = 1) -> AMGBatch:
    """Set up one hierarchy per matrix (same config) and upload them as a batch
    for ``nrhs`` right-hand sides per system."""
    config = config or AMGConfig()
    mats = [as_csr(A) for A in matrices]
    keep = []
    n = np.array([A.n for A in mats], dtype=np.int32)
    nnz = np.array([A.nnz for A in mats], dtype=np.int32)
    rps = (ctypes.c_void_p * len(mats))()
    cvs = (ctypes.c_void_p * len(mats))()
    nzs = (ctypes.c_void_p * len(mats))()
    for s, A in enumerate(mats):
        rp = np.ascontiguousarray(A.rowptr, np.int32)
        cv = np.ascontiguousarray(A.colval, np.int32)
        nz = np.ascontiguousarray(A.nzval, np.float64)
        keep += [rp, cv, nz]
        rps[s], cvs[s], nzs[s] = rp.ctypes.data, cv.ctypes.data, nz.ctypes.data
    c = _Config(config)
    h = check(_lib.lib().draugr_hip_batch_setup(len(mats), n.ctypes.data_as(ctypes.c_void_p),
                                                nnz.ctypes.data_as(ctypes.c_void_p), ctypes.cast(rps, ctypes.c_void_p),
                                                ctypes.cast(cvs, ctypes.c_void_p), ctypes.cast(nzs, ctypes.c_void_p),
                                                c.handle, 0, nrhs), "draugr_hip_batch_setup")
    return AMGBatch(h, config, nrhs)
