"""Host-side mirror of Draugr.jl's solve / resetup API over the HIP C-ABI.

Same names, argument meaning and error behaviour as the reference
(``amg_setup`` src/setup.jl:47-87, ``amg_cycle!`` src/cycle.jl:12,
``amg_solve!`` src/cycle.jl:110, ``amg_resetup!`` src/resetup.jl:20,
``DraugrPreconditioner`` / ``ldiv!`` src/preconditioner.jl:29-102); Julia's
``!`` becomes a trailing underscore.  Vectors are numpy arrays (host, FP64;
copied to HBM for the call, as the reference C-ABI borrows host pointers) or
torch CUDA tensors (FP64, already in HBM: zero-copy, enqueued on torch's
current stream).
"""
from __future__ import annotations

import collections
import ctypes
from typing import Optional, Tuple

import numpy as np

from . import _lib
from ._lib import DraugrError, check
from .config import AMGConfig
from .problems import CSR

# level_op codes of draugr_hip_level_op
OP_SPMV, OP_RESID, OP_JACOBI, OP_ZERO_JACOBI_RESID, OP_RESTRICT, OP_PROLONG, OP_GS, OP_COARSE, OP_GALERKIN, \
    OP_INVDIAG = range(10)
PREC = {"fp64": 0, "fp32": 1, "bf16": 2}


def _is_torch(a) -> bool:
    return type(a).__module__.startswith("torch")


def _ptr(a):
    if _is_torch(a):
        return ctypes.c_void_p(a.data_ptr())
    return a.ctypes.data_as(ctypes.c_void_p)


def _torch_stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def as_csr(A) -> CSR:
    """Accept a ``problems.CSR``, a scipy sparse matrix, or (rowptr, colval, nzval)."""
    if isinstance(A, CSR):
        return A
    if isinstance(A, tuple):
        return CSR(np.asarray(A[0], np.int32), np.asarray(A[1], np.int32), np.asarray(A[2], np.float64))
    try:
        import scipy.sparse as sp
        if sp.issparse(A):
            A = sp.csr_matrix(A)
            return CSR(A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64))
    except ImportError:
        pass
    raise TypeError("unsupported matrix type %r" % type(A))


class _Config:
    """A C-ABI config handle made from an AMGConfig (freed with the object)."""

    def __init__(self, config: AMGConfig, json_text: Optional[str] = None):
        text = json_text if json_text is not None else config.to_json()
        self.handle = check(_lib.lib().draugr_amg_config_from_json(text.encode()), "draugr_amg_config_from_json")

    def __del__(self):
        try:
            _lib.lib().draugr_amg_config_free(self.handle)
        except Exception:
            pass


class AMGHierarchy:
    """Handle of a hierarchy built on the host and resident in HBM."""

    def __init__(self, handle: int, config: AMGConfig, n: int, uploaded: bool):
        self.handle = handle
        self.config = config
        self.n = n
        self.uploaded = uploaded

    def __del__(self):
        try:
            if self.handle > 0:
                _lib.lib().draugr_amg_free(self.handle)
        except Exception:
            pass

    def free(self):
        if self.handle > 0:
            check(_lib.lib().draugr_amg_free(self.handle), "draugr_amg_free")
            self.handle = -1

    # ── inspection ──────────────────────────────────────────────────────────
    def info(self) -> dict:
        out = (ctypes.c_int64 * 8)()
        m = check(_lib.lib().draugr_hip_hierarchy_info(self.handle, out, 8), "draugr_hip_hierarchy_info")
        keys = ["nlevels", "n_coarse", "n_fine", "setup_us", "uploaded", "total_nnz", "total_rows"]
        return dict(zip(keys, [int(out[i]) for i in range(m)]))

    @property
    def nlevels(self) -> int:
        return self.info()["nlevels"]

    def level_info(self, l: int) -> dict:
        out = (ctypes.c_int64 * 10)()
        m = check(_lib.lib().draugr_hip_level_info(self.handle, l, out, 10), "draugr_hip_level_info")
        keys = ["n", "nnz", "n_coarse", "nnz_P", "ntriples", "num_colors", "has_rmap", "nblk", "has16", "nnz_sorted"]
        return dict(zip(keys, [int(out[i]) for i in range(m)]))

    def _export(self, l, which, count, dtype):
        a = np.empty(count, dtype=dtype)
        check(_lib.lib().draugr_hip_level_export(self.handle, l, which, a.ctypes.data_as(ctypes.c_void_p)),
              "draugr_hip_level_export")
        return a

    def level(self, l: int, with_rmap: bool = True) -> dict:
        """Host copy of level l's arrays (0-based), as set up."""
        li = self.level_info(l)
        n, nnz, nc, nP, ntr = li["n"], li["nnz"], li["n_coarse"], li["nnz_P"], li["ntriples"]
        d = {
            "A_rp": self._export(l, 0, n + 1, np.int32), "A_cv": self._export(l, 1, nnz, np.int32),
            "A_nz": self._export(l, 2, nnz, np.float64),
            "P_rp": self._export(l, 3, n + 1, np.int32), "P_cv": self._export(l, 4, nP, np.int32),
            "P_nz": self._export(l, 5, nP, np.float64),
            "Pt_off": self._export(l, 6, nc + 1, np.int32), "Pt_rows": self._export(l, 7, nP, np.int32),
            "Pt_pnz": self._export(l, 8, nP, np.int32),
            "n": n, "n_coarse": nc,
        }
        if li["has_rmap"] and with_rmap:
            nnzc = self.level_info(l + 1)["nnz"] if l + 1 < self.nlevels else self.coarsest()["nnz"]
            d["R_off"] = self._export(l, 9, nnzc + 1, np.int64)
            d["R_pi"] = self._export(l, 10, ntr, np.int32)
            d["R_ai"] = self._export(l, 11, ntr, np.int32)
            d["R_pj"] = self._export(l, 12, ntr, np.int32)
        if li["num_colors"] > 0:
            d["c_off"] = self._export(l, 13, li["num_colors"] + 1, np.int32)
            d["c_order"] = self._export(l, 14, n, np.int32)
        return d

    def coarsest(self) -> dict:
        nc = self.info()["n_coarse"]
        rp = self._export(-1, 0, nc + 1, np.int32)
        nnz = int(rp[-1])
        return {"rp": rp, "cv": self._export(-1, 1, nnz, np.int32), "nz": self._export(-1, 2, nnz, np.float64),
                "dense": self._export(-1, 15, nc * nc, np.float64).reshape(nc, nc, order="F"),
                "n": nc, "nnz": nnz}

    def summary(self) -> str:
        L = _lib.lib()
        m = check(L.draugr_hip_summary(self.handle, None, 0), "draugr_hip_summary")
        buf = ctypes.create_string_buffer(m + 1)
        L.draugr_hip_summary(self.handle, buf, m + 1)
        return buf.value.decode()

    # ── device-side measurement helpers ─────────────────────────────────────
    def cycle_bytes(self, config: AMGConfig) -> float:
        c = _Config(config)
        v = ctypes.c_double(0.0)
        check(_lib.lib().draugr_hip_cycle_bytes(self.handle, c.handle, ctypes.byref(v)), "draugr_hip_cycle_bytes")
        return v.value

    def op_bytes(self, level: int, which: int, precision: str, as_stored: bool = True) -> float:
        """Bytes of one pass over operator ``which`` (0 A, 1 P, 2 P^T) of
        ``level`` at ``precision``: indices as the tuned kernel stores them
        (``as_stored``) or int32 (the algorithmic count)."""
        v = ctypes.c_double(0.0)
        p = {"fp64": 0, "fp32": 1, "bf16": 2}[precision]
        check(_lib.lib().draugr_hip_op_bytes(self.handle, level, which, p, int(as_stored), ctypes.byref(v)),
              "draugr_hip_op_bytes")
        return v.value

    def tune_report(self) -> str:
        L = _lib.lib()
        m = check(L.draugr_hip_tune_report(self.handle, None, 0), "draugr_hip_tune_report")
        buf = ctypes.create_string_buffer(m + 1)
        L.draugr_hip_tune_report(self.handle, buf, m + 1)
        return buf.value.decode()

    def resident_bytes(self) -> int:
        v = ctypes.c_int64(0)
        check(_lib.lib().draugr_hip_resident_bytes(self.handle, ctypes.byref(v)), "draugr_hip_resident_bytes")
        return v.value

    def level_op(self, l, op, prec, in1=None, in2=None, out=None, out2=None):
        """Run one hot-path kernel on torch CUDA tensors (test hook)."""
        p = lambda t: None if t is None else ctypes.c_void_p(t.data_ptr())  # noqa: E731
        check(_lib.lib().draugr_hip_level_op(self.handle, l, op, PREC.get(prec, prec), p(in1), p(in2), p(out),
                                             p(out2), _torch_stream()), "draugr_hip_level_op")

    def level_op_repeat(self, l, op, prec, reps, in1=None, in2=None, out=None, out2=None):
        """level_op `reps` times back to back on torch's current stream (measurement hook)."""
        p = lambda t: None if t is None else ctypes.c_void_p(t.data_ptr())  # noqa: E731
        check(_lib.lib().draugr_hip_level_op_repeat(self.handle, l, op, PREC.get(prec, prec), p(in1), p(in2), p(out),
                                                    p(out2), reps, _torch_stream()), "draugr_hip_level_op_repeat")

    def precondition_repeat(self, b, config: AMGConfig, reps: int):
        """The PCG's preconditioner application ``reps`` times back to back on
        torch's current stream, the residual set from the float64 CUDA tensor b
        (measurement hook)."""
        c = _Config(config)
        check(_lib.lib().draugr_hip_precondition_repeat(self.handle, c.handle, _dptr(b, self.n, "b"), reps,
                                                        _torch_stream()), "draugr_hip_precondition_repeat")

    def device_values(self, l, which, count):
        a = np.empty(count)
        check(_lib.lib().draugr_hip_export_device_values(self.handle, l, which, a.ctypes.data_as(ctypes.c_void_p)),
              "draugr_hip_export_device_values")
        return a


def _setup_args(A: CSR):
    return (A.n, A.nnz, A.rowptr.ctypes.data_as(ctypes.c_void_p), A.colval.ctypes.data_as(ctypes.c_void_p),
            A.nzval.ctypes.data_as(ctypes.c_void_p))


def amg_setup(A, config: Optional[AMGConfig] = None, allow_partial_resetup: bool = True,
              upload: bool = True) -> AMGHierarchy:
    """Build the hierarchy (host setup) and upload it once to HBM
    (src/setup.jl:47-87).  ``upload=False`` keeps it on the host (no GPU needed)."""
    config = config or AMGConfig()
    A = as_csr(A)
    A = CSR(np.ascontiguousarray(A.rowptr, np.int32), np.ascontiguousarray(A.colval, np.int32),
            np.ascontiguousarray(A.nzval, np.float64))
    c = _Config(config)
    fn = _lib.lib().draugr_amg_setup if upload else _lib.lib().draugr_hip_setup_host
    h = check(fn(*_setup_args(A), c.handle, 0, 1 if allow_partial_resetup else 0),
              "draugr_amg_setup" if upload else "draugr_hip_setup_host")
    return AMGHierarchy(h, config, A.n, upload)


def upload(h: AMGHierarchy) -> AMGHierarchy:
    check(_lib.lib().draugr_hip_upload(h.handle), "draugr_hip_upload")
    h.uploaded = True
    return h


_HANDLES: "collections.OrderedDict[str, _Config]" = collections.OrderedDict()


def _cached_config(config: AMGConfig) -> _Config:
    """Config handle of the per-call entry points, reused for equal configs
    (a few dozen kept) so that a cycle or solve call does not parse JSON."""
    key = config.to_json()
    c = _HANDLES.get(key)
    if c is None:
        c = _Config(config, key)
        _HANDLES[key] = c
        if len(_HANDLES) > 32:
            _HANDLES.popitem(last=False)
    else:
        _HANDLES.move_to_end(key)
    return c


def _vec_call(kind, x, b, h: AMGHierarchy, config: AMGConfig, tol=0.0, maxiter=0, zero_guess=False):
    c = _cached_config(config)
    L = _lib.lib()
    rel = ctypes.c_double(0.0)
    if len(x) != h.n or len(b) != h.n:
        raise DraugrError("DimensionMismatch: vector length does not match the hierarchy")
    if _is_torch(x):
        import torch
        if x.dtype != torch.float64 or b.dtype != torch.float64 or not x.is_cuda or not b.is_cuda:
            raise TypeError("device vectors must be float64 CUDA tensors")
        if not (x.is_contiguous() and b.is_contiguous()):
            raise TypeError("device vectors must be contiguous")
        st = _torch_stream()
        if kind == "cycle":
            r = L.draugr_hip_cycle_device(h.handle, h.n, _ptr(x), _ptr(b), c.handle, int(zero_guess), st)
        elif kind == "solve":
            r = L.draugr_hip_solve_device(h.handle, h.n, _ptr(x), _ptr(b), c.handle, tol, maxiter, st,
                                          ctypes.byref(rel))
        else:
            r = L.draugr_hip_pcg_device(h.handle, h.n, _ptr(x), _ptr(b), c.handle, tol, maxiter, st,
                                        ctypes.byref(rel))
    else:
        if not (isinstance(x, np.ndarray) and x.dtype == np.float64 and x.flags.c_contiguous):
            raise TypeError("host vectors must be contiguous float64 numpy arrays")
        b = np.ascontiguousarray(b, dtype=np.float64)
        if kind == "cycle":
            r = L.draugr_amg_cycle(h.handle, h.n, _ptr(x), _ptr(b), c.handle)
        elif kind == "solve":
            r = L.draugr_amg_solve(h.handle, h.n, _ptr(x), _ptr(b), c.handle, tol, maxiter)
        else:
            r = L.draugr_hip_pcg(h.handle, h.n, _ptr(x), _ptr(b), c.handle, tol, maxiter, ctypes.byref(rel))
    check(r, "amg_" + kind)
    return r, rel.value


def amg_cycle_(x, b, h: AMGHierarchy, config: Optional[AMGConfig] = None, zero_guess: bool = False):
    """One V/W cycle improving x for A x = b, in place (src/cycle.jl:12-28)."""
    _vec_call("cycle", x, b, h, config or h.config, zero_guess=zero_guess)
    return x


def amg_solve_(x, b, h: AMGHierarchy, config: Optional[AMGConfig] = None, tol: float = 1e-10,
               maxiter: int = 100) -> Tuple[object, int]:
    """Stationary AMG iteration, true-residual stopping test (src/cycle.jl:110-160)."""
    it, _ = _vec_call("solve", x, b, h, config or h.config, tol, maxiter)
    return x, it


def pcg_solve_(x, b, h: AMGHierarchy, config: Optional[AMGConfig] = None, tol: float = 1e-8,
               maxiter: int = 200) -> Tuple[object, int, float]:
    """FP64 PCG with one cycle (precision from config.precision) as preconditioner.
    Returns (x, iterations, final recurrence relative residual)."""
    it, rel = _vec_call("pcg", x, b, h, config or h.config, tol, maxiter)
    return x, it, rel


def _dptr(t, n, what):
    """Device address of a contiguous float64 CUDA tensor of n entries; a CPU,
    non-float64 or strided tensor is refused before any device call."""
    import torch
    if not _is_torch(t) or not t.is_cuda or t.dtype != torch.float64 or not t.is_contiguous():
        raise TypeError("%s must be a contiguous float64 CUDA tensor" % what)
    if t.numel() != n:
        raise DraugrError("DimensionMismatch: %s has %d entries, expected %d" % (what, t.numel(), n))
    return ctypes.c_void_p(t.data_ptr())


def amg_resetup_(h: AMGHierarchy, A_new, config: Optional[AMGConfig] = None, partial: bool = True,
                 allow_partial_resetup: bool = True) -> AMGHierarchy:
    """Numeric (partial) or full resetup (src/resetup.jl:20-70).  ``A_new`` may
    also be a float64 CUDA tensor of new values (same pattern, partial only)."""
    config = config or h.config
    L = _lib.lib()
    if _is_torch(A_new):
        if not partial:
            raise DraugrError("a full resetup needs the matrix, not only its values")
        check(L.draugr_hip_resetup_device(h.handle, A_new.numel(), _dptr(A_new, A_new.numel(), "nzval"),
                                          _torch_stream()), "draugr_hip_resetup_device")
        return h
    A = as_csr(A_new)
    c = _Config(config)
    check(L.draugr_amg_resetup(h.handle, *_setup_args(A), c.handle, 0, int(partial), int(allow_partial_resetup)),
          "draugr_amg_resetup")
    if not partial:
        h.config = config
        h.n = A.n
    return h


class DraugrPreconditioner:
    """Preconditioner adapter (src/preconditioner.jl:29-102): update builds the
    hierarchy once, then resets it up; apply is z = 0, one cycle."""

    def __init__(self, config: Optional[AMGConfig] = None, **kwargs):
        self.config = config or AMGConfig(**kwargs)
        self.hierarchy: Optional[AMGHierarchy] = None
        self.dim = None

    def update_(self, A):
        A = as_csr(A)
        if self.hierarchy is None:
            self.hierarchy = amg_setup(A, self.config)
        else:
            amg_resetup_(self.hierarchy, A, self.config)
        self.dim = (A.n, A.n)
        return self

    def apply_(self, x, y):
        if _is_torch(x):
            amg_cycle_(x, y, self.hierarchy, self.config, zero_guess=True)
        else:
            x[:] = 0.0
            amg_cycle_(x, y, self.hierarchy, self.config)
        return x

    def nrows(self) -> int:
        return 0 if self.dim is None else self.dim[0]

    def ldiv_(self, x, b):
        return self.apply_(x, b)


def aslinearoperator(prec: DraugrPreconditioner):
    """The preconditioner as a ``scipy.sparse.linalg.LinearOperator`` (z = V(0; r)
    on host vectors), the role ldiv! plays for LinearSolve.jl
    (ext/DraugrLinearSolveExt.jl:37-39) and apply! for Jutul
    (ext/DraugrJutulExt.jl:172-176): ``scipy.sparse.linalg.cg(A, b, M=...)``."""
    import scipy.sparse.linalg as spla
    n = prec.nrows()

    def mv(r):
        z = np.zeros(n)
        prec.apply_(z, np.ascontiguousarray(np.asarray(r, dtype=np.float64).ravel()))
        return z

    return spla.LinearOperator((n, n), matvec=mv, dtype=np.float64)


preconditioner_update_ = DraugrPreconditioner.update_
preconditioner_apply_ = DraugrPreconditioner.apply_
ldiv_ = DraugrPreconditioner.ldiv_


def device_count() -> int:
    return int(_lib.lib().draugr_hip_device_count())
