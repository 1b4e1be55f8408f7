"""ORACLE — test infrastructure only.

FP64 CPU restatement of Draugr.jl's AMG solve / cycle / numeric-resetup path
(C code in ``oracle.c``, built to ``oracle/_build/liboracle.so``), wrapped for
numpy.  Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg
of ``bench.py`` use it, as the checker.  The product never imports it.

The hierarchy (levels, P, transpose and triple maps, colouring) is an input:
parity is defined on a fixed hierarchy.  ``levels`` is a list of dicts with
numpy arrays (0-based indices):

  A_rp, A_cv, A_nz, P_rp, P_cv, P_nz, Pt_off, Pt_rows, Pt_pnz,
  optional R_off (int64), R_pi, R_ai, R_pj, c_off, c_order

plus the coarsest sparse operator (coarse_rp, coarse_cv, coarse_nz).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

SMOOTHER_CODES = {"jacobi": 0, "colored_gs": 1, "spai0": 3, "l1_jacobi": 5, "l1_colored_gs": 6, "chebyshev": 8,
                  "ilu0": 9}

_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_f64p = ctypes.POINTER(ctypes.c_double)


class _OLevel(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int32), ("nc", ctypes.c_int32),
        ("a_rp", _i32p), ("a_cv", _i32p), ("a_nz", _f64p),
        ("p_rp", _i32p), ("p_cv", _i32p), ("p_nz", _f64p),
        ("t_off", _i32p), ("t_rows", _i32p), ("t_pnz", _i32p),
        ("r_off", _i64p), ("r_pi", _i32p), ("r_ai", _i32p), ("r_pj", _i32p),
        ("ncolors", ctypes.c_int32), ("c_off", _i32p), ("c_order", _i32p),
        ("dinv", _f64p), ("r", _f64p), ("tmp", _f64p), ("xc", _f64p), ("bc", _f64p),
        ("cheb_lmin", ctypes.c_double), ("cheb_lmax", ctypes.c_double), ("cheb_d", _f64p),
        ("ilu_diag", _i32p), ("ilu_l", _f64p), ("ilu_u", _f64p),
    ]


class _OHier(ctypes.Structure):
    _fields_ = [
        ("nlevels", ctypes.c_int32), ("lv", ctypes.POINTER(_OLevel)),
        ("ncoarse", ctypes.c_int32), ("coarse_lu", _f64p), ("coarse_piv", _i32p),
        ("coarse_b", _f64p), ("coarse_x", _f64p),
        ("ac_rp", _i32p), ("ac_cv", _i32p), ("ac_nz", _f64p),
        ("smoother", ctypes.c_int32), ("omega", ctypes.c_double),
        ("w_cycle", ctypes.c_int32), ("pre", ctypes.c_int32), ("post", ctypes.c_int32),
        ("gs_sym", ctypes.c_int32),
    ]


def build():
    """Compile the oracle's C restatement (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.o_pcg.restype = ctypes.c_int32
        L.o_solve.restype = ctypes.c_int32
        L.o_lu.restype = ctypes.c_int32
        L.o_resetup.restype = ctypes.c_int32
        _lib = L
    return _lib


def _p(a, t):
    if a is None:
        return ctypes.cast(None, t)
    return a.ctypes.data_as(t)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ── single kernels (numpy in / numpy out) ───────────────────────────────────
def spmv(rp, cv, nz, x):
    rp, cv, nz, x = _i32(rp), _i32(cv), _f64(nz), _f64(x)
    y = np.empty(len(rp) - 1)
    lib().o_spmv(ctypes.c_int32(len(y)), _p(rp, _i32p), _p(cv, _i32p), _p(nz, _f64p), _p(x, _f64p), _p(y, _f64p))
    return y


def residual(rp, cv, nz, x, b):
    rp, cv, nz, x, b = _i32(rp), _i32(cv), _f64(nz), _f64(x), _f64(b)
    r = np.empty(len(rp) - 1)
    lib().o_residual(ctypes.c_int32(len(r)), _p(rp, _i32p), _p(cv, _i32p), _p(nz, _f64p), _p(x, _f64p),
                     _p(b, _f64p), _p(r, _f64p))
    return r


def invdiag(rp, cv, nz, l1):
    """mode (l1): 0 plain 1 / a_ii, 1 l1, 2 SPAI(0) a_ii / sum a_ij^2."""
    rp, cv, nz = _i32(rp), _i32(cv), _f64(nz)
    d = np.empty(len(rp) - 1)
    lib().o_invdiag(ctypes.c_int32(len(d)), _p(rp, _i32p), _p(cv, _i32p), _p(nz, _f64p), ctypes.c_int32(int(l1)),
                    _p(d, _f64p))
    return d


def cheb_smooth(rp, cv, nz, x, b, dinv, lmin, lmax, steps=1):
    """Chebyshev smoothing of x (degree 3) with the given bounds; returns the new x."""
    x = _f64(x).copy()
    n = len(x)
    rp, cv, nz, b, dinv = _i32(rp), _i32(cv), _f64(nz), _f64(b), _f64(dinv)
    r, d = np.zeros(n), np.zeros(n)
    lib().o_cheb_smooth(ctypes.c_int32(n), _p(rp, _i32p), _p(cv, _i32p), _p(nz, _f64p), _p(x, _f64p),
                        _p(b, _f64p), _p(dinv, _f64p), ctypes.c_double(lmin), ctypes.c_double(lmax),
                        ctypes.c_int32(steps), _p(r, _f64p), _p(d, _f64p))
    return x


def ilu0_diag(rp, cv):
    """First stored diagonal position of every row (-1 when absent)."""
    rp, cv = _i32(rp), _i32(cv)
    n = len(rp) - 1
    d = np.full(n, -1, dtype=np.int32)
    rows = np.repeat(np.arange(n), np.diff(rp))
    hit = np.nonzero(cv == rows)[0]
    first = np.unique(rows[hit], return_index=True)
    d[first[0]] = hit[first[1]]
    return d


def ilu0_factor(rp, cv, nz):
    """ILU(0) factors (L strict lower, U upper with diagonal) on A's pattern."""
    rp, cv, nz = _i32(rp), _i32(cv), _f64(nz)
    diag = ilu0_diag(rp, cv)
    L, U = np.zeros(len(nz)), np.zeros(len(nz))
    lib().o_ilu0_factor(ctypes.c_int32(len(rp) - 1), _p(rp, _i32p), _p(cv, _i32p), _p(nz, _f64p), _p(diag, _i32p),
                        _p(L, _f64p), _p(U, _f64p))
    return diag, L, U


def ilu0_smooth(rp, cv, nz, x, b, steps=1):
    """ILU(0) smoothing of x with the oracle's own factors; returns the new x."""
    rp, cv, nz, b = _i32(rp), _i32(cv), _f64(nz), _f64(b)
    diag, L, U = ilu0_factor(rp, cv, nz)
    x = _f64(x).copy()
    t = np.zeros(len(x))
    lib().o_ilu0_smooth(ctypes.c_int32(len(x)), _p(rp, _i32p), _p(cv, _i32p), _p(nz, _f64p), _p(diag, _i32p),
                        _p(L, _f64p), _p(U, _f64p), _p(x, _f64p), _p(b, _f64p), _p(t, _f64p), ctypes.c_int32(steps))
    return x


def power_rho(rp, cv, nz, dinv, v0, niter=10):
    """Spectral-radius estimate of D^-1 A by power iteration from the unit vector
    v0 (src/smoothers.jl:866-885): w = (A v) .* dinv, lambda = ||w||,
    v = w / lambda while lambda > eps.  numpy restatement."""
    import scipy.sparse as sp
    n = len(rp) - 1
    A = sp.csr_matrix((_f64(nz), _i32(cv), _i32(rp)), shape=(n, n))
    v = _f64(v0).copy()
    lam = 1.0
    for _ in range(niter):
        w = (A @ v) * dinv
        lam = float(np.sqrt(np.dot(w, w)))
        if lam > np.finfo(float).eps:
            v = w / lam
    return lam


def jacobi_step(rp, cv, nz, x, b, d, w):
    rp, cv, nz, x, b, d = _i32(rp), _i32(cv), _f64(nz), _f64(x), _f64(b), _f64(d)
    xo = np.empty(len(rp) - 1)
    lib().o_jacobi_step(ctypes.c_int32(len(xo)), _p(rp, _i32p), _p(cv, _i32p), _p(nz, _f64p), _p(x, _f64p),
                        _p(b, _f64p), _p(d, _f64p), ctypes.c_double(w), _p(xo, _f64p))
    return xo


def gs_sweep(rp, cv, nz, x, b, d, c_off, c_order):
    rp, cv, nz, b, d = _i32(rp), _i32(cv), _f64(nz), _f64(b), _f64(d)
    c_off, c_order = _i32(c_off), _i32(c_order)
    x = _f64(x).copy()
    lib().o_gs_sweep(ctypes.c_int32(len(x)), _p(rp, _i32p), _p(cv, _i32p), _p(nz, _f64p), _p(x, _f64p),
                     _p(b, _f64p), _p(d, _f64p), ctypes.c_int32(len(c_off) - 1), _p(c_off, _i32p),
                     _p(c_order, _i32p))
    return x


def restrict(off, rows, pnz, pv, r):
    off, rows, pnz, pv, r = _i32(off), _i32(rows), _i32(pnz), _f64(pv), _f64(r)
    bc = np.empty(len(off) - 1)
    lib().o_restrict(ctypes.c_int32(len(bc)), _p(off, _i32p), _p(rows, _i32p), _p(pnz, _i32p), _p(pv, _f64p),
                     _p(r, _f64p), _p(bc, _f64p))
    return bc


def prolong(rp, cv, pv, xc, x):
    rp, cv, pv, xc = _i32(rp), _i32(cv), _f64(pv), _f64(xc)
    x = _f64(x).copy()
    lib().o_prolong(ctypes.c_int32(len(rp) - 1), _p(rp, _i32p), _p(cv, _i32p), _p(pv, _f64p), _p(xc, _f64p),
                    _p(x, _f64p))
    return x


def galerkin(off, pi, ai, pj, P, A):
    off = np.ascontiguousarray(off, dtype=np.int64)
    pi, ai, pj, P, A = _i32(pi), _i32(ai), _i32(pj), _f64(P), _f64(A)
    out = np.empty(len(off) - 1)
    lib().o_galerkin(ctypes.c_int64(len(out)), _p(off, _i64p), _p(pi, _i32p), _p(ai, _i32p), _p(pj, _i32p),
                     _p(P, _f64p), _p(A, _f64p), _p(out, _f64p))
    return out


def lu_solve_dense(A_dense, b):
    """Solve with the oracle's partial-pivoting LU (A_dense: 2-D numpy array)."""
    n = A_dense.shape[0]
    lu = np.asfortranarray(A_dense, dtype=np.float64).copy(order="F")
    piv = np.zeros(n, dtype=np.int32)
    code = lib().o_lu(ctypes.c_int32(n), lu.ctypes.data_as(_f64p), _p(piv, _i32p))
    if code:
        raise ZeroDivisionError("singular matrix (zero pivot at step %d)" % (code - 1))
    x = np.empty(n)
    bb = _f64(b)
    lib().o_lu_solve(ctypes.c_int32(n), lu.ctypes.data_as(_f64p), _p(piv, _i32p), _p(bb, _f64p), _p(x, _f64p))
    return x


# ── whole hierarchy ──────────────────────────────────────────────────────────
class OracleHierarchy:
    """A fixed hierarchy on which the oracle runs cycles, solves and resetups."""

    def __init__(self, levels, coarse_rp, coarse_cv, coarse_nz, smoother="l1_jacobi", omega=2.0 / 3.0,
                 cycle="v", pre=1, post=1, cheb=None, gs_symmetric=False):
        self._keep = []
        keep = self._keep.append
        L = len(levels)
        self.levels = []
        arr = (_OLevel * max(L, 1))()
        for l, lv in enumerate(levels):
            d = {}
            d["A_rp"], d["A_cv"] = _i32(lv["A_rp"]), _i32(lv["A_cv"])
            d["A_nz"] = _f64(lv["A_nz"]).copy()
            d["P_rp"], d["P_cv"], d["P_nz"] = _i32(lv["P_rp"]), _i32(lv["P_cv"]), _f64(lv["P_nz"])
            d["Pt_off"], d["Pt_rows"], d["Pt_pnz"] = _i32(lv["Pt_off"]), _i32(lv["Pt_rows"]), _i32(lv["Pt_pnz"])
            n = len(d["A_rp"]) - 1
            nc = len(d["Pt_off"]) - 1
            for k in ("R_pi", "R_ai", "R_pj", "c_off", "c_order"):
                d[k] = _i32(lv[k]) if lv.get(k) is not None else None
            d["R_off"] = np.ascontiguousarray(lv["R_off"], dtype=np.int64) if lv.get("R_off") is not None else None
            d["dinv"], d["r"], d["tmp"] = np.zeros(n), np.zeros(n), np.zeros(n)
            d["xc"], d["bc"] = np.zeros(nc), np.zeros(nc)
            self.levels.append(d)
            keep(d)
            o = arr[l]
            o.n, o.nc = n, nc
            o.a_rp, o.a_cv, o.a_nz = _p(d["A_rp"], _i32p), _p(d["A_cv"], _i32p), _p(d["A_nz"], _f64p)
            o.p_rp, o.p_cv, o.p_nz = _p(d["P_rp"], _i32p), _p(d["P_cv"], _i32p), _p(d["P_nz"], _f64p)
            o.t_off, o.t_rows, o.t_pnz = _p(d["Pt_off"], _i32p), _p(d["Pt_rows"], _i32p), _p(d["Pt_pnz"], _i32p)
            o.r_off = _p(d["R_off"], _i64p)
            o.r_pi, o.r_ai, o.r_pj = _p(d["R_pi"], _i32p), _p(d["R_ai"], _i32p), _p(d["R_pj"], _i32p)
            o.ncolors = (len(d["c_off"]) - 1) if d["c_off"] is not None else 0
            o.c_off, o.c_order = _p(d["c_off"], _i32p), _p(d["c_order"], _i32p)
            o.dinv, o.r, o.tmp = _p(d["dinv"], _f64p), _p(d["r"], _f64p), _p(d["tmp"], _f64p)
            o.xc, o.bc = _p(d["xc"], _f64p), _p(d["bc"], _f64p)
            d["cheb_d"] = np.zeros(n)
            o.cheb_d = _p(d["cheb_d"], _f64p)
            d["ilu_diag"] = ilu0_diag(d["A_rp"], d["A_cv"])
            d["ilu_l"], d["ilu_u"] = np.zeros(len(d["A_cv"])), np.zeros(len(d["A_cv"]))
            o.ilu_diag = _p(d["ilu_diag"], _i32p)
            o.ilu_l, o.ilu_u = _p(d["ilu_l"], _f64p), _p(d["ilu_u"], _f64p)
            if cheb is not None:
                o.cheb_lmin, o.cheb_lmax = float(cheb[l][0]), float(cheb[l][1])
        self._arr = arr
        nco = len(coarse_rp) - 1
        self.coarse = {"rp": _i32(coarse_rp), "cv": _i32(coarse_cv), "nz": _f64(coarse_nz).copy()}
        self.coarse_lu = np.zeros(nco * nco)
        self.coarse_piv = np.zeros(max(nco, 1), dtype=np.int32)
        self.coarse_b, self.coarse_x = np.zeros(max(nco, 1)), np.zeros(max(nco, 1))
        h = _OHier()
        h.nlevels, h.lv = L, ctypes.cast(arr, ctypes.POINTER(_OLevel))
        h.ncoarse = nco
        h.coarse_lu, h.coarse_piv = _p(self.coarse_lu, _f64p), _p(self.coarse_piv, _i32p)
        h.coarse_b, h.coarse_x = _p(self.coarse_b, _f64p), _p(self.coarse_x, _f64p)
        h.ac_rp, h.ac_cv, h.ac_nz = (_p(self.coarse["rp"], _i32p), _p(self.coarse["cv"], _i32p),
                                     _p(self.coarse["nz"], _f64p))
        h.smoother = SMOOTHER_CODES[smoother]
        h.omega = omega
        h.w_cycle = 1 if str(cycle).lower() == "w" else 0
        h.pre, h.post = pre, post
        h.gs_sym = 1 if gs_symmetric else 0
        self._h = h
        self.n = (len(self.levels[0]["A_rp"]) - 1) if L else nco
        lib().o_prepare(ctypes.byref(h))
        self._factor_coarse()

    def _factor_coarse(self):
        nco = self._h.ncoarse
        self.coarse_lu[:] = 0.0
        c = self.coarse
        lu = self.coarse_lu.reshape(nco, nco, order="F") if nco else self.coarse_lu
        for i in range(nco):
            for k in range(c["rp"][i], c["rp"][i + 1]):
                lu[i, c["cv"][k]] = c["nz"][k]
        code = lib().o_lu(ctypes.c_int32(nco), _p(self.coarse_lu, _f64p), _p(self.coarse_piv, _i32p))
        if code:
            raise ZeroDivisionError("coarsest matrix singular")

    def set_cheb(self, cheb):
        """Chebyshev bounds (lambda_min, lambda_max) per level."""
        for l, (lo, hi) in enumerate(cheb):
            self._arr[l].cheb_lmin, self._arr[l].cheb_lmax = float(lo), float(hi)

    def set_options(self, cycle=None, pre=None, post=None):
        if cycle is not None:
            self._h.w_cycle = 1 if str(cycle).lower() == "w" else 0
        if pre is not None:
            self._h.pre = pre
        if post is not None:
            self._h.post = post

    def finest(self):
        if self.levels:
            d = self.levels[0]
            return d["A_rp"], d["A_cv"], d["A_nz"]
        return self.coarse["rp"], self.coarse["cv"], self.coarse["nz"]

    def cycle(self, x, b):
        x = _f64(x).copy()
        b = _f64(b)
        lib().o_cycle(ctypes.byref(self._h), _p(x, _f64p), _p(b, _f64p))
        return x

    def solve(self, x, b, tol=1e-10, maxiter=100):
        x = _f64(x).copy()
        b = _f64(b)
        r = np.zeros(len(x))
        rel = ctypes.c_double(0.0)
        it = lib().o_solve(ctypes.byref(self._h), ctypes.c_int32(len(x)), _p(x, _f64p), _p(b, _f64p),
                           ctypes.c_double(tol), ctypes.c_int32(maxiter), _p(r, _f64p), ctypes.byref(rel))
        return x, int(it), rel.value

    def pcg(self, x, b, tol=1e-8, maxiter=200):
        rp, cv, nz = self.finest()
        x = _f64(x).copy()
        b = _f64(b)
        work = np.zeros(4 * len(x))
        rel = ctypes.c_double(0.0)
        it = lib().o_pcg(ctypes.byref(self._h), ctypes.c_int32(len(x)), _p(rp, _i32p), _p(cv, _i32p),
                         _p(nz, _f64p), _p(x, _f64p), _p(b, _f64p), ctypes.c_double(tol), ctypes.c_int32(maxiter),
                         _p(work, _f64p), ctypes.byref(rel))
        return x, int(it), rel.value

    def resetup(self, new_nz):
        """Partial resetup with new fine-level values (same pattern)."""
        if self.levels:
            self.levels[0]["A_nz"][:] = _f64(new_nz)
        else:
            self.coarse["nz"][:] = _f64(new_nz)
        if self.levels:
            code = lib().o_resetup(ctypes.byref(self._h))
            if code:
                raise ZeroDivisionError("coarsest matrix singular after resetup")
        else:
            self._factor_coarse()

    def level_values(self, l):
        return self.levels[l]["A_nz"] if l < len(self.levels) else self.coarse["nz"]

    def dinv(self, l):
        return self.levels[l]["dinv"]
