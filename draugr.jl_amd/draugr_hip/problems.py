"""Synthetic operators and right-hand sides for tests and benchmarks.

* ``poisson1d`` / ``poisson2d``: the reference test generators
  (test/runtests.jl:13-48), lexicographic ``idx = i + nx*j``, Dirichlet rows
  eliminated, as 0-based CSR with sorted columns.
* ``heterogeneous3d``: 7-point cell-centred diffusion on an m^3 grid,
  ``idx = i + m*(j + m*k)``, log10 k ~ N(0, sigma^2) i.i.d., harmonic-mean
  face transmissibility T = 2 k_a k_b / (k_a + k_b), Dirichlet on all faces
  through a half-cell term 2 k added to the diagonal.  SPD M-matrix.
* ``splitmix_uniform`` / ``splitmix_normal``: a portable counter-based
  generator (SplitMix64 -> 53-bit uniforms -> Box-Muller), so every run and
  every host produces the same inputs from the same seed.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z):
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def splitmix_uniform(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """n uniforms in [0, 1): u_i from counter i (SplitMix64 state seed + (i+1) * gamma)."""
    with np.errstate(over="ignore"):
        idx = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + idx * _GAMMA
        return (_mix(z) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def splitmix_normal(seed: int, n: int) -> np.ndarray:
    """n standard normals by Box-Muller on consecutive uniform pairs."""
    m = (n + 1) // 2
    u = splitmix_uniform(seed, 2 * m)
    u1, u2 = u[0::2], u[1::2]
    rad = np.sqrt(-2.0 * np.log1p(-u1))
    out = np.empty(2 * m)
    out[0::2] = rad * np.cos(2.0 * np.pi * u2)
    out[1::2] = rad * np.sin(2.0 * np.pi * u2)
    return out[:n]


@dataclass
class CSR:
    """0-based CSR with int32 indices and float64 values (square)."""
    rowptr: np.ndarray
    colval: np.ndarray
    nzval: np.ndarray

    @property
    def n(self) -> int:
        return len(self.rowptr) - 1

    @property
    def nnz(self) -> int:
        return len(self.colval)

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.nzval, self.colval, self.rowptr), shape=(self.n, self.n))

    def matvec(self, x):
        return self.to_scipy() @ x


def _from_coo(n, rows, cols, vals) -> CSR:
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    rp = np.zeros(n + 1, dtype=np.int64)
    np.add.at(rp, rows + 1, 1)
    rp = np.cumsum(rp)
    return CSR(rp.astype(np.int32), cols.astype(np.int32), vals.astype(np.float64))


def poisson1d(n: int) -> CSR:
    i = np.arange(n)
    rows = np.concatenate([i, i[1:], i[:-1]])
    cols = np.concatenate([i, i[1:] - 1, i[:-1] + 1])
    vals = np.concatenate([np.full(n, 2.0), np.full(n - 1, -1.0), np.full(n - 1, -1.0)])
    return _from_coo(n, rows, cols, vals)


def poisson2d(nx: int, ny: int | None = None) -> CSR:
    ny = nx if ny is None else ny
    n = nx * ny
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    ii, jj = ii.ravel(), jj.ravel()
    idx = ii + nx * jj
    R, C, V = [idx], [idx], [np.full(n, 4.0)]
    for di, dj in ((-1, 0), (1, 0), (0, -1), (0, 1)):
        m = (ii + di >= 0) & (ii + di < nx) & (jj + dj >= 0) & (jj + dj < ny)
        R.append(idx[m])
        C.append((ii[m] + di) + nx * (jj[m] + dj))
        V.append(np.full(int(m.sum()), -1.0))
    return _from_coo(n, np.concatenate(R), np.concatenate(C), np.concatenate(V))


def log_permeability(m: int, sigma_log10: float, seed: int) -> np.ndarray:
    """Natural log of the cell permeability: ln k = ln(10) * sigma * N(0,1)."""
    return np.log(10.0) * sigma_log10 * splitmix_normal(seed, m ** 3)


def heterogeneous3d(m: int, logk: np.ndarray) -> CSR:
    """7-point operator of the cell permeabilities exp(logk) on an m^3 grid."""
    n = m ** 3
    k = np.exp(np.asarray(logk, dtype=np.float64)).reshape(m, m, m)  # [kz, jy, ix]
    diag = np.zeros((m, m, m))
    rows, cols, vals = [], [], []
    idx = np.arange(n, dtype=np.int64).reshape(m, m, m)
    for axis in range(3):
        sl_a = [slice(None)] * 3
        sl_b = [slice(None)] * 3
        sl_a[axis] = slice(0, m - 1)
        sl_b[axis] = slice(1, m)
        ka, kb = k[tuple(sl_a)], k[tuple(sl_b)]
        T = 2.0 * ka * kb / (ka + kb)
        diag[tuple(sl_a)] += T
        diag[tuple(sl_b)] += T
        ia, ib = idx[tuple(sl_a)].ravel(), idx[tuple(sl_b)].ravel()
        t = T.ravel()
        rows += [ia, ib]
        cols += [ib, ia]
        vals += [-t, -t]
        # Dirichlet half-cell terms on the two boundary faces of this axis
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[axis] = 0
        hi[axis] = m - 1
        diag[tuple(lo)] += 2.0 * k[tuple(lo)]
        diag[tuple(hi)] += 2.0 * k[tuple(hi)]
    rows.append(idx.ravel())
    cols.append(idx.ravel())
    vals.append(diag.ravel())
    return _from_coo(n, np.concatenate(rows), np.concatenate(cols), np.concatenate(vals))


def heterogeneous3d_faces(m: int):
    """Face list of the m^3 grid: (cell_a, cell_b) index arrays (a < b), for
    the adjoint's chain rule from matrix entries to cell permeabilities."""
    idx = np.arange(m ** 3, dtype=np.int64).reshape(m, m, m)
    A, B = [], []
    for axis in range(3):
        sl_a = [slice(None)] * 3
        sl_b = [slice(None)] * 3
        sl_a[axis] = slice(0, m - 1)
        sl_b[axis] = slice(1, m)
        A.append(idx[tuple(sl_a)].ravel())
        B.append(idx[tuple(sl_b)].ravel())
    return np.concatenate(A), np.concatenate(B)
