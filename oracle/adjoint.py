"""ORACLE — test infrastructure only.

FP64 CPU restatement of the reverse-mode rule of x = A(k)^-1 b for the
cell-centred 7-point diffusion operator (the reference has no AD rules; the
rule is the textbook adjoint of a linear solve):

  * ``assemble_7pt``: the operator of permeabilities k on an m^3 grid
    (cell c = i + m (j + m z)), rows ordered z-1, y-1, x-1, diag, x+1, y+1,
    z+1; harmonic-mean face transmissibility T = 2 k_a k_b / (k_a + k_b) with
    a the lower cell; Dirichlet faces add 2 k_c to the diagonal; the diagonal
    is summed per axis (z, y, x) as: face to +1, face to -1, lower boundary,
    upper boundary.
  * ``matrix_grad``: dL/dA_ij = -lambda_i x_j on the pattern.
  * ``grad_logk_7pt``: dL/dlog k_c = sum over faces f = (c, o) of
    -(l_c - l_o)(x_c - x_o) * 2 k_c k_o^2 / (k_c + k_o)^2, faces in the order
    z-1, y-1, x-1, x+1, y+1, z+1, plus -l_c x_c 2 k_c per Dirichlet face.
  * ``adjoint``: x = A^-1 b and lambda = A^-1 dL/dx by the oracle's FP64 PCG
    on a given hierarchy (A symmetric), then the gradients above.

Pinned in tests by central finite differences on direct (scipy) solves and by
the independent generator of the operator's values.
"""
from __future__ import annotations

import numpy as np


def _grid(m):
    c = np.arange(m ** 3, dtype=np.int64)
    return c, c % m, (c // m) % m, c // (m * m)


def _neighbours(m):
    """(offset, valid mask) of the six faces in the order z-1, y-1, x-1, x+1, y+1, z+1."""
    c, i, j, z = _grid(m)
    mm = m * m
    return [(-mm, z > 0), (-m, j > 0), (-1, i > 0), (1, i < m - 1), (m, j < m - 1), (mm, z < m - 1)]


def _trans(ka, kb):
    return 2.0 * ka * kb / (ka + kb)


def assemble_7pt(m: int, k):
    """(rowptr, colval, nzval) of the 7-point operator of permeabilities k."""
    k = np.asarray(k)
    k = k if np.iscomplexobj(k) else k.astype(np.float64)  # complex k: complex-step derivatives
    n = m ** 3
    c, i, j, z = _grid(m)
    nb = _neighbours(m)
    T = {}
    for off, ok in nb:
        o = np.where(ok, c + off, c)
        lo, hi = (o, c) if off < 0 else (c, o)
        T[off] = np.where(ok, _trans(k[lo], k[hi]), 0.0)
    mm = m * m
    d = np.zeros(n, dtype=k.dtype)
    for (off_p, off_m), coord in (((mm, -mm), z), ((m, -m), j), ((1, -1), i)):
        d = d + T[off_p]
        d = d + T[off_m]
        d = d + np.where(coord == 0, 2.0 * k, 0.0)
        d = d + np.where(coord == m - 1, 2.0 * k, 0.0)
    slots_v = np.stack([-T[-mm], -T[-m], -T[-1], d, -T[1], -T[m], -T[mm]], axis=1)
    slots_c = np.stack([c - mm, c - m, c - 1, c, c + 1, c + m, c + mm], axis=1)
    valid = np.stack([nb[0][1], nb[1][1], nb[2][1], np.ones(n, bool), nb[3][1], nb[4][1], nb[5][1]], axis=1)
    counts = valid.sum(axis=1)
    rp = np.zeros(n + 1, dtype=np.int32)
    rp[1:] = np.cumsum(counts)
    return rp, slots_c[valid].astype(np.int32), slots_v[valid]


def matrix_grad(rp, cv, x, lam):
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    return -np.asarray(lam)[rows] * np.asarray(x)[cv]


def grad_logk_7pt(m: int, k, x, lam):
    k, x, lam = (np.asarray(a, dtype=np.float64) for a in (k, x, lam))
    c, i, j, z = _grid(m)
    acc = np.zeros(m ** 3)
    for off, ok in _neighbours(m):
        o = np.where(ok, c + off, c)
        dT = -(lam - lam[o]) * (x - x[o])
        s = k + k[o]
        dlog = 2.0 * k * k[o] * k[o] / (s * s)
        acc = acc + np.where(ok, dT * dlog, 0.0)
    nbd = ((z == 0).astype(int) + (z == m - 1) + (j == 0) + (j == m - 1) + (i == 0) + (i == m - 1))
    return acc + -lam * x * (2.0 * k) * nbd


def adjoint(hier, m: int, k, b, dloss_dx, tol=1e-12, maxiter=1000):
    """Gradients of L(x), x = A(k)^-1 b, on the oracle hierarchy ``hier``
    (set up on a 7-point operator of the same grid; its values are replaced).
    ``dloss_dx(x)`` gives dL/dx.  Returns (x, dL/dlog k, dL/db, dL/dA values)."""
    rp, cv, nz = assemble_7pt(m, k)
    hier.resetup(nz)
    x, _, _ = hier.pcg(np.zeros(m ** 3), b, tol=tol, maxiter=maxiter)
    lam, _, _ = hier.pcg(np.zeros(m ** 3), dloss_dx(x), tol=tol, maxiter=maxiter)
    return x, grad_logk_7pt(m, k, x, lam), lam, matrix_grad(rp, cv, x, lam)
