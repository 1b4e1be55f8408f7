"""Independent dense restatement of the AMG cycle for tiny problems (numpy).

A second, deliberately different implementation of the cycle semantics of
src/cycle.jl:63-99 (matrices as dense arrays, smoothing as matrix formulas,
coarse solve by numpy.linalg.solve), used to pin the C oracle: a dropped term,
a wrong sign or a transposed operand in either shows up as a mismatch.
"""
import numpy as np


def dense(rp, cv, nz, ncol=None):
    n = len(rp) - 1
    M = np.zeros((n, ncol if ncol is not None else n))
    for i in range(n):
        for k in range(rp[i], rp[i + 1]):
            M[i, cv[k]] += nz[k]
    return M


def invdiag(A, l1, spai0=False):
    if spai0:
        s = (A * A).sum(axis=1)
        return np.where(s > 0, np.diag(A) / np.where(s == 0, 1, s), 0.0)
    if l1:
        s = np.abs(A).sum(axis=1)
        return np.where(s > np.finfo(float).eps, 1.0 / np.where(s == 0, 1, s), 0.0)
    d = np.diag(A).copy()
    thr = np.finfo(float).eps * np.maximum(1.0, np.abs(A).sum(axis=1))
    return np.where(np.abs(d) > thr, 1.0 / np.where(d == 0, 1, d), 0.0)


def ilu0(A):
    """ILU(0) of a dense matrix on its own nonzero pattern (IKJ order, the
    reference's safeguards: src/smoothers.jl:1016-1079)."""
    n = A.shape[0]
    pat = A != 0
    np.fill_diagonal(pat, True)
    U = A.astype(float).copy()
    L = np.zeros_like(U)
    eps = np.finfo(float).eps
    rn = np.abs(A).sum(axis=1)
    for i in range(n):
        for k in np.nonzero(pat[i, :i])[0]:
            if abs(U[k, k]) < eps * max(1.0, rn[k]):
                L[i, k] = 0.0
                U[i, k] = 0.0
                continue
            lik = U[i, k] / U[k, k]
            if abs(lik) > 1e8:
                lik = lik * (1e8 / abs(lik))
            L[i, k] = lik
            U[i, k] = 0.0
            for j in np.nonzero(pat[k, k + 1:])[0] + k + 1:
                if pat[i, j]:
                    U[i, j] -= lik * U[k, j]
        thr = eps * max(1.0, rn[i])
        if abs(U[i, i]) < thr:
            U[i, i] = thr
    return L, np.triu(U)


class DenseAMG:
    def __init__(self, levels, coarse, smoother, omega=2.0 / 3.0, cycle="v", pre=1, post=1, cheb=None,
                 gs_symmetric=False):
        self.A = [dense(l["A_rp"], l["A_cv"], l["A_nz"]) for l in levels]
        self.P = [dense(l["P_rp"], l["P_cv"], l["P_nz"], len(l["Pt_off"]) - 1) for l in levels]
        self.colors = [(l.get("c_off"), l.get("c_order")) for l in levels]
        self.Ac = dense(coarse["rp"], coarse["cv"], coarse["nz"])
        self.smoother = smoother
        self.l1 = smoother.startswith("l1")
        self.D = [invdiag(A, self.l1, smoother == "spai0") for A in self.A]
        self.LU = [ilu0(A) for A in self.A] if smoother == "ilu0" else None
        self.cheb = cheb
        self.omega, self.cycle_type, self.pre, self.post = omega, cycle, pre, post
        self.gs_symmetric = gs_symmetric

    def smooth(self, l, x, b, steps, post=False):
        A, D = self.A[l], self.D[l]
        if self.smoother == "chebyshev":
            lo, hi = self.cheb[l]
            th, de = (hi + lo) / 2, (hi - lo) / 2
            for _ in range(steps):
                d = D * (b - A @ x) / th
                x = x + d
                so = th / de
                for _k in range(2):
                    sn = 1 / (2 * th / de - so)
                    d = 2 * sn / de * D * (b - A @ x) + sn * so * d
                    x = x + d
                    so = sn
            return x
        if self.smoother == "ilu0":
            Lf, Uf = self.LU[l]
            M = (np.eye(len(x)) + Lf) @ Uf
            for _ in range(steps):
                x = x + np.linalg.solve(M, b - A @ x)
            return x
        if self.smoother == "spai0":
            for _ in range(steps):
                x = x + D * (b - A @ x)
            return x
        if "gs" in self.smoother:
            off, order = self.colors[l]
            colours = list(range(len(off) - 1))
            if post and self.gs_symmetric:
                colours.reverse()
            for _ in range(steps):
                for c in colours:
                    rows = order[off[c]:off[c + 1]]
                    r = b[rows] - A[rows] @ x
                    x = x.copy()
                    x[rows] += D[rows] * r
            return x
        for _ in range(steps):
            x = x + self.omega * D * (b - A @ x)
        return x

    def descend(self, l, x, b):
        x = self.smooth(l, x, b, self.pre)
        r = b - self.A[l] @ x
        bc = self.P[l].T @ r
        xc = np.zeros(len(bc))
        if l + 1 < len(self.A):
            for _ in range(2 if self.cycle_type == "w" else 1):
                xc = self.descend(l + 1, xc, bc)
        else:
            xc = np.linalg.solve(self.Ac, bc)
        x = x + self.P[l] @ xc
        return self.smooth(l, x, b, self.post, post=True)

    def cycle(self, x, b):
        if not self.A:
            return np.linalg.solve(self.Ac, b)
        return self.descend(0, np.array(x, dtype=float), np.asarray(b, dtype=float))
