/*
 * ORACLE — test infrastructure only.  FP64 CPU restatement of Draugr.jl's AMG
 * solve / cycle / numeric-resetup path, used by tests/, the smoke check and
 * the CPU-baseline leg of bench.py as the checker.  It is never linked into
 * or called by the product library.
 *
 * Every function follows the reference routine cited beside it, in the
 * reference's own summation order (sequential over a row, left to right), so
 * that its results are the reference's results on the same hierarchy.  The
 * hierarchy itself (A_l, P_l, transpose map, triple map, colouring) is input.
 * Parallelism: OpenMP over rows (order inside a row unchanged), so results do
 * not depend on the thread count.
 *
 * Pinning: see oracle/README.md (which test checks which function against
 * what).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define EPS64 2.220446049250313e-16

typedef struct {
    int32_t n, nc;                 /* rows of A_l, columns of P_l */
    const int32_t *a_rp, *a_cv;    /* A_l */
    double *a_nz;
    const int32_t *p_rp, *p_cv;    /* P_l */
    const double *p_nz;
    const int32_t *t_off, *t_rows, *t_pnz; /* transpose map of P_l */
    const int64_t *r_off;          /* triple map of A_{l+1} (may be NULL) */
    const int32_t *r_pi, *r_ai, *r_pj;
    int32_t ncolors;
    const int32_t *c_off, *c_order; /* colouring (may be NULL) */
    double *dinv;                  /* smoother inverse diagonal (n) */
    double *r, *tmp, *xc, *bc;     /* workspaces: n, n, nc, nc */
    double cheb_lmin, cheb_lmax;   /* Chebyshev bounds of D^-1 A (input) */
    double *cheb_d;                /* Chebyshev direction vector (n) */
    int32_t *ilu_diag;             /* ILU(0): diagonal positions (n), factors on A's pattern (nnz) */
    double *ilu_l, *ilu_u;
} OLevel;

typedef struct {
    int32_t nlevels;
    OLevel *lv;
    int32_t ncoarse;
    double *coarse_lu;             /* column-major LU factors (ncoarse^2) */
    int32_t *coarse_piv;
    double *coarse_b, *coarse_x;
    const int32_t *ac_rp, *ac_cv;  /* coarsest sparse operator */
    double *ac_nz;
    int32_t smoother;              /* 0 jacobi, 1 colored gs, 3 spai0, 5 l1 jacobi, 6 l1 colored gs, 8 chebyshev, 9 ilu0 */
    double omega;
    int32_t w_cycle, pre, post;
    int32_t gs_sym;                /* coloured GS post-sweeps in reverse colour order (no reference counterpart) */
} OHier;

/* y = A x, one row at a time (src/csr_matrix.jl:150-159). */
void o_spmv(int32_t n, const int32_t *rp, const int32_t *cv, const double *nz, const double *x, double *y) {
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < n; ++i) {
        double s = 0.0;
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k) s += nz[k] * x[cv[k]];
        y[i] = s;
    }
}

/* r = b - (sum_j a_ij x_j) (src/cycle.jl:32-43). */
void o_residual(int32_t n, const int32_t *rp, const int32_t *cv, const double *nz, const double *x,
                const double *b, double *r) {
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < n; ++i) {
        double s = 0.0;
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k) s += nz[k] * x[cv[k]];
        r[i] = b[i] - s;
    }
}

/* Smoother inverse diagonal: plain (src/smoothers.jl:47-63) or l1
 * (src/smoothers.jl:816-827). */
void o_invdiag(int32_t n, const int32_t *rp, const int32_t *cv, const double *nz, int32_t mode, double *d) {
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < n; ++i) {
        double nrm = 0.0, dg = 0.0;
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
            nrm += mode == 2 ? nz[k] * nz[k] : fabs(nz[k]);
            if (cv[k] == i) dg = nz[k];
        }
        if (mode == 1) {
            d[i] = nrm > EPS64 ? 1.0 / nrm : 0.0;
        } else if (mode == 2) { /* SPAI(0), src/smoothers.jl:517-532 */
            d[i] = nrm > 0.0 ? dg / nrm : 0.0;
        } else {
            double thr = EPS64 * (nrm > 1.0 ? nrm : 1.0);
            d[i] = fabs(dg) > thr ? 1.0 / dg : 0.0;
        }
    }
}

/* One weighted Jacobi step out of place: the residual accumulated by
 * successive subtraction from b_i (src/smoothers.jl:78-94). */
void o_jacobi_step(int32_t n, const int32_t *rp, const int32_t *cv, const double *nz, const double *x,
                   const double *b, const double *d, double w, double *xo) {
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < n; ++i) {
        double ri = b[i];
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k) ri -= nz[k] * x[cv[k]];
        xo[i] = x[i] + w * d[i] * ri;
    }
}

/* One coloured Gauss-Seidel sweep in place, colours ascending
 * (src/smoothers.jl:203-241), or descending (reverse != 0: the transposed
 * sweep of the symmetric option, which the reference does not have).  Rows of
 * one colour are independent. */
static void o_gs_sweep_dir(const int32_t *rp, const int32_t *cv, const double *nz, double *x, const double *b,
                           const double *d, int32_t ncolors, const int32_t *coff, const int32_t *order, int reverse) {
    for (int32_t cc = 0; cc < ncolors; ++cc) {
        const int32_t c = reverse ? ncolors - 1 - cc : cc;
#pragma omp parallel for schedule(static)
        for (int32_t t = coff[c]; t < coff[c + 1]; ++t) {
            const int32_t i = order[t];
            double ri = b[i];
            for (int32_t k = rp[i]; k < rp[i + 1]; ++k) ri -= nz[k] * x[cv[k]];
            x[i] += d[i] * ri;
        }
    }
}

void o_gs_sweep(int32_t n, const int32_t *rp, const int32_t *cv, const double *nz, double *x, const double *b,
                const double *d, int32_t ncolors, const int32_t *coff, const int32_t *order) {
    (void)n;
    o_gs_sweep_dir(rp, cv, nz, x, b, d, ncolors, coff, order, 0);
}

/* bc_J = sum over the transpose map of P[p_nz_idx] * r[fine_row]
 * (src/prolongation.jl:805-827). */
void o_restrict(int32_t nc, const int32_t *off, const int32_t *rows, const int32_t *pnz, const double *pv,
                const double *r, double *bc) {
#pragma omp parallel for schedule(static)
    for (int32_t J = 0; J < nc; ++J) {
        double acc = 0.0;
        for (int32_t k = off[J]; k < off[J + 1]; ++k) acc += pv[pnz[k]] * r[rows[k]];
        bc[J] = acc;
    }
}

/* x_i += P_iJ xc_J, accumulated into x one entry at a time
 * (src/prolongation.jl:742-759). */
void o_prolong(int32_t n, const int32_t *rp, const int32_t *cv, const double *pv, const double *xc, double *x) {
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < n; ++i)
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k) x[i] += pv[k] * xc[cv[k]];
}

/* Ac[k] = sum_t P[pi_t] * A[ai_t] * P[pj_t] (src/galerkin.jl:212-222). */
void o_galerkin(int64_t nnzc, const int64_t *off, const int32_t *pi, const int32_t *ai, const int32_t *pj,
                const double *P, const double *A, double *Ac) {
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < nnzc; ++k) {
        double acc = 0.0;
        for (int64_t t = off[k]; t < off[k + 1]; ++t) acc += P[pi[t]] * A[ai[t]] * P[pj[t]];
        Ac[k] = acc;
    }
}

/* Dense LU with partial pivoting, column-major (LAPACK getrf semantics:
 * largest |a|, first index on ties).  Returns 0, or k+1 for a zero pivot. */
int32_t o_lu(int32_t n, double *a, int32_t *piv) {
    for (int32_t k = 0; k < n; ++k) {
        int32_t p = k;
        double mx = fabs(a[(size_t)k * n + k]);
        for (int32_t r = k + 1; r < n; ++r)
            if (fabs(a[(size_t)k * n + r]) > mx) { mx = fabs(a[(size_t)k * n + r]); p = r; }
        piv[k] = p;
        if (a[(size_t)k * n + p] == 0.0) return k + 1;
        if (p != k)
            for (int32_t c = 0; c < n; ++c) {
                double t = a[(size_t)c * n + k];
                a[(size_t)c * n + k] = a[(size_t)c * n + p];
                a[(size_t)c * n + p] = t;
            }
        const double dkk = a[(size_t)k * n + k];
        for (int32_t r = k + 1; r < n; ++r) a[(size_t)k * n + r] /= dkk;
        for (int32_t c = k + 1; c < n; ++c) {
            const double u = a[(size_t)c * n + k];
            for (int32_t r = k + 1; r < n; ++r) a[(size_t)c * n + r] -= a[(size_t)k * n + r] * u;
        }
    }
    return 0;
}

/* x = A^-1 b from the LU factors, in the loop order of reference LAPACK
 * getrs for one right-hand side: row interchanges (laswp), then the
 * column-oriented unit-lower and upper solves of reference BLAS trsm
 * ('L','L','N','U' and 'L','U','N','N'). */
void o_lu_solve(int32_t n, const double *lu, const int32_t *piv, const double *b, double *x) {
    if (x != b) memcpy(x, b, (size_t)n * sizeof(double));
    for (int32_t k = 0; k < n; ++k) {
        double t = x[k];
        x[k] = x[piv[k]];
        x[piv[k]] = t;
    }
    for (int32_t j = 0; j < n; ++j)
        for (int32_t i = j + 1; i < n; ++i) x[i] -= lu[(size_t)j * n + i] * x[j];
    for (int32_t j = n - 1; j >= 0; --j) {
        x[j] /= lu[(size_t)j * n + j];
        for (int32_t i = 0; i < j; ++i) x[i] -= lu[(size_t)j * n + i] * x[j];
    }
}

/* Chebyshev smoothing, degree 3 (src/smoothers.jl:952-990): per step
 * r = b - A x; d = dinv r / theta; x += d; then twice
 * sigma' = 1 / (2 theta / delta - sigma); d = (2 sigma' / delta) dinv r + sigma' sigma d; x += d. */
void o_cheb_smooth(int32_t n, const int32_t *rp, const int32_t *cv, const double *nz, double *x, const double *b,
                   const double *dinv, double lmin, double lmax, int32_t steps, double *r, double *d) {
    const double theta = (lmax + lmin) / 2.0, delta = (lmax - lmin) / 2.0;
    for (int32_t s = 0; s < steps; ++s) {
        o_residual(n, rp, cv, nz, x, b, r);
        const double inv_theta = 1.0 / theta;
#pragma omp parallel for schedule(static)
        for (int32_t i = 0; i < n; ++i) {
            d[i] = dinv[i] * r[i] * inv_theta;
            x[i] += d[i];
        }
        double sigma_old = theta / delta;
        for (int k = 1; k < 3; ++k) {
            o_residual(n, rp, cv, nz, x, b, r);
            const double sigma_new = 1.0 / (2.0 * theta / delta - sigma_old);
            const double cr = 2.0 * sigma_new / delta, cd = sigma_new * sigma_old;
#pragma omp parallel for schedule(static)
            for (int32_t i = 0; i < n; ++i) {
                d[i] = cr * dinv[i] * r[i] + cd * d[i];
                x[i] += d[i];
            }
            sigma_old = sigma_new;
        }
    }
}

/* ILU(0) factorization on A's pattern (src/smoothers.jl:1016-1079): IKJ order,
 * a pivot below eps max(1, ||a_k||_1) zeroes the entry, multipliers clamped
 * to 1e8 in magnitude, tiny diagonals raised to the threshold.  diag: the
 * first stored diagonal position of each row. */
void o_ilu0_factor(int32_t n, const int32_t *rp, const int32_t *cv, const double *nz, const int32_t *diag, double *l,
                   double *u) {
    const int32_t nnz = rp[n];
    double *rn = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    for (int32_t k = 0; k < nnz; ++k) { u[k] = nz[k]; l[k] = 0.0; }
    for (int32_t i = 0; i < n; ++i) {
        double s = 0.0;
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k) s += fabs(nz[k]);
        rn[i] = s;
    }
    for (int32_t i = 0; i < n; ++i) {
        for (int32_t q = rp[i]; q < diag[i]; ++q) {
            const int32_t k = cv[q];
            const double ukk = u[diag[k]];
            if (fabs(ukk) < EPS64 * (rn[k] > 1.0 ? rn[k] : 1.0)) { l[q] = 0.0; u[q] = 0.0; continue; }
            double lik = u[q] / ukk;
            if (fabs(lik) > 1e8) lik = lik * (1e8 / fabs(lik));
            l[q] = lik;
            u[q] = 0.0;
            for (int32_t qk = diag[k] + 1; qk < rp[k + 1]; ++qk) {
                /* (i, cv[qk]) in row i, by bisection (columns sorted) */
                int32_t lo = rp[i], hi = rp[i + 1] - 1, hit = -1;
                while (lo <= hi) {
                    const int32_t mid = (lo + hi) >> 1;
                    if (cv[mid] == cv[qk]) { hit = mid; break; }
                    if (cv[mid] < cv[qk]) lo = mid + 1; else hi = mid - 1;
                }
                if (hit >= 0) u[hit] -= lik * u[qk];
            }
        }
        const double thr = EPS64 * (rn[i] > 1.0 ? rn[i] : 1.0);
        if (fabs(u[diag[i]]) < thr) u[diag[i]] = thr;
    }
    free(rn);
}

/* ILU(0) smoothing (src/smoothers.jl:1356-1433) with sequential substitutions
 * (any order that respects the dependencies gives the same values). */
void o_ilu0_smooth(int32_t n, const int32_t *rp, const int32_t *cv, const double *nz, const int32_t *diag,
                   const double *l, const double *u, double *x, const double *b, double *t, int32_t steps) {
    for (int32_t s = 0; s < steps; ++s) {
        o_residual(n, rp, cv, nz, x, b, t);
        for (int32_t i = 0; i < n; ++i)
            for (int32_t k = rp[i]; k < diag[i]; ++k) t[i] -= l[k] * t[cv[k]];
        for (int32_t i = n - 1; i >= 0; --i) {
            for (int32_t k = diag[i] + 1; k < rp[i + 1]; ++k) t[i] -= u[k] * t[cv[k]];
            const double uii = u[diag[i]];
            t[i] = fabs(uii) > EPS64 ? t[i] / uii : 0.0;
        }
        for (int32_t i = 0; i < n; ++i)
            if (isfinite(t[i])) x[i] += t[i];
    }
}

static void o_smooth(const OHier *h, OLevel *L, double *x, const double *b, int32_t steps, int post) {
    if (h->smoother == 9) {
        o_ilu0_smooth(L->n, L->a_rp, L->a_cv, L->a_nz, L->ilu_diag, L->ilu_l, L->ilu_u, x, b, L->tmp, steps);
        return;
    }
    if (h->smoother == 8) {
        o_cheb_smooth(L->n, L->a_rp, L->a_cv, L->a_nz, x, b, L->dinv, L->cheb_lmin, L->cheb_lmax, steps, L->tmp,
                      L->cheb_d);
        return;
    }
    const int gs = (h->smoother == 1 || h->smoother == 6);
    if (gs) {
        for (int32_t s = 0; s < steps; ++s)
            o_gs_sweep_dir(L->a_rp, L->a_cv, L->a_nz, x, b, L->dinv, L->ncolors, L->c_off, L->c_order,
                           post && h->gs_sym);
        return;
    }
    /* ping-pong with a copy back on odd step counts (src/smoothers.jl:103-124) */
    double *src = x, *dst = L->tmp;
    const double w = h->smoother == 3 ? 1.0 : h->omega; /* SPAI(0): x + m r (src/smoothers.jl:547-558) */
    for (int32_t s = 0; s < steps; ++s) {
        o_jacobi_step(L->n, L->a_rp, L->a_cv, L->a_nz, src, b, L->dinv, w, dst);
        double *t = src;
        src = dst;
        dst = t;
    }
    if (steps % 2) memcpy(x, L->tmp, (size_t)L->n * sizeof(double));
}

/* _vcycle_descend! (src/cycle.jl:63-99). */
static void o_descend(const OHier *h, int32_t l, double *x, const double *b) {
    OLevel *L = &h->lv[l];
    o_smooth(h, L, x, b, h->pre, 0);
    o_residual(L->n, L->a_rp, L->a_cv, L->a_nz, x, b, L->r);
    o_restrict(L->nc, L->t_off, L->t_rows, L->t_pnz, L->p_nz, L->r, L->bc);
    memset(L->xc, 0, (size_t)L->nc * sizeof(double));
    if (l < h->nlevels - 1) {
        const int reps = h->w_cycle ? 2 : 1;
        for (int rep = 0; rep < reps; ++rep) o_descend(h, l + 1, L->xc, L->bc);
    } else {
        memcpy(h->coarse_b, L->bc, (size_t)L->nc * sizeof(double));
        o_lu_solve(h->ncoarse, h->coarse_lu, h->coarse_piv, h->coarse_b, h->coarse_x);
        memcpy(L->xc, h->coarse_x, (size_t)L->nc * sizeof(double));
    }
    o_prolong(L->n, L->p_rp, L->p_cv, L->p_nz, L->xc, x);
    o_smooth(h, L, x, b, h->post, 1);
}

/* amg_cycle! (src/cycle.jl:12-28). */
void o_cycle(const OHier *h, double *x, const double *b) {
    if (h->nlevels == 0) {
        o_lu_solve(h->ncoarse, h->coarse_lu, h->coarse_piv, b, x);
        return;
    }
    o_descend(h, 0, x, b);
}

static double o_dot(int32_t n, const double *a, const double *b) {
    double s = 0.0;
    for (int32_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

static double o_nrm2(int32_t n, const double *a) {
    /* BLAS nrm2: scaled sum of squares, robust to overflow */
    double scale = 0.0, ssq = 1.0;
    for (int32_t i = 0; i < n; ++i) {
        if (a[i] != 0.0) {
            const double ax = fabs(a[i]);
            if (scale < ax) {
                ssq = 1.0 + ssq * (scale / ax) * (scale / ax);
                scale = ax;
            } else {
                ssq += (ax / scale) * (ax / scale);
            }
        }
    }
    return scale * sqrt(ssq);
}

/* amg_solve! (src/cycle.jl:110-160): returns iterations, rel residual out. */
int32_t o_solve(const OHier *h, int32_t n, double *x, const double *b, double tol, int32_t maxiter,
                double *rvec, double *relres) {
    const double bnorm = o_nrm2(n, b);
    *relres = 0.0;
    if (bnorm == 0.0) {
        memset(x, 0, (size_t)n * sizeof(double));
        return 0;
    }
    if (h->nlevels == 0) {
        o_cycle(h, x, b);
        return 1;
    }
    const OLevel *L = &h->lv[0];
    for (int32_t it = 1; it <= maxiter; ++it) {
        o_cycle(h, x, b);
        o_residual(n, L->a_rp, L->a_cv, L->a_nz, x, b, rvec);
        *relres = o_nrm2(n, rvec) / bnorm;
        if (*relres < tol) return it;
    }
    return maxiter;
}

/* Preconditioned CG (Hestenes-Stiefel), z = cycle(0; r) as
 * preconditioner_apply! (src/preconditioner.jl:76-80); stop when the
 * recurrence residual satisfies ||r||/||b|| < tol.  work: 4n doubles. */
int32_t o_pcg(const OHier *h, int32_t n, const int32_t *rp, const int32_t *cv, const double *nz, double *x,
              const double *b, double tol, int32_t maxiter, double *work, double *relres) {
    double *r = work, *z = work + n, *p = work + 2 * (size_t)n, *q = work + 3 * (size_t)n;
    const double bb = o_dot(n, b, b);
    *relres = 0.0;
    if (bb == 0.0) {
        memset(x, 0, (size_t)n * sizeof(double));
        return 0;
    }
    o_residual(n, rp, cv, nz, x, b, r);
    double rr = o_dot(n, r, r);
    *relres = sqrt(rr / bb);
    if (*relres < tol) return 0;
    memset(z, 0, (size_t)n * sizeof(double));
    o_cycle(h, z, r);
    double rz = o_dot(n, r, z);
    memcpy(p, z, (size_t)n * sizeof(double));
    for (int32_t it = 1; it <= maxiter; ++it) {
        o_spmv(n, rp, cv, nz, p, q);
        const double alpha = rz / o_dot(n, p, q);
#pragma omp parallel for schedule(static)
        for (int32_t i = 0; i < n; ++i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * q[i];
        }
        rr = o_dot(n, r, r);
        *relres = sqrt(rr / bb);
        if (*relres < tol) return it;
        memset(z, 0, (size_t)n * sizeof(double));
        o_cycle(h, z, r);
        const double rz_new = o_dot(n, r, z);
        const double beta = rz_new / rz;
        rz = rz_new;
#pragma omp parallel for schedule(static)
        for (int32_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    }
    return maxiter;
}

/* Partial resetup (src/resetup.jl:27-45) on the oracle hierarchy: new fine
 * values already in lv[0].a_nz; recompute dinv, Galerkin level by level, the
 * coarsest operator and its LU.  Returns 0 or a zero-pivot code. */
int32_t o_resetup(OHier *h) {
    const int mode = (h->smoother == 5 || h->smoother == 6) ? 1 : (h->smoother == 3 ? 2 : 0);
    for (int32_t l = 0; l < h->nlevels; ++l) {
        OLevel *L = &h->lv[l];
        o_invdiag(L->n, L->a_rp, L->a_cv, L->a_nz, mode, L->dinv);
        if (h->smoother == 9) o_ilu0_factor(L->n, L->a_rp, L->a_cv, L->a_nz, L->ilu_diag, L->ilu_l, L->ilu_u);
        double *out = (l + 1 < h->nlevels) ? h->lv[l + 1].a_nz : h->ac_nz;
        const int32_t nrow_next = (l + 1 < h->nlevels) ? h->lv[l + 1].n : h->ncoarse;
        const int32_t *rp_next = (l + 1 < h->nlevels) ? h->lv[l + 1].a_rp : h->ac_rp;
        o_galerkin(rp_next[nrow_next], L->r_off, L->r_pi, L->r_ai, L->r_pj, L->p_nz, L->a_nz, out);
    }
    const int32_t nc = h->ncoarse;
    memset(h->coarse_lu, 0, (size_t)nc * nc * sizeof(double));
    for (int32_t i = 0; i < nc; ++i)
        for (int32_t k = h->ac_rp[i]; k < h->ac_rp[i + 1]; ++k) h->coarse_lu[(size_t)h->ac_cv[k] * nc + i] = h->ac_nz[k];
    return o_lu(nc, h->coarse_lu, h->coarse_piv);
}

/* Inverse diagonals of all levels (the smoother build of src/setup.jl:150). */
void o_prepare(OHier *h) {
    const int mode = (h->smoother == 5 || h->smoother == 6) ? 1 : (h->smoother == 3 ? 2 : 0);
    for (int32_t l = 0; l < h->nlevels; ++l) {
        OLevel *L = &h->lv[l];
        o_invdiag(L->n, L->a_rp, L->a_cv, L->a_nz, mode, L->dinv);
        if (h->smoother == 9) o_ilu0_factor(L->n, L->a_rp, L->a_cv, L->a_nz, L->ilu_diag, L->ilu_l, L->ilu_u);
    }
}
