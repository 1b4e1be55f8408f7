// Host-side symbolic Galerkin product, transpose map, colouring and the dense
// coarse factorisation (setup phase).
#include "types.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace draugr {

// Transpose structure of P: coarse column J -> fine rows (ascending) and the
// index of P[i, J] in P's value array; src/prolongation.jl:767-797.
void build_transpose_map(const CSR& P, TransposeMap& t) {
    const int32_t nc = P.ncol;
    t.offsets.assign(nc + 1, 0);
    for (int32_t k = 0; k < (int32_t)P.cv.size(); ++k) t.offsets[P.cv[k] + 1]++;
    for (int32_t J = 0; J < nc; ++J) t.offsets[J + 1] += t.offsets[J];
    t.fine_rows.resize(P.cv.size());
    t.p_nz_idx.resize(P.cv.size());
    std::vector<int32_t> pos(t.offsets.begin(), t.offsets.end() - 1);
    for (int32_t i = 0; i < P.nrow; ++i)
        for (int32_t k = P.rp[i]; k < P.rp[i + 1]; ++k) {
            const int32_t J = P.cv[k];
            t.fine_rows[pos[J]] = i;
            t.p_nz_idx[pos[J]] = k;
            pos[J]++;
        }
}

// Coarse operator Ac = P^T A P: pattern, values and (optionally) the triple map
// grouped by destination coarse nonzero; src/galerkin.jl:11-183.
//
// The pattern is the set of every generated (I, J) pair, structural zeros
// included, columns sorted within rows.  The triples of one coarse nonzero keep
// their generation order (fine row i, then P entry of i, then A entry (i, j),
// then P entry of j), which is also the order the values are summed in.  The
// work is organised per coarse row I through the transpose of P, so rows are
// independent and run in parallel.
void coarse_sparsity(const CSR& A, const CSR& P, int32_t nc, CSR& Ac, RestrictionMap* rmap) {
    TransposeMap T;
    build_transpose_map(P, T);
    Ac.nrow = Ac.ncol = nc;
    Ac.rp.assign(nc + 1, 0);

    int nthreads = 1;
#ifdef _OPENMP
    nthreads = omp_get_max_threads();
#endif
    // pass A: distinct columns per coarse row
    {
        std::vector<std::vector<int32_t>> marks(nthreads);
#pragma omp parallel
        {
            int tid = 0;
#ifdef _OPENMP
            tid = omp_get_thread_num();
#endif
            std::vector<int32_t>& mark = marks[tid];
            mark.assign(nc, -1);
#pragma omp for schedule(dynamic, 256)
            for (int32_t I = 0; I < nc; ++I) {
                int32_t cnt = 0;
                for (int32_t t = T.offsets[I]; t < T.offsets[I + 1]; ++t) {
                    const int32_t i = T.fine_rows[t];
                    for (int32_t a = A.rp[i]; a < A.rp[i + 1]; ++a) {
                        const int32_t j = A.cv[a];
                        for (int32_t q = P.rp[j]; q < P.rp[j + 1]; ++q) {
                            const int32_t J = P.cv[q];
                            if (mark[J] != I) { mark[J] = I; ++cnt; }
                        }
                    }
                }
                Ac.rp[I + 1] = cnt;
            }
        }
    }
    for (int32_t I = 0; I < nc; ++I) Ac.rp[I + 1] += Ac.rp[I];
    const int64_t nnzc = Ac.rp[nc];
    Ac.cv.resize(nnzc);
    Ac.nz.assign(nnzc, 0.0);
    std::vector<int64_t> tcount(rmap ? nnzc + 1 : 0, 0);

    // pass B: sorted columns, triple counts per coarse nonzero, values
#pragma omp parallel
    {
        std::vector<int32_t> pos(nc, -1);
#pragma omp for schedule(dynamic, 256)
        for (int32_t I = 0; I < nc; ++I) {
            const int32_t r0 = Ac.rp[I];
            int32_t len = 0;
            for (int32_t t = T.offsets[I]; t < T.offsets[I + 1]; ++t) {
                const int32_t i = T.fine_rows[t];
                for (int32_t a = A.rp[i]; a < A.rp[i + 1]; ++a) {
                    const int32_t j = A.cv[a];
                    for (int32_t q = P.rp[j]; q < P.rp[j + 1]; ++q) {
                        const int32_t J = P.cv[q];
                        if (pos[J] < 0) { pos[J] = 0; Ac.cv[r0 + len++] = J; }
                    }
                }
            }
            std::sort(Ac.cv.begin() + r0, Ac.cv.begin() + r0 + len);
            for (int32_t k = 0; k < len; ++k) pos[Ac.cv[r0 + k]] = r0 + k;
            for (int32_t t = T.offsets[I]; t < T.offsets[I + 1]; ++t) {
                const int32_t i = T.fine_rows[t];
                const double pi = P.nz[T.p_nz_idx[t]];
                for (int32_t a = A.rp[i]; a < A.rp[i + 1]; ++a) {
                    const int32_t j = A.cv[a];
                    const double pa = pi * A.nz[a];
                    for (int32_t q = P.rp[j]; q < P.rp[j + 1]; ++q) {
                        const int32_t k = pos[P.cv[q]];
                        Ac.nz[k] += pa * P.nz[q];
                        if (rmap) tcount[k + 1]++;
                    }
                }
            }
            for (int32_t k = 0; k < len; ++k) pos[Ac.cv[r0 + k]] = -1;
        }
    }
    if (!rmap) return;

    // pass C: triple map in generation order within each coarse nonzero
    for (int64_t k = 0; k < nnzc; ++k) tcount[k + 1] += tcount[k];
    const int64_t ntr = tcount[nnzc];
    rmap->nz_offsets = tcount;
    rmap->pi.resize(ntr);
    rmap->ai.resize(ntr);
    rmap->pj.resize(ntr);
#pragma omp parallel
    {
        std::vector<int32_t> pos(nc, -1);
        std::vector<int64_t> fill;
#pragma omp for schedule(dynamic, 256)
        for (int32_t I = 0; I < nc; ++I) {
            const int32_t r0 = Ac.rp[I], len = Ac.rp[I + 1] - r0;
            fill.assign(rmap->nz_offsets.begin() + r0, rmap->nz_offsets.begin() + r0 + len);
            for (int32_t k = 0; k < len; ++k) pos[Ac.cv[r0 + k]] = k;
            for (int32_t t = T.offsets[I]; t < T.offsets[I + 1]; ++t) {
                const int32_t i = T.fine_rows[t];
                const int32_t pnz_i = T.p_nz_idx[t];
                for (int32_t a = A.rp[i]; a < A.rp[i + 1]; ++a) {
                    const int32_t j = A.cv[a];
                    for (int32_t q = P.rp[j]; q < P.rp[j + 1]; ++q) {
                        const int64_t dst = fill[pos[P.cv[q]]]++;
                        rmap->pi[dst] = pnz_i;
                        rmap->ai[dst] = a;
                        rmap->pj[dst] = q;
                    }
                }
            }
            for (int32_t k = 0; k < len; ++k) pos[Ac.cv[r0 + k]] = -1;
        }
    }
}

// Greedy colouring in natural row order (smallest colour unused by already
// coloured neighbours), then rows grouped by colour; src/smoothers.jl:136-195.
void greedy_coloring(const CSR& A, Coloring& c) {
    const int32_t n = A.nrow;
    c.colors.assign(n, 0);
    c.num_colors = 0;
    std::vector<int32_t> seen;  // seen[color] == i  <=> colour used by a neighbour of i
    for (int32_t i = 0; i < n; ++i) {
        for (int32_t k = A.rp[i]; k < A.rp[i + 1]; ++k) {
            const int32_t j = A.cv[k];
            if (j != i && c.colors[j] > 0) {
                if ((int32_t)seen.size() <= c.colors[j]) seen.resize(c.colors[j] + 1, -1);
                seen[c.colors[j]] = i;
            }
        }
        int32_t col = 1;
        while (col < (int32_t)seen.size() && seen[col] == i) ++col;
        c.colors[i] = col;
        c.num_colors = std::max(c.num_colors, col);
    }
    c.color_offsets.assign(c.num_colors + 1, 0);
    for (int32_t i = 0; i < n; ++i) c.color_offsets[c.colors[i]]++;
    for (int32_t k = 0; k < c.num_colors; ++k) c.color_offsets[k + 1] += c.color_offsets[k];
    c.color_order.resize(n);
    std::vector<int32_t> pos(c.color_offsets.begin(), c.color_offsets.end() - 1);
    for (int32_t i = 0; i < n; ++i) c.color_order[pos[c.colors[i] - 1]++] = i;
}

// Gaussian elimination with partial pivoting (largest |a| in the column, first
// on ties, as LAPACK getrf), then the inverse from the LU factors.
bool dense_lu_inverse(int n, const double* A, double* inv) {
    std::vector<double> LU(A, A + (size_t)n * n);
    std::vector<int> piv(n);
    auto at = [&](int r, int c) -> double& { return LU[(size_t)c * n + r]; };
    for (int k = 0; k < n; ++k) {
        int p = k;
        double mx = std::fabs(at(k, k));
        for (int r = k + 1; r < n; ++r)
            if (std::fabs(at(r, k)) > mx) { mx = std::fabs(at(r, k)); p = r; }
        piv[k] = p;
        if (at(p, k) == 0.0) return false;
        if (p != k)
            for (int c = 0; c < n; ++c) std::swap(at(k, c), at(p, c));
        const double d = at(k, k);
        for (int r = k + 1; r < n; ++r) at(r, k) /= d;
        for (int c = k + 1; c < n; ++c) {
            const double u = at(k, c);
            if (u == 0.0) continue;
            for (int r = k + 1; r < n; ++r) at(r, c) -= at(r, k) * u;
        }
    }
    // columns of the inverse: solve L U x = P e_c
    std::vector<double> x(n);
    for (int c = 0; c < n; ++c) {
        std::fill(x.begin(), x.end(), 0.0);
        x[c] = 1.0;
        for (int k = 0; k < n; ++k) std::swap(x[k], x[piv[k]]);
        for (int r = 0; r < n; ++r)
            for (int k = 0; k < r; ++k) x[r] -= at(r, k) * x[k];
        for (int r = n - 1; r >= 0; --r) {
            for (int k = r + 1; k < n; ++k) x[r] -= at(r, k) * x[k];
            x[r] /= at(r, r);
        }
        std::copy(x.begin(), x.end(), inv + (size_t)c * n);
    }
    return true;
}

}  // namespace draugr
