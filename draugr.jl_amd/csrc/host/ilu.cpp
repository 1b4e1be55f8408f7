// ILU(0) of a level operator for the level-scheduled ILU(0) smoother
// (src/smoothers.jl:1016-1098 factorization and diagonal indices,
// :1228-1307 level scheduling).  Host code: the reference factorizes on the
// CPU as well (build_ilu0_smoother / update_smoother!).
#include <algorithm>
#include <cmath>
#include <limits>
#include <string>

#include "types.hpp"

namespace draugr {

namespace {

double safe_threshold(double scale) { return std::numeric_limits<double>::epsilon() * std::max(1.0, scale); }

// position of column col in [start, stop) of a sorted row, or -1
// (src/csr_matrix.jl:237-251)
int32_t find_in_row(const std::vector<int32_t>& cv, int32_t start, int32_t stop, int32_t col) {
    int32_t lo = start, hi = stop - 1;
    while (lo <= hi) {
        const int32_t mid = (lo + hi) >> 1;
        const int32_t c = cv[mid];
        if (c == col) return mid;
        if (c < col) lo = mid + 1;
        else hi = mid - 1;
    }
    return -1;
}

// rows grouped by level (1-based levels), ascending within a level:
// rows of level c are order[offsets[c-1] .. offsets[c])
void level_order(const std::vector<int32_t>& level, int nlev, std::vector<int32_t>& order,
                 std::vector<int32_t>& offsets) {
    const int32_t n = (int32_t)level.size();
    std::vector<int32_t> cnt(nlev + 1, 0);
    for (int32_t i = 0; i < n; ++i) ++cnt[level[i]];
    offsets.assign(nlev + 1, 0);
    for (int c = 1; c <= nlev; ++c) offsets[c] = offsets[c - 1] + cnt[c];
    std::vector<int32_t> pos(offsets.begin(), offsets.end() - 1);
    order.assign(n, 0);
    for (int32_t i = 0; i < n; ++i) order[pos[level[i] - 1]++] = i;
}

}  // namespace

void ilu0_structure(const CSR& A, IluStructure& s) {
    const int32_t n = A.nrow;
    s.diag_idx.assign(n, -1);
    for (int32_t i = 0; i < n; ++i)
        for (int32_t k = A.rp[i]; k < A.rp[i + 1]; ++k)
            if (A.cv[k] == i) { s.diag_idx[i] = k; break; }  // first match, as the reference
    for (int32_t i = 0; i < n; ++i)
        if (s.diag_idx[i] < 0) throw Error("ILU(0) needs a stored diagonal entry in every row (row " + std::to_string(i) + ")");
    std::vector<int32_t> fl(n), bl(n);
    int nf = 0, nb = 0;
    for (int32_t i = 0; i < n; ++i) {
        int m = 0;
        for (int32_t k = A.rp[i]; k < s.diag_idx[i]; ++k) m = std::max(m, fl[A.cv[k]]);
        fl[i] = m + 1;
        nf = std::max(nf, fl[i]);
    }
    for (int32_t i = n - 1; i >= 0; --i) {
        int m = 0;
        for (int32_t k = s.diag_idx[i] + 1; k < A.rp[i + 1]; ++k) m = std::max(m, bl[A.cv[k]]);
        bl[i] = m + 1;
        nb = std::max(nb, bl[i]);
    }
    level_order(fl, nf, s.fwd_order, s.fwd_offsets);
    level_order(bl, nb, s.bwd_order, s.bwd_offsets);
}

void ilu0_factorize(const CSR& A, const std::vector<int32_t>& diag_idx, std::vector<double>& L,
                    std::vector<double>& U) {
    const int32_t n = A.nrow;
    U = A.nz;
    L.assign(A.nz.size(), 0.0);
    const double max_factor = 1e8;
    std::vector<double> row_norms(n);
    for (int32_t i = 0; i < n; ++i) {
        double s = 0.0;
        for (int32_t k = A.rp[i]; k < A.rp[i + 1]; ++k) s += std::fabs(A.nz[k]);
        row_norms[i] = s;
    }
    for (int32_t i = 0; i < n; ++i) {
        for (int32_t nz = A.rp[i]; nz < diag_idx[i]; ++nz) {
            const int32_t k = A.cv[nz];
            const double u_kk = U[diag_idx[k]];
            if (std::fabs(u_kk) < safe_threshold(row_norms[k])) {
                L[nz] = 0.0;
                U[nz] = 0.0;
                continue;
            }
            double l_ik = U[nz] / u_kk;
            const double ln = std::fabs(l_ik);
            if (ln > max_factor) l_ik = l_ik * (max_factor / ln);
            L[nz] = l_ik;
            U[nz] = 0.0;
            for (int32_t nzk = diag_idx[k] + 1; nzk < A.rp[k + 1]; ++nzk) {
                const int32_t nzij = find_in_row(A.cv, A.rp[i], A.rp[i + 1], A.cv[nzk]);
                if (nzij >= 0) U[nzij] -= l_ik * U[nzk];
            }
        }
        const double thr = safe_threshold(row_norms[i]);
        if (std::fabs(U[diag_idx[i]]) < thr) U[diag_idx[i]] = thr;
    }
}

}  // namespace draugr
