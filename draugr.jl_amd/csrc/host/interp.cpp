// Host-side construction of the prolongation P (setup phase).
//
// Each routine restates the reference routine cited beside it.  P is returned
// as a 0-based CSR (n_fine x n_coarse) with columns ascending in every row,
// the order the reference's COO->CSR conversion produces
// (src/prolongation.jl:713-734).
#include "types.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <map>
#include <numeric>
#include <utility>

namespace draugr {

namespace {

constexpr double kEps = std::numeric_limits<double>::epsilon();

inline double sgn(double v) { return v > 0 ? 1.0 : (v < 0 ? -1.0 : 0.0); }

// Row-wise builder: entries of one row are collected, then sorted by column.
struct RowBuilder {
    CSR P;
    std::vector<std::pair<int32_t, double>> row;
    RowBuilder(int32_t nf, int32_t nc) {
        P.nrow = nf;
        P.ncol = nc;
        P.rp.reserve(nf + 1);
        P.rp.push_back(0);
    }
    void add(int32_t j, double v) { row.emplace_back(j, v); }
    void end_row() {
        std::stable_sort(row.begin(), row.end(),
                         [](const auto& a, const auto& b) { return a.first < b.first; });
        for (auto& e : row) {
            P.cv.push_back(e.first);
            P.nz.push_back(e.second);
        }
        P.rp.push_back((int32_t)P.cv.size());
        row.clear();
    }
};

double diag_first(const CSR& A, int32_t i) {
    for (int32_t k = A.rp[i]; k < A.rp[i + 1]; ++k)
        if (A.cv[k] == i) return A.nz[k];
    return 0.0;
}

// Nearest coarse point by |a_ij| over C neighbours; src/prolongation.jl:695-710.
int32_t nearest_coarse(const CSR& A, int32_t i, const std::vector<int8_t>& cf,
                       const std::vector<int32_t>& cmap) {
    int32_t best = -1;
    double bv = 0.0;
    for (int32_t k = A.rp[i]; k < A.rp[i + 1]; ++k) {
        const int32_t j = A.cv[k];
        if (j == i) continue;
        if (cf[j] == 1 && std::fabs(A.nz[k]) > bv) { bv = std::fabs(A.nz[k]); best = cmap[j]; }
    }
    return best >= 0 ? best : 0;
}

// Direct interpolation; src/prolongation.jl:219-335.
CSR interp_direct(const CSR& A, const std::vector<int8_t>& cf, const std::vector<int32_t>& cmap,
                  int32_t nc, const std::vector<uint8_t>& s) {
    RowBuilder rb(A.nrow, nc);
    std::vector<std::pair<int32_t, double>> strong_c;
    for (int32_t i = 0; i < A.nrow; ++i) {
        if (cf[i] == 1) {
            rb.add(cmap[i], 1.0);
            rb.end_row();
            continue;
        }
        double aii = 0.0;
        for (int32_t k = A.rp[i]; k < A.rp[i + 1]; ++k)
            if (A.cv[k] == i) aii = A.nz[k];
        const double ds = sgn(aii);
        double sum_nonc = 0.0;
        strong_c.clear();
        for (int32_t k = A.rp[i]; k < A.rp[i + 1]; ++k) {
            const int32_t j = A.cv[k];
            if (j == i) continue;
            const bool interp = s[k] && cf[j] == 1 && (ds == 0.0 || sgn(A.nz[k]) != ds);
            if (interp) strong_c.emplace_back(cmap[j], A.nz[k]);
            else sum_nonc += A.nz[k];
        }
        const double d = aii + sum_nonc;
        if (strong_c.empty()) {
            int32_t best = -1;
            double bv = 0.0;
            for (int32_t k = A.rp[i]; k < A.rp[i + 1]; ++k) {
                const int32_t j = A.cv[k];
                if (j == i) continue;
                if (cf[j] == 1 && std::fabs(A.nz[k]) > bv) { bv = std::fabs(A.nz[k]); best = cmap[j]; }
            }
            rb.add(best >= 0 ? best : 0, 1.0);
        } else {
            const double ad = std::fabs(d);
            const double thr = kEps * std::max(1.0, ad);
            for (auto& e : strong_c) rb.add(e.first, ad > thr ? -e.second / d : 0.0);
        }
        rb.end_row();
    }
    return rb.P;
}

// Standard (classical RS) interpolation; src/prolongation.jl:349-456.
CSR interp_standard(const CSR& A, const std::vector<int8_t>& cf, const std::vector<int32_t>& cmap,
                    int32_t nc, const std::vector<uint8_t>& s) {
    RowBuilder rb(A.nrow, nc);
    for (int32_t i = 0; i < A.nrow; ++i) {
        if (cf[i] == 1) {
            rb.add(cmap[i], 1.0);
            rb.end_row();
            continue;
        }
        double aii = 0.0, sum_weak = 0.0;
        std::map<int32_t, double> strong_c;
        std::vector<std::pair<int32_t, double>> strong_f;
        for (int32_t k = A.rp[i]; k < A.rp[i + 1]; ++k) {
            const int32_t j = A.cv[k];
            if (j == i) aii = A.nz[k];
            else if (s[k] && cf[j] == 1) strong_c[cmap[j]] += A.nz[k];
            else if (s[k] && cf[j] == -1) strong_f.emplace_back(j, A.nz[k]);
            else sum_weak += A.nz[k];
        }
        double d = aii + sum_weak;
        std::map<int32_t, double> contrib(strong_c.begin(), strong_c.end());
        for (auto& f : strong_f) {
            const int32_t k = f.first;
            const double aik = f.second;
            const double dk = diag_first(A, k);
            const double sg = dk < 0 ? -1.0 : 1.0;
            double sum_ck = 0.0;
            std::vector<std::pair<int32_t, double>> cvals;  // unique coarse columns in row k
            for (int32_t t = A.rp[k]; t < A.rp[k + 1]; ++t) {
                const int32_t j2 = A.cv[t];
                if (j2 == k) continue;
                const double akj = A.nz[t];
                if (cf[j2] == 1) {
                    const int32_t cm2 = cmap[j2];
                    if (strong_c.count(cm2) && sg * akj < 0) {
                        cvals.emplace_back(cm2, akj);
                        sum_ck += akj;
                    }
                }
                if (j2 == i && sg * akj < 0) sum_ck += akj;
            }
            if (std::fabs(sum_ck) > kEps) {
                const double dist = aik / sum_ck;
                for (auto& e : cvals) contrib[e.first] += dist * e.second;
                for (int32_t t = A.rp[k]; t < A.rp[k + 1]; ++t)
                    if (A.cv[t] == i && sg * A.nz[t] < 0) { d += dist * A.nz[t]; break; }
            } else {
                d += aik;
            }
        }
        if (contrib.empty()) {
            rb.add(nearest_coarse(A, i, cf, cmap), 1.0);
        } else {
            for (auto& e : contrib) rb.add(e.first, std::fabs(d) > kEps ? -e.second / d : 0.0);
        }
        rb.end_row();
    }
    return rb.P;
}

}  // namespace

// Extended+i interpolation of one F row i (src/prolongation.jl:467-690): the
// row's entries (coarse column, weight) in the reference's keep order (the
// caller sorts them by column).  w.pmark / w.fmark are n-long work arrays
// holding -1 where unused; pmark is restored on return.
void ext_i_row(const CSR& A, const std::vector<int8_t>& cf, const std::vector<int32_t>& cmap,
               const std::vector<uint8_t>& s, const InterpSpec& ip, int32_t i, ExtIWork& w,
               std::vector<std::pair<int32_t, double>>& out) {
    std::vector<int32_t>& pmark = w.pmark;
    std::vector<int32_t>& fmark = w.fmark;
    std::vector<int32_t>& chat = w.chat;
    std::vector<double>& pdata = w.pdata;
    std::vector<int32_t>& keep = w.keep;
    out.clear();
    chat.clear();
    // phase 1: C-hat = strong C neighbours and strong C neighbours of strong F neighbours
    for (int32_t k = A.rp[i]; k < A.rp[i + 1]; ++k) {
        const int32_t j = A.cv[k];
        if (!(j != i && s[k])) continue;
        if (cf[j] == 1) {
            if (pmark[j] < 0) { pmark[j] = (int32_t)chat.size(); chat.push_back(j); }
        } else if (cf[j] == -1) {
            fmark[j] = i;
            for (int32_t k2 = A.rp[j]; k2 < A.rp[j + 1]; ++k2) {
                const int32_t m = A.cv[k2];
                if (m != j && s[k2] && cf[m] == 1 && pmark[m] < 0) {
                    pmark[m] = (int32_t)chat.size();
                    chat.push_back(m);
                }
            }
        }
    }
    const int32_t nch = (int32_t)chat.size();
    if (nch == 0) {
        out.emplace_back(nearest_coarse(A, i, cf, cmap), 1.0);
        return;
    }
    // phase 2: weights
    pdata.assign(nch, 0.0);
    double diag = 0.0;
    for (int32_t k = A.rp[i]; k < A.rp[i + 1]; ++k) {
        const int32_t j = A.cv[k];
        const double aij = A.nz[k];
        if (j == i) { diag += aij; continue; }
        if (pmark[j] >= 0) {
            pdata[pmark[j]] += aij;
        } else if (fmark[j] == i) {
            const double dj = diag_first(A, j);
            const double sg = dj < 0 ? -1.0 : 1.0;
            double sum = 0.0;
            for (int32_t t = A.rp[j]; t < A.rp[j + 1]; ++t) {
                const int32_t m = A.cv[t];
                if (m == j) continue;
                if (sg * A.nz[t] < 0 && (pmark[m] >= 0 || m == i)) sum += A.nz[t];
            }
            if (std::fabs(sum) > kEps) {
                const double dist = aij / sum;
                for (int32_t t = A.rp[j]; t < A.rp[j + 1]; ++t) {
                    const int32_t m = A.cv[t];
                    if (m == j) continue;
                    const double ajm = A.nz[t];
                    if (sg * ajm < 0) {
                        if (pmark[m] >= 0) pdata[pmark[m]] += dist * ajm;
                        else if (m == i) diag += dist * ajm;
                    }
                }
            } else {
                diag += aij;
            }
        } else {
            diag += aij;
        }
    }
    // phase 3: normalise
    if (std::fabs(diag) > kEps)
        for (auto& v : pdata) v /= -diag;
    // phase 4: truncation by factor, then by element count
    keep.resize(nch);
    std::iota(keep.begin(), keep.end(), 0);
    auto mag = [&](int32_t idx) { return std::pow(std::fabs(pdata[idx]), (double)ip.norm_p); };
    if (ip.trunc_factor > 0) {
        double mx = 0.0;
        for (int32_t idx = 0; idx < nch; ++idx) mx = std::max(mx, mag(idx));
        const double thr = ip.trunc_factor * mx;
        keep.clear();
        for (int32_t idx = 0; idx < nch; ++idx)
            if (mag(idx) >= thr) keep.push_back(idx);
    }
    if (ip.max_elements > 0 && (int32_t)keep.size() > ip.max_elements) {
        std::stable_sort(keep.begin(), keep.end(), [&](int32_t a, int32_t b) {
            return std::fabs(pdata[a]) > std::fabs(pdata[b]);
        });
        keep.resize(ip.max_elements);
    }
    double scale = 1.0;
    if (ip.rescale && (int32_t)keep.size() < nch) {
        std::vector<uint8_t> kept(nch, 0);
        for (auto idx : keep) kept[idx] = 1;
        double removed = 0.0;
        for (int32_t idx = 0; idx < nch; ++idx)
            if (!kept[idx]) removed += pdata[idx];
        const double den = 1.0 - removed;
        if (std::fabs(den) > 1e-12) scale = 1.0 / den;
    }
    for (auto idx : keep) out.emplace_back(cmap[chat[idx]], pdata[idx] * scale);
    for (auto c : chat) pmark[c] = -1;
}

namespace {

CSR interp_ext_i(const CSR& A, const std::vector<int8_t>& cf, const std::vector<int32_t>& cmap,
                 int32_t nc, const std::vector<uint8_t>& s, const InterpSpec& ip) {
    const int32_t n = A.nrow;
    RowBuilder rb(n, nc);
    ExtIWork w(n);
    std::vector<std::pair<int32_t, double>> out;
    for (int32_t i = 0; i < n; ++i) {
        if (cf[i] == 1) {
            rb.add(cmap[i], 1.0);
            rb.end_row();
            continue;
        }
        ext_i_row(A, cf, cmap, s, ip, i, w, out);
        for (auto& e : out) rb.add(e.first, e.second);
        rb.end_row();
    }
    return rb.P;
}

}  // namespace

// Piecewise-constant P from an aggregation; src/prolongation.jl:8-22.
CSR prolongation_from_aggregates(const std::vector<int32_t>& agg, int32_t nc) {
    CSR P;
    P.nrow = (int32_t)agg.size();
    P.ncol = nc;
    P.rp.resize(P.nrow + 1);
    std::iota(P.rp.begin(), P.rp.end(), 0);
    P.cv = agg;
    P.nz.assign(P.nrow, 1.0);
    return P;
}

// Smoothed aggregation P = (I - omega D^-1 A) P_tent; src/prolongation.jl:38-120.
CSR smooth_prolongation(const CSR& A, const CSR& Pt, double omega) {
    const int32_t n = Pt.nrow;
    std::vector<double> dinv(n);
    for (int32_t i = 0; i < n; ++i) {
        const double d = diag_first(A, i);
        const double ad = std::fabs(d);
        dinv[i] = ad > kEps * std::max(1.0, ad) ? 1.0 / d : 0.0;
    }
    RowBuilder rb(n, Pt.ncol);
    std::map<int32_t, double> row;
    for (int32_t i = 0; i < n; ++i) {
        row.clear();
        for (int32_t p = Pt.rp[i]; p < Pt.rp[i + 1]; ++p) row[Pt.cv[p]] += Pt.nz[p];
        const double factor = -omega * dinv[i];
        for (int32_t k = A.rp[i]; k < A.rp[i + 1]; ++k) {
            const int32_t j = A.cv[k];
            if (j < 0 || j >= n) continue;
            const double w = factor * A.nz[k];
            for (int32_t p = Pt.rp[j]; p < Pt.rp[j + 1]; ++p) row[Pt.cv[p]] += w * Pt.nz[p];
        }
        bool any = false;
        for (auto& e : row)
            if (std::fabs(e.second) > kEps) { rb.add(e.first, e.second); any = true; }
        if (!any && Pt.rp[i] < Pt.rp[i + 1]) rb.add(Pt.cv[Pt.rp[i]], Pt.nz[Pt.rp[i]]);
        rb.end_row();
    }
    return rb.P;
}

// Drop |p_ij| < tol * max_j |p_ij| per row (keep the largest if all drop);
// src/prolongation.jl:134-182.  Used for aggregation filtering and for the
// interpolation truncation of src/setup.jl:348-384.
CSR filter_prolongation(const CSR& P, double tol) {
    RowBuilder rb(P.nrow, P.ncol);
    for (int32_t i = 0; i < P.nrow; ++i) {
        const int32_t s = P.rp[i], e = P.rp[i + 1];
        if (s < e) {
            double mx = 0.0;
            for (int32_t k = s; k < e; ++k) mx = std::max(mx, std::fabs(P.nz[k]));
            const double thr = tol * mx;
            int cnt = 0;
            for (int32_t k = s; k < e; ++k)
                if (std::fabs(P.nz[k]) >= thr) { rb.add(P.cv[k], P.nz[k]); ++cnt; }
            if (cnt == 0) {
                int32_t bk = s;
                double bv = 0.0;
                for (int32_t k = s; k < e; ++k)
                    if (std::fabs(P.nz[k]) > bv) { bv = std::fabs(P.nz[k]); bk = k; }
                rb.add(P.cv[bk], P.nz[bk]);
            }
        }
        rb.end_row();
    }
    return rb.P;
}

// Interpolation from a C/F splitting; dispatch of src/prolongation.jl:196-201.
CSR interpolation_cf(const CSR& A, const std::vector<int8_t>& cf, int32_t nc,
                     const InterpSpec& ip, double theta) {
    std::vector<int32_t> cmap(A.nrow, -1);
    int32_t c = 0;
    for (int32_t i = 0; i < A.nrow; ++i)
        if (cf[i] == 1) cmap[i] = c++;
    const std::vector<uint8_t> s = strength_absolute(A, theta);
    switch (ip.type) {
        case Interp::Direct: return interp_direct(A, cf, cmap, nc, s);
        case Interp::Standard: return interp_standard(A, cf, cmap, nc, s);
        case Interp::ExtendedI: return interp_ext_i(A, cf, cmap, nc, s, ip);
    }
    throw Error("unknown interpolation type");
}

}  // namespace draugr
