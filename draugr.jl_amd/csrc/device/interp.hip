// Strength of connection and extended+i interpolation on the GPU at setup;
// src/strength.jl:24-61 and src/prolongation.jl:467-690.
//
// One thread per fine row, the row's arithmetic in the host's order (the
// host's ext_i_row), so P is bit-identical to the host's:
//   * C-hat (strong C neighbours, then strong C neighbours of each strong F
//     neighbour, first appearance order) is kept in a private list of at most
//     kChatCap entries; position lookups are linear searches of it, which give
//     the same positions as the host's mark array;
//   * weights are accumulated in row order from 0.0 (no contraction);
//   * truncation by factor compares |w| (norm_p = 1, the reference default;
//     other exponents use the host routine for the whole level), truncation
//     by count ranks |w| descending with ties kept in C-hat order, as a stable
//     sort does, and the rescale sums dropped weights in C-hat order;
//   * entries are written in ascending column order.
// A row whose C-hat would exceed kChatCap, or that is (or has a strong F
// neighbour that is) longer than kRowCap entries, is computed by the host's
// ext_i_row instead (flagged by the counting pass): no row is ever truncated
// by the device's capacity, and a thread's work stays bounded.
#include <hip/hip_runtime.h>

#include <cmath>
#include <limits>
#include <string>

#include "dbuf.hpp"
#include "interp.hpp"

namespace draugr {

namespace {

constexpr int kIB = 256;
constexpr int kChatCap = 64;
constexpr int kRowCap = 256;  // longest row (own or strong F neighbour's) a device thread takes

inline unsigned grid_of(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kIB - 1) / kIB); }

// s[k] = (j != i) && |a_ij| >= theta * max_{j != i} |a_ij|  (src/strength.jl:38-61)
__global__ __launch_bounds__(kIB) void k_strength(int32_t n, const int32_t* __restrict__ rp,
                                                  const int32_t* __restrict__ cv, const double* __restrict__ nz,
                                                  double theta, uint8_t* __restrict__ s) {
    const int32_t i = blockIdx.x * kIB + threadIdx.x;
    if (i >= n) return;
    double mx = 0.0;
    for (int32_t k = rp[i]; k < rp[i + 1]; ++k)
        if (cv[k] != i) {
            const double v = fabs(nz[k]);
            mx = (mx < v) ? v : mx;
        }
    const double thr = theta * mx;
    for (int32_t k = rp[i]; k < rp[i + 1]; ++k) s[k] = (cv[k] != i) && (fabs(nz[k]) >= thr);
}

struct ExtIParams {
    double trunc_factor;
    int max_elements;
    int rescale;
};

__device__ inline int find_pos(const int32_t* chat, int nch, int32_t j) {
    for (int p = 0; p < nch; ++p)
        if (chat[p] == j) return p;
    return -1;
}

// cnt[i] = entries of row i of P, or -1 when its C-hat exceeds kChatCap
// (WRITE = false); entries written at prp[i] (WRITE = true, flagged rows skipped)
template <bool WRITE>
__global__ __launch_bounds__(kIB) void k_ext_i(int32_t n, const int32_t* __restrict__ rp,
                                               const int32_t* __restrict__ cv, const double* __restrict__ nz,
                                               const uint8_t* __restrict__ s, const int8_t* __restrict__ cf,
                                               const int32_t* __restrict__ cmap, ExtIParams ip,
                                               int32_t* __restrict__ cnt, const int32_t* __restrict__ prp,
                                               int32_t* __restrict__ pcv, double* __restrict__ pnz) {
    const int32_t i = blockIdx.x * kIB + threadIdx.x;
    if (i >= n) return;
    if (cf[i] == 1) {
        if (WRITE) {
            pcv[prp[i]] = cmap[i];
            pnz[prp[i]] = 1.0;
        } else {
            cnt[i] = 1;
        }
        return;
    }
    if (WRITE && cnt[i] < 0) return;
    constexpr double eps = std::numeric_limits<double>::epsilon();
    int32_t chat[kChatCap];
    double pdata[kChatCap];
    int nch = 0;
    const int32_t r0 = rp[i], r1 = rp[i + 1];
    if (r1 - r0 > kRowCap) {
        if (!WRITE) cnt[i] = -1;
        return;
    }
    // phase 1
    for (int32_t k = r0; k < r1; ++k) {
        const int32_t j = cv[k];
        if (!(j != i && s[k])) continue;
        if (cf[j] == 1) {
            if (find_pos(chat, nch, j) < 0) {
                if (nch == kChatCap) { if (!WRITE) cnt[i] = -1; return; }
                chat[nch++] = j;
            }
        } else if (cf[j] == -1) {
            if (rp[j + 1] - rp[j] > kRowCap) {
                if (!WRITE) cnt[i] = -1;
                return;
            }
            for (int32_t k2 = rp[j]; k2 < rp[j + 1]; ++k2) {
                const int32_t m = cv[k2];
                if (m != j && s[k2] && cf[m] == 1 && find_pos(chat, nch, m) < 0) {
                    if (nch == kChatCap) { if (!WRITE) cnt[i] = -1; return; }
                    chat[nch++] = m;
                }
            }
        }
    }
    if (nch == 0) {  // nearest coarse point by |a_ij| (src/prolongation.jl:695-710)
        int32_t best = -1;
        double bv = 0.0;
        for (int32_t k = r0; k < r1; ++k) {
            const int32_t j = cv[k];
            if (j == i) continue;
            if (cf[j] == 1 && fabs(nz[k]) > bv) { bv = fabs(nz[k]); best = cmap[j]; }
        }
        if (WRITE) {
            pcv[prp[i]] = best >= 0 ? best : 0;
            pnz[prp[i]] = 1.0;
        } else {
            cnt[i] = 1;
        }
        return;
    }
    // phase 2
    for (int p = 0; p < nch; ++p) pdata[p] = 0.0;
    double diag = 0.0;
    for (int32_t k = r0; k < r1; ++k) {
        const int32_t j = cv[k];
        const double aij = nz[k];
        if (j == i) { diag += aij; continue; }
        const int pj = find_pos(chat, nch, j);
        if (pj >= 0) {
            pdata[pj] += aij;
            continue;
        }
        // j is a strong F neighbour when some entry (i, j) of the row is strong and j is F
        bool strong_f = false;
        if (cf[j] == -1) {
            if (s[k]) strong_f = true;
            else
                for (int32_t k2 = r0; k2 < r1; ++k2)
                    if (cv[k2] == j && s[k2]) { strong_f = true; break; }
        }
        if (!strong_f) { diag += aij; continue; }
        const int32_t t0 = rp[j], t1 = rp[j + 1];
        double dj = 0.0;
        for (int32_t t = t0; t < t1; ++t)
            if (cv[t] == j) { dj = nz[t]; break; }
        const double sg = dj < 0 ? -1.0 : 1.0;
        double sum = 0.0;
        for (int32_t t = t0; t < t1; ++t) {
            const int32_t m = cv[t];
            if (m == j) continue;
            if (sg * nz[t] < 0 && (m == i || find_pos(chat, nch, m) >= 0)) sum += nz[t];
        }
        if (fabs(sum) > eps) {
            const double dist = aij / sum;
            for (int32_t t = t0; t < t1; ++t) {
                const int32_t m = cv[t];
                if (m == j) continue;
                const double ajm = nz[t];
                if (sg * ajm < 0) {
                    const int pm = find_pos(chat, nch, m);
                    if (pm >= 0) pdata[pm] += dist * ajm;
                    else if (m == i) diag += dist * ajm;
                }
            }
        } else {
            diag += aij;
        }
    }
    // phase 3
    if (fabs(diag) > eps)
        for (int p = 0; p < nch; ++p) pdata[p] /= -diag;
    // phase 4: kept set as a bit mask over C-hat positions
    uint64_t keep = (nch == 64) ? ~0ull : ((1ull << nch) - 1);
    if (ip.trunc_factor > 0) {
        double mx = 0.0;
        for (int p = 0; p < nch; ++p) {
            const double v = fabs(pdata[p]);
            mx = (mx < v) ? v : mx;
        }
        const double thr = ip.trunc_factor * mx;
        keep = 0;
        for (int p = 0; p < nch; ++p)
            if (fabs(pdata[p]) >= thr) keep |= 1ull << p;
    }
    int nkeep = __popcll(keep);
    if (ip.max_elements > 0 && nkeep > ip.max_elements) {
        uint64_t top = 0;
        for (int p = 0; p < nch; ++p) {
            if (!((keep >> p) & 1)) continue;
            const double v = fabs(pdata[p]);
            int rank = 0;
            for (int q = 0; q < nch; ++q) {
                if (!((keep >> q) & 1) || q == p) continue;
                const double w = fabs(pdata[q]);
                if (w > v || (w == v && q < p)) ++rank;
            }
            if (rank < ip.max_elements) top |= 1ull << p;
        }
        keep = top;
        nkeep = ip.max_elements;
    }
    double scale = 1.0;
    if (ip.rescale && nkeep < nch) {
        double removed = 0.0;
        for (int p = 0; p < nch; ++p)
            if (!((keep >> p) & 1)) removed += pdata[p];
        const double den = 1.0 - removed;
        if (fabs(den) > 1e-12) scale = 1.0 / den;
    }
    if (!WRITE) {
        cnt[i] = nkeep;
        return;
    }
    // ascending coarse column (cmap is increasing, so C-hat order by node id)
    const int32_t base = prp[i];
    for (int p = 0; p < nch; ++p) {
        if (!((keep >> p) & 1)) continue;
        const int32_t c = cmap[chat[p]];
        int rank = 0;
        for (int q = 0; q < nch; ++q)
            if (((keep >> q) & 1) && cmap[chat[q]] < c) ++rank;
        pcv[base + rank] = c;
        pnz[base + rank] = pdata[p] * scale;
    }
}

// Truncation of P's rows relative to the row's largest |p_ij|, keeping the
// largest entry when nothing passes (src/prolongation.jl:134-182, as the
// host's filter_prolongation): count pass (WRITE = false), then the kept
// entries in their order.
template <bool WRITE>
__global__ __launch_bounds__(kIB) void k_filter(int32_t n, double tol, const int32_t* __restrict__ rp,
                                                const int32_t* __restrict__ cv, const double* __restrict__ nz,
                                                int32_t* __restrict__ cnt, const int32_t* __restrict__ orp,
                                                int32_t* __restrict__ ocv, double* __restrict__ onz) {
    const int32_t i = blockIdx.x * kIB + threadIdx.x;
    if (i >= n) return;
    const int32_t s = rp[i], e = rp[i + 1];
    if (s >= e) {
        if (!WRITE) cnt[i] = 0;
        return;
    }
    double mx = 0.0;
    for (int32_t k = s; k < e; ++k) {
        const double v = fabs(nz[k]);
        mx = (mx < v) ? v : mx;
    }
    const double thr = tol * mx;
    int32_t c = 0;
    int32_t o = WRITE ? orp[i] : 0;
    for (int32_t k = s; k < e; ++k)
        if (fabs(nz[k]) >= thr) {
            if (WRITE) {
                ocv[o] = cv[k];
                onz[o] = nz[k];
                ++o;
            }
            ++c;
        }
    if (c == 0) {
        int32_t bk = s;
        double bv = 0.0;
        for (int32_t k = s; k < e; ++k)
            if (fabs(nz[k]) > bv) { bv = fabs(nz[k]); bk = k; }
        if (WRITE) {
            ocv[o] = cv[bk];
            onz[o] = nz[bk];
        }
        c = 1;
    }
    if (!WRITE) cnt[i] = c;
}

template <typename T>
void download(std::vector<T>& dst, const T* src, size_t n, hipStream_t st) {
    dst.resize(n);
    if (n) DRAUGR_HIP_CHECK(hipMemcpyAsync(dst.data(), src, n * sizeof(T), hipMemcpyDeviceToHost, st));
}

}  // namespace

CSR interpolation_device(const CSR& A, const std::vector<int8_t>& cf, int32_t nc, const InterpSpec& ip, double theta,
                         bool truncate, int32_t min_rows) {
    if (ip.type != Interp::ExtendedI || ip.norm_p != 1 || A.nrow < min_rows) {
        CSR P = interpolation_cf(A, cf, nc, ip, theta);
        return truncate ? filter_prolongation(P, ip.trunc_factor) : P;
    }
    const int32_t n = A.nrow;
    const int64_t nnz = A.nnz();
    std::vector<int32_t> cmap(n, -1);
    {
        int32_t c = 0;
        for (int32_t i = 0; i < n; ++i)
            if (cf[i] == 1) cmap[i] = c++;
        if (c != nc) throw Error("interpolation_device: the splitting has " + std::to_string(c) + " C points, not " +
                                 std::to_string(nc));
    }
    hipStream_t st;
    DRAUGR_HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct StreamOwner {
        hipStream_t s;
        ~StreamOwner() { (void)hipStreamDestroy(s); }
    } owner{st};
    DBuf<int32_t> arp, acv, dcmap, cnt, prp, pcv, frp, fcv;
    DBuf<double> anz, pnz, fnz;
    DBuf<uint8_t> s;
    DBuf<int8_t> dcf;
    struct SyncOnExit {  // runs before the buffers above are freed, also when a step throws
        hipStream_t s;
        ~SyncOnExit() { (void)hipStreamSynchronize(s); }
    } sync_on_exit{st};
    arp.upload(A.rp, st);
    acv.upload(A.cv, st);
    anz.upload(A.nz, st);
    dcf.upload(cf, st);
    dcmap.upload(cmap, st);
    s.alloc(std::max<int64_t>(nnz, 1));
    cnt.alloc(std::max(n, 1));
    const ExtIParams prm{ip.trunc_factor, ip.max_elements, ip.rescale ? 1 : 0};
    k_strength<<<grid_of(n), kIB, 0, st>>>(n, arp.p, acv.p, anz.p, theta, s.p);
    k_ext_i<false><<<grid_of(n), kIB, 0, st>>>(n, arp.p, acv.p, anz.p, s.p, dcf.p, dcmap.p, prm, cnt.p, nullptr,
                                                nullptr, nullptr);
    std::vector<int32_t> h_cnt;
    download(h_cnt, cnt.p, (size_t)n, st);
    DRAUGR_HIP_CHECK(hipStreamSynchronize(st));
    DRAUGR_HIP_CHECK(hipGetLastError());

    // rows whose C-hat exceeds the device capacity: the host's routine
    std::vector<int32_t> big;
    for (int32_t i = 0; i < n; ++i)
        if (h_cnt[i] < 0) big.push_back(i);
    std::vector<std::vector<std::pair<int32_t, double>>> big_rows(big.size());
    if (!big.empty()) {
        std::vector<uint8_t> h_s;
        download(h_s, s.p, (size_t)nnz, st);
        DRAUGR_HIP_CHECK(hipStreamSynchronize(st));
        ExtIWork w(n);
        for (size_t b = 0; b < big.size(); ++b) {
            auto& row = big_rows[b];
            ext_i_row(A, cf, cmap, h_s, ip, big[b], w, row);
            std::stable_sort(row.begin(), row.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
            h_cnt[big[b]] = (int32_t)row.size();
        }
    }
    CSR P;
    P.nrow = n;
    P.ncol = nc;
    P.rp.assign((size_t)n + 1, 0);
    for (int32_t i = 0; i < n; ++i) P.rp[i + 1] = P.rp[i] + h_cnt[i];
    const int32_t nnzP = P.rp[n];
    for (int32_t i : big) h_cnt[i] = -1;  // the write pass skips these rows
    cnt.upload(h_cnt, st);
    prp.upload(P.rp, st);
    pcv.alloc(std::max(nnzP, 1));
    pnz.alloc(std::max(nnzP, 1));
    k_ext_i<true><<<grid_of(n), kIB, 0, st>>>(n, arp.p, acv.p, anz.p, s.p, dcf.p, dcmap.p, prm, cnt.p, prp.p, pcv.p,
                                               pnz.p);
    // the host-computed rows into their places
    std::vector<int32_t> big_cv;
    std::vector<double> big_nz;
    for (const auto& row : big_rows)
        for (const auto& e : row) {
            big_cv.push_back(e.first);
            big_nz.push_back(e.second);
        }
    for (size_t b = 0, o = 0; b < big.size(); o += big_rows[b].size(), ++b) {
        const size_t len = big_rows[b].size();
        if (!len) continue;
        DRAUGR_HIP_CHECK(hipMemcpyAsync(pcv.p + P.rp[big[b]], big_cv.data() + o, len * sizeof(int32_t),
                                        hipMemcpyHostToDevice, st));
        DRAUGR_HIP_CHECK(hipMemcpyAsync(pnz.p + P.rp[big[b]], big_nz.data() + o, len * sizeof(double),
                                        hipMemcpyHostToDevice, st));
    }
    if (truncate) {
        k_filter<false><<<grid_of(n), kIB, 0, st>>>(n, ip.trunc_factor, prp.p, pcv.p, pnz.p, cnt.p, nullptr, nullptr,
                                                    nullptr);
        download(h_cnt, cnt.p, (size_t)n, st);
        DRAUGR_HIP_CHECK(hipStreamSynchronize(st));
        for (int32_t i = 0; i < n; ++i) P.rp[i + 1] = P.rp[i] + h_cnt[i];
        const int32_t nnzF = P.rp[n];
        frp.upload(P.rp, st);
        fcv.alloc(std::max(nnzF, 1));
        fnz.alloc(std::max(nnzF, 1));
        k_filter<true><<<grid_of(n), kIB, 0, st>>>(n, ip.trunc_factor, prp.p, pcv.p, pnz.p, nullptr, frp.p, fcv.p,
                                                   fnz.p);
        download(P.cv, fcv.p, (size_t)nnzF, st);
        download(P.nz, fnz.p, (size_t)nnzF, st);
    } else {
        download(P.cv, pcv.p, (size_t)nnzP, st);
        download(P.nz, pnz.p, (size_t)nnzP, st);
    }
    DRAUGR_HIP_CHECK(hipStreamSynchronize(st));
    DRAUGR_HIP_CHECK(hipGetLastError());
    return P;
}

}  // namespace draugr
