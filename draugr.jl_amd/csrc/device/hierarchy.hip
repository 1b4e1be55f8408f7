// Device-resident hierarchy and the stream-ordered drivers of the hot path:
// V/W cycle (src/cycle.jl:12-99), stationary solve (src/cycle.jl:110-160),
// FP64 PCG around a FP64 / FP32 / BF16-storage cycle, and the partial numeric
// resetup (src/resetup.jl:20-45).  No host synchronisation happens inside a
// cycle; the solvers read one scalar per iteration back to the host.
#include "hierarchy.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <type_traits>
#include <unordered_map>

namespace draugr {

using namespace dev;

namespace {

constexpr int kRedBlocks = 4096;   // grid of the vector reductions
constexpr int kCoarseLdsMax = 60 * 1024;
constexpr int kLuOneGroupMax = 256;  // larger coarsest operators are factored column by column over the chip

template <class F>
__global__ __launch_bounds__(kBlock) void k_map(int64_t n, F f) {
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) f(i);
}

template <class F>
void launch_map(int64_t n, F f, hipStream_t st) {
    if (n <= 0) return;
    int64_t nb = (n + kBlock - 1) / kBlock;
    if (nb > 16384) nb = 16384;
    hipLaunchKernelGGL(k_map<F>, dim3((unsigned)nb), dim3(kBlock), 0, st, n, f);
}

// Block sum in LDS (fixed tree order) -> partial[block] (plain store).  The
// partials are summed by k_finalize, a one-workgroup kernel launched right
// after: the kernel boundary makes them visible without per-block fences (an
// agent-scope release fence per block writes back L2 on every XCD and costs
// far more than the extra launch).  Deterministic for a fixed grid size.
__device__ __forceinline__ void block_sum_store(double v, double* sh, double* partial, int block) {
    sh[threadIdx.x] = v;
    __syncthreads();
#pragma unroll
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[block] = sh[0];
}

template <class FIN>
__global__ __launch_bounds__(kBlock) void k_finalize(const double* __restrict__ partial, int nblocks, FIN fin) {
    __shared__ double sh[kBlock];
    double acc = 0.0;
    // eight partials in flight per thread, added in the same order as one at a time
    int i = threadIdx.x;
    for (; i + 7 * kBlock < nblocks; i += 8 * kBlock) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = partial[i + u * kBlock];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc += v[u];
    }
    for (; i < nblocks; i += kBlock) acc += partial[i];
    sh[threadIdx.x] = acc;
    __syncthreads();
#pragma unroll
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) fin(sh[0]);
}

// Fused vector pass + global sum: every element i contributes f(i).
template <class F>
__global__ __launch_bounds__(kBlock) void k_reduce(int64_t n, F f, double* partial) {
    __shared__ double sh[kBlock];
    double acc = 0.0;
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) acc += f(i);
    block_sum_store(acc, sh, partial, blockIdx.x);
}

// Grid of the vector reductions over n elements (at most kRedBlocks).
inline int red_blocks(int64_t n) {
    const int64_t nb = (n + kBlock - 1) / kBlock;
    return (int)std::max<int64_t>(1, std::min<int64_t>(nb, kRedBlocks));
}

template <class F, class FIN>
void launch_reduce(int64_t n, F f, FIN fin, double* partial, hipStream_t st) {
    const int nb = red_blocks(n);
    hipLaunchKernelGGL((k_reduce<F>), dim3((unsigned)nb), dim3(kBlock), 0, st, n, f, partial);
    hipLaunchKernelGGL((k_finalize<FIN>), dim3(1), dim3(kBlock), 0, st, partial, nb, fin);
}

// The vector pass and its per-block partials only (red_blocks(n) of them),
// for a sum the host finishes.
template <class F>
void launch_reduce_partials(int64_t n, F f, double* partial, hipStream_t st) {
    hipLaunchKernelGGL((k_reduce<F>), dim3((unsigned)red_blocks(n)), dim3(kBlock), 0, st, n, f, partial);
}

// CSR-stream pass whose per-row epilogue also returns a contribution to a
// global sum (q = A p fused with p.q in PCG).
template <int PER, bool C16, bool SRT, typename TV, typename TX, class XF, class EPI>
__global__ __launch_bounds__(kBlock) void k_stream_reduce(const BlkHdr* __restrict__ hdr, const int* __restrict__ rp,
                                                          const int* __restrict__ cv,
                                                          const uint16_t* __restrict__ c16, const TV* __restrict__ av,
                                                          XF xf, EPI epi, double* partial,
                                                          const uint32_t* __restrict__ cs) {
    __shared__ TX prod[PER * kBlock];
    __shared__ double sh[kBlock];
    double dot = 0.0;
    auto pref = [](int) { return NoPre{}; };
    auto initf = [](NoPre) { return (TX)0; };
    auto accf = [](TX a, TX p) { return a + p; };
    auto finf = [&](int row, TX a, NoPre) { dot = epi(row, a); };
    stream_block<PER, C16, TV, TX, XF, decltype(pref), decltype(initf), decltype(accf), decltype(finf), SRT>(
        xcd_block(), hdr, rp, cv, c16, av, xf, prod, pref, initf, accf, finf, cs);
    block_sum_store(dot, sh, partial, blockIdx.x);
}

// The same pass over the pattern form: the same row blocks, rows on the
// same threads and partial per block, so p.q is the same bits.
template <int KW, typename TV, typename TX, class XF, class EPI>
__global__ __launch_bounds__(kBlock) void k_pattern_reduce(const BlkHdr* __restrict__ hdr, int npat, int w,
                                                           const uint16_t* __restrict__ pid,
                                                           const int* __restrict__ tab, const TV* __restrict__ tv,
                                                           XF xf, EPI epi, double* partial) {
    __shared__ int stab[kPatMax * (kPatSlots + 1)];
    __shared__ double sh[kBlock];
    double dot = 0.0;
    auto pref = [](int) { return NoPre{}; };
    auto initf = [](NoPre) { return (TX)0; };
    auto accf = [](TX a, TX p) { return a + p; };
    auto finf = [&](int row, TX a, NoPre) { dot = epi(row, a); };
    pattern_block<KW, TV, TX, XF, decltype(pref), decltype(initf), decltype(accf), decltype(finf)>(
        xcd_block(), hdr, npat, w, pid, tab, tv, xf, stab, pref, initf, accf, finf);
    block_sum_store(dot, sh, partial, blockIdx.x);
}

// Forms of the PCG's q = A p pass (TV = double): the row-order stream kernel
// (16-bit columns where the operator has them), the column-sorted form
// (values from M.sv64), the pattern form (values from M.tv64).
enum PcgForm : int { kPcgStream = 0, kPcgSorted = 1, kPcgPattern = 2 };

template <typename TV, typename TX, class XF, class EPI, class FIN>
void launch_stream_reduce(const DevCSR& M, const TV* av, XF xf, EPI epi, FIN fin, double* partial, hipStream_t st,
                          int form = kPcgStream) {
#define DRAUGR_SR(PER, C, S, COLS, VALS, CS)                                                                   \
    hipLaunchKernelGGL((k_stream_reduce<PER, C, S, TV, TX, XF, EPI>), dim3(M.nblk), dim3(kBlock), 0, st,         \
                       S ? M.shdr.p : M.hdr.p, M.rp.p, M.cv.p, COLS, VALS, xf, epi, partial, CS)
    if (form == kPcgPattern && M.hasT) {
        const TV* tv = reinterpret_cast<const TV*>(M.tv64.p);
        if (M.tw <= 8)
            hipLaunchKernelGGL((k_pattern_reduce<8, TV, TX, XF, EPI>), dim3(M.nblk), dim3(kBlock), 0, st, M.hdr.p,
                               M.npat, M.tw, M.tpid.p, M.ttab.p, tv, xf, epi, partial);
        else
            hipLaunchKernelGGL((k_pattern_reduce<16, TV, TX, XF, EPI>), dim3(M.nblk), dim3(kBlock), 0, st, M.hdr.p,
                               M.npat, M.tw, M.tpid.p, M.ttab.p, tv, xf, epi, partial);
    } else if (form == kPcgSorted && M.hasS) {
        const TV* sv = reinterpret_cast<const TV*>(M.sv64.p);
        switch (M.per) {
            case 1: DRAUGR_SR(1, true, true, nullptr, sv, M.scs.p); break;
            case 2: DRAUGR_SR(2, true, true, nullptr, sv, M.scs.p); break;
            case 4: DRAUGR_SR(4, true, true, nullptr, sv, M.scs.p); break;
            case 8: DRAUGR_SR(8, true, true, nullptr, sv, M.scs.p); break;
            default: DRAUGR_SR(16, true, true, nullptr, sv, M.scs.p); break;
        }
    } else if (M.has16) {
        switch (M.per) {
            case 1: DRAUGR_SR(1, true, false, M.c16.p, av, nullptr); break;
            case 2: DRAUGR_SR(2, true, false, M.c16.p, av, nullptr); break;
            case 4: DRAUGR_SR(4, true, false, M.c16.p, av, nullptr); break;
            case 8: DRAUGR_SR(8, true, false, M.c16.p, av, nullptr); break;
            default: DRAUGR_SR(16, true, false, M.c16.p, av, nullptr); break;
        }
    } else {
        switch (M.per) {
            case 1: DRAUGR_SR(1, false, false, M.c16.p, av, nullptr); break;
            case 2: DRAUGR_SR(2, false, false, M.c16.p, av, nullptr); break;
            case 4: DRAUGR_SR(4, false, false, M.c16.p, av, nullptr); break;
            case 8: DRAUGR_SR(8, false, false, M.c16.p, av, nullptr); break;
            default: DRAUGR_SR(16, false, false, M.c16.p, av, nullptr); break;
        }
    }
#undef DRAUGR_SR
    hipLaunchKernelGGL((k_finalize<FIN>), dim3(1), dim3(kBlock), 0, st, partial, M.nblk, fin);
}

// Inverse diagonal for the smoothers: plain (src/smoothers.jl:47-63) or l1
// (src/smoothers.jl:816-827); one thread per row, row-order sums.
template <int MODE>  // 0 plain, 1 l1, 2 SPAI(0): m_i = a_ii / sum_j a_ij^2 (src/smoothers.jl:517-532)
__global__ __launch_bounds__(kBlock) void k_invdiag(int n, const int* __restrict__ rp, const int* __restrict__ cv,
                                                    const double* __restrict__ av, double* __restrict__ d) {
    const int row = blockIdx.x * kBlock + threadIdx.x;
    if (row >= n) return;
    double nrm = 0.0, dg = 0.0;
    for (int k = rp[row]; k < rp[row + 1]; ++k) {
        const double v = av[k];
        nrm += MODE == 2 ? v * v : fabs(v);
        if (cv[k] == row) dg = v;  // the last diagonal match wins, as in the reference
    }
    const double eps = 2.220446049250313e-16;
    if (MODE == 1) {
        d[row] = nrm > eps ? 1.0 / nrm : 0.0;
    } else if (MODE == 2) {
        d[row] = nrm > 0.0 ? dg / nrm : 0.0;
    } else {
        const double thr = eps * fmax(1.0, nrm);
        d[row] = fabs(dg) > thr ? 1.0 / dg : 0.0;
    }
}

// The l1 inverse diagonal over the operator's row blocks: a block's |a_k|
// staged in LDS with coalesced loads, then each row's sum on its own thread in
// row order (the same sums as k_invdiag<1>); a row longer than the stage
// (alone in its block) is summed by one thread from memory.
__global__ __launch_bounds__(kBlock) void k_invdiag_l1_blocks(const BlkHdr* __restrict__ hdr,
                                                              const int* __restrict__ rp,
                                                              const double* __restrict__ av, double* __restrict__ d) {
    __shared__ double term[kStreamNnz];
    const int b = gridDim.x >= 64 ? xcd_block() : (int)blockIdx.x;
    const BlkHdr h0 = hdr[b];
    const int r0 = h0.r0, r1 = h0.r1, z0 = h0.z0, nz = hdr[b + 1].z0 - z0;
    const int t = (int)threadIdx.x;
    const double eps = 2.220446049250313e-16;
    if (nz > kStreamNnz) {
        if (t == 0) {
            double nrm = 0.0;
            for (int k = 0; k < nz; ++k) nrm += fabs(av[z0 + k]);
            d[r0] = nrm > eps ? 1.0 / nrm : 0.0;
        }
        return;
    }
    const int row = r0 + t;
    const bool own = row < r1;
    const int rs = own ? rp[row] - z0 : 0, re = own ? rp[row + 1] - z0 : 0;
    for (int k = t; k < nz; k += kBlock) term[k] = fabs(av[z0 + k]);
    __syncthreads();
    if (own) {
        double nrm = 0.0;
        for (int k = rs; k < re; ++k) nrm += term[k];
        d[row] = nrm > eps ? 1.0 / nrm : 0.0;
    }
}

// ILU(0) triangular solves by level schedule (src/smoothers.jl:1383-1411):
// rows of one level are independent; each row subtracts in column order.
template <typename TF, typename TX>
__global__ __launch_bounds__(kBlock) void k_ilu_fwd(int cnt, const int* __restrict__ list, const int* __restrict__ rp,
                                                    const int* __restrict__ cv, const int* __restrict__ diag,
                                                    const TF* __restrict__ L, TX* t) {
    const int q = blockIdx.x * kBlock + threadIdx.x;
    if (q >= cnt) return;
    const int i = list[q];
    TX acc = t[i];
    for (int k = rp[i]; k < diag[i]; ++k) acc -= (TX)L[k] * t[cv[k]];
    t[i] = acc;
}
template <typename TF, typename TX>
__global__ __launch_bounds__(kBlock) void k_ilu_bwd(int cnt, const int* __restrict__ list, const int* __restrict__ rp,
                                                    const int* __restrict__ cv, const int* __restrict__ diag,
                                                    const TF* __restrict__ U, TX* t, TX eps) {
    const int q = blockIdx.x * kBlock + threadIdx.x;
    if (q >= cnt) return;
    const int i = list[q];
    TX acc = t[i];
    const int dk = diag[i];
    for (int k = dk + 1; k < rp[i + 1]; ++k) acc -= (TX)U[k] * t[cv[k]];
    const TX u = (TX)U[dk];
    t[i] = fabs(u) > eps ? acc / u : (TX)0;
}

// Galerkin numeric product into the fixed coarse pattern:
// Ac[k] = sum_t P[pi_t] * A[ai_t] * P[pj_t] over the triples of nonzero k in
// generation order (src/galerkin.jl:194-222).  One thread per coarse nonzero.
// A workgroup takes kBlock consecutive coarse nonzeros, whose triples are one
// contiguous range of the map.  The range is processed kGalChunk triples at a
// time: every thread loads kGalPer triples of the chunk (coalesced, indices
// past its end clamped to its last entry and their products dropped), issues
// their three gathers together and stages P[pi] * A[ai] * P[pj] in LDS; then
// each thread adds the products of its own coarse nonzero that fall in the
// chunk, in triple order, to its running sum.  Same products, same order as
// one thread walking its triples, so the values are the same bits; a thread
// walking its triples alone would touch one cache line per coarse nonzero of
// its wavefront on every step of the map loads.
constexpr int kGalPer = 8;
constexpr int kGalChunk = kGalPer * kBlock;
__global__ __launch_bounds__(kBlock) void k_galerkin(int64_t nnzc, const int64_t* __restrict__ off,
                                                     const int* __restrict__ pi, const int* __restrict__ ai,
                                                     const int* __restrict__ pj, const double* __restrict__ P,
                                                     const double* __restrict__ A, double* __restrict__ Ac) {
    __shared__ double prod[kGalChunk];
    const int b = gridDim.x >= 64 ? xcd_block() : (int)blockIdx.x;
    const int t = (int)threadIdx.x;
    const int64_t k0 = (int64_t)b * kBlock;
    const int64_t k = k0 + t;
    const int64_t kend = k0 + kBlock < nnzc ? k0 + kBlock : nnzc;
    const int64_t T0 = off[k0], T1 = off[kend];
    const bool own = k < nnzc;
    const int64_t ts = own ? off[k] : 0, te = own ? off[k + 1] : 0;
    double acc = 0.0;
    for (int64_t c0 = T0; c0 < T1; c0 += kGalChunk) {
        const int cn = (int)(T1 - c0 < kGalChunk ? T1 - c0 : kGalChunk);
        int ip[kGalPer], ia[kGalPer], jp[kGalPer];
#pragma unroll
        for (int u = 0; u < kGalPer; ++u) {
            const int64_t q = c0 + min(t + u * kBlock, cn - 1);
            ip[u] = pi[q];
            ia[u] = ai[q];
            jp[u] = pj[q];
        }
        double vp[kGalPer], va[kGalPer], vq[kGalPer];
#pragma unroll
        for (int u = 0; u < kGalPer; ++u) {
            vp[u] = P[ip[u]];
            va[u] = A[ia[u]];
            vq[u] = P[jp[u]];
        }
#pragma unroll
        for (int u = 0; u < kGalPer; ++u)
            if (t + u * kBlock < cn) prod[t + u * kBlock] = vp[u] * va[u] * vq[u];
        __syncthreads();
        if (own) {
            const int64_t a = ts > c0 ? ts : c0;
            const int64_t e = te < c0 + cn ? te : c0 + cn;
            for (int64_t q = a; q < e; ++q) acc += prod[q - c0];
        }
        __syncthreads();
    }
    if (own) Ac[k] = acc;
}

// Dense LU with partial pivoting of the coarsest operator, one workgroup
// (n_c is small: <= max_coarse_size normally).  Pivot: largest |a| in the
// column, first index on ties (LAPACK getf2); whole rows are swapped.
// flag = 1 if a pivot is exactly zero.
__global__ __launch_bounds__(kBlock) void k_dense_lu(int n, double* __restrict__ a, int* __restrict__ piv,
                                                     int* __restrict__ flag) {
    __shared__ double sv[kBlock];
    __shared__ int si[kBlock];
    auto at = [&](int r, int c) -> double& { return a[(size_t)c * n + r]; };
    for (int k = 0; k < n; ++k) {
        double best = -1.0;
        int bi = n;
        for (int r = k + threadIdx.x; r < n; r += kBlock) {
            const double v = fabs(at(r, k));
            if (v > best) { best = v; bi = r; }
        }
        sv[threadIdx.x] = best;
        si[threadIdx.x] = bi;
        __syncthreads();
        for (int s = kBlock / 2; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) {
                const double ov = sv[threadIdx.x + s];
                const int oi = si[threadIdx.x + s];
                if (ov > sv[threadIdx.x] || (ov == sv[threadIdx.x] && oi < si[threadIdx.x])) {
                    sv[threadIdx.x] = ov;
                    si[threadIdx.x] = oi;
                }
            }
            __syncthreads();
        }
        const int p = si[0];
        __syncthreads();
        if (threadIdx.x == 0) piv[k] = p;
        if (at(p, k) == 0.0) {
            if (threadIdx.x == 0) *flag = 1;
            return;
        }
        if (p != k)
            for (int c = threadIdx.x; c < n; c += kBlock) {
                const double tmp = at(k, c);
                at(k, c) = at(p, c);
                at(p, c) = tmp;
            }
        __syncthreads();
        const double dkk = at(k, k);
        for (int r = k + 1 + threadIdx.x; r < n; r += kBlock) at(r, k) /= dkk;
        __syncthreads();
        const int m = n - k - 1;
        for (int64_t e = threadIdx.x; e < (int64_t)m * m; e += kBlock) {
            const int r = k + 1 + (int)(e % m), c = k + 1 + (int)(e / m);
            at(r, c) -= at(r, k) * at(k, c);
        }
        __syncthreads();
    }
    // the row interchanges as one permutation: (P b)[i] = b[perm[i]]
    if (threadIdx.x == 0) {
        int* perm = piv + n;
        for (int i = 0; i < n; ++i) perm[i] = i;
        for (int k = 0; k < n; ++k) {
            const int t = perm[k];
            perm[k] = perm[piv[k]];
            perm[piv[k]] = t;
        }
    }
}

// The same factorisation for a large coarsest operator, one launch pair per
// column: k_lu_pivot (one workgroup: pivot search, row swap, column scale)
// then k_lu_update (the trailing update spread over the chip).  Every entry
// sees the operations of k_dense_lu in the same order, so the factors are the
// same bits; a zero pivot sets the flag and the remaining launches return.
__global__ __launch_bounds__(kBlock) void k_lu_pivot(int n, int k, double* __restrict__ a, int* __restrict__ piv,
                                                     int* __restrict__ flag) {
    __shared__ double sv[kBlock];
    __shared__ int si[kBlock];
    if (*flag) return;
    auto at = [&](int r, int c) -> double& { return a[(size_t)c * n + r]; };
    double best = -1.0;
    int bi = n;
    for (int r = k + threadIdx.x; r < n; r += kBlock) {
        const double v = fabs(at(r, k));
        if (v > best) { best = v; bi = r; }
    }
    sv[threadIdx.x] = best;
    si[threadIdx.x] = bi;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            const double ov = sv[threadIdx.x + s];
            const int oi = si[threadIdx.x + s];
            if (ov > sv[threadIdx.x] || (ov == sv[threadIdx.x] && oi < si[threadIdx.x])) {
                sv[threadIdx.x] = ov;
                si[threadIdx.x] = oi;
            }
        }
        __syncthreads();
    }
    const int p = si[0];
    if (threadIdx.x == 0) piv[k] = p;
    if (at(p, k) == 0.0) {
        if (threadIdx.x == 0) *flag = 1;
        return;
    }
    if (p != k)
        for (int c = threadIdx.x; c < n; c += kBlock) {
            const double tmp = at(k, c);
            at(k, c) = at(p, c);
            at(p, c) = tmp;
        }
    __syncthreads();
    const double dkk = at(k, k);
    for (int r = k + 1 + threadIdx.x; r < n; r += kBlock) at(r, k) /= dkk;
}

__global__ __launch_bounds__(kBlock) void k_lu_update(int n, int k, double* __restrict__ a,
                                                      const int* __restrict__ flag) {
    if (*flag) return;
    const int64_t m = n - k - 1;
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < m * m; e += stride) {
        const int r = k + 1 + (int)(e % m), c = k + 1 + (int)(e / m);
        a[(size_t)c * n + r] -= a[(size_t)k * n + r] * a[(size_t)c * n + k];
    }
}

// the row interchanges as one permutation: (P b)[i] = b[perm[i]]
__global__ void k_lu_perm(int n, const int* __restrict__ piv, int* __restrict__ perm, const int* __restrict__ flag) {
    if (*flag || threadIdx.x != 0) return;
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int k = 0; k < n; ++k) {
        const int t = perm[k];
        perm[k] = perm[piv[k]];
        perm[piv[k]] = t;
    }
}

// x = A_c^-1 b from the LU factors in the order of the reference LAPACK
// getrs for one right-hand side: row interchanges, then column-oriented unit
// lower and upper triangular solves (the loops of reference BLAS trsm).  FP64
// arithmetic; b and x in the cycle's vector type.  Factors staged in LDS when
// they fit.
// xg non-null: the vector is kept in that global workspace (n doubles)
// instead of LDS, for a coarsest level too large for LDS; same operations.
template <typename TX>
__global__ __launch_bounds__(kBlock) void k_coarse_lu_solve(int n, const double* __restrict__ lu,
                                                            const int* __restrict__ piv, const TX* __restrict__ b,
                                                            TX* __restrict__ x, int stage, double* xg) {
    extern __shared__ double sm[];
    double* xs = xg ? xg : sm;
    const double* L = lu;
    if (stage) {
        double* ls = sm + n;
        for (int e = threadIdx.x; e < n * n; e += kBlock) ls[e] = lu[e];
        L = ls;
    }
    for (int i = threadIdx.x; i < n; i += kBlock) xs[i] = (double)b[i];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 0; k < n; ++k) {
            const int p = piv[k];
            const double t = xs[k];
            xs[k] = xs[p];
            xs[p] = t;
        }
    __syncthreads();
    for (int j = 0; j < n; ++j) {
        const double xj = xs[j];
        for (int i = j + 1 + threadIdx.x; i < n; i += kBlock) xs[i] -= L[(size_t)j * n + i] * xj;
        __syncthreads();
    }
    for (int j = n - 1; j >= 0; --j) {
        if (threadIdx.x == 0) xs[j] /= L[(size_t)j * n + j];
        __syncthreads();
        const double xj = xs[j];
        for (int i = threadIdx.x; i < j; i += kBlock) xs[i] -= L[(size_t)j * n + i] * xj;
        __syncthreads();
    }
    for (int i = threadIdx.x; i < n; i += kBlock) x[i] = (TX)xs[i];
}

// The same solve for n_c <= 64 in one wavefront without barriers: lane i holds
// row i of the factors in registers and x_i (wave_lu_solve); per-element
// operation order is that of k_coarse_lu_solve; the row interchanges are
// applied as one permutation.
template <typename TX>
__global__ __launch_bounds__(64) void k_coarse_lu_solve_wave(int n, const double* __restrict__ lu,
                                                             const int* __restrict__ perm,
                                                             const TX* __restrict__ b, TX* __restrict__ x) {
    const int i = threadIdx.x;
    double row[64];
    wave_lu_row(row, n, i, [&](int j) { return lu[(size_t)j * n + i]; });
    const double xi = wave_lu_solve(i < n ? (double)b[perm[i]] : 0.0, row, n, i);
    if (i < n) x[i] = (TX)xi;
}

template <typename TV> const TV* vals(const DevCSR& m);
template <> const double* vals<double>(const DevCSR& m) { return m.v64.p; }
template <> const float* vals<float>(const DevCSR& m) { return m.v32.p; }
template <> const bf16_t* vals<bf16_t>(const DevCSR& m) { return m.v16.p; }
// the values of the column-sorted form
template <typename TV> const TV* svals(const DevCSR& m);
template <> const double* svals<double>(const DevCSR& m) { return m.sv64.p; }
template <> const float* svals<float>(const DevCSR& m) { return m.sv32.p; }
template <> const bf16_t* svals<bf16_t>(const DevCSR& m) { return m.sv16.p; }
// the values of the pattern form
template <typename TV> const TV* tvals(const DevCSR& m);
template <> const double* tvals<double>(const DevCSR& m) { return m.tv64.p; }
template <> const float* tvals<float>(const DevCSR& m) { return m.tv32.p; }
template <> const bf16_t* tvals<bf16_t>(const DevCSR& m) { return m.tv16.p; }

// Transfer operators use FP32 values when the cycle stores A in BF16.
template <typename TV> using TransferT = TransferOf<TV>;

template <typename TX> TX* r_of(DevLevel& l);
template <> double* r_of<double>(DevLevel& l) { return l.r64.p; }
template <> float* r_of<float>(DevLevel& l) { return l.r32.p; }
template <typename TX> TX* t_of(DevLevel& l);
template <> double* t_of<double>(DevLevel& l) { return l.t64.p; }
template <> float* t_of<float>(DevLevel& l) { return l.t32.p; }
template <typename TX> TX* xc_of(DevLevel& l);
template <> double* xc_of<double>(DevLevel& l) { return l.xc64.p; }
template <> float* xc_of<float>(DevLevel& l) { return l.xc32.p; }
template <typename TX> TX* bc_of(DevLevel& l);
template <> double* bc_of<double>(DevLevel& l) { return l.bc64.p; }
template <> float* bc_of<float>(DevLevel& l) { return l.bc32.p; }
template <typename TX> const TX* d_of(const DevLevel& l);
template <> const double* d_of<double>(const DevLevel& l) { return l.d64.p; }
template <> const float* d_of<float>(const DevLevel& l) { return l.d32.p; }

template <typename TV> constexpr int pidx();
template <> constexpr int pidx<double>() { return 0; }
template <> constexpr int pidx<float>() { return 1; }
template <> constexpr int pidx<bf16_t>() { return 2; }

// ── One-pass sweeps for the small levels ────────────────────────────────────
// On a small level every pass costs about one launch's latency, so two
// passes in one launch halve it, at the price of recomputing values a
// neighbour would otherwise read from memory.  The down sweep is k_down_stream;
// the up sweep is a sparse pass over A whose gathers form the prolonged
// iterate on the fly (XProlong / EpiJacobiP).  Both repeat the separate
// passes' arithmetic exactly (same products, same order), so the results are
// the same bits; the autotuner decides per level whether they pay.
//
// Down: x1 = (w dinv) b for every row (the first Jacobi step from zero) and,
// for every coarse row I, bc_I = sum_i P_iI r_i over its fine rows with
// r_i = b_i - sum_k a_ik x1_k formed on the fly (a fine row is formed once
// per coarse row it feeds).  A block takes consecutive coarse rows with at
// most kBlock pairs (I, i) and kDownChunk products in all: the products of
// every pair's row are formed by all threads into LDS, then one thread per
// pair sums its row in order, then one thread per coarse row sums its pairs.
// When x1c is set, the coarse level's own first step x1c = (w dinv_c) bc is
// written too.
constexpr int kDownChunk = 4096;
constexpr int64_t kDownMaxNnz = 4 << 20;  // larger levels keep the separate passes

template <typename TV, typename TP, typename TX>
__global__ __launch_bounds__(kBlock) void k_down_stream(int nb, int n, const int* __restrict__ dblk,
                                                        const int* __restrict__ arp, const int* __restrict__ acv,
                                                        const TV* __restrict__ av, const int* __restrict__ trp,
                                                        const int* __restrict__ tcv, const TP* __restrict__ tv,
                                                        const TX* __restrict__ b, const TX* __restrict__ d, TX w,
                                                        TX* __restrict__ x1, TX* __restrict__ bc,
                                                        const TX* __restrict__ dc, TX* __restrict__ x1c) {
    __shared__ TX prod[kDownChunk];
    __shared__ int woff[kBlock + 1];
    __shared__ int qrow[kBlock], qbase[kBlock];
    __shared__ TX rq[kBlock];
    const int blk = blockIdx.x;
    const int t = (int)threadIdx.x;
    for (int i = blk * kBlock + t; i < n; i += nb * kBlock) x1[i] = (w * d[i]) * b[i];
    const int I0 = dblk[blk], I1 = dblk[blk + 1];
    const int q0 = trp[I0], nq = trp[I1] - q0;
    int len = 0;
    if (t < nq) {
        const int i = tcv[q0 + t];
        const int s = arp[i];
        qrow[t] = i;
        qbase[t] = s;
        len = arp[i + 1] - s;
    }
    woff[t + 1] = len;
    if (t == 0) woff[0] = 0;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {  // inclusive scan of the row lengths
        const int v = t + 1 - off >= 1 ? woff[t + 1 - off] : 0;
        __syncthreads();
        woff[t + 1] += v;
        __syncthreads();
    }
    const int W = nq > 0 ? woff[nq] : 0;
    for (int p = t; p < W; p += kBlock) {
        int lo = 0, hi = nq - 1;  // the pair whose products hold p: last q with woff[q] <= p
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (woff[mid] <= p) lo = mid;
            else hi = mid - 1;
        }
        const int k = qbase[lo] + (p - woff[lo]);
        const int j = acv[k];
        prod[p] = ld<TX, TV>(av, k) * ((w * d[j]) * b[j]);
    }
    __syncthreads();
    if (t < nq) {
        TX s = 0;
        for (int p = woff[t]; p < woff[t + 1]; ++p) s = s + prod[p];
        const TX ri = b[qrow[t]] - s;
        rq[t] = ld<TX, TP>(tv, q0 + t) * ri;
    }
    __syncthreads();
    const int I = I0 + t;
    if (I < I1) {
        TX acc = 0;
        for (int q = trp[I] - q0; q < trp[I + 1] - q0; ++q) acc = acc + rq[q];
        bc[I] = acc;
        if (x1c) x1c[I] = (w * dc[I]) * acc;
    }
}

template <typename TV, typename TX, class XF, class EPI>
void op_kind(int kind, const DevCSR& M, XF xf, EPI epi, hipStream_t st) {
    if (kind == kOpRowSeq) launch_rowseq<TV, TX>(M.nrow, M.rp.p, M.cv.p, vals<TV>(M), xf, epi, st);
    else if (kind == kOpPipe) launch_pipe<TV, TX>(M.nblk, M.per, M.blk.p, M.bz.p, M.rp.p, M.cv.p, vals<TV>(M), xf, epi, st);
    else if (kind == kOpStream16 && M.has16)
        launch_stream<TV, TX>(M.nblk, M.per, M.hdr.p, M.rp.p, M.cv.p, vals<TV>(M), xf, epi, st, M.c16.p);
    else if (kind == kOpSorted && M.hasS && svals<TV>(M))
        launch_stream<TV, TX>(M.nblk, M.per, M.shdr.p, M.rp.p, M.cv.p, svals<TV>(M), xf, epi, st, nullptr, M.scs.p);
    else if (kind == kOpStream16Pipe && M.has16)
        launch_stream_pipe<TV, TX>(M.nblk, M.per, M.hdr.p, M.rp.p, M.cv.p, M.c16.p, vals<TV>(M), xf, epi, st, nullptr);
    else if (kind == kOpSortedPipe && M.hasS && svals<TV>(M))
        launch_stream_pipe<TV, TX>(M.nblk, M.per, M.shdr.p, M.rp.p, M.cv.p, nullptr, svals<TV>(M), xf, epi, st,
                                   M.scs.p);
    else if (kind == kOpSortedQuad && M.hasS && svals<TV>(M))
        launch_stream_quad<TV, TX>(M.nblk, M.per, M.shdr.p, M.rp.p, M.cv.p, M.scs.p, svals<TV>(M), xf, epi, st);
    else if (kind == kOpPattern && M.hasT && tvals<TV>(M))
        launch_pattern<TV, TX>(M.nblk, M.hdr.p, M.npat, M.tw, M.tpid.p, M.ttab.p, tvals<TV>(M), xf, epi, st);
    else if (kind == kOpVector && sizeof(TX) == 4) {
        if constexpr (sizeof(TX) == 4)  // FP32 / BF16 only: not the reference's sequential sum order
            launch_vector<TV, TX>(M.nrow, M.nnz, M.rp.p, M.cv.p, vals<TV>(M), xf, epi, st);
    }
    else launch_stream<TV, TX>(M.nblk, M.per, M.hdr.p, M.rp.p, M.cv.p, vals<TV>(M), xf, epi, st);
}

// One pass of operator M with the kernel the autotuner picked for this
// value type (CSR-stream until tuned).
template <typename TV, typename TX, class XF, class EPI>
void op(const DevCSR& M, XF xf, EPI epi, hipStream_t st) {
    op_kind<TV, TX>(M.kind[pidx<TV>()], M, xf, epi, st);
}

// Column-sorted form: a bitonic sort in LDS of each block's keys (column
// offset << 12 | position in the block), unique per entry, giving cs = column
// offset | position << kSortColBits and the values in that order; blocks
// without a sorted base keep row order (cs = position << kSortColBits, unread:
// those blocks take their columns from the 32-bit array).  A sorted block
// holds at most its budget (<= kStreamNnz = 2^12 entries) and spans fewer than
// 2^kSortColBits columns, so a key fits 32 bits.
static_assert(kSortColBits + 12 == 32, "column offset and position share one word");
__global__ __launch_bounds__(kBlock) void k_sort_blocks(const BlkHdr* __restrict__ shdr, const int* __restrict__ cv,
                                                        const double* __restrict__ v, uint32_t* __restrict__ cs,
                                                        double* __restrict__ sv) {
    __shared__ uint32_t key[kStreamNnz];
    const int b = blockIdx.x;
    const int z0 = shdr[b].z0, nz = shdr[b + 1].z0 - z0, cb = shdr[b].cb;
    const int t = (int)threadIdx.x;
    if (cb < 0 || nz > kStreamNnz) {
        for (int k = t; k < nz; k += kBlock) {
            cs[z0 + k] = (uint32_t)(k & 0xFFF) << kSortColBits;
            sv[z0 + k] = v[z0 + k];
        }
        return;
    }
    int m = 1;
    while (m < nz) m <<= 1;
    for (int k = t; k < m; k += kBlock)
        key[k] = k < nz ? ((uint32_t)(cv[z0 + k] - cb) << 12) | (uint32_t)k : 0xFFFFFFFFu;
    __syncthreads();
    for (int size = 2; size <= m; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = t; i < m; i += kBlock) {
                const int j = i ^ stride;
                if (j > i) {
                    const uint32_t a = key[i], c = key[j];
                    if ((a > c) == ((i & size) == 0)) {
                        key[i] = c;
                        key[j] = a;
                    }
                }
            }
            __syncthreads();
        }
    for (int k = t; k < nz; k += kBlock) {
        const uint32_t kk = key[k];
        const int pos = (int)(kk & 0xFFFu);
        cs[z0 + k] = (kk >> 12) | ((uint32_t)pos << kSortColBits);
        sv[z0 + k] = v[z0 + pos];
    }
}

// sv = the values in the sorted form's order (after new values in v), and
// its FP32 / BF16 copies when they exist (null otherwise).  A sorted block
// (<= kStreamNnz entries) first copies its values into LDS with coalesced
// loads, then writes each entry from its position there; blocks without a
// sorted base (row order, possibly one row longer than 65536 entries) copy.
__global__ __launch_bounds__(kBlock) void k_permute_blocks(const BlkHdr* __restrict__ shdr,
                                                           const uint32_t* __restrict__ cs,
                                                           const double* __restrict__ v, double* __restrict__ sv,
                                                           float* __restrict__ sv32, bf16_t* __restrict__ sv16) {
    __shared__ double sh[kStreamNnz];
    const int b = blockIdx.x;
    const int z0 = shdr[b].z0, nz = shdr[b + 1].z0 - z0;
    const int t = (int)threadIdx.x;
    auto put = [&](int k, double x) {
        sv[z0 + k] = x;
        if (sv32) sv32[z0 + k] = (float)x;
        if (sv16) sv16[z0 + k] = float_to_bf16((float)x);
    };
    if (shdr[b].cb < 0) {
        for (int k = t; k < nz; k += kBlock) put(k, v[z0 + k]);
        return;
    }
    for (int k = t; k < nz; k += kBlock) sh[k] = v[z0 + k];
    __syncthreads();
    for (int k = t; k < nz; k += kBlock) put(k, sh[cs[z0 + k] >> kSortColBits]);
}

// Entries allocated for an array of the column-sorted form: a multiple of 4,
// so that the four-entry loads of the last block stay inside it.
inline size_t sorted_alloc(int64_t nnz) { return (size_t)((nnz + 3) & ~(int64_t)3); }

// The pattern form's values from the FP64 values v (row-order CSR): slot
// (row, j) holds entry rp[row] + j, zero past the row's end, with the FP32
// copy in the same pass; then the BF16 copy (paired layout) when it exists.
void fill_pattern(const DevCSR& d, hipStream_t st) {
    const int n = d.nrow, w = d.tw;
    const int* rp = d.rp.p;
    const double* v = d.v64.p;
    double* t64 = d.tv64.p;
    float* t32 = d.tv32.p;
    bf16_t* t16 = d.tv16.p;
    auto value = [=] __device__(int row, int j) -> double {
        if (row >= n) return 0.0;
        const int k = rp[row] + j;
        return k < rp[row + 1] ? v[k] : 0.0;
    };
    launch_map(pattern_slots(n, w), [=] __device__(int64_t s) {
        const int lane = (int)(s & (kPatSlice - 1));
        const int64_t q = s / kPatSlice;
        const double x = value((int)(q / w) * kPatSlice + lane, (int)(q % w));
        t64[s] = x;
        if (t32) t32[s] = (float)x;
    }, st);
    if (t16) {
        const int wh = (w + 1) >> 1;
        launch_map(pattern_slots2(n, w), [=] __device__(int64_t e) {
            const int64_t pair = e >> 1;
            const int lane = (int)(pair & (kPatSlice - 1));
            const int64_t q = pair / kPatSlice;
            t16[e] = float_to_bf16((float)value((int)(q / wh) * kPatSlice + lane, 2 * (int)(q % wh) + (int)(e & 1)));
        }, st);
    }
}

// The pattern form of h, when its rows use at most kPatMax distinct column
// patterns relative to the row (col - row, in CSR order), none longer than
// kPatSlots, and the 64-row slices pad the values by at most 1/8.  Rows are
// matched against the previous row's pattern first (a stencil's interior
// rows repeat it), then by a hash of the offsets.
void build_pattern(DevCSR& d, const CSR& h, hipStream_t st) {
    d.hasT = false;
    d.tw = d.npat = 0;
    const int n = h.nrow;
    int w = 0;
    for (int i = 0; i < n; ++i) w = std::max(w, h.rp[i + 1] - h.rp[i]);
    const bool eligible = n > 0 && w > 0 && w <= kPatSlots && pattern_slots(n, w) * 8 <= (int64_t)h.nnz() * 9;
    std::vector<int32_t> tab;
    std::vector<uint16_t> pid(eligible ? n : 0);
    auto same = [&](int p, int i) {
        const int32_t* e = tab.data() + (size_t)p * (w + 1);
        const int z0 = h.rp[i], len = h.rp[i + 1] - z0;
        if (e[0] != len) return false;
        for (int j = 0; j < len; ++j)
            if (e[1 + j] != h.cv[z0 + j] - i) return false;
        return true;
    };
    std::unordered_map<uint64_t, std::vector<int>> seen;
    bool ok = eligible;
    int npat = 0, prev = -1;
    for (int i = 0; ok && i < n; ++i) {
        if (prev >= 0 && same(prev, i)) {
            pid[i] = (uint16_t)prev;
            continue;
        }
        const int z0 = h.rp[i], len = h.rp[i + 1] - z0;
        uint64_t key = 1469598103934665603ull ^ (uint64_t)len;
        for (int j = 0; j < len; ++j) key = (key ^ (uint32_t)(h.cv[z0 + j] - i)) * 1099511628211ull;
        std::vector<int>& cand = seen[key];
        int found = -1;
        for (int p : cand)
            if (same(p, i)) { found = p; break; }
        if (found < 0) {
            if (npat == kPatMax) { ok = false; break; }
            found = npat++;
            tab.push_back(len);
            for (int j = 0; j < w; ++j) tab.push_back(j < len ? h.cv[z0 + j] - i : 0);
            cand.push_back(found);
        }
        pid[i] = (uint16_t)found;
        prev = found;
    }
    if (!ok) {
        d.tpid.release();
        d.ttab.release();
        d.tv64.release();
        d.tv32.release();
        d.tv16.release();
        return;
    }
    d.hasT = true;
    d.tw = w;
    d.npat = npat;
    d.tpid.upload(pid, st);
    d.ttab.upload(tab, st);
    d.tv64.alloc(pattern_slots(n, w));
    fill_pattern(d, st);
    DRAUGR_HIP_CHECK(hipStreamSynchronize(st));  // pid, tab are pageable and local
}

void upload_csr(DevCSR& d, const CSR& h, hipStream_t st) {
    d.nrow = h.nrow;
    d.ncol = h.ncol;
    d.nnz = h.nnz();
    d.rp.upload(h.rp, st);
    d.cv.upload(h.cv, st);
    d.v64.upload(h.nz, st);
    const int budget = stream_budget(h.nnz(), h.nrow);
    const std::vector<int32_t> blk = build_row_blocks(h.rp, budget);
    std::vector<int32_t> bz(blk.size());
    for (size_t b = 0; b < blk.size(); ++b) bz[b] = h.rp[blk[b]];
    d.nblk = (int)blk.size() - 1;
    d.per = budget / kBlock;
    d.blk.upload(blk, st);
    d.bz.upload(bz, st);
    // 16-bit column offsets from a per-block base for every block that the
    // stream kernel loads branch-free and that spans fewer than 65536 columns;
    // the others keep cbase = -1 and are read through the 32-bit columns
    std::vector<int32_t> cbase(d.nblk + 1, -1);
    std::vector<uint16_t> c16;
    int narrow = 0;
    for (int b = 0; b < d.nblk; ++b) {
        const int z0 = bz[b], z1 = bz[b + 1];
        if (z1 - z0 > budget || z1 == z0) continue;
        const auto mm = std::minmax_element(h.cv.begin() + z0, h.cv.begin() + z1);
        if ((int64_t)*mm.second - *mm.first > 65535) continue;
        cbase[b] = *mm.first;
        ++narrow;
    }
    const bool fits = narrow > 0;
    d.has16 = fits;
    d.nnz16 = 0;
    for (int b = 0; b < d.nblk; ++b)
        if (cbase[b] >= 0) d.nnz16 += bz[b + 1] - bz[b];
    if (fits) {
        c16.assign(h.nnz(), 0);
        for (int b = 0; b < d.nblk; ++b) {
            if (cbase[b] < 0) continue;
            for (int k = bz[b]; k < bz[b + 1]; ++k) c16[k] = (uint16_t)(h.cv[k] - cbase[b]);
        }
        d.c16.upload(c16, st);
    } else {
        d.c16.release();
    }
    std::vector<BlkHdr> hdr(d.nblk + 1);
    for (int b = 0; b <= d.nblk; ++b)
        hdr[b] = BlkHdr{blk[b], bz[b], fits ? cbase[b] : -1, blk[std::min(b + 1, d.nblk)]};
    d.hdr.upload(hdr, st);
    // the column-sorted form: every branch-free block spanning fewer than
    // 2^kSortColBits columns, with its own headers
    std::vector<BlkHdr> shdr(d.nblk + 1);
    d.nnzS = 0;
    for (int b = 0; b <= d.nblk; ++b) {
        int32_t base = -1;
        if (b < d.nblk) {
            const int z0 = bz[b], z1 = bz[b + 1];
            if (z1 > z0 && z1 - z0 <= budget) {
                const auto mm = std::minmax_element(h.cv.begin() + z0, h.cv.begin() + z1);
                if ((int64_t)*mm.second - *mm.first < ((int64_t)1 << kSortColBits)) {
                    base = *mm.first;
                    d.nnzS += z1 - z0;
                }
            }
        }
        shdr[b] = BlkHdr{blk[b], bz[b], base, blk[std::min(b + 1, d.nblk)]};
    }
    d.hasS = d.nnzS > 0;
    if (d.hasS) {
        d.shdr.upload(shdr, st);
        d.scs.alloc(sorted_alloc(d.nnz));
        d.sv64.alloc(sorted_alloc(d.nnz));
        hipLaunchKernelGGL(k_sort_blocks, dim3(d.nblk), dim3(kBlock), 0, st, d.shdr.p, d.cv.p, d.v64.p, d.scs.p,
                           d.sv64.p);
    } else {
        d.shdr.release();
        d.scs.release();
        d.sv64.release();
    }
    DRAUGR_HIP_CHECK(hipStreamSynchronize(st));  // blk, c16, hdr are pageable and local
    build_pattern(d, h, st);
}

bool is_l1(Smoother s) { return s == Smoother::L1Jacobi || s == Smoother::L1ColoredGS; }
bool is_gs(Smoother s) { return s == Smoother::ColoredGS || s == Smoother::L1ColoredGS; }

}  // namespace

// ─────────────────────────────────────────────────────────────────────────────
DeviceHierarchy::DeviceHierarchy(const HostHierarchy& h, const Config& cfg)
    : forced_tail_(cfg.tail), smoother_(cfg.smoother), omega_(cfg.jacobi_omega), seed_(cfg.seed),
      forced_kind_(cfg.kernel), forced_fuse_(cfg.fuse) {
    if (!(smoother_ == Smoother::Jacobi || smoother_ == Smoother::L1Jacobi || is_gs(smoother_) ||
          smoother_ == Smoother::SPAI0 || smoother_ == Smoother::Chebyshev || smoother_ == Smoother::ILU0))
        throw Error("smoother not available in the HIP backend (supported: jacobi, l1_jacobi, colored_gs, "
                    "l1_colored_gs, spai0, chebyshev, ilu0)");
    DRAUGR_HIP_CHECK(hipGetDevice(&device_));
    try {
        build(h);
    } catch (...) {
        release_handles();  // the destructor does not run for a constructor that throws
        throw;
    }
}

void DeviceHierarchy::build(const HostHierarchy& h) {
    DRAUGR_HIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    DRAUGR_HIP_CHECK(hipEventCreateWithFlags(&last_ev_, hipEventDisableTiming));
    DRAUGR_HIP_CHECK(hipEventCreateWithFlags(&pcg_ev_, hipEventDisableTiming));
    hipStream_t st = stream_;
    n0_ = h.levels.empty() ? h.A_coarsest.nrow : h.levels[0].A.nrow;
    L_.resize(h.levels.size());
    int max_blocks = kRedBlocks;
    for (size_t l = 0; l < h.levels.size(); ++l) {
        const HostLevel& H = h.levels[l];
        DevLevel& D = L_[l];
        upload_csr(D.A, H.A, st);
        upload_csr(D.P, H.P, st);
        CSR Pt;  // P^T as CSR with the values of P gathered through the transpose map
        Pt.nrow = H.P.ncol;
        Pt.ncol = H.P.nrow;
        Pt.rp = H.Pt.offsets;
        Pt.cv = H.Pt.fine_rows;
        Pt.nz.resize(H.Pt.p_nz_idx.size());
        for (size_t k = 0; k < Pt.nz.size(); ++k) Pt.nz[k] = H.P.nz[H.Pt.p_nz_idx[k]];
        upload_csr(D.Pt, Pt, st);
        D.pt_pnz.upload(H.Pt.p_nz_idx, st);
        D.has_rmap = H.has_rmap;
        if (H.has_rmap) D.rmap = device_rmap(H.R, st);
        if (is_gs(smoother_)) {
            D.ncolors = H.coloring.num_colors;
            D.color_off = H.coloring.color_offsets;
            D.color_order.upload(H.coloring.color_order, st);
        }
        const int32_t n = H.A.nrow, nc = H.P.ncol;
        if (smoother_ == Smoother::ILU0) {
            ilu0_structure(H.A, D.ilu);
            D.h_rp = H.A.rp;
            D.h_cv = H.A.cv;
            D.ilu_diag.upload(D.ilu.diag_idx, st);
            D.ilu_fwd.upload(D.ilu.fwd_order, st);
            D.ilu_bwd.upload(D.ilu.bwd_order, st);
            ilu_factor((int)l, H.A.nz, st);
        }
        if (smoother_ == Smoother::Chebyshev) {
            // start vector of the power iteration: N(0,1) from the setup seed, normalised
            std::vector<double> v(n);
            Rng rng(seed_ + 0x632BE59BD9B4E019ull * (uint64_t)(l + 1));
            for (int32_t i = 0; i < n; i += 2) {
                const double u1 = rng.uniform(), u2 = rng.uniform();
                const double rad = std::sqrt(-2.0 * std::log1p(-u1));
                v[i] = rad * std::cos(6.283185307179586 * u2);
                if (i + 1 < n) v[i + 1] = rad * std::sin(6.283185307179586 * u2);
            }
            double nrm = 0.0;
            for (double a : v) nrm += a * a;
            nrm = std::sqrt(nrm);
            for (double& a : v) a /= nrm;
            D.pv0.upload(v, st);
            DRAUGR_HIP_CHECK(hipStreamSynchronize(st));
            D.pv.alloc(n);
            D.pw.alloc(n);
            D.cd64.alloc(n);
        }
        D.d64.alloc(n);
        D.r64.alloc(n);
        D.t64.alloc(n);
        D.xc64.alloc(nc);
        D.bc64.alloc(nc);
        max_blocks = std::max(max_blocks, D.A.nblk);
    }
    upload_csr(Ac_, h.A_coarsest, st);
    max_blocks = std::max(max_blocks, Ac_.nblk);
    nc_ = h.n_coarse;
    clu_.alloc((size_t)nc_ * nc_);
    if ((size_t)nc_ * sizeof(double) > (size_t)kCoarseLdsMax) cx_.alloc(nc_);  // coarse solve vector
    flag_.alloc(1 + 2 * (size_t)nc_);  // [0] singular flag, [1..n] pivots, [n+1..2n] permutation
    const int64_t n0 = n0_;
    w_r_.alloc(n0);
    w_p0_.alloc(n0);
    w_p1_.alloc(n0);
    w_q_.alloc(n0);
    w_z_.alloc(n0);
    w_x_.alloc(n0);
    scal_.alloc(S_COUNT_);
    partial_.alloc(max_blocks);
    DRAUGR_HIP_CHECK(hipMemsetAsync(scal_.p, 0, S_COUNT_ * sizeof(double), st));
    DRAUGR_HIP_CHECK(hipHostMalloc((void**)&pinned_, (kPinPart + kRedBlocks) * sizeof(double),
                                   hipHostMallocMapped | hipHostMallocCoherent));
    DRAUGR_HIP_CHECK(hipHostGetDevicePointer((void**)&pinned_dev_, pinned_, 0));
    for (int i = 0; i < kPinPart + kRedBlocks; ++i) pinned_[i] = 0.0;
    for (int l = 0; l < nlevels(); ++l) compute_dinv(l, st);
    if (smoother_ == Smoother::Chebyshev) chebyshev_bounds(st);
    coarse_factor(st);
    DRAUGR_HIP_CHECK(hipStreamSynchronize(st));
    int fl = 0;
    DRAUGR_HIP_CHECK(hipMemcpy(&fl, flag_.p, sizeof(int), hipMemcpyDeviceToHost));
    if (fl) throw Error("coarsest-level matrix is singular (dense LU found a zero pivot)");
}

void DeviceHierarchy::begin(hipStream_t st) {
    if (last_valid_ && last_st_ != st) DRAUGR_HIP_CHECK(hipStreamWaitEvent(st, last_ev_, 0));
}

void DeviceHierarchy::end(hipStream_t st) {
    DRAUGR_HIP_CHECK(hipEventRecord(last_ev_, st));
    last_st_ = st;
    last_valid_ = true;
}

DeviceHierarchy::~DeviceHierarchy() { release_handles(); }

void DeviceHierarchy::release_handles() {
    if (last_valid_ && last_ev_) (void)hipEventSynchronize(last_ev_);
    for (auto& kv : graphs_) (void)hipGraphExecDestroy(kv.second);
    graphs_.clear();
    if (cap_) (void)hipStreamDestroy(cap_);
    if (last_ev_) (void)hipEventDestroy(last_ev_);
    if (pcg_ev_) (void)hipEventDestroy(pcg_ev_);
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (pinned_) (void)hipHostFree(pinned_);
    if (stream_) (void)hipStreamDestroy(stream_);
    cap_ = stream_ = nullptr;
    last_ev_ = pcg_ev_ = nullptr;
    pinned_ = nullptr;
    last_valid_ = false;
}

// The sum k_finalize forms from nblocks partials (pinned_ + kPinPart), in its
// order: lane t adds partials t, t + 256, ... one at a time, then the lanes'
// sums are added in the same tree, so the value is the same bits.
double DeviceHierarchy::host_finalize(int nblocks) const {
    const volatile double* p = pinned_ + kPinPart;
    double acc[kBlock];
    for (int t = 0; t < kBlock; ++t) {
        double a = 0.0;
        for (int i = t; i < nblocks; i += kBlock) a += p[i];
        acc[t] = a;
    }
    for (int s = kBlock / 2; s > 0; s >>= 1)
        for (int t = 0; t < s; ++t) acc[t] += acc[t + s];
    return acc[0];
}

int64_t DeviceHierarchy::bytes_resident() const {
    int64_t b = 0;
    auto csr = [&](const DevCSR& m) {
        b += m.rp.bytes() + m.cv.bytes() + m.v64.bytes() + m.v32.bytes() + m.v16.bytes() + m.blk.bytes();
        b += m.c16.bytes() + m.hdr.bytes() + m.shdr.bytes() + m.scs.bytes() + m.sv64.bytes() + m.sv32.bytes() +
             m.sv16.bytes();
        b += m.tpid.bytes() + m.ttab.bytes() + m.tv64.bytes() + m.tv32.bytes() + m.tv16.bytes();
    };
    for (auto& l : L_) {
        csr(l.A); csr(l.P); csr(l.Pt);
        b += l.pt_pnz.bytes();
        if (l.rmap) b += l.rmap->off.bytes() + l.rmap->pi.bytes() + l.rmap->ai.bytes() + l.rmap->pj.bytes();
        b += l.d64.bytes() + l.d32.bytes() + l.color_order.bytes();
        b += l.r64.bytes() + l.t64.bytes() + l.xc64.bytes() + l.bc64.bytes();
        b += l.r32.bytes() + l.t32.bytes() + l.xc32.bytes() + l.bc32.bytes();
    }
    csr(Ac_);
    b += clu_.bytes();
    b += w_r_.bytes() + w_p0_.bytes() + w_p1_.bytes() + w_q_.bytes() + w_z_.bytes() + w_x_.bytes();
    b += w_r32_.bytes() + w_z32_.bytes();
    return b;
}

void DeviceHierarchy::compute_dinv(int l, hipStream_t st) {
    DevLevel& D = L_[l];
    const int n = D.A.nrow;
    if (n == 0) return;
    if (is_l1(smoother_) && D.A.nblk > 0)
        hipLaunchKernelGGL(k_invdiag_l1_blocks, dim3(D.A.nblk), dim3(kBlock), 0, st, D.A.hdr.p, D.A.rp.p, D.A.v64.p,
                           D.d64.p);
    else
        invdiag_device(n, D.A.rp.p, D.A.cv.p, D.A.v64.p, D.d64.p,
                       is_l1(smoother_) ? 1 : (smoother_ == Smoother::SPAI0 ? 2 : 0), st);
    if (D.d32.p) {
        const double* s = D.d64.p;
        float* d = D.d32.p;
        launch_map(n, [=] __device__(int64_t i) { d[i] = (float)s[i]; }, st);
    }
}

// Spectral radius of D^-1 A by 10 power-iteration steps from the level's
// fixed random unit vector (src/smoothers.jl:866-885): w = (A v) .* dinv,
// lambda = ||w||, v = w / lambda while lambda > eps.  FP64, on the device; the
// norm is the deterministic two-kernel sum of squares.
double DeviceHierarchy::estimate_rho(int l, hipStream_t st) {
    DevLevel& D = L_[l];
    const int n = D.A.nrow;
    if (n == 0) return 0.0;
    double* v = D.pv.p;
    double* w = D.pw.p;
    double* scal = scal_.p;
    DRAUGR_HIP_CHECK(hipMemcpyAsync(v, D.pv0.p, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, st));
    for (int it = 0; it < 10; ++it) {
        op<double, double>(D.A, XVec<double>{v}, EpiStoreScaled<double>{w, D.d64.p}, st);
        launch_reduce(n, [=] __device__(int64_t i) { return w[i] * w[i]; },
                      [=] __device__(double s) { scal[S_RHO] = sqrt(s); }, partial_.p, st);
        launch_map(n, [=] __device__(int64_t i) {
            const double lam = scal[S_RHO];
            if (lam > 2.220446049250313e-16) v[i] = w[i] / lam;
        }, st);
    }
    return reduce_to_host(st, S_RHO);
}

// Chebyshev bounds per level: lambda_max = 1.1 rho, lambda_min = lambda_max / 30
// (src/smoothers.jl:897-903, 910-920).
void DeviceHierarchy::chebyshev_bounds(hipStream_t st) {
    for (int l = 0; l < nlevels(); ++l) {
        const double rho = estimate_rho(l, st);
        L_[l].cheb_lmax = 1.1 * rho;
        L_[l].cheb_lmin = L_[l].cheb_lmax / 30.0;
    }
}

void DeviceHierarchy::galerkin(int l, hipStream_t st) {
    DevLevel& D = L_[l];
    if (!D.has_rmap || !D.rmap) throw Error("partial resetup needs the restriction maps (setup with allow_partial_resetup=1)");
    DevCSR& out = (l + 1 < nlevels()) ? L_[l + 1].A : Ac_;
    const int64_t nnzc = out.nnz;
    if (nnzc == 0) return;
    const int64_t nb = (nnzc + kBlock - 1) / kBlock;
    const DeviceRMap& R = *D.rmap;
    hipLaunchKernelGGL(k_galerkin, dim3((unsigned)nb), dim3(kBlock), 0, st, nnzc, R.off.p, R.pi.p, R.ai.p, R.pj.p,
                       D.P.v64.p, D.A.v64.p, out.v64.p);
}

// Coarsest operator: dense from its CSR values, then LU (src/setup.jl:177-209,
// src/resetup.jl:123-158 and :44), all on the device.
void dense_lu_device(int n, double* a, int* piv, int* flag, hipStream_t st, bool stepwise) {
    if (n <= 0) return;
    if (!stepwise) {
        hipLaunchKernelGGL(k_dense_lu, dim3(1), dim3(kBlock), 0, st, n, a, piv, flag);
        return;
    }
    for (int k = 0; k < n; ++k) {
        hipLaunchKernelGGL(k_lu_pivot, dim3(1), dim3(kBlock), 0, st, n, k, a, piv, flag);
        const int64_t m = n - k - 1;
        if (m > 0) {
            const int64_t nb = std::min<int64_t>((m * m + kBlock - 1) / kBlock, 4096);
            hipLaunchKernelGGL(k_lu_update, dim3((unsigned)nb), dim3(kBlock), 0, st, n, k, a, flag);
        }
    }
    hipLaunchKernelGGL(k_lu_perm, dim3(1), dim3(1), 0, st, n, piv, piv + n, flag);
}

void invdiag_device(int n, const int* rp, const int* cv, const double* av, double* d, int mode, hipStream_t st) {
    if (n <= 0) return;
    const int nb = (n + kBlock - 1) / kBlock;
    if (mode == 1) hipLaunchKernelGGL(k_invdiag<1>, dim3(nb), dim3(kBlock), 0, st, n, rp, cv, av, d);
    else if (mode == 2) hipLaunchKernelGGL(k_invdiag<2>, dim3(nb), dim3(kBlock), 0, st, n, rp, cv, av, d);
    else hipLaunchKernelGGL(k_invdiag<0>, dim3(nb), dim3(kBlock), 0, st, n, rp, cv, av, d);
}

void DeviceHierarchy::coarse_factor(hipStream_t st, bool stepwise) {
    const int n = nc_;
    if (n == 0) return;
    double* dn = clu_.p;
    launch_map((int64_t)n * n, [=] __device__(int64_t i) { dn[i] = 0.0; }, st);
    const int* rp = Ac_.rp.p;
    const int* cv = Ac_.cv.p;
    const double* av = Ac_.v64.p;
    launch_map(n, [=] __device__(int64_t i) {
        for (int k = rp[i]; k < rp[i + 1]; ++k) dn[(size_t)cv[k] * n + i] = av[k];
    }, st);
    int* fl = flag_.p;
    launch_map(1, [=] __device__(int64_t) { fl[0] = 0; }, st);
    dense_lu_device(n, clu_.p, flag_.p + 1, flag_.p, st, stepwise || n > kLuOneGroupMax);
}

// the reduced-precision and column-sorted copies of a matrix's values
// The reduced-precision and column-sorted copies of M's values after new
// FP64 values: one pass over the values for the row-order copies, one for the
// sorted form and its copies.
static void refresh_mirrors(DevCSR& M, hipStream_t st) {
    const double* s = M.v64.p;
    if (M.v32.p || M.v16.p) {
        float* d32 = M.v32.p;
        bf16_t* d16 = M.v16.p;
        launch_map(M.nnz, [=] __device__(int64_t i) {
            const double x = s[i];
            if (d32) d32[i] = (float)x;
            if (d16) d16[i] = float_to_bf16((float)x);
        }, st);
    }
    if (M.hasS && M.nblk > 0)
        hipLaunchKernelGGL(k_permute_blocks, dim3(M.nblk), dim3(kBlock), 0, st, M.shdr.p, M.scs.p, M.v64.p, M.sv64.p,
                           M.sv32.p, M.sv16.p);
    if (M.hasT) fill_pattern(M, st);
}

void DeviceHierarchy::refresh_level_mirrors(int l, hipStream_t st) { refresh_mirrors(L_[l].A, st); }

void DeviceHierarchy::ensure_precision(Precision p) {
    if (p == Precision::FP64) return;
    if (have32_ && (p != Precision::BF16 || have16_)) return;
    // one-time creation of the mirrors: order against all work in flight
    DRAUGR_HIP_CHECK(hipDeviceSynchronize());
    hipStream_t st = stream_;
    if (!have32_) {
        for (auto& D : L_) {
            D.A.v32.alloc(D.A.nnz);
            D.P.v32.alloc(D.P.nnz);
            D.Pt.v32.alloc(D.Pt.nnz);
            D.d32.alloc(D.A.nrow);
            D.r32.alloc(D.A.nrow);
            D.t32.alloc(D.A.nrow);
            D.xc32.alloc(D.P.ncol);
            D.bc32.alloc(D.P.ncol);
            if (smoother_ == Smoother::Chebyshev) D.cd32.alloc(D.A.nrow);
            if (smoother_ == Smoother::ILU0) {
                D.ilu_L32.alloc(D.A.nnz);
                D.ilu_U32.alloc(D.A.nnz);
                for (auto pr : {std::make_pair(D.ilu_L64.p, D.ilu_L32.p), std::make_pair(D.ilu_U64.p, D.ilu_U32.p)}) {
                    const double* s = pr.first;
                    float* d = pr.second;
                    launch_map(D.A.nnz, [=] __device__(int64_t i) { d[i] = (float)s[i]; }, st);
                }
            }
            for (DevCSR* m : {&D.A, &D.P, &D.Pt}) {
                const double* s = m->v64.p;
                float* d = m->v32.p;
                launch_map(m->nnz, [=] __device__(int64_t i) { d[i] = (float)s[i]; }, st);
                if (m->hasS) {
                    m->sv32.alloc(sorted_alloc(m->nnz));
                    const double* ss = m->sv64.p;
                    float* ds = m->sv32.p;
                    launch_map(m->nnz, [=] __device__(int64_t i) { ds[i] = (float)ss[i]; }, st);
                }
                if (m->hasT) {
                    m->tv32.alloc(pattern_slots(m->nrow, m->tw));
                    fill_pattern(*m, st);
                }
            }
            const double* s = D.d64.p;
            float* d = D.d32.p;
            launch_map(D.A.nrow, [=] __device__(int64_t i) { d[i] = (float)s[i]; }, st);
        }
        w_r32_.alloc(n0_);
        w_z32_.alloc(n0_);
        have32_ = true;
    }
    if (p == Precision::BF16 && !have16_) {
        for (auto& D : L_) {
            D.A.v16.alloc(D.A.nnz);
            const double* s = D.A.v64.p;
            bf16_t* d = D.A.v16.p;
            launch_map(D.A.nnz, [=] __device__(int64_t i) { d[i] = float_to_bf16((float)s[i]); }, st);
            if (D.A.hasS) {
                D.A.sv16.alloc(sorted_alloc(D.A.nnz));
                const double* ss = D.A.sv64.p;
                bf16_t* ds = D.A.sv16.p;
                launch_map(D.A.nnz, [=] __device__(int64_t i) { ds[i] = float_to_bf16((float)ss[i]); }, st);
            }
            if (D.A.hasT) {
                D.A.tv16.alloc(pattern_slots2(D.A.nrow, D.A.tw));
                fill_pattern(D.A, st);  // also rewrites the FP64 / FP32 copies with the same values
            }
        }
        have16_ = true;
    }
    DRAUGR_HIP_CHECK(hipStreamSynchronize(st));
}

// ── smoothing (src/smoothers.jl:78-124, 203-241, 296-313, 835-854) ──────────
// Jacobi steps ping-pong between x and alt (no copy back: the caller tracks
// which buffer holds the iterate).  A first step from a zero guess is
// x1 = w d b, fused with the residual that follows it when fuse_resid.
// Blocks of the one-pass down sweep of level l (built once, from the row
// pointers): consecutive coarse rows with at most kBlock rows, kBlock pairs
// and kDownChunk products.  False when a coarse row alone exceeds a block, or
// the level is too large for the one-pass form to pay.
bool DeviceHierarchy::down_blocks(int l) {
    DevLevel& D = L_[l];
    if (D.down_state != 0) return D.down_state > 0;
    D.down_state = -1;
    if (D.A.nnz > kDownMaxNnz) return false;
    const int n = D.A.nrow, nc = D.P.ncol;
    std::vector<int32_t> arp(n + 1), trp(nc + 1), tcv(D.Pt.nnz);
    DRAUGR_HIP_CHECK(hipMemcpy(arp.data(), D.A.rp.p, arp.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    DRAUGR_HIP_CHECK(hipMemcpy(trp.data(), D.Pt.rp.p, trp.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (!tcv.empty())
        DRAUGR_HIP_CHECK(hipMemcpy(tcv.data(), D.Pt.cv.p, tcv.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    std::vector<int32_t> blk{0};
    int rows = 0, pairs = 0;
    int64_t work = 0;
    for (int I = 0; I < nc; ++I) {
        const int np = trp[I + 1] - trp[I];
        int64_t w = 0;
        for (int q = trp[I]; q < trp[I + 1]; ++q) w += arp[tcv[q] + 1] - arp[tcv[q]];
        if (np > kBlock || w > kDownChunk) return false;
        if (rows + 1 > kBlock || pairs + np > kBlock || work + w > kDownChunk) {
            blk.push_back(I);
            rows = pairs = 0;
            work = 0;
        }
        ++rows;
        pairs += np;
        work += w;
    }
    if (nc > 0) blk.push_back(nc);
    D.ndblk = (int)blk.size() - 1;
    if (D.ndblk < 1) return false;
    D.dblk.upload(blk, stream_);
    DRAUGR_HIP_CHECK(hipStreamSynchronize(stream_));  // blk is pageable and local
    D.down_state = 1;
    return true;
}

template <typename TV, typename TX>
void DeviceHierarchy::launch_down(int l, const TX* b, TX* x1, TX* bc, const TX* dc, TX* x1c, hipStream_t st) {
    typedef typename TransferT<TV>::type TP;
    DevLevel& D = L_[l];
    hipLaunchKernelGGL((k_down_stream<TV, TP, TX>), dim3(D.ndblk), dim3(kBlock), 0, st, D.ndblk, D.A.nrow, D.dblk.p,
                       D.A.rp.p, D.A.cv.p, vals<TV>(D.A), D.Pt.rp.p, D.Pt.cv.p, vals<TP>(D.Pt), b, d_of<TX>(D),
                       (TX)weight(), x1, bc, dc, x1c);
}

// Whether level l's first pre-smoothing step from a zero guess is the plain
// map x1 = (w dinv) b (a Jacobi-type smoother not fused with the residual):
// then the parent's restriction can write x1 along with b.
template <typename TV>
bool DeviceHierarchy::zero_step_is_map(int l, int pre) const {
    if (l < 0 || l >= nlevels() || pre < 1) return false;
    if (l == tail_lt_[pidx<TV>()]) return false;  // the tail forms its own first step
    if (!(smoother_ == Smoother::Jacobi || smoother_ == Smoother::L1Jacobi || smoother_ == Smoother::SPAI0))
        return false;
    return !(pre == 1 && (L_[l].fuse_zero[pidx<TV>()] || L_[l].fuse_down[pidx<TV>()]));
}

// x1_ready: alt already holds the first step from a zero guess (written by
// the parent's restriction).
template <typename TV, typename TX>
void DeviceHierarchy::smooth(int l, TX*& x, TX*& alt, const TX* b, int steps, bool& zero, bool fuse_resid,
                             bool& resid_done, hipStream_t st, bool x1_ready, bool reverse) {
    DevLevel& D = L_[l];
    const int n = D.A.nrow;
    const TX* d = d_of<TX>(D);
    const TX w = (TX)weight();
    if (smoother_ == Smoother::Chebyshev) {
        chebyshev<TV, TX>(l, x, alt, b, steps, zero, st);
        return;
    }
    if (smoother_ == Smoother::ILU0) {
        ilu_smooth<TV, TX>(l, x, alt, b, steps, zero, st);
        return;
    }
    if (is_gs(smoother_)) {
        if (zero) {
            TX* xx = x;
            launch_map(n, [=] __device__(int64_t i) { xx[i] = 0; }, st);
            zero = false;
        }
        for (int s = 0; s < steps; ++s)
            for (int cc = 0; cc < D.ncolors; ++cc) {
                const int c = reverse ? D.ncolors - 1 - cc : cc;  // reverse: the transposed sweep
                const int off = D.color_off[c], cnt = D.color_off[c + 1] - off;
                if (cnt == 0) continue;
                hipLaunchKernelGGL((k_gs_color<TV, TX>), dim3((cnt + kBlock - 1) / kBlock), dim3(kBlock), 0, st, cnt,
                                   D.color_order.p + off, D.A.rp.p, D.A.cv.p, vals<TV>(D.A), x, b, d);
            }
        return;
    }
    for (int s = 0; s < steps; ++s) {
        if (zero) {
            if (fuse_resid && steps == 1 && D.fuse_zero[pidx<TV>()]) {
                op<TV, TX>(D.A, XZeroJacobi<TX>{d, b, w}, EpiZeroJacobiResid<TX>{b, d, alt, r_of<TX>(D), w}, st);
                resid_done = true;
            } else if (!x1_ready) {
                TX* o = alt;
                launch_map(n, [=] __device__(int64_t i) { o[i] = (w * d[i]) * b[i]; }, st);
            }
            zero = false;
        } else {
            op<TV, TX>(D.A, XVec<TX>{x}, EpiJacobi<TX>{x, b, d, alt, w}, st);
        }
        std::swap(x, alt);
    }
    if (zero) {
        TX* xx = x;
        launch_map(n, [=] __device__(int64_t i) { xx[i] = 0; }, st);
        zero = false;
    }
}

template <typename TX>
void DeviceHierarchy::coarse_solve(TX* x, const TX* b, hipStream_t st, bool global_x) {
    const int n = nc_;
    if (n == 0) return;
    if (n <= 64 && !global_x) {
        hipLaunchKernelGGL(k_coarse_lu_solve_wave<TX>, dim3(1), dim3(64), 0, st, n, clu_.p, flag_.p + 1 + n, b, x);
        return;
    }
    const size_t full = (size_t)(n + (size_t)n * n) * sizeof(double);
    const int stage = full <= (size_t)kCoarseLdsMax ? 1 : 0;
    if (stage || (!global_x && (size_t)n * sizeof(double) <= (size_t)kCoarseLdsMax)) {
        const size_t lds = stage ? full : (size_t)n * sizeof(double);
        hipLaunchKernelGGL(k_coarse_lu_solve<TX>, dim3(1), dim3(kBlock), lds, st, n, clu_.p, flag_.p + 1, b, x,
                           stage, (double*)nullptr);
        return;
    }
    // the vector does not fit in LDS (or the test hook asks for this form): a
    // global workspace, allocated with the hierarchy when the level needs it
    if (!cx_.p) cx_.alloc(n);
    hipLaunchKernelGGL(k_coarse_lu_solve<TX>, dim3(1), dim3(kBlock), 0, st, n, clu_.p, flag_.p + 1, b, x, 0, cx_.p);
}

// ── the LDS-resident last level ─────────────────────────────────────────────
// It qualifies with a Jacobi-type smoother, at least two AMG levels (its
// parent launches it), a coarsest level of at most 64 rows and its image plus
// vectors within the workgroup's LDS.
template <typename TV, typename TX>
bool DeviceHierarchy::tail_fits(TailLayout* lay) const {
    typedef typename TransferT<TV>::type TP;
    const int L = nlevels();
    if (L < 2 || nc_ < 1 || nc_ > 64) return false;
    if (!(smoother_ == Smoother::Jacobi || smoother_ == Smoother::L1Jacobi || smoother_ == Smoother::SPAI0))
        return false;
    const DevLevel& D = L_[L - 1];
    if (D.A.nrow >= 65536 || D.A.nnz > (1 << 20) || D.P.nnz > (1 << 20)) return false;
    const TailLayout y = tail_layout(D.A.nrow, nc_, (int)D.A.nnz, (int)D.P.nnz, (int)sizeof(TV), (int)sizeof(TP),
                                     (int)sizeof(TX));
    if (y.lds_bytes > kTailLds) return false;
    if (lay) *lay = y;
    return true;
}

// Pack the last level's image for precision TV: A, P, P^T (16-bit columns),
// the smoother's inverse diagonal, the coarsest LU factors and permutation,
// copied on the device from the hierarchy's own arrays (again after every
// resetup, since values, diagonal and factors change).
template <typename TV, typename TX>
void DeviceHierarchy::pack_tail(hipStream_t st) {
    typedef typename TransferT<TV>::type TP;
    const int pi = pidx<TV>();
    TailLayout y;
    if (!tail_fits<TV, TX>(&y)) return;
    const DevLevel& D = L_[nlevels() - 1];
    tail_img_[pi].alloc(y.img_bytes);
    tail_lay_[pi] = y;
    uint8_t* im = tail_img_[pi].p;
    auto copy_i32 = [&](int off, const int* s, int64_t cnt) {
        int* o = reinterpret_cast<int*>(im + off);
        launch_map(cnt, [=] __device__(int64_t i) { o[i] = s[i]; }, st);
    };
    auto copy_c16 = [&](int off, const int* s, int64_t cnt) {
        uint16_t* o = reinterpret_cast<uint16_t*>(im + off);
        launch_map(cnt, [=] __device__(int64_t i) { o[i] = (uint16_t)s[i]; }, st);
    };
    auto copy_v = [&](int off, auto* s, int64_t cnt) {
        typedef typename std::remove_const<typename std::remove_pointer<decltype(s)>::type>::type T;
        T* o = reinterpret_cast<T*>(im + off);
        launch_map(cnt, [=] __device__(int64_t i) { o[i] = s[i]; }, st);
    };
    copy_i32(y.a_rp, D.A.rp.p, y.n + 1);
    copy_c16(y.a_cv, D.A.cv.p, y.nnz_a);
    copy_v(y.a_v, vals<TV>(D.A), y.nnz_a);
    copy_i32(y.p_rp, D.P.rp.p, y.n + 1);
    copy_c16(y.p_cv, D.P.cv.p, y.nnz_p);
    copy_v(y.p_v, vals<TP>(D.P), y.nnz_p);
    copy_i32(y.t_rp, D.Pt.rp.p, y.nc + 1);
    copy_c16(y.t_cv, D.Pt.cv.p, y.nnz_p);
    copy_v(y.t_v, vals<TP>(D.Pt), y.nnz_p);
    copy_v(y.d, d_of<TX>(D), y.n);
    copy_v(y.lu, (const double*)clu_.p, (int64_t)y.nc * y.nc);
    copy_i32(y.perm, flag_.p + 1 + nc_, y.nc);
}

// after a resetup, on the stream that produced the new values
void DeviceHierarchy::refresh_tails(hipStream_t st) {
    if (tail_img_[0].p) pack_tail<double, double>(st);
    if (tail_img_[1].p) pack_tail<float, float>(st);
    if (tail_img_[2].p) pack_tail<bf16_t, float>(st);
}

// rin non-null: the parent's restriction of rin is formed in the launch (else
// the tail reads the parent's bc); xpar non-null: the parent's prolongation
// into xpar is applied in the launch (else the parent applies it).
template <typename TV, typename TX>
void DeviceHierarchy::launch_tail(const CycleOpts& o, hipStream_t st, const TX* rin, TX* xpar) {
    typedef typename TransferT<TV>::type TP;
    const int pi = pidx<TV>();
    DevLevel& parent = L_[nlevels() - 2];
    TailRun<TX> run{tail_img_[pi].p, bc_of<TX>(parent), xc_of<TX>(parent), (TX)weight(), o.pre, o.post,
                    tail_lay_[pi]};
    if (rin) {
        run.trp = parent.Pt.rp.p;
        run.tcv = parent.Pt.cv.p;
        run.tvv = vals<TP>(parent.Pt);
        run.rin = rin;
    }
    if (xpar) {
        run.prp = parent.P.rp.p;
        run.pcv = parent.P.cv.p;
        run.pvv = vals<TP>(parent.P);
        run.xpar = xpar;
        run.npar = parent.A.nrow;
    }
    hipLaunchKernelGGL((k_tail_lds<TV, TP, TX>), dim3(1), dim3(kTailThreads), 0, st, run);
}

// Whether the LDS-resident last level pays: the zero-guess cycle (the
// preconditioner's form) with and without it, best of 5 x 4 back-to-back
// cycles; either way gives the same values.
template <typename TV, typename TX>
void DeviceHierarchy::tune_tail() {
    const int pi = pidx<TV>();
    tail_lt_[pi] = -1;
    if (forced_tail_ == 0 || !tail_fits<TV, TX>(nullptr)) return;
    pack_tail<TV, TX>(stream_);  // the tuner has synchronized the device
    DRAUGR_HIP_CHECK(hipStreamSynchronize(stream_));
    if (forced_tail_ == 1) {
        tail_lt_[pi] = nlevels() - 1;
        return;
    }
    hipStream_t st = stream_;
    hipEvent_t e0, e1;
    DRAUGR_HIP_CHECK(hipEventCreate(&e0));
    DRAUGR_HIP_CHECK(hipEventCreate(&e1));
    TX* x = reinterpret_cast<TX*>(w_q_.p);
    TX* b = reinterpret_cast<TX*>(w_z_.p);
    launch_map(n0_, [=] __device__(int64_t i) { b[i] = (TX)1; }, st);
    CycleOpts o;
    o.prec = pi == 0 ? Precision::FP64 : (pi == 1 ? Precision::FP32 : Precision::BF16);
    auto time_cycle = [&]() {
        cycle_top<TV, TX>(x, b, true, o, st);
        float best = 1e30f;
        for (int rep = 0; rep < 5; ++rep) {
            DRAUGR_HIP_CHECK(hipEventRecord(e0, st));
            for (int k = 0; k < 4; ++k) cycle_top<TV, TX>(x, b, true, o, st);
            DRAUGR_HIP_CHECK(hipEventRecord(e1, st));
            DRAUGR_HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0.f;
            DRAUGR_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            best = std::min(best, ms);
        }
        return best * 1e3f / 4.0f;
    };
    notail_us_[pi] = time_cycle();
    tail_lt_[pi] = nlevels() - 1;
    tail_us_[pi] = time_cycle();
    if (!(tail_us_[pi] < notail_us_[pi])) tail_lt_[pi] = -1;
    DRAUGR_HIP_CHECK(hipStreamSynchronize(st));
    DRAUGR_HIP_CHECK(hipEventDestroy(e0));
    DRAUGR_HIP_CHECK(hipEventDestroy(e1));
}

// Autotuner: for every level, time the candidate kernels of each operator
// pass (CSR-stream with LDS-staged row blocks vs thread-per-row) and the two
// forms of the zero-guess smoothing step (fused with the residual vs a map
// followed by the residual), on the level's real operators, and keep the
// fastest.  All candidates compute the same values in the same order, so the
// choice never changes a result.  Runs once per cycle precision, on first use.
template <typename TV, typename TX>
void DeviceHierarchy::tune() {
    typedef typename TransferT<TV>::type TP;
    const int pi = pidx<TV>();
    // nothing left to time when both the kernel and the fusions are fixed
    if (forced_kind_ >= 0 && forced_fuse_ >= 0) {
        for (auto& D : L_) {
            D.A.kind[pi] = D.Pt.kind[pi] = D.P.kind[pi] = forced_kind_;
            D.fuse_zero[pi] = D.fuse_up[pi] = forced_fuse_ == 1;
        }
        for (int l = 0; l < nlevels(); ++l) L_[l].fuse_down[pi] = forced_fuse_ == 1 && down_blocks(l);
        tuned_[pi] = true;
        tune_tail<TV, TX>();
        return;
    }
    DRAUGR_HIP_CHECK(hipDeviceSynchronize());
    hipStream_t st = stream_;
    hipEvent_t e0, e1;
    DRAUGR_HIP_CHECK(hipEventCreate(&e0));
    DRAUGR_HIP_CHECK(hipEventCreate(&e1));
    // best of 5 repetitions of 4 back-to-back launches: the per-launch time a
    // cycle sees, with the events' own overhead and the noise spread thin
    auto timeit = [&](auto&& fn) {
        fn();
        float best = 1e30f;
        for (int rep = 0; rep < 5; ++rep) {
            DRAUGR_HIP_CHECK(hipEventRecord(e0, st));
            for (int k = 0; k < 4; ++k) fn();
            DRAUGR_HIP_CHECK(hipEventRecord(e1, st));
            DRAUGR_HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0.f;
            DRAUGR_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            best = std::min(best, ms);
        }
        return best;
    };
    TX* s1 = reinterpret_cast<TX*>(w_q_.p);  // scratch (finest size, >= any level)
    TX* s2 = reinterpret_cast<TX*>(w_z_.p);
    for (int l = 0; l < nlevels(); ++l) {
        DevLevel& D = L_[l];
        const int n = D.A.nrow, nc = D.P.ncol;
        TX* xin = t_of<TX>(D);
        TX* bin = r_of<TX>(D);
        TX* xcv = xc_of<TX>(D);
        launch_map(n, [=] __device__(int64_t i) { xin[i] = (TX)1; bin[i] = (TX)1; s1[i] = (TX)0; s2[i] = (TX)0; }, st);
        launch_map(nc, [=] __device__(int64_t i) { xcv[i] = (TX)1; }, st);
        const TX* d = d_of<TX>(D);
        const TX w = (TX)weight();
        float tA[kOpKinds], tP[kOpKinds], tPt[kOpKinds];
        for (int k = 0; k < kOpKinds; ++k) {
            if (forced_kind_ >= 0) {  // fixed kernel: only the fusions are timed
                tA[k] = tPt[k] = tP[k] = k == forced_kind_ ? 0.f : 1e30f;
                continue;
            }
            const float skip = 1e30f;  // the 16-bit and sorted forms only where the operator has them
            auto absent = [&](const DevCSR& M) {
                return ((k == kOpStream16 || k == kOpStream16Pipe) && !M.has16) ||
                       ((k == kOpSorted || k == kOpSortedPipe || k == kOpSortedQuad) && !M.hasS) ||
                       (k == kOpPattern && !M.hasT) ||
                       (k == kOpVector && sizeof(TX) == 8);  // FP64 keeps the sequential row sums
            };
            tA[k] = absent(D.A) ? skip : timeit([&] {
                op_kind<TV, TX>(k, D.A, XVec<TX>{xin}, EpiJacobi<TX>{xin, bin, d, s1, w}, st);
            });
            tPt[k] = absent(D.Pt) ? skip : timeit([&] {
                op_kind<TP, TX>(k, D.Pt, XVec<TX>{xin}, EpiStore<TX>{s2}, st);
            });
            tP[k] = absent(D.P) ? skip : timeit([&] {
                op_kind<TP, TX>(k, D.P, XVec<TX>{xcv}, EpiAdd<TX>{s2}, st);
            });
        }
        auto best = [](const float* t) { return (int)(std::min_element(t, t + kOpKinds) - t); };
        for (int k = 0; k < kOpKinds; ++k) D.tA_us[pi][k] = tA[k] < 1e29f ? tA[k] * 1e3f / 4.0f : -1.0f;
        D.A.kind[pi] = best(tA);
        D.Pt.kind[pi] = best(tPt);
        D.P.kind[pi] = best(tP);
        const float fused = timeit([&] {
            op<TV, TX>(D.A, XZeroJacobi<TX>{d, bin, w}, EpiZeroJacobiResid<TX>{bin, d, s1, s2, w}, st);
        });
        // below the finest level the split form's map rides along with the
        // parent's restriction, so only its residual pass is counted there
        const bool map_folded = l >= 1 && (smoother_ == Smoother::Jacobi || smoother_ == Smoother::L1Jacobi ||
                                           smoother_ == Smoother::SPAI0);
        const float split = timeit([&] {
            if (!map_folded) launch_map(n, [=] __device__(int64_t i) { s1[i] = (w * d[i]) * bin[i]; }, st);
            op<TV, TX>(D.A, XVec<TX>{s1}, EpiResid<TX>{bin, s2}, st);
        });
        D.fuse_zero[pi] = forced_fuse_ >= 0 ? forced_fuse_ == 1 : fused < split;
        // fused down (zero step + residual + restriction) and up (prolongation +
        // first post-smoothing step) passes against their separate launches
        D.fuse_down[pi] = D.fuse_up[pi] = false;
        const bool jac = smoother_ == Smoother::Jacobi || smoother_ == Smoother::L1Jacobi ||
                         smoother_ == Smoother::SPAI0;
        if (jac) {
            const float restrict_t = timeit([&] { op<TP, TX>(D.Pt, XVec<TX>{s2}, EpiStore<TX>{s1}, st); });
            if (down_blocks(l)) {
                const float down_t = timeit([&] {
                    launch_down<TV, TX>(l, bin, s1, s2, (const TX*)nullptr, (TX*)nullptr, st);
                });
                D.fuse_down[pi] =
                    forced_fuse_ >= 0 ? forced_fuse_ == 1 : down_t < std::min(fused, split) + restrict_t;
            }
            const float sep_up = timeit([&] {
                op<TP, TX>(D.P, XVec<TX>{xcv}, EpiAdd<TX>{s2}, st);
                op<TV, TX>(D.A, XVec<TX>{s2}, EpiJacobi<TX>{s2, bin, d, s1, w}, st);
            });
            const XProlong<TP, TX> xpf{s2, D.P.rp.p, D.P.cv.p, vals<TP>(D.P), xcv};
            const float up_t = timeit([&] {
                op<TV, TX>(D.A, xpf, EpiJacobiP<TP, TX>{xpf, bin, d, s1, w}, st);
            });
            D.fuse_up[pi] = forced_fuse_ >= 0 ? forced_fuse_ == 1 : up_t < sep_up;
        }
    }
    DRAUGR_HIP_CHECK(hipStreamSynchronize(st));
    DRAUGR_HIP_CHECK(hipEventDestroy(e0));
    DRAUGR_HIP_CHECK(hipEventDestroy(e1));
    tuned_[pi] = true;
    tune_tail<TV, TX>();
}

std::string DeviceHierarchy::tune_report() const {
    std::string out;
    const char* names[3] = {"fp64", "fp32", "bf16"};
    char buf[160];
    for (int p = 0; p < 3; ++p) {
        if (!tuned_[p]) continue;
        for (int l = 0; l < nlevels(); ++l) {
            const DevLevel& D = L_[l];
            const char* kn[kOpKinds] = {"stream",       "rowseq",     "pipe",       "stream16",  "sorted",
                                        "stream16pipe", "sortedpipe", "sortedquad", "vector",
                                        "pattern"};
            std::snprintf(buf, sizeof buf, "%s level %d: A=%s Pt=%s P=%s zero-guess=%s down=%s up=%s\n", names[p], l,
                          kn[D.A.kind[p]], kn[D.Pt.kind[p]], kn[D.P.kind[p]], D.fuse_zero[p] ? "fused" : "split",
                          D.fuse_down[p] ? "one-pass" : (D.down_state < 0 ? "separate(only)" : "separate"),
                          D.fuse_up[p] ? "one-pass" : "separate");
            out += buf;
            if (p == 0 && D.A.hasT) {
                std::snprintf(buf, sizeof buf, "level %d: A has the pattern form (%d row patterns, width %d)\n", l,
                              D.A.npat, D.A.tw);
                out += buf;
            }
            if (D.A.nnz >= (1 << 20) && forced_kind_ < 0) {  // the big levels: every kind's time for the Jacobi pass
                std::string tl = std::string(names[p]) + " level " + std::to_string(l) + " Jacobi pass us:";
                for (int k = 0; k < kOpKinds; ++k) {
                    if (D.tA_us[p][k] < 0.f) continue;
                    std::snprintf(buf, sizeof buf, " %s %.1f", kn[k], D.tA_us[p][k]);
                    tl += buf;
                }
                out += tl + "\n";
            }
        }
    }
    for (int p = 0; p < 3; ++p) {
        if (!tuned_[p]) continue;
        if (tail_lt_[p] >= 0)
            std::snprintf(buf, sizeof buf, "%s tail: level %d and the coarse solve in one launch from LDS "
                          "(zero-guess cycle %.1f us, %.1f us without)\n", names[p], tail_lt_[p], tail_us_[p],
                          notail_us_[p]);
        else if (notail_us_[p] > 0.f)
            std::snprintf(buf, sizeof buf, "%s tail: separate launches (zero-guess cycle %.1f us, %.1f us with "
                          "the LDS-resident last level)\n", names[p], notail_us_[p], tail_us_[p]);
        else
            std::snprintf(buf, sizeof buf, "%s tail: separate launches (the last level does not qualify or the "
                          "choice is fixed)\n", names[p]);
        out += buf;
    }
    if (pcg_form_ >= 0) {  // the form of the PCG's q = A p pass, once a PCG has run
        const DevCSR& A0 = nlevels() > 0 ? L_[0].A : Ac_;
        out += pcg_form_ == kPcgSorted    ? "pcg q = A p: sorted\n"
               : pcg_form_ == kPcgPattern ? "pcg q = A p: pattern\n"
               : A0.has16                 ? "pcg q = A p: stream16\n"
                                          : "pcg q = A p: stream\n";
    }
    return out;
}

// ILU(0) smoothing (src/smoothers.jl:1356-1433): per step tmp = b - A x (row
// sum, then subtraction), forward substitution with L, backward with U by
// level schedule, x += tmp where tmp is finite.  In place on x.
template <typename TV, typename TX>
void DeviceHierarchy::ilu_smooth(int l, TX* x, TX* tmp, const TX* b, int steps, bool& zero, hipStream_t st) {
    typedef typename TransferT<TV>::type TF;
    DevLevel& D = L_[l];
    const int n = D.A.nrow;
    if (zero) {
        launch_map(n, [=] __device__(int64_t i) { x[i] = 0; }, st);
        zero = false;
    }
    const TF* Lf = sizeof(TX) == 8 ? (const TF*)D.ilu_L64.p : (const TF*)D.ilu_L32.p;
    const TF* Uf = sizeof(TX) == 8 ? (const TF*)D.ilu_U64.p : (const TF*)D.ilu_U32.p;
    const TX eps = sizeof(TX) == 8 ? (TX)2.220446049250313e-16 : (TX)1.1920928955078125e-07;
    for (int s = 0; s < steps; ++s) {
        op<TV, TX>(D.A, XVec<TX>{x}, EpiResid<TX>{b, tmp}, st);
        const auto& I = D.ilu;
        for (size_t c = 1; c < I.fwd_offsets.size(); ++c) {
            const int off = I.fwd_offsets[c - 1], cnt = I.fwd_offsets[c] - off;
            hipLaunchKernelGGL((k_ilu_fwd<TF, TX>), dim3((cnt + kBlock - 1) / kBlock), dim3(kBlock), 0, st, cnt,
                               D.ilu_fwd.p + off, D.A.rp.p, D.A.cv.p, D.ilu_diag.p, Lf, tmp);
        }
        for (size_t c = 1; c < I.bwd_offsets.size(); ++c) {
            const int off = I.bwd_offsets[c - 1], cnt = I.bwd_offsets[c] - off;
            hipLaunchKernelGGL((k_ilu_bwd<TF, TX>), dim3((cnt + kBlock - 1) / kBlock), dim3(kBlock), 0, st, cnt,
                               D.ilu_bwd.p + off, D.A.rp.p, D.A.cv.p, D.ilu_diag.p, Uf, tmp, eps);
        }
        const TX* t = tmp;
        launch_map(n, [=] __device__(int64_t i) {
            const TX v = t[i];
            if (isfinite(v)) x[i] += v;
        }, st);
    }
}

void DeviceHierarchy::ilu_factor(int l, const std::vector<double>& values, hipStream_t st) {
    DevLevel& D = L_[l];
    CSR A;
    A.nrow = A.ncol = D.A.nrow;
    A.rp = D.h_rp;
    A.cv = D.h_cv;
    A.nz = values;
    std::vector<double> Lv, Uv;
    ilu0_factorize(A, D.ilu.diag_idx, Lv, Uv);
    D.ilu_L64.upload(Lv, st);
    D.ilu_U64.upload(Uv, st);
    DRAUGR_HIP_CHECK(hipStreamSynchronize(st));  // host vectors are local
    if (D.ilu_L32.p) {
        for (auto pr : {std::make_pair(D.ilu_L64.p, D.ilu_L32.p), std::make_pair(D.ilu_U64.p, D.ilu_U32.p)}) {
            const double* s = pr.first;
            float* d = pr.second;
            launch_map(D.A.nnz, [=] __device__(int64_t i) { d[i] = (float)s[i]; }, st);
        }
    }
}

// Chebyshev smoothing (src/smoothers.jl:952-990): per step, degree 3 passes
// of r = b - A x fused with the three-term update of d and x (out of place:
// x and alt are swapped after each pass).  The recurrence scalars are formed
// in FP64 as the reference forms them, then rounded to the cycle precision.
template <typename TV, typename TX>
void DeviceHierarchy::chebyshev(int l, TX*& x, TX*& alt, const TX* b, int steps, bool& zero, hipStream_t st) {
    DevLevel& D = L_[l];
    const int n = D.A.nrow;
    if (zero) {
        TX* xx = x;
        launch_map(n, [=] __device__(int64_t i) { xx[i] = 0; }, st);
        zero = false;
    }
    const TX* dv = d_of<TX>(D);
    TX* dd = sizeof(TX) == 8 ? (TX*)D.cd64.p : (TX*)D.cd32.p;
    const double theta = (D.cheb_lmax + D.cheb_lmin) / 2.0;
    const double delta = (D.cheb_lmax - D.cheb_lmin) / 2.0;
    constexpr int kDegree = 3;  // build_chebyshev_smoother's default (src/smoothers.jl:890)
    for (int s = 0; s < steps; ++s) {
        op<TV, TX>(D.A, XVec<TX>{x}, EpiCheb0<TX>{x, b, dv, dd, alt, (TX)(1.0 / theta)}, st);
        std::swap(x, alt);
        double sigma_old = theta / delta;
        for (int k = 1; k < kDegree; ++k) {
            const double sigma_new = 1.0 / (2.0 * theta / delta - sigma_old);
            const double cr = 2.0 * sigma_new / delta;
            const double cd = sigma_new * sigma_old;
            op<TV, TX>(D.A, XVec<TX>{x}, EpiChebK<TX>{x, b, dv, dd, alt, (TX)cr, (TX)cd}, st);
            std::swap(x, alt);
            sigma_old = sigma_new;
        }
    }
}

// One visit of level l (src/cycle.jl:63-99).
template <typename TV, typename TX>
void DeviceHierarchy::cycle_level(int l, TX* x, const TX* b, bool zero, const CycleOpts& o, hipStream_t st,
                                  bool x1_ready) {
    typedef typename TransferT<TV>::type TP;
    const int pi = pidx<TV>();
    DevLevel& D = L_[l];
    TX* xcur = x;
    TX* xalt = t_of<TX>(D);
    TX* r = r_of<TX>(D);
    TX* xc = xc_of<TX>(D);
    TX* bc = bc_of<TX>(D);
    const bool jac = smoother_ == Smoother::Jacobi || smoother_ == Smoother::L1Jacobi || smoother_ == Smoother::SPAI0;
    // the child's first step from zero rides along with the restriction
    const bool child_x1 = zero_step_is_map<TV>(l + 1, o.pre);
    TX* x1c = child_x1 ? t_of<TX>(L_[l + 1]) : nullptr;
    const TX* dc = child_x1 ? d_of<TX>(L_[l + 1]) : nullptr;
    // the next level is the LDS-resident tail: this level's restriction and
    // prolongation ride in its launch unless a one-pass sweep does them
    const bool tail_next = l + 1 < nlevels() && l + 1 == tail_lt_[pi] && !o.w_cycle && jac;
    const bool down_fused = zero && jac && o.pre == 1 && D.fuse_down[pi];
    const bool up_fused = jac && o.post >= 1 && D.fuse_up[pi];
    const bool fold_in = tail_next && !down_fused;
    const bool fold_out = tail_next && !up_fused;
    if (down_fused) {
        // first step from zero, residual and restriction in one launch
        launch_down<TV, TX>(l, b, xalt, bc, dc, x1c, st);
        std::swap(xcur, xalt);
    } else {
        bool resid_done = false;
        smooth<TV, TX>(l, xcur, xalt, b, o.pre, zero, true, resid_done, st, x1_ready && zero);
        if (!resid_done) op<TV, TX>(D.A, XVec<TX>{xcur}, EpiResid<TX>{b, r}, st);
        if (fold_in) {
            // the restriction is formed by the tail's launch
        } else if (child_x1) {
            op<TP, TX>(D.Pt, XVec<TX>{r}, EpiStoreJac0<TX>{bc, dc, x1c, (TX)weight()}, st);
        } else {
            op<TP, TX>(D.Pt, XVec<TX>{r}, EpiStore<TX>{bc}, st);
        }
    }
    if (tail_next) {
        // the last level and the coarse solve in one launch
        launch_tail<TV, TX>(o, st, fold_in ? r : nullptr, fold_out ? xcur : nullptr);
    } else if (l + 1 < nlevels()) {
        const int reps = o.w_cycle ? 2 : 1;
        for (int rep = 0; rep < reps; ++rep) cycle_level<TV, TX>(l + 1, xc, bc, rep == 0, o, st, rep == 0 && child_x1);
    } else {
        coarse_solve<TX>(xc, bc, st);
    }
    bool z = false, dummy = false;
    if (jac && o.post >= 1 && D.fuse_up[pi]) {
        // prolongation and the first post-smoothing step in one launch
        const XProlong<TP, TX> xp{xcur, D.P.rp.p, D.P.cv.p, vals<TP>(D.P), xc};
        op<TV, TX>(D.A, xp, EpiJacobiP<TP, TX>{xp, b, d_of<TX>(D), xalt, (TX)weight()}, st);
        std::swap(xcur, xalt);
        smooth<TV, TX>(l, xcur, xalt, b, o.post - 1, z, false, dummy, st, false, o.gs_sym);
    } else {
        if (!fold_out) op<TP, TX>(D.P, XVec<TX>{xc}, EpiAdd<TX>{xcur}, st);
        smooth<TV, TX>(l, xcur, xalt, b, o.post, z, false, dummy, st, false, o.gs_sym);
    }
    if (xcur != x) {
        const TX* s = xcur;
        launch_map(D.A.nrow, [=] __device__(int64_t i) { x[i] = s[i]; }, st);
    }
}

template <typename TV, typename TX>
void DeviceHierarchy::cycle_top(TX* x, const TX* b, bool zero, const CycleOpts& o, hipStream_t st) {
    if (!tuned_[pidx<TV>()]) tune<TV, TX>();
    if (nlevels() == 0) {
        coarse_solve<TX>(x, b, st);
        return;
    }
    cycle_level<TV, TX>(0, x, b, zero, o, st, zero && o.x1_ready0);
}

// The graph holds the cycle's launches with their arguments (buffers of this
// hierarchy and the caller's x, b), so it stays valid across resetups, which
// rewrite values in place.  The tuner and the precision mirrors must exist
// before capture (both synchronize).  The cycle itself contains no memset
// nodes (buffers are cleared by fill kernels).
template <typename TV, typename TX>
void DeviceHierarchy::precondition(TX* x, const TX* b, const CycleOpts& o, hipStream_t st, bool capture) {
    if (!tuned_[pidx<TV>()]) tune<TV, TX>();
    if (!o.graphs || nlevels() == 0) {
        cycle_top<TV, TX>(x, b, true, o, st);
        return;
    }
    const auto key = std::make_tuple(pidx<TV>(), (int)o.w_cycle, o.pre, o.post, (int)o.x1_ready0, (int)o.gs_sym,
                                     (const void*)x, (const void*)b);
    auto it = graphs_.find(key);
    if (it == graphs_.end() && !capture) {  // no graph yet and none may be made: the same launches, eagerly
        cycle_top<TV, TX>(x, b, true, o, st);
        return;
    }
    if (it == graphs_.end()) {
        if (!cap_) DRAUGR_HIP_CHECK(hipStreamCreateWithFlags(&cap_, hipStreamNonBlocking));
        hipGraph_t g = nullptr;
        DRAUGR_HIP_CHECK(hipStreamBeginCapture(cap_, hipStreamCaptureModeThreadLocal));
        try {
            cycle_top<TV, TX>(x, b, true, o, cap_);
        } catch (...) {
            (void)hipStreamEndCapture(cap_, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        DRAUGR_HIP_CHECK(hipStreamEndCapture(cap_, &g));
        hipGraphExec_t ex = nullptr;
        const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        DRAUGR_HIP_CHECK(e);
        it = graphs_.emplace(key, ex).first;
    }
    DRAUGR_HIP_CHECK(hipGraphLaunch(it->second, st));
}

void DeviceHierarchy::cycle(double* x, const double* b, bool zero, const CycleOpts& o, hipStream_t st) {
    if (o.prec == Precision::FP64 && zero && o.graphs && nlevels() > 0) {
        // x = V(0; b) replayed from the graph on the hierarchy's own vectors (one
        // graph whatever the caller's pointers); copies do not change values
        double* r = w_r_.p;
        double* z = w_z_.p;
        const int n = n0_;
        launch_map(n, [=] __device__(int64_t i) { r[i] = b[i]; }, st);
        precondition<double, double>(z, r, o, st);
        launch_map(n, [=] __device__(int64_t i) { x[i] = z[i]; }, st);
        return;
    }
    if (o.prec == Precision::FP64 || nlevels() == 0) {
        // no AMG levels: the cycle is the direct solve (src/cycle.jl:18-24), kept in FP64
        cycle_top<double, double>(x, b, zero, o, st);
        return;
    }
    ensure_precision(o.prec);
    const int n = n0_;
    float* r32 = w_r32_.p;
    float* z32 = w_z32_.p;
    if (zero) {
        launch_map(n, [=] __device__(int64_t i) { r32[i] = (float)b[i]; }, st);
    } else {
        // correction form: x += V(0; b - A x), the same affine map in exact arithmetic
        double* r = w_r_.p;
        op<double, double>(L_[0].A, XVec<double>{x}, EpiResid<double>{b, r}, st);
        launch_map(n, [=] __device__(int64_t i) { r32[i] = (float)r[i]; }, st);
    }
    if (o.prec == Precision::FP32) precondition<float, float>(z32, r32, o, st);
    else precondition<bf16_t, float>(z32, r32, o, st);
    if (zero) launch_map(n, [=] __device__(int64_t i) { x[i] = (double)z32[i]; }, st);
    else launch_map(n, [=] __device__(int64_t i) { x[i] += (double)z32[i]; }, st);
}

void DeviceHierarchy::precondition_repeat(const double* b, const CycleOpts& o, int reps, hipStream_t st) {
    if (nlevels() == 0) throw Error("precondition_repeat: the hierarchy has no AMG levels");
    const int n = n0_;
    if (o.prec == Precision::FP64) {
        double* r = w_r_.p;
        double* z = w_z_.p;
        launch_map(n, [=] __device__(int64_t i) { r[i] = b[i]; }, st);
        for (int k = 0; k < reps; ++k) precondition<double, double>(z, r, o, st);
        return;
    }
    ensure_precision(o.prec);
    float* r32 = w_r32_.p;
    float* z32 = w_z32_.p;
    launch_map(n, [=] __device__(int64_t i) { r32[i] = (float)b[i]; }, st);
    for (int k = 0; k < reps; ++k) {
        if (o.prec == Precision::FP32) precondition<float, float>(z32, r32, o, st);
        else precondition<bf16_t, float>(z32, r32, o, st);
    }
}

double DeviceHierarchy::reduce_to_host(hipStream_t st, int slot) { return reduce_to_host(st, scal_.p, slot); }

double DeviceHierarchy::reduce_to_host(hipStream_t st, const double* scal, int slot) {
    DRAUGR_HIP_CHECK(hipMemcpyAsync(pinned_ + slot, scal + slot, sizeof(double), hipMemcpyDeviceToHost, st));
    DRAUGR_HIP_CHECK(hipStreamSynchronize(st));
    return pinned_[slot];
}

// Stationary AMG iteration with the true-residual stopping test
// (src/cycle.jl:110-160).  In FP32/BF16 cycle mode each iteration is the
// correction form x += V(0; b - A x) with the residual in FP64.
int DeviceHierarchy::solve(double* x, const double* b, double tol, int maxiter, const CycleOpts& o,
                           hipStream_t st, double* relres, int verbose) {
    const int n = n0_;
    double* scal = scal_.p;
    launch_reduce(n, [=] __device__(int64_t i) { return b[i] * b[i]; },
                  [=] __device__(double s) { scal[S_BB] = s; }, partial_.p, st);
    const double bnorm = std::sqrt(reduce_to_host(st, S_BB));
    if (relres) *relres = 0.0;
    if (bnorm == 0.0) {
        launch_map(n, [=] __device__(int64_t i) { x[i] = 0.0; }, st);
        DRAUGR_HIP_CHECK(hipStreamSynchronize(st));
        return 0;
    }
    if (nlevels() == 0) {
        cycle(x, b, true, o, st);
        DRAUGR_HIP_CHECK(hipStreamSynchronize(st));
        return 1;
    }
    DevLevel& D = L_[0];
    double* r = w_r_.p;
    double rel = 1.0;
    for (int it = 1; it <= maxiter; ++it) {
        cycle(x, b, false, o, st);
        op<double, double>(D.A, XVec<double>{x}, EpiResid<double>{b, r}, st);
        launch_reduce(n, [=] __device__(int64_t i) { return r[i] * r[i]; },
                      [=] __device__(double s) { scal[S_RR] = s; }, partial_.p, st);
        rel = std::sqrt(reduce_to_host(st, S_RR)) / bnorm;
        if (relres) *relres = rel;
        if (verbose >= 2) std::printf("  AMG iter %4d: relative residual = %.6e\n", it, rel);
        // a non-finite residual never passes the test: the iteration runs on to
        // maxiter and returns it, as src/cycle.jl:135-159 does
        if (rel < tol) return it;
    }
    return maxiter;
}

// The form of the PCG's q = A p pass: timed once among the forms the
// operator has (4 back-to-back launches, best of 3, on scratch vectors with
// the pass's two gathers per nonzero); a forced kernel kind decides instead.
// Every form gives the same values.
int DeviceHierarchy::pcg_form(const DevCSR& A, hipStream_t st) {
    if (pcg_form_ >= 0) return pcg_form_;
    pcg_form_ = kPcgStream;
    if (!A.hasS && !A.hasT) return pcg_form_;
    if (forced_kind_ >= 0) {
        if (A.hasS && (forced_kind_ == kOpSorted || forced_kind_ == kOpSortedPipe || forced_kind_ == kOpSortedQuad))
            pcg_form_ = kPcgSorted;
        else if (A.hasT && forced_kind_ == kOpPattern)
            pcg_form_ = kPcgPattern;
        return pcg_form_;
    }
    const double* u = w_z_.p;
    const double* v = w_p0_.p;
    double* q = w_q_.p;
    auto xf = [=] __device__(int j) -> double { return u[j] + 0.5 * v[j]; };
    auto epi = [=] __device__(int i, double s) -> double {
        q[i] = s;
        return s;
    };
    auto fin = [=] __device__(double) {};
    hipEvent_t e0, e1;
    DRAUGR_HIP_CHECK(hipEventCreate(&e0));
    DRAUGR_HIP_CHECK(hipEventCreate(&e1));
    float best[3] = {1e30f, 1e30f, 1e30f};
    const bool have[3] = {true, A.hasS, A.hasT};
    for (int rep = 0; rep < 3; ++rep)
        for (int s = 0; s < 3; ++s) {
            if (!have[s]) continue;
            DRAUGR_HIP_CHECK(hipEventRecord(e0, st));
            for (int k = 0; k < 4; ++k) launch_stream_reduce<double, double>(A, A.v64.p, xf, epi, fin, partial_.p, st, s);
            DRAUGR_HIP_CHECK(hipEventRecord(e1, st));
            DRAUGR_HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0.f;
            DRAUGR_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            best[s] = std::min(best[s], ms);
        }
    DRAUGR_HIP_CHECK(hipEventDestroy(e0));
    DRAUGR_HIP_CHECK(hipEventDestroy(e1));
    pcg_form_ = (int)(std::min_element(best, best + 3) - best);
    return pcg_form_;
}

// FP64 preconditioned CG (Hestenes-Stiefel) with z = V(0; r).  Stopping test
// on the recurrence residual ||r||/||b|| < tol.  Kernel fusion: p = z + beta p
// is formed inside the q = A p pass (ping-pong p), the pass also reduces p.q,
// sets alpha on the device and applies the previous iteration's x update; the
// r update reduces r.r.  The cycle for the next z is enqueued before the host
// reads r.r, so the GPU never idles on the per-iteration convergence check
// (the last cycle is speculative).
//
// ws (allocation-free entry point): the device scalars and the reduction
// partials live in caller-provided memory instead of the hierarchy's, and the
// call allocates nothing: the precision mirrors, the autotuner's choices and
// the form of the q = A p pass must already exist (a previous call of the
// allocating PCG with the same precision made them), cycles without a graph
// captured for them are launched eagerly (same kernels, same values), and the
// scalars are cleared by a fill kernel.
int DeviceHierarchy::pcg(double* x, const double* b, double tol, int maxiter, const CycleOpts& o,
                         hipStream_t st, double* relres, int verbose, const PcgScratch* ws) {
    const int n = n0_;
    double* scal = ws ? ws->scal : scal_.p;
    double* part = ws ? ws->partial : partial_.p;
    double* r = w_r_.p;
    double* q = w_q_.p;
    double* z = w_z_.p;
    double* pa = w_p0_.p;
    double* pb = w_p1_.p;
    const bool mixed = o.prec != Precision::FP64 && nlevels() > 0;
    const DevCSR& A = nlevels() > 0 ? L_[0].A : Ac_;
    if (nlevels() > 0) {
        // CG needs a symmetric preconditioner: the cycle is symmetric with as
        // many post- as pre-smoothing steps and a smoother that is its own
        // transpose (Jacobi-type, Chebyshev) or is paired with its transpose
        // (coloured GS with the post-sweep in reverse colour order)
        if (o.pre != o.post)
            throw Error("PCG needs a symmetric preconditioner: pre_smoothing_steps (" + std::to_string(o.pre) +
                        ") must equal post_smoothing_steps (" + std::to_string(o.post) + ")");
        if (smoother_ == Smoother::ILU0)
            throw Error("PCG needs a symmetric preconditioner: the ilu0 smoother is not (use amg_solve or "
                        "another smoother)");
        if (is_gs(smoother_) && !o.gs_sym)
            throw Error("PCG needs a symmetric preconditioner: coloured Gauss-Seidel sweeps the colours in the "
                        "same order before and after (set gs_symmetric for a reverse-order post-sweep)");
    }
    if (ws) {
        const int pi = mixed ? (o.prec == Precision::FP32 ? 1 : 2) : 0;
        const bool mirrors = !mixed || (have32_ && (o.prec != Precision::BF16 || have16_));
        if (!mirrors || (nlevels() > 0 && !tuned_[pi]) || pcg_form_ < 0)
            throw Error("allocation-free PCG: the hierarchy is not prepared for this precision (call the "
                        "allocating PCG once with the same config first)");
        const int64_t need = std::max<int64_t>(kRedBlocks, A.nblk);
        if (need > ws->partial_cap)
            throw Error("allocation-free PCG: the finest operator needs " + std::to_string(need) +
                        " reduction partials; the static scratch holds " + std::to_string(ws->partial_cap));
        launch_map(S_COUNT_, [=] __device__(int64_t i) { scal[i] = 0.0; }, st);
    }
    if (mixed) ensure_precision(o.prec);
    float* r32 = w_r32_.p;
    float* z32 = w_z32_.p;
    launch_reduce(n, [=] __device__(int64_t i) { return b[i] * b[i]; },
                  [=] __device__(double s) { scal[S_BB] = s; }, part, st);
    const double bb = reduce_to_host(st, scal, S_BB);
    if (relres) *relres = 0.0;
    if (bb == 0.0) {
        launch_map(n, [=] __device__(int64_t i) { x[i] = 0.0; }, st);
        DRAUGR_HIP_CHECK(hipStreamSynchronize(st));
        return 0;
    }
    const int form = pcg_form(A, st);
    op<double, double>(A, XVec<double>{x}, EpiResid<double>{b, r}, st);  // r = b - A x
    launch_reduce(n, [=] __device__(int64_t i) { return r[i] * r[i]; },
                  [=] __device__(double s) { scal[S_RR] = s; }, part, st);
    double rel = std::sqrt(reduce_to_host(st, scal, S_RR) / bb);
    if (relres) *relres = rel;
    if (rel < tol) return 0;
    launch_map(n, [=] __device__(int64_t i) { pa[i] = 0.0; }, st);
    const bool capture = ws == nullptr;
    auto apply_prec = [&](const CycleOpts& oc) {
        if (mixed) {
            if (oc.prec == Precision::FP32) precondition<float, float>(z32, r32, oc, st, capture);
            else precondition<bf16_t, float>(z32, r32, oc, st, capture);
        } else {
            precondition<double, double>(z, r, oc, st, capture);
        }
    };
    if (mixed) launch_map(n, [=] __device__(int64_t i) { r32[i] = (float)r[i]; }, st);
    apply_prec(o);
    // From here on, when the finest level's first smoothing step from zero is
    // the plain map x1 = (w dinv) r, the residual update below writes x1 too
    // and the preconditioner starts from it (the same values, one pass fewer).
    CycleOpts oit = o;
    if (nlevels() > 0) {
        if (!mixed) oit.x1_ready0 = zero_step_is_map<double>(0, o.pre);
        else if (o.prec == Precision::FP32) oit.x1_ready0 = zero_step_is_map<float>(0, o.pre);
        else oit.x1_ready0 = zero_step_is_map<bf16_t>(0, o.pre);
    }
    const bool x1_0 = oit.x1_ready0;
    float* t32 = (mixed && x1_0) ? L_[0].t32.p : nullptr;
    double* t64 = (!mixed && x1_0) ? L_[0].t64.p : nullptr;
    const float* d32 = (mixed && x1_0) ? L_[0].d32.p : nullptr;
    const double* d64 = (!mixed && x1_0) ? L_[0].d64.p : nullptr;
    const float w32 = (float)weight();
    const double w64 = weight();
    // rz = r.z ; beta = 0 for the first direction
    if (mixed)
        launch_reduce(n, [=] __device__(int64_t i) { return r[i] * (double)z32[i]; },
                      [=] __device__(double s) { scal[S_RZ] = s; scal[S_BETA] = 0.0; }, part, st);
    else
        launch_reduce(n, [=] __device__(int64_t i) { return r[i] * z[i]; },
                      [=] __device__(double s) { scal[S_RZ] = s; scal[S_BETA] = 0.0; }, part, st);
    hipEvent_t ev = pcg_ev_;
    int result = maxiter;
    try {
        for (int it = 1; it <= maxiter; ++it) {
            // p_new = z + beta p_old; q = A p_new; alpha = rz / (p_new . q).
            // The pass also applies the previous iteration's x += alpha p_old
            // (it reads p_old anyway; alpha still holds that iteration's value
            // until this pass's finalize): the same update, one pass later.
            const double* betap = scal + S_BETA;
            const double* alphap = scal + S_ALPHA;
            const bool xupd = it > 1;
            auto fin_pq = [=] __device__(double pq) { scal[S_PQ] = pq; scal[S_ALPHA] = scal[S_RZ] / pq; };
            if (mixed) {
                auto xf = [=] __device__(int j) -> double { return (double)z32[j] + *betap * pa[j]; };
                auto epi = [=] __device__(int i, double s) -> double {
                    const double po = pa[i];
                    const double pn = (double)z32[i] + *betap * po;
                    if (xupd) x[i] += *alphap * po;
                    pb[i] = pn;
                    q[i] = s;
                    return pn * s;
                };
                launch_stream_reduce<double, double>(A, A.v64.p, xf, epi, fin_pq, part, st, form);
            } else {
                auto xf = [=] __device__(int j) -> double { return z[j] + *betap * pa[j]; };
                auto epi = [=] __device__(int i, double s) -> double {
                    const double po = pa[i];
                    const double pn = z[i] + *betap * po;
                    if (xupd) x[i] += *alphap * po;
                    pb[i] = pn;
                    q[i] = s;
                    return pn * s;
                };
                launch_stream_reduce<double, double>(A, A.v64.p, xf, epi, fin_pq, part, st, form);
            }
            // r -= alpha q; r.r  (and the FP32 copy of r).  The per-block
            // partials of r.r go straight into mapped pinned host memory and
            // the host sums them (host_finalize): no finalize launch, and no
            // copy is enqueued for the convergence check.
            double* hp = pinned_dev_ + kPinPart;
            const int nbr = red_blocks(n);
            if (mixed)
                launch_reduce_partials(n,
                              [=] __device__(int64_t i) {
                                  const double a = *alphap;
                                  const double ri = r[i] - a * q[i];
                                  r[i] = ri;
                                  const float rf = (float)ri;
                                  r32[i] = rf;
                                  if (x1_0) t32[i] = (w32 * d32[i]) * rf;
                                  return ri * ri;
                              },
                              hp, st);
            else
                launch_reduce_partials(n,
                              [=] __device__(int64_t i) {
                                  const double a = *alphap;
                                  const double ri = r[i] - a * q[i];
                                  r[i] = ri;
                                  if (x1_0) t64[i] = (w64 * d64[i]) * ri;
                                  return ri * ri;
                              },
                              hp, st);
            DRAUGR_HIP_CHECK(hipEventRecord(ev, st));
            // speculative: next preconditioned residual and beta
            apply_prec(oit);
            auto fin_rz = [=] __device__(double s) {
                scal[S_BETA] = s / scal[S_RZ];
                scal[S_RZ] = s;
            };
            if (mixed)
                launch_reduce(n, [=] __device__(int64_t i) { return r[i] * (double)z32[i]; }, fin_rz, part, st);
            else
                launch_reduce(n, [=] __device__(int64_t i) { return r[i] * z[i]; }, fin_rz, part, st);
            std::swap(pa, pb);
            DRAUGR_HIP_CHECK(hipEventSynchronize(ev));  // the r update has completed: its partials are visible
            const double rr = host_finalize(nbr);
            rel = std::sqrt(rr / bb);
            if (relres) *relres = rel;
            if (verbose >= 2) std::printf("  PCG iter %4d: relative residual = %.6e\n", it, rel);
            if (rel < tol) { result = it; break; }
            if (!std::isfinite(rel)) throw Error("PCG diverged (non-finite residual)");
        }
        if (maxiter >= 1) {  // the last iteration's x += alpha p (p in pa after the swap)
            const double* alphap = scal + S_ALPHA;
            const double* pl = pa;
            launch_map(n, [=] __device__(int64_t i) { x[i] += *alphap * pl[i]; }, st);
        }
    } catch (...) {
        (void)hipStreamSynchronize(st);
        throw;
    }
    DRAUGR_HIP_CHECK(hipStreamSynchronize(st));
    return result;
}

// ── Allocation-free PCG: its scalars and reduction partials in a static
//    device array ──────────────────────────────────────────────────────────
// scratch[0, kScratchScal) holds the device scalars (S_COUNT_ used; the size
// keeps the partials 256-byte aligned), scratch[kScratchScal, +kScratchPartials)
// the per-block partial sums: 513 KiB per GPU, held from the library's load
// until the process ends.  The finest operator may have at most
// kScratchPartials row blocks (about 16 M rows).  Host code uses the array only
// through the device address hipGetSymbolAddress gives for it on each device.
namespace {
constexpr int kScratchScal = 64;
constexpr int64_t kScratchPartials = 65536;
constexpr int kMaxDevices = 64;
static_assert(kScratchScal >= S_COUNT_, "scalar slots");
static __device__ __align__(256) double scratch[kScratchScal + kScratchPartials];

std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up on the first
// call per device (with that device current) and kept.
double* static_scratch_current_device() {
    int dev = 0;
    DRAUGR_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw Error("allocation-free PCG: device index out of range");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            throw Error(std::string("allocation-free PCG: hipGetSymbolAddress(scratch) failed: ") +
                        hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<double*>(ptr);
    }
    return g_scratch_ptr[dev];
}
}  // namespace

int DeviceHierarchy::pcg_no_alloc(double* x, const double* b, double tol, int maxiter, const CycleOpts& o,
                                  hipStream_t st, double* relres, int verbose) {
    int dev = 0;
    DRAUGR_HIP_CHECK(hipGetDevice(&dev));
    if (dev != device_)
        throw Error("allocation-free PCG: the current HIP device (" + std::to_string(dev) +
                    ") is not the hierarchy's (" + std::to_string(device_) + ")");
    double* base = static_scratch_current_device();
    const PcgScratch ws{base, base + kScratchScal, kScratchPartials};
    return pcg(x, b, tol, maxiter, o, st, relres, verbose, &ws);
}

// Partial resetup (src/resetup.jl:27-45): new fine values -> smoother ->
// Galerkin level by level -> coarsest dense operator and its factorisation.
void DeviceHierarchy::resetup(const double* nzval, hipStream_t st) {
    if (nlevels() == 0) {
        DRAUGR_HIP_CHECK(hipMemcpyAsync(Ac_.v64.p, nzval, Ac_.nnz * sizeof(double), hipMemcpyDeviceToDevice, st));
    } else {
        DRAUGR_HIP_CHECK(
            hipMemcpyAsync(L_[0].A.v64.p, nzval, L_[0].A.nnz * sizeof(double), hipMemcpyDeviceToDevice, st));
        compute_dinv(0, st);
        for (int l = 0; l < nlevels(); ++l) {
            galerkin(l, st);
            if (l + 1 < nlevels()) compute_dinv(l + 1, st);
        }
        for (int l = 0; l < nlevels(); ++l) refresh_level_mirrors(l, st);
        if (smoother_ == Smoother::Chebyshev) chebyshev_bounds(st);
        if (smoother_ == Smoother::ILU0) {
            // the factorization runs on the host, as the reference's update_smoother! (src/smoothers.jl:1331-1337)
            DRAUGR_HIP_CHECK(hipStreamSynchronize(st));
            for (int l = 0; l < nlevels(); ++l) {
                std::vector<double> v(L_[l].A.nnz);
                DRAUGR_HIP_CHECK(hipMemcpy(v.data(), L_[l].A.v64.p, v.size() * sizeof(double), hipMemcpyDeviceToHost));
                ilu_factor(l, v, st);
            }
        }
    }
    refresh_mirrors(Ac_, st);  // the PCG's operator when there are no AMG levels
    coarse_factor(st);
    refresh_tails(st);
    DRAUGR_HIP_CHECK(hipStreamSynchronize(st));
    int fl = 0;
    DRAUGR_HIP_CHECK(hipMemcpy(&fl, flag_.p, sizeof(int), hipMemcpyDeviceToHost));
    if (fl) throw Error("coarsest-level matrix is singular after resetup (dense LU found a zero pivot)");
}

// Algorithmic bytes of one cycle from a zero guess, counting every array a
// kernel must touch once (value size sv, vector size sx, int32 indices).
// Column index bytes of one pass as stored: 2 per nonzero of a 16-bit block
// plus the block bases when the tuned kernel reads the 16-bit form, else 4 per
// nonzero.  The algorithmic count of the cycle uses 4 per nonzero whatever
// the form (int32 indices: the work, not the storage, defines it).
static double index_bytes(const DevCSR& M, int pi) {
    if ((M.kind[pi] == kOpStream16 || M.kind[pi] == kOpStream16Pipe) && M.has16)
        return 2.0 * M.nnz16 + 4.0 * (M.nnz - M.nnz16) + 16.0 * M.nblk;  // + the block headers
    if ((M.kind[pi] == kOpSorted || M.kind[pi] == kOpSortedPipe || M.kind[pi] == kOpSortedQuad) &&
        M.hasS)  // one column / position word per sorted entry, 32-bit columns elsewhere
        return 4.0 * M.nnzS + 4.0 * (M.nnz - M.nnzS) + 16.0 * M.nblk;
    return 4.0 * M.nnz;
}

double DeviceHierarchy::op_bytes(int l, int which, Precision p, bool as_stored) const {
    if (l < 0 || l >= nlevels()) throw Error("level index out of range");
    const int pi = (int)p;
    const DevLevel& D = L_[l];
    const DevCSR& M = which == 0 ? D.A : (which == 1 ? D.P : D.Pt);
    const double sv = which == 0 ? (p == Precision::FP64 ? 8 : (p == Precision::FP32 ? 4 : 2))
                                 : (p == Precision::FP64 ? 8 : 4);
    return M.nnz * sv + (as_stored ? index_bytes(M, pi) : 4.0 * M.nnz) + (M.nrow + 1) * 4.0;
}

double DeviceHierarchy::cycle_bytes(const CycleOpts& o) const {
    const double sx = o.prec == Precision::FP64 ? 8 : 4;
    double total = 0.0, visits = 1.0;
    for (int l = 0; l < nlevels(); ++l) {
        const DevLevel& D = L_[l];
        const double n = D.A.nrow, nc = D.P.ncol;
        const double BA = op_bytes(l, 0, o.prec, false);
        double lv = 0.0;
        if (is_gs(smoother_)) {
            lv += n * sx;  // zero fill
            lv += (o.pre + o.post) * (BA + 4 * n * sx);
            lv += BA + 3 * n * sx;  // residual
        } else if (smoother_ == Smoother::ILU0) {
            // zero fill, then per step: residual, both triangular sweeps (each reads A's pattern and a factor), update
            const double sf = o.prec == Precision::FP64 ? 8 : 4;  // factors are FP64 or FP32
            const double BT = D.A.nnz * (sf + 4) + (n + 1) * 4 + n * 4;  // one sweep's factor, pattern, schedule
            lv += n * sx + (o.pre + o.post) * (BA + 3 * n * sx + 2 * BT + 4 * n * sx + 3 * n * sx);
            lv += BA + 3 * n * sx;  // residual
        } else if (smoother_ == Smoother::Chebyshev) {
            // zero fill, then 3 passes per step: x, b, dinv (and d) read, d and x_out written
            lv += n * sx + (o.pre + o.post) * (3 * BA + 5 * n * sx + 2 * 6 * n * sx);
            lv += BA + 3 * n * sx;  // residual
        } else {
            // pre-smoothing from zero (fused with the residual when pre == 1)
            if (o.pre == 1) lv += BA + 4 * n * sx;
            else if (o.pre == 0) lv += n * sx + BA + 3 * n * sx;
            else lv += 3 * n * sx + (o.pre - 1) * (BA + 4 * n * sx) + BA + 3 * n * sx;
            lv += o.post * (BA + 4 * n * sx);
            if ((o.pre + o.post) % 2) lv += 2 * n * sx;
        }
        lv += op_bytes(l, 2, o.prec, false) + n * sx + nc * sx;      // restrict
        lv += op_bytes(l, 1, o.prec, false) + nc * sx + 2 * n * sx;  // prolong
        total += visits * lv;
        if (o.w_cycle) visits *= 2;
    }
    total += visits * ((double)nc_ * nc_ * 8 + 2.0 * nc_ * sx);
    return total;
}

// ── per-kernel hooks for parity tests and measurement ───────────────────────
void DeviceHierarchy::level_op(int lvl, int op_code, int prec, const void* in1, const void* in2, void* out,
                               void* out2, hipStream_t st) {
    if (op_code == 7 || op_code == 12) {  // 12: the form with the vector in a global workspace
        if (prec == 0) coarse_solve<double>((double*)out, (const double*)in1, st, op_code == 12);
        else coarse_solve<float>((float*)out, (const float*)in1, st, op_code == 12);
        return;
    }
    if (op_code == 13) {  // refactor the coarsest operator column by column (the large-operator form)
        coarse_factor(st, true);
        return;
    }
    if (lvl < 0 || lvl >= nlevels()) throw Error("level index out of range");
    DevLevel& D = L_[lvl];
    if (op_code == 8) { galerkin(lvl, st); return; }
    if (op_code == 9) {
        compute_dinv(lvl, st);
        DRAUGR_HIP_CHECK(hipMemcpyAsync(out, D.d64.p, D.A.nrow * sizeof(double), hipMemcpyDeviceToDevice, st));
        return;
    }
    auto run = [&](auto tv_tag, auto tx_tag) {
        typedef decltype(tv_tag) TV;
        typedef decltype(tx_tag) TX;
        typedef typename TransferT<TV>::type TP;
        const TX* d = d_of<TX>(D);
        const TX w = (TX)weight();
        const TX* a = (const TX*)in1;
        const TX* bb = (const TX*)in2;
        TX* o = (TX*)out;
        switch (op_code) {
            case 0: op<TV, TX>(D.A, XVec<TX>{a}, EpiStore<TX>{o}, st); break;
            case 1: op<TV, TX>(D.A, XVec<TX>{a}, EpiResid<TX>{bb, o}, st); break;
            case 2: op<TV, TX>(D.A, XVec<TX>{a}, EpiJacobi<TX>{a, bb, d, o, w}, st); break;
            case 3: op<TV, TX>(D.A, XZeroJacobi<TX>{d, bb, w}, EpiZeroJacobiResid<TX>{bb, d, o, (TX*)out2, w}, st);
                    break;
            case 4: op<TP, TX>(D.Pt, XVec<TX>{a}, EpiStore<TX>{o}, st); break;
            case 5: op<TP, TX>(D.P, XVec<TX>{a}, EpiAdd<TX>{o}, st); break;
            case 6:
                for (int c = 0; c < D.ncolors; ++c) {
                    const int off = D.color_off[c], cnt = D.color_off[c + 1] - off;
                    if (cnt == 0) continue;
                    hipLaunchKernelGGL((k_gs_color<TV, TX>), dim3((cnt + kBlock - 1) / kBlock), dim3(kBlock), 0, st,
                                       cnt, D.color_order.p + off, D.A.rp.p, D.A.cv.p, vals<TV>(D.A), o, bb, d);
                }
                break;
            case 10: {  // one Chebyshev smoothing step in place on out (out holds x on entry)
                if (smoother_ != Smoother::Chebyshev) throw Error("level op 10 needs the chebyshev smoother");
                TX* xp = o;
                TX* ap = t_of<TX>(D);
                bool z = false;
                chebyshev<TV, TX>(lvl, xp, ap, bb, 1, z, st);
                if (xp != o) {
                    const TX* s = xp;
                    launch_map(D.A.nrow, [=] __device__(int64_t i) { o[i] = s[i]; }, st);
                }
                break;
            }
            case 11: {  // one ILU(0) smoothing step in place on out (out holds x), rhs in2
                if (smoother_ != Smoother::ILU0) throw Error("level op 11 needs the ilu0 smoother");
                bool z = false;
                ilu_smooth<TV, TX>(lvl, o, t_of<TX>(D), bb, 1, z, st);
                break;
            }
            default: throw Error("unknown level op");
        }
    };
    if (prec == 0) run(double(), double());
    else if (prec == 1) { ensure_precision(Precision::FP32); run(float(), float()); }
    else { ensure_precision(Precision::BF16); run(bf16_t(), float()); }
}

void DeviceHierarchy::export_level_values(int lvl, int which, double* dst) {
    DRAUGR_HIP_CHECK(hipDeviceSynchronize());
    if (which == 2) {
        DRAUGR_HIP_CHECK(hipMemcpy(dst, clu_.p, clu_.bytes(), hipMemcpyDeviceToHost));
        return;
    }
    if (which == 3) {
        DRAUGR_HIP_CHECK(hipMemcpy(dst, Ac_.v64.p, Ac_.v64.bytes(), hipMemcpyDeviceToHost));
        return;
    }
    if (lvl < 0 || lvl >= nlevels()) throw Error("level index out of range");
    const DevLevel& D = L_[lvl];
    if (which == 0) DRAUGR_HIP_CHECK(hipMemcpy(dst, D.A.v64.p, D.A.v64.bytes(), hipMemcpyDeviceToHost));
    else if (which == 1) DRAUGR_HIP_CHECK(hipMemcpy(dst, D.d64.p, D.d64.bytes(), hipMemcpyDeviceToHost));
    else if (which == 4) { dst[0] = D.cheb_lmin; dst[1] = D.cheb_lmax; }
    else if (which == 6 || which == 7) {
        const DBuf<double>& f = which == 6 ? D.ilu_L64 : D.ilu_U64;
        if (!f.p) throw Error("no ILU(0) factors (smoother is not ilu0)");
        DRAUGR_HIP_CHECK(hipMemcpy(dst, f.p, f.bytes(), hipMemcpyDeviceToHost));
    }
    else if (which == 5) {
        if (!D.pv0.p) throw Error("no power-iteration start vector (smoother is not chebyshev)");
        DRAUGR_HIP_CHECK(hipMemcpy(dst, D.pv0.p, D.pv0.bytes(), hipMemcpyDeviceToHost));
    } else throw Error("unknown export selector");
}

}  // namespace draugr
