// Batched path (see batch.hpp): packed block-diagonal levels, K-wide SpMM
// passes of the V(1,1) cycle, batched coarsest solves, FP64 PCG with one
// stopping test per (system, right-hand side), hipGraph replay.
//
// Every row (i, j) of a pass is summed in the reference's sequential order
// with the same epilogue arithmetic as the single-system kernels, so system
// s, right-hand side j of a batched cycle equals the single-system cycle of
// that system on that right-hand side bit for bit (FP64 and FP32).
#include "batch.hpp"

#include <algorithm>
#include <cstdio>
#include <cmath>
#include <numeric>
#include <string>

namespace draugr {

namespace {

using dev::kBlock;
using dev::ld;
using dev::ldi;
using dev::xcd_block;

constexpr int kChunks = 16;  // row groups per block of the PCG reductions

__device__ __forceinline__ size_t kidx(int i, int j, int ks) { return ((size_t)i << ks) + (size_t)j; }

// four consecutive right-hand sides of one row (16- / 32-byte aligned: K is
// a multiple of 4 and vectors start on 256-byte boundaries)
__device__ __forceinline__ void ld4(const float* p, float (&o)[4]) {
    const float4 v = *reinterpret_cast<const float4*>(p);
    o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
}
__device__ __forceinline__ void ld4(const double* p, double (&o)[4]) {
    const double2 a = *reinterpret_cast<const double2*>(p);
    const double2 c = *reinterpret_cast<const double2*>(p + 2);
    o[0] = a.x; o[1] = a.y; o[2] = c.x; o[3] = c.y;
}

// ── K-wide operands and epilogues (same arithmetic as kernels.hpp) ──────────
template <typename TX> struct KX {
    const TX* x; int ks;
    __device__ TX operator()(int c, int j) const { return x[kidx(c, j, ks)]; }
    __device__ void get4(int c, int j0, TX (&o)[4]) const { ld4(x + kidx(c, j0, ks), o); }
};
template <typename TX> struct KXZero {  // x1_c = w d_c b_c (first Jacobi step from zero)
    const TX* d; const TX* b; TX w; int ks;
    __device__ TX operator()(int c, int j) const { return (w * d[c]) * b[kidx(c, j, ks)]; }
    __device__ void get4(int c, int j0, TX (&o)[4]) const {
        const TX t = w * d[c];
        ld4(b + kidx(c, j0, ks), o);
#pragma unroll
        for (int u = 0; u < 4; ++u) o[u] = t * o[u];
    }
};
template <typename TX> struct KStore {
    TX* y; int ks;
    __device__ TX init(int, int) const { return 0; }
    __device__ TX acc(TX a, TX p) const { return a + p; }
    __device__ void fin(int i, int j, TX a) const { y[kidx(i, j, ks)] = a; }
};
template <typename TX> struct KZeroJacobiResid {
    const TX* b; const TX* d; TX* x1; TX* r; TX w; int ks;
    __device__ TX init(int, int) const { return 0; }
    __device__ TX acc(TX a, TX p) const { return a + p; }
    __device__ void fin(int i, int j, TX a) const {
        const size_t q = kidx(i, j, ks);
        const TX bi = b[q];
        x1[q] = (w * d[i]) * bi;
        r[q] = bi - a;
    }
};
template <typename TX> struct KAdd {
    TX* x; int ks;
    __device__ TX init(int i, int j) const { return x[kidx(i, j, ks)]; }
    __device__ TX acc(TX a, TX p) const { return a + p; }
    __device__ void fin(int i, int j, TX a) const { x[kidx(i, j, ks)] = a; }
};
template <typename TX> struct KJacobi {
    const TX* x; const TX* b; const TX* d; TX* xo; TX w; int ks;
    __device__ TX init(int i, int j) const { return b[kidx(i, j, ks)]; }
    __device__ TX acc(TX a, TX p) const { return a - p; }
    __device__ void fin(int i, int j, TX a) const {
        const size_t q = kidx(i, j, ks);
        xo[q] = x[q] + (w * d[i]) * a;
    }
};

// One thread per (row i, rhs j): kBlock >> ks rows per block, the K lanes of
// a row adjacent (their gathers x[c*K + j] are one contiguous segment; the
// row's column and value loads are shared by the K lanes).
template <typename TV, typename TX, class XF, class EPI>
__global__ __launch_bounds__(kBlock) void k_kpass(int n, int ks, const int* __restrict__ rp,
                                                  const int* __restrict__ cv, const TV* __restrict__ av, XF xf,
                                                  EPI epi) {
    const int b = gridDim.x >= 64 ? xcd_block() : (int)blockIdx.x;
    const int i = (b << (8 - ks)) + ((int)threadIdx.x >> ks);
    const int j = (int)threadIdx.x & ((1 << ks) - 1);
    if (i >= n) return;
    TX a = epi.init(i, j);
    const int e = rp[i + 1];
#pragma unroll 4
    for (int k = rp[i]; k < e; ++k) a = epi.acc(a, ld<TX, TV>(av, k) * xf(ldi(cv, k), j));
    epi.fin(i, j, a);
}

// K >= 4: one thread per (row, four consecutive right-hand sides), which
// gathers its four x values with one vector load; each (row, rhs) is still
// summed on its own in row order.
template <typename TV, typename TX, class XF, class EPI>
__global__ __launch_bounds__(kBlock) void k_kpass4(int n, int ks, const int* __restrict__ rp,
                                                   const int* __restrict__ cv, const TV* __restrict__ av, XF xf,
                                                   EPI epi) {
    const int q = ks - 2;  // log2(threads per row)
    const int b = gridDim.x >= 64 ? xcd_block() : (int)blockIdx.x;
    const int i = (b << (8 - q)) + ((int)threadIdx.x >> q);
    const int j0 = ((int)threadIdx.x & ((1 << q) - 1)) << 2;
    if (i >= n) return;
    TX a[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) a[u] = epi.init(i, j0 + u);
    const int e = rp[i + 1];
#pragma unroll 4
    for (int k = rp[i]; k < e; ++k) {
        const TX v = ld<TX, TV>(av, k);
        TX g[4];
        xf.get4(ldi(cv, k), j0, g);
#pragma unroll
        for (int u = 0; u < 4; ++u) a[u] = epi.acc(a[u], v * g[u]);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) epi.fin(i, j0 + u, a[u]);
}

// K >= 8: one thread per (row, LR = 8 or 16 consecutive right-hand sides):
// fewer threads share (and re-load) a row's column and value, each gathering
// LR / 4 vector loads per product; every (row, rhs) still summed on its own
// in row order, so all mappings give the same values.
template <int LR, typename TV, typename TX, class XF, class EPI>
__global__ __launch_bounds__(kBlock) void k_kpassR(int n, int ks, const int* __restrict__ rp,
                                                   const int* __restrict__ cv, const TV* __restrict__ av, XF xf,
                                                   EPI epi) {
    constexpr int lg = LR == 8 ? 3 : 4;
    const int q = ks - lg;  // log2(threads per row)
    const int b = gridDim.x >= 64 ? xcd_block() : (int)blockIdx.x;
    const int i = (b << (8 - q)) + ((int)threadIdx.x >> q);
    const int j0 = ((int)threadIdx.x & ((1 << q) - 1)) << lg;
    if (i >= n) return;
    TX a[LR];
#pragma unroll
    for (int u = 0; u < LR; ++u) a[u] = epi.init(i, j0 + u);
    const int e = rp[i + 1];
#pragma unroll 2
    for (int k = rp[i]; k < e; ++k) {
        const TX v = ld<TX, TV>(av, k);
        const int c = ldi(cv, k);
#pragma unroll
        for (int h = 0; h < LR / 4; ++h) {
            TX g[4];
            xf.get4(c, j0 + 4 * h, g);
#pragma unroll
            for (int u = 0; u < 4; ++u) a[4 * h + u] = epi.acc(a[4 * h + u], v * g[u]);
        }
    }
#pragma unroll
    for (int u = 0; u < LR; ++u) epi.fin(i, j0 + u, a[u]);
}

// LDS-staged forms of the one and four rhs-per-thread mappings: the
// workgroup's rows are contiguous, so their nonzeros are one range; its
// columns and values are first copied into LDS with coalesced loads, then
// every thread walks its row there.  In the plain kernels each load of a
// column or value touches one cache line per row of the wavefront (the rows
// are a row length apart), as many line requests as the gathers make; here
// the range's lines are each requested once.  Same products, same row-order
// sums.  cap = the most nonzeros of any workgroup's rows (the stage's size).
constexpr int kStageBytes = 32 * 1024;

template <typename TX>
__device__ __forceinline__ int* stage_cols(unsigned char* base, int cap) {
    return reinterpret_cast<int*>(base + (((size_t)cap * sizeof(TX) + 15) & ~(size_t)15));
}

template <typename TV, typename TX>
__device__ __forceinline__ void stage_range(int z0, int nz, const int* __restrict__ cv, const TV* __restrict__ av,
                                            TX* sv, int* sc) {
#pragma unroll 4
    for (int k = (int)threadIdx.x; k < nz; k += kBlock) {
        sv[k] = ld<TX, TV>(av, z0 + k);
        sc[k] = ldi(cv, z0 + k);
    }
}

template <typename TV, typename TX, class XF, class EPI>
__global__ __launch_bounds__(kBlock) void k_kpass_s(int n, int ks, const int* __restrict__ rp,
                                                    const int* __restrict__ cv, const TV* __restrict__ av, XF xf,
                                                    EPI epi, int cap) {
    extern __shared__ __align__(16) unsigned char kst[];
    TX* sv = reinterpret_cast<TX*>(kst);
    int* sc = stage_cols<TX>(kst, cap);
    const int rows = kBlock >> ks;
    const int b = gridDim.x >= 64 ? xcd_block() : (int)blockIdx.x;
    const int i0 = b * rows;
    const int i = i0 + ((int)threadIdx.x >> ks);
    const int j = (int)threadIdx.x & ((1 << ks) - 1);
    const int z0 = rp[i0];
    const bool own = i < n;
    const int rs = own ? rp[i] - z0 : 0, re = own ? rp[i + 1] - z0 : 0;
    stage_range<TV, TX>(z0, rp[min(i0 + rows, n)] - z0, cv, av, sv, sc);
    __syncthreads();
    if (!own) return;
    TX a = epi.init(i, j);
#pragma unroll 4
    for (int k = rs; k < re; ++k) a = epi.acc(a, sv[k] * xf(sc[k], j));
    epi.fin(i, j, a);
}

template <typename TV, typename TX, class XF, class EPI>
__global__ __launch_bounds__(kBlock) void k_kpass4_s(int n, int ks, const int* __restrict__ rp,
                                                     const int* __restrict__ cv, const TV* __restrict__ av, XF xf,
                                                     EPI epi, int cap) {
    extern __shared__ __align__(16) unsigned char kst[];
    TX* sv = reinterpret_cast<TX*>(kst);
    int* sc = stage_cols<TX>(kst, cap);
    const int q = ks - 2;
    const int rows = kBlock >> q;
    const int b = gridDim.x >= 64 ? xcd_block() : (int)blockIdx.x;
    const int i0 = b * rows;
    const int i = i0 + ((int)threadIdx.x >> q);
    const int j0 = ((int)threadIdx.x & ((1 << q) - 1)) << 2;
    const int z0 = rp[i0];
    const bool own = i < n;
    const int rs = own ? rp[i] - z0 : 0, re = own ? rp[i + 1] - z0 : 0;
    stage_range<TV, TX>(z0, rp[min(i0 + rows, n)] - z0, cv, av, sv, sc);
    __syncthreads();
    if (!own) return;
    TX a[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) a[u] = epi.init(i, j0 + u);
#pragma unroll 4
    for (int k = rs; k < re; ++k) {
        const TX v = sv[k];
        TX g[4];
        xf.get4(sc[k], j0, g);
#pragma unroll
        for (int u = 0; u < 4; ++u) a[u] = epi.acc(a[u], v * g[u]);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) epi.fin(i, j0 + u, a[u]);
}

// The most nonzeros of `rows` consecutive rows starting at a multiple of rows.
inline int32_t stage_cap(const std::vector<int32_t>& rp, int32_t nrow, int rows) {
    int32_t c = 0;
    for (int32_t i0 = 0; i0 < nrow; i0 += rows) c = std::max(c, rp[std::min(i0 + rows, nrow)] - rp[i0]);
    return c;
}

// Right-hand sides per thread of a K-wide pass: rpt in {1, 4, 8, 16}, at
// most K (the per-level choice of the batch's autotuner; every choice gives
// the same values).
inline int rpt_fit(int rpt, int K) {
    int r = std::min(rpt, K);
    if (r >= 16) return 16;
    if (r >= 8) return 8;
    if (r >= 4) return 4;
    return 1;
}

// rpt < 0: the LDS-staged kernel of mapping -rpt (1 or 4) when the stage
// (caps[0] / caps[1] nonzeros) fits kStageBytes, else the plain kernel.
template <typename TV, typename TX, class XF, class EPI>
void kpass(int n, int ks, const int* rp, const int* cv, const TV* av, XF xf, EPI epi, hipStream_t st, int rpt = 4,
           const int32_t* caps = nullptr) {
    if (n <= 0) return;
    if (rpt < 0) {
        rpt = -rpt;
        const int r = rpt_fit(rpt, 1 << ks);
        const int cap = caps && (r == 1 || r == 4) ? caps[r == 4 ? 1 : 0] : -1;
        const size_t lds = cap >= 0 ? ((size_t)cap * sizeof(TX) + 15) / 16 * 16 + (size_t)cap * 4 : 0;
        if (cap >= 0 && lds <= (size_t)kStageBytes) {
            const int rows = r == 4 ? kBlock >> (ks - 2) : kBlock >> ks;
            const int nb = (n + rows - 1) / rows;
            if (r == 4)
                hipLaunchKernelGGL((k_kpass4_s<TV, TX, XF, EPI>), dim3(nb), dim3(kBlock), lds, st, n, ks, rp, cv, av,
                                   xf, epi, cap);
            else
                hipLaunchKernelGGL((k_kpass_s<TV, TX, XF, EPI>), dim3(nb), dim3(kBlock), lds, st, n, ks, rp, cv, av,
                                   xf, epi, cap);
            return;
        }
    }
    const int r = rpt_fit(rpt, 1 << ks);
    const int lr = r == 16 ? 4 : (r == 8 ? 3 : (r == 4 ? 2 : 0));
    const int rows = kBlock >> (ks - lr);
    const int nb = (n + rows - 1) / rows;
    if (r == 16)
        hipLaunchKernelGGL((k_kpassR<16, TV, TX, XF, EPI>), dim3(nb), dim3(kBlock), 0, st, n, ks, rp, cv, av, xf, epi);
    else if (r == 8)
        hipLaunchKernelGGL((k_kpassR<8, TV, TX, XF, EPI>), dim3(nb), dim3(kBlock), 0, st, n, ks, rp, cv, av, xf, epi);
    else if (r == 4)
        hipLaunchKernelGGL((k_kpass4<TV, TX, XF, EPI>), dim3(nb), dim3(kBlock), 0, st, n, ks, rp, cv, av, xf, epi);
    else
        hipLaunchKernelGGL((k_kpass<TV, TX, XF, EPI>), dim3(nb), dim3(kBlock), 0, st, n, ks, rp, cv, av, xf, epi);
}

// Coarsest solves, one wavefront per (system, rhs), n <= 64: LAPACK getrs
// order, the single-system wave kernel's solve (wave_lu_solve).
template <typename TX>
__global__ __launch_bounds__(64) void k_kcoarse_wave(const CoarseDesc* __restrict__ cd, int ks,
                                                     const TX* __restrict__ b, TX* __restrict__ x) {
    const CoarseDesc c = cd[blockIdx.x];
    const int j = blockIdx.y, i = threadIdx.x, n = c.n;
    double row[64];
    dev::wave_lu_row(row, n, i, [&](int q) { return c.lu[(size_t)q * n + i]; });
    const double xi = dev::wave_lu_solve(i < n ? (double)b[kidx(c.row0 + c.perm[i], j, ks)] : 0.0, row, n, i);
    if (i < n) x[kidx(c.row0 + i, j, ks)] = (TX)xi;
}

// Larger coarsest levels: the same operation order on one thread, the
// intermediate vector in LDS (n doubles).
template <typename TX>
__global__ __launch_bounds__(64) void k_kcoarse_seq(const CoarseDesc* __restrict__ cd, int ks,
                                                    const TX* __restrict__ b, TX* __restrict__ x) {
    extern __shared__ double y[];
    const CoarseDesc c = cd[blockIdx.x];
    const int j = blockIdx.y, n = c.n;
    if (threadIdx.x != 0) return;
    for (int i = 0; i < n; ++i) {
        double acc = (double)b[kidx(c.row0 + c.perm[i], j, ks)];
        for (int q = 0; q < i; ++q) acc -= c.lu[(size_t)q * n + i] * y[q];
        y[i] = acc;
    }
    for (int i = n - 1; i >= 0; --i) {
        double acc = y[i];
        for (int q = n - 1; q > i; --q) acc -= c.lu[(size_t)q * n + i] * y[q];
        y[i] = acc / c.lu[(size_t)i * n + i];
    }
    for (int i = 0; i < n; ++i) x[kidx(c.row0 + i, j, ks)] = (TX)y[i];
}

template <class F>
__global__ __launch_bounds__(kBlock) void k_kmap(int64_t n, F f) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) f(i);
}
template <class F> void kmap(int64_t n, F f, hipStream_t st) {
    if (n <= 0) return;
    const int64_t nb = std::min<int64_t>((n + kBlock - 1) / kBlock, 16384);
    hipLaunchKernelGGL((k_kmap<F>), dim3((unsigned)nb), dim3(kBlock), 0, st, n, f);
}

// ── PCG reductions: block b covers rows [br0[b], br1[b]) of system bs[b] ────
__device__ __forceinline__ void kblock_sum(double v, double* sh, int ks, double* part, int b) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = (kBlock >> ks) >> 1; s > 0; s >>= 1) {
        if (((int)threadIdx.x >> ks) < s) sh[threadIdx.x] += sh[threadIdx.x + (s << ks)];
        __syncthreads();
    }
    if (((int)threadIdx.x >> ks) == 0) part[kidx(b, threadIdx.x, ks)] = sh[threadIdx.x];
}

// q = A p_new, p_new = z + beta p_old (per system and rhs); partial p_new . q
template <typename TZ>
__global__ __launch_bounds__(kBlock) void k_kpq(int ks, const int* __restrict__ br0, const int* __restrict__ br1,
                                                const int* __restrict__ bsys, const int* __restrict__ rp,
                                                const int* __restrict__ cv, const double* __restrict__ av,
                                                const TZ* __restrict__ z, const double* __restrict__ pa,
                                                double* __restrict__ pb, double* __restrict__ q,
                                                const double* __restrict__ beta, double* __restrict__ part) {
    __shared__ double sh[kBlock];
    const int b = blockIdx.x;
    const int r1 = br1[b], s = bsys[b];
    const int j = (int)threadIdx.x & ((1 << ks) - 1);
    const double bt = beta[kidx(s, j, ks)];
    double dot = 0.0;
    for (int i = br0[b] + ((int)threadIdx.x >> ks); i < r1; i += (kBlock >> ks)) {
        const size_t ii = kidx(i, j, ks);
        const double pn = (double)z[ii] + bt * pa[ii];
        double a = 0.0;
        const int e = rp[i + 1];
        for (int k = rp[i]; k < e; ++k) {
            const size_t cc = kidx(cv[k], j, ks);
            a += av[k] * ((double)z[cc] + bt * pa[cc]);
        }
        pb[ii] = pn;
        q[ii] = a;
        dot += pn * a;
    }
    kblock_sum(dot, sh, ks, part, b);
}

// K >= 4: the same pass with one thread per (row, four consecutive rhs) and
// vector gathers.  Each (row, rhs) product sum is formed as in k_kpq; the
// block's partial p.q per rhs is summed in another (fixed) order.
// STAGE: the block's nonzero range is first copied into LDS (cap entries;
// the staged pass kernels' reason), and the row walks read it there.
template <typename TZ, bool STAGE>
__global__ __launch_bounds__(kBlock) void k_kpq4(int ks, const int* __restrict__ br0, const int* __restrict__ br1,
                                                 const int* __restrict__ bsys, const int* __restrict__ rp,
                                                 const int* __restrict__ cv, const double* __restrict__ av,
                                                 const TZ* __restrict__ z, const double* __restrict__ pa,
                                                 double* __restrict__ pb, double* __restrict__ q,
                                                 const double* __restrict__ beta, double* __restrict__ part,
                                                 int cap) {
    __shared__ double sh[kBlock * 4];
    extern __shared__ __align__(16) unsigned char kst[];
    const int qs = ks - 2;                   // log2(threads per row)
    const int lanes = kBlock >> qs;          // rows in flight per block
    const int K = 1 << ks;
    const int b = blockIdx.x;
    const int r1 = br1[b], s = bsys[b];
    const int lr = (int)threadIdx.x >> qs;
    const int j0 = ((int)threadIdx.x & ((1 << qs) - 1)) << 2;
    double bt[4], dot[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        bt[u] = beta[kidx(s, j0 + u, ks)];
        dot[u] = 0.0;
    }
    const double* sv = av;  // the values and columns the row walks read (LDS copies when STAGE)
    const int* sc = cv;
    int zs = 0;
    if constexpr (STAGE) {
        double* tv = reinterpret_cast<double*>(kst);
        int* tc = stage_cols<double>(kst, cap);
        zs = rp[br0[b]];
        stage_range<double, double>(zs, rp[r1] - zs, cv, av, tv, tc);
        __syncthreads();
        sv = tv;
        sc = tc;
    }
    for (int i = br0[b] + lr; i < r1; i += lanes) {
        const size_t ii = kidx(i, j0, ks);
        TZ zi[4];
        double pi4[4], pn[4], a[4];
        ld4(z + ii, zi);
        ld4(pa + ii, pi4);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            pn[u] = (double)zi[u] + bt[u] * pi4[u];
            a[u] = 0.0;
        }
        const int e = rp[i + 1] - zs;
        for (int k = rp[i] - zs; k < e; ++k) {
            const size_t cc = kidx(sc[k], j0, ks);
            const double v = sv[k];
            TZ zc[4];
            double pc[4];
            ld4(z + cc, zc);
            ld4(pa + cc, pc);
#pragma unroll
            for (int u = 0; u < 4; ++u) a[u] += v * ((double)zc[u] + bt[u] * pc[u]);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            pb[ii + u] = pn[u];
            q[ii + u] = a[u];
            dot[u] += pn[u] * a[u];
        }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) sh[lr * K + j0 + u] = dot[u];
    __syncthreads();
    for (int h = lanes >> 1; h > 0; h >>= 1) {
        if (lr < h)
#pragma unroll
            for (int u = 0; u < 4; ++u) sh[lr * K + j0 + u] += sh[(lr + h) * K + j0 + u];
        __syncthreads();
    }
    if (lr == 0)
#pragma unroll
        for (int u = 0; u < 4; ++u) part[kidx(b, j0 + u, ks)] = sh[j0 + u];
}

// x += alpha p; r -= alpha q (and the FP32 copy of r); partial r.r
template <bool MIXED>
__global__ __launch_bounds__(kBlock) void k_kupdate(int ks, const int* __restrict__ br0, const int* __restrict__ br1,
                                                    const int* __restrict__ bsys, const double* __restrict__ alpha,
                                                    double* __restrict__ x, const double* __restrict__ pb,
                                                    double* __restrict__ r, const double* __restrict__ q,
                                                    float* __restrict__ r32, double* __restrict__ part) {
    __shared__ double sh[kBlock];
    const int b = blockIdx.x;
    const int r1 = br1[b], s = bsys[b];
    const int j = (int)threadIdx.x & ((1 << ks) - 1);
    const double a = alpha[kidx(s, j, ks)];
    double acc = 0.0;
    for (int i = br0[b] + ((int)threadIdx.x >> ks); i < r1; i += (kBlock >> ks)) {
        const size_t ii = kidx(i, j, ks);
        x[ii] += a * pb[ii];
        const double ri = r[ii] - a * q[ii];
        r[ii] = ri;
        if (MIXED) r32[ii] = (float)ri;
        acc += ri * ri;
    }
    kblock_sum(acc, sh, ks, part, b);
}

// r = b - A x (row-order sum, then subtraction) and partial r.r
template <bool MIXED>
__global__ __launch_bounds__(kBlock) void k_kresid_rr(int ks, const int* __restrict__ br0,
                                                      const int* __restrict__ br1, const int* __restrict__ rp,
                                                      const int* __restrict__ cv, const double* __restrict__ av,
                                                      const double* __restrict__ x, const double* __restrict__ bv,
                                                      double* __restrict__ r, float* __restrict__ r32,
                                                      double* __restrict__ part) {
    __shared__ double sh[kBlock];
    const int b = blockIdx.x;
    const int r1 = br1[b];
    const int j = (int)threadIdx.x & ((1 << ks) - 1);
    double acc = 0.0;
    for (int i = br0[b] + ((int)threadIdx.x >> ks); i < r1; i += (kBlock >> ks)) {
        const size_t ii = kidx(i, j, ks);
        double a = 0.0;
        const int e = rp[i + 1];
        for (int k = rp[i]; k < e; ++k) a += av[k] * x[kidx(cv[k], j, ks)];
        const double ri = bv[ii] - a;
        r[ii] = ri;
        if (MIXED) r32[ii] = (float)ri;
        acc += ri * ri;
    }
    kblock_sum(acc, sh, ks, part, b);
}

// partial u . v (u FP64, v FP64 or FP32)
template <typename TV>
__global__ __launch_bounds__(kBlock) void k_kdot(int ks, const int* __restrict__ br0, const int* __restrict__ br1,
                                                 const double* __restrict__ u, const TV* __restrict__ v,
                                                 double* __restrict__ part) {
    __shared__ double sh[kBlock];
    const int b = blockIdx.x;
    const int r1 = br1[b];
    const int j = (int)threadIdx.x & ((1 << ks) - 1);
    double acc = 0.0;
    for (int i = br0[b] + ((int)threadIdx.x >> ks); i < r1; i += (kBlock >> ks)) {
        const size_t ii = kidx(i, j, ks);
        acc += u[ii] * (double)v[ii];
    }
    kblock_sum(acc, sh, ks, part, b);
}

// x = 0 for a (system, rhs) whose b is zero (the reference returns x = 0 then)
__global__ __launch_bounds__(kBlock) void k_kzero_bb(int ks, const int* __restrict__ br0, const int* __restrict__ br1,
                                                     const int* __restrict__ bsys, const double* __restrict__ bb,
                                                     double* __restrict__ x) {
    const int b = blockIdx.x;
    const int r1 = br1[b], s = bsys[b];
    const int j = (int)threadIdx.x & ((1 << ks) - 1);
    if (bb[kidx(s, j, ks)] != 0.0) return;
    for (int i = br0[b] + ((int)threadIdx.x >> ks); i < r1; i += (kBlock >> ks)) x[kidx(i, j, ks)] = 0.0;
}

enum FinMode : int { F_BB = 0, F_RR0, F_PQ, F_RR, F_RZ0, F_RZ };
struct KState {
    double *bb, *rr, *rz, *pq, *alpha, *beta;
    int *active, *iters;
    const double* tol;  // in device memory, so that a captured graph serves every tolerance
};

// Per (system, rhs): the sum of the system's block partials (fixed order for
// a fixed launch shape), then the PCG scalar update of that mode.
__global__ __launch_bounds__(kBlock) void k_kfinal(int mode, int ks, const int* __restrict__ sysb,
                                                   const double* __restrict__ part, KState S) {
    __shared__ double sh[kBlock];
    const int s = blockIdx.x, K = 1 << ks;
    const int b0 = sysb[s], b1 = sysb[s + 1];
    for (int j = 0; j < K; ++j) {
        double acc = 0.0;
        for (int b = b0 + (int)threadIdx.x; b < b1; b += kBlock) acc += part[kidx(b, j, ks)];
        sh[threadIdx.x] = acc;
        __syncthreads();
        for (int w = kBlock / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            const int q = s * K + j;
            const double v = sh[0];
            switch (mode) {
                case F_BB: S.bb[q] = v; break;
                case F_RR0:
                    S.rr[q] = v;
                    S.iters[q] = 0;
                    S.active[q] = (S.bb[q] != 0.0 && !(sqrt(v / S.bb[q]) < *S.tol)) ? 1 : 0;
                    break;
                case F_PQ:
                    S.pq[q] = v;
                    S.alpha[q] = S.active[q] ? S.rz[q] / v : 0.0;
                    break;
                case F_RR:
                    S.rr[q] = v;
                    if (S.active[q]) {
                        S.iters[q] += 1;
                        S.active[q] = (sqrt(v / S.bb[q]) < *S.tol) ? 0 : 1;
                    }
                    break;
                case F_RZ0:
                    S.rz[q] = v;
                    S.beta[q] = 0.0;
                    break;
                default:  // F_RZ
                    S.beta[q] = S.active[q] ? v / S.rz[q] : 0.0;
                    S.rz[q] = v;
                    break;
            }
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(kBlock) void k_kcount(int m, const int* __restrict__ active, int* __restrict__ out) {
    __shared__ int sh[kBlock];
    int c = 0;
    for (int i = threadIdx.x; i < m; i += kBlock) c += active[i];
    sh[threadIdx.x] = c;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

template <typename TX> const TX* vals(const DBuf<double>& v64, const DBuf<float>& v32);
template <> const double* vals<double>(const DBuf<double>& v64, const DBuf<float>&) { return v64.p; }
template <> const float* vals<float>(const DBuf<double>&, const DBuf<float>& v32) { return v32.p; }

void upload_cast(DBuf<float>& dst, const DBuf<double>& src, int64_t n, hipStream_t st) {
    dst.alloc(n);
    const double* s = src.p;
    float* d = dst.p;
    kmap(n, [=] __device__(int64_t i) { d[i] = (float)s[i]; }, st);
}

}  // namespace

// ─────────────────────────────────────────────────────────────────────────────
BatchedHierarchy::BatchedHierarchy(const std::vector<const HostHierarchy*>& hs, const Config& cfg, int nrhs) {
    if (hs.empty()) throw Error("ArgumentError: a batch needs at least one system");
    if (nrhs < 1 || nrhs > 256 || (nrhs & (nrhs - 1)))
        throw Error("ArgumentError: the number of right-hand sides must be a power of two in 1..256");
    if (!(cfg.smoother == Smoother::Jacobi || cfg.smoother == Smoother::L1Jacobi))
        throw Error("the batched path supports the jacobi and l1_jacobi smoothers");
    K_ = nrhs;
    forced_rpt_ = cfg.batch_rpt;
    while ((1 << ks_) < K_) ++ks_;
    l1_ = cfg.smoother == Smoother::L1Jacobi;
    omega_ = cfg.jacobi_omega;
    const int S = (int)hs.size();
    n0_.resize(S);
    off0_.resize(S + 1, 0);
    sys_levels_.resize(S);
    sys_coarse_.resize(S);
    for (int s = 0; s < S; ++s) {
        const HostHierarchy& h = *hs[s];
        if (h.levels.empty())
            throw Error("ArgumentError: system " + std::to_string(s) +
                        " has no AMG level (n <= max_coarse_size); solve it directly");
        sys_levels_[s] = (int)h.levels.size();
        sys_coarse_[s] = h.n_coarse;
        n0_[s] = h.levels[0].A.nrow;
        off0_[s + 1] = off0_[s] + n0_[s];
        lmax_ = std::max(lmax_, sys_levels_[s]);
    }
    N0_ = off0_[S];
    if (N0_ * K_ >= ((int64_t)1 << 31)) throw Error("batch too large: rows x right-hand sides must stay below 2^31");
    std::vector<int> ord(S);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return sys_levels_[a] > sys_levels_[b]; });
    auto rows_at = [&](int s, int l) -> int32_t {
        return l < sys_levels_[s] ? hs[s]->levels[l].A.nrow : hs[s]->n_coarse;
    };
    // level lists and offsets
    std::vector<std::vector<int>> V(lmax_ + 1);
    std::vector<std::vector<int64_t>> off(lmax_ + 1, std::vector<int64_t>(S, -1));
    L_.resize(lmax_ + 1);
    for (int l = 0; l <= lmax_; ++l) {
        int64_t acc = 0, accA = 0;
        bool prefix = true;
        for (int t = 0; t < S; ++t) {
            const int s = l == 0 ? t : ord[t];
            if (sys_levels_[s] < l) continue;
            V[l].push_back(s);
            off[l][s] = acc;
            acc += rows_at(s, l);
            if (sys_levels_[s] > l) {
                if (!prefix) throw Error("internal: smoothing systems are not a prefix of the level");
                accA += rows_at(s, l);
            } else {
                prefix = false;
            }
        }
        if (acc >= ((int64_t)1 << 31)) throw Error("batch level too large for int32 indices");
        L_[l].nV = (int32_t)acc;
        L_[l].nA = (int32_t)accA;
    }
    hipStream_t st = nullptr;
    // packed operators
    std::vector<int32_t> rp0;  // level 0's row offsets (the PCG pass's stage size)
    for (int l = 0; l < lmax_; ++l) {
        PLevel& D = L_[l];
        std::vector<int32_t> arp{0}, acv, prp{0}, pcv, trp{0}, tcv;
        std::vector<double> anz, pnz, tnz;
        for (int s : V[l]) {
            if (sys_levels_[s] <= l) continue;
            const HostLevel& H = hs[s]->levels[l];
            const int32_t o = (int32_t)off[l][s], oc = (int32_t)off[l + 1][s];
            for (int32_t i = 0; i < H.A.nrow; ++i) {
                for (int32_t k = H.A.rp[i]; k < H.A.rp[i + 1]; ++k) {
                    acv.push_back(H.A.cv[k] + o);
                    anz.push_back(H.A.nz[k]);
                }
                arp.push_back((int32_t)acv.size());
                for (int32_t k = H.P.rp[i]; k < H.P.rp[i + 1]; ++k) {
                    pcv.push_back(H.P.cv[k] + oc);
                    pnz.push_back(H.P.nz[k]);
                }
                prp.push_back((int32_t)pcv.size());
            }
        }
        for (int s : V[l + 1]) {
            const HostLevel& H = hs[s]->levels[l];
            const int32_t o = (int32_t)off[l][s];
            const int32_t nc = H.P.ncol;
            for (int32_t c = 0; c < nc; ++c) {
                for (int32_t k = H.Pt.offsets[c]; k < H.Pt.offsets[c + 1]; ++k) {
                    tcv.push_back(H.Pt.fine_rows[k] + o);
                    tnz.push_back(H.P.nz[H.Pt.p_nz_idx[k]]);
                }
                trp.push_back((int32_t)tcv.size());
            }
        }
        auto up = [&](PCSR& M, int32_t nrow, int32_t ncol, const std::vector<int32_t>& rp,
                      const std::vector<int32_t>& cv, const std::vector<double>& nz) {
            M.nrow = nrow;
            M.ncol = ncol;
            M.nnz = (int64_t)cv.size();
            M.rp.upload(rp, st);
            M.cv.upload(cv, st);
            M.v64.upload(nz, st);
            M.cap[0] = stage_cap(rp, nrow, kBlock >> ks_);
            M.cap[1] = ks_ >= 2 ? stage_cap(rp, nrow, kBlock >> (ks_ - 2)) : -1;
            DRAUGR_HIP_CHECK(hipStreamSynchronize(st));
            upload_cast(M.v32, M.v64, M.nnz, st);
        };
        up(D.A, D.nA, D.nV, arp, acv, anz);
        if (l == 0) rp0 = arp;
        up(D.P, D.nA, L_[l + 1].nV, prp, pcv, pnz);
        up(D.Pt, L_[l + 1].nV, D.nA, trp, tcv, tnz);
        D.d64.alloc(D.nA);
        invdiag_device(D.nA, D.A.rp.p, D.A.cv.p, D.A.v64.p, D.d64.p, l1_ ? 1 : 0, st);
        upload_cast(D.d32, D.d64, D.nA, st);
    }
    // coarsest levels: dense LU per system on the device
    lu_.resize(S);
    piv_.resize(S);
    DBuf<int32_t> flags;
    flags.upload(std::vector<int32_t>(S, 0), st);
    for (int s = 0; s < S; ++s) {
        const HostHierarchy& h = *hs[s];
        const int n = h.n_coarse;
        const int l = sys_levels_[s];
        if (n == 0) continue;
        lu_[s].upload(h.coarse_dense, st);
        piv_[s].alloc(2 * (size_t)n);
        dense_lu_device(n, lu_[s].p, piv_[s].p, flags.p + s, st);
        const CoarseDesc c{(int32_t)off[l][s], n, lu_[s].p, piv_[s].p + n};
        if (n <= 64) L_[l].coarse_wave.push_back(c);
        else L_[l].coarse_seq.push_back(c);
    }
    for (auto& D : L_) {
        if (!D.coarse_wave.empty()) D.cw.upload(D.coarse_wave, st);
        if (!D.coarse_seq.empty()) {
            for (const auto& c : D.coarse_seq)
                if ((size_t)c.n * sizeof(double) > 60 * 1024) throw Error("coarsest level too large for the batched direct solve");
            D.cs.upload(D.coarse_seq, st);
        }
    }
    // level-0 reduction blocks, aligned to systems
    {
        const int chunk = (kBlock >> ks_) * kChunks;
        std::vector<int32_t> r0, r1, bs, sb{0};
        for (int s = 0; s < S; ++s) {
            for (int64_t a = off0_[s]; a < off0_[s + 1]; a += chunk) {
                r0.push_back((int32_t)a);
                r1.push_back((int32_t)std::min<int64_t>(a + chunk, off0_[s + 1]));
                bs.push_back(s);
            }
            sb.push_back((int32_t)r0.size());
        }
        nblk0_ = (int)r0.size();
        // the most nonzeros of one reduction block's rows (-1: no staging)
        pq_cap_ = rp0.empty() ? -1 : 0;
        for (size_t k = 0; k < r0.size() && pq_cap_ >= 0; ++k)
            pq_cap_ = r1[k] < (int32_t)rp0.size() ? std::max(pq_cap_, rp0[r1[k]] - rp0[r0[k]]) : -1;
        br0_.upload(r0, st);
        br1_.upload(r1, st);
        bs_.upload(bs, st);
        sysb_.upload(sb, st);
    }
    DRAUGR_HIP_CHECK(hipDeviceSynchronize());
    std::vector<int32_t> fl(S);
    DRAUGR_HIP_CHECK(hipMemcpy(fl.data(), flags.p, S * sizeof(int32_t), hipMemcpyDeviceToHost));
    for (int s = 0; s < S; ++s)
        if (fl[s]) throw Error("SingularException: the coarsest matrix of system " + std::to_string(s) + " is singular");
}

BatchedHierarchy::~BatchedHierarchy() {
    (void)hipDeviceSynchronize();
    for (auto& kv : graphs_) (void)hipGraphExecDestroy(kv.second);
    if (cap_) (void)hipStreamDestroy(cap_);
    if (pinned_) (void)hipHostFree(pinned_);
}

int64_t BatchedHierarchy::bytes_resident() const {
    int64_t b = 0;
    auto csr = [&](const PCSR& m) { b += m.rp.bytes() + m.cv.bytes() + m.v64.bytes() + m.v32.bytes(); };
    for (const auto& D : L_) {
        csr(D.A);
        csr(D.P);
        csr(D.Pt);
        b += D.d64.bytes() + D.d32.bytes() + D.x64.bytes() + D.r64.bytes() + D.b64.bytes() + D.f64.bytes() +
             D.x32.bytes() + D.r32.bytes() + D.b32.bytes() + D.f32.bytes();
    }
    for (const auto& l : lu_) b += l.bytes();
    b += w_r_.bytes() + w_p0_.bytes() + w_p1_.bytes() + w_q_.bytes() + w_z_.bytes() + w_r32_.bytes() +
         w_z32_.bytes() + part_.bytes();
    return b;
}

void BatchedHierarchy::check_opts(const CycleOpts& o) const {
    if (o.w_cycle || o.pre != 1 || o.post != 1)
        throw Error("the batched path runs the V(1,1) cycle (cycle \"v\", one pre- and one post-smoothing step)");
    if (o.prec == Precision::BF16) throw Error("the batched path supports precision fp64 and fp32");
}

template <typename TX>
void BatchedHierarchy::ensure_vectors() {
    const int pi = sizeof(TX) == 8 ? 0 : 1;
    if (have_vec_[pi]) return;
    for (int l = 0; l <= lmax_; ++l) {
        PLevel& D = L_[l];
        const int64_t m = (int64_t)D.nV * K_;
        if (sizeof(TX) == 8) {
            if (l < lmax_) { D.x64.alloc(m); D.r64.alloc(m); }
            if (l > 0) { D.b64.alloc(m); D.f64.alloc(m); }
        } else {
            if (l < lmax_) { D.x32.alloc(m); D.r32.alloc(m); }
            if (l > 0) { D.b32.alloc(m); D.f32.alloc(m); }
        }
    }
    have_vec_[pi] = true;
}

// One V(1,1) cycle from zero for all systems and right-hand sides:
// f0 = V(0; b0) (src/cycle.jl:63-99 per system).
template <typename TX>
void BatchedHierarchy::cycle_t(TX* f0, const TX* b0, hipStream_t st) {
    const int ks = ks_;
    const int pi = sizeof(TX) == 8 ? 0 : 1;
    const TX w = (TX)omega_;
    auto xb = [](PLevel& D) -> TX* { if constexpr (sizeof(TX) == 8) return D.x64.p; else return D.x32.p; };
    auto rb = [](PLevel& D) -> TX* { if constexpr (sizeof(TX) == 8) return D.r64.p; else return D.r32.p; };
    auto bb = [](PLevel& D) -> TX* { if constexpr (sizeof(TX) == 8) return D.b64.p; else return D.b32.p; };
    auto fb = [](PLevel& D) -> TX* { if constexpr (sizeof(TX) == 8) return D.f64.p; else return D.f32.p; };
    auto db = [](PLevel& D) -> const TX* { if constexpr (sizeof(TX) == 8) return D.d64.p; else return D.d32.p; };
    auto coarse = [&](PLevel& C, const TX* b, TX* f) {
        if (!C.coarse_wave.empty())
            hipLaunchKernelGGL(k_kcoarse_wave<TX>, dim3((unsigned)C.coarse_wave.size(), K_), dim3(64), 0, st, C.cw.p,
                               ks, b, f);
        if (!C.coarse_seq.empty()) {
            int nmax = 0;
            for (const auto& c : C.coarse_seq) nmax = std::max(nmax, c.n);
            hipLaunchKernelGGL(k_kcoarse_seq<TX>, dim3((unsigned)C.coarse_seq.size(), K_), dim3(64),
                               (size_t)nmax * sizeof(double), st, C.cs.p, ks, b, f);
        }
    };
    for (int l = 0; l < lmax_; ++l) {
        PLevel& D = L_[l];
        PLevel& C = L_[l + 1];
        const TX* b = l == 0 ? b0 : bb(D);
        const TX* d = db(D);
        const int* rs = D.rpt[pi];
        kpass<TX, TX>(D.nA, ks, D.A.rp.p, D.A.cv.p, vals<TX>(D.A.v64, D.A.v32), KXZero<TX>{d, b, w, ks},
                      KZeroJacobiResid<TX>{b, d, xb(D), rb(D), w, ks}, st, rs[0], D.A.cap);
        kpass<TX, TX>(C.nV, ks, D.Pt.rp.p, D.Pt.cv.p, vals<TX>(D.Pt.v64, D.Pt.v32), KX<TX>{rb(D), ks},
                      KStore<TX>{bb(C), ks}, st, rs[1], D.Pt.cap);
        coarse(C, bb(C), fb(C));
    }
    for (int l = lmax_ - 1; l >= 0; --l) {
        PLevel& D = L_[l];
        PLevel& C = L_[l + 1];
        const TX* b = l == 0 ? b0 : bb(D);
        TX* f = l == 0 ? f0 : fb(D);
        const int* rs = D.rpt[pi];
        kpass<TX, TX>(D.nA, ks, D.P.rp.p, D.P.cv.p, vals<TX>(D.P.v64, D.P.v32), KX<TX>{fb(C), ks},
                      KAdd<TX>{xb(D), ks}, st, rs[2], D.P.cap);
        kpass<TX, TX>(D.nA, ks, D.A.rp.p, D.A.cv.p, vals<TX>(D.A.v64, D.A.v32), KX<TX>{xb(D), ks},
                      KJacobi<TX>{xb(D), b, db(D), f, w, ks}, st, rs[3], D.A.cap);
    }
}

// Autotuner of the K-wide passes: for every level and each of its four passes
// (zero-guess step + residual, restriction, prolongation, Jacobi step), the
// right-hand sides per thread (1, 4, 8 or 16, at most K) that time fastest
// (best of 3 x 3 back-to-back launches on the level's real operators and
// workspaces, before any graph is captured).  Every choice gives the same
// values.  A forced choice (config batch_rhs_per_thread) skips the timing.
template <typename TX>
void BatchedHierarchy::tune() {
    const int pi = sizeof(TX) == 8 ? 0 : 1;
    if (tuned_[pi]) return;
    if (forced_rpt_ != 0) {  // negative: the LDS-staged kernel of that mapping
        for (auto& D : L_)
            for (int k = 0; k < 4; ++k)
                D.rpt[pi][k] = forced_rpt_ > 0 ? rpt_fit(forced_rpt_, K_) : -rpt_fit(-forced_rpt_, K_);
        tuned_[pi] = true;
        return;
    }
    DRAUGR_HIP_CHECK(hipDeviceSynchronize());
    hipStream_t st = nullptr;
    DRAUGR_HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    hipEvent_t e0, e1;
    DRAUGR_HIP_CHECK(hipEventCreate(&e0));
    DRAUGR_HIP_CHECK(hipEventCreate(&e1));
    auto timeit = [&](auto&& fn) {
        fn();
        float best = 1e30f;
        for (int rep = 0; rep < 3; ++rep) {
            DRAUGR_HIP_CHECK(hipEventRecord(e0, st));
            for (int k = 0; k < 3; ++k) fn();
            DRAUGR_HIP_CHECK(hipEventRecord(e1, st));
            DRAUGR_HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0.f;
            DRAUGR_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            best = std::min(best, ms);
        }
        return best;
    };
    const int ks = ks_;
    const TX w = (TX)omega_;
    auto xb = [](PLevel& D) -> TX* { if constexpr (sizeof(TX) == 8) return D.x64.p; else return D.x32.p; };
    auto rb = [](PLevel& D) -> TX* { if constexpr (sizeof(TX) == 8) return D.r64.p; else return D.r32.p; };
    auto bb = [](PLevel& D) -> TX* { if constexpr (sizeof(TX) == 8) return D.b64.p; else return D.b32.p; };
    auto fb = [](PLevel& D) -> TX* { if constexpr (sizeof(TX) == 8) return D.f64.p; else return D.f32.p; };
    auto db = [](PLevel& D) -> const TX* { if constexpr (sizeof(TX) == 8) return D.d64.p; else return D.d32.p; };
    const int cands[6] = {1, 4, 8, 16, -1, -4};  // negative: the LDS-staged kernel of that mapping
    for (int l = 0; l < lmax_; ++l) {
        PLevel& D = L_[l];
        PLevel& C = L_[l + 1];
        // finite inputs in the workspaces the passes read (level 0's right-hand
        // side and f are the level's own r and x here)
        TX* x = xb(D);
        TX* r = rb(D);
        TX* bc = bb(C);
        TX* fc = fb(C);
        const int64_t m = (int64_t)D.nV * K_, mc = (int64_t)C.nV * K_;
        kmap(m, [=] __device__(int64_t i) { x[i] = (TX)1; r[i] = (TX)1; }, st);
        kmap(mc, [=] __device__(int64_t i) { bc[i] = (TX)0; fc[i] = (TX)1; }, st);
        const TX* d = db(D);
        float best[4] = {1e30f, 1e30f, 1e30f, 1e30f};
        for (int c : cands) {
            if (std::abs(c) > K_) continue;
            const float t[4] = {
                timeit([&] {
                    kpass<TX, TX>(D.nA, ks, D.A.rp.p, D.A.cv.p, vals<TX>(D.A.v64, D.A.v32), KXZero<TX>{d, r, w, ks},
                                  KZeroJacobiResid<TX>{r, d, x, x, w, ks}, st, c, D.A.cap);
                }),
                timeit([&] {
                    kpass<TX, TX>(C.nV, ks, D.Pt.rp.p, D.Pt.cv.p, vals<TX>(D.Pt.v64, D.Pt.v32), KX<TX>{r, ks},
                                  KStore<TX>{bc, ks}, st, c, D.Pt.cap);
                }),
                timeit([&] {
                    kpass<TX, TX>(D.nA, ks, D.P.rp.p, D.P.cv.p, vals<TX>(D.P.v64, D.P.v32), KX<TX>{fc, ks},
                                  KAdd<TX>{x, ks}, st, c, D.P.cap);
                }),
                timeit([&] {
                    kpass<TX, TX>(D.nA, ks, D.A.rp.p, D.A.cv.p, vals<TX>(D.A.v64, D.A.v32), KX<TX>{r, ks},
                                  KJacobi<TX>{r, r, d, x, w, ks}, st, c, D.A.cap);
                })};
            for (int k = 0; k < 4; ++k)
                if (t[k] < best[k]) {
                    best[k] = t[k];
                    D.rpt[pi][k] = c;
                    D.rpt_us[pi][k] = t[k] * 1e3f / 3.0f;
                }
        }
    }
    DRAUGR_HIP_CHECK(hipStreamSynchronize(st));
    DRAUGR_HIP_CHECK(hipEventDestroy(e0));
    DRAUGR_HIP_CHECK(hipEventDestroy(e1));
    DRAUGR_HIP_CHECK(hipStreamDestroy(st));
    tuned_[pi] = true;
}

// One K-wide Jacobi / l1-Jacobi step on packed level l with the tuned mapping
// (a measurement and test hook): out = x + w d (b - A x) for the nA rows of
// the systems that smooth on level l, vectors n x K row-major.
void BatchedHierarchy::jacobi_pass(int l, Precision p, const void* x, const void* b, void* out, hipStream_t st) {
    if (l < 0 || l >= lmax_) throw Error("batched level index out of range");
    if (p == Precision::BF16) throw Error("the batched path supports precision fp64 and fp32");
    PLevel& D = L_[l];
    const int ks = ks_;
    if (p == Precision::FP64) {
        ensure_vectors<double>();
        tune<double>();
        kpass<double, double>(D.nA, ks, D.A.rp.p, D.A.cv.p, D.A.v64.p, KX<double>{(const double*)x, ks},
                              KJacobi<double>{(const double*)x, (const double*)b, D.d64.p, (double*)out,
                                              (double)omega_, ks},
                              st, D.rpt[0][3], D.A.cap);
    } else {
        ensure_vectors<float>();
        tune<float>();
        kpass<float, float>(D.nA, ks, D.A.rp.p, D.A.cv.p, D.A.v32.p, KX<float>{(const float*)x, ks},
                            KJacobi<float>{(const float*)x, (const float*)b, D.d32.p, (float*)out, (float)omega_, ks},
                            st, D.rpt[1][3], D.A.cap);
    }
}

void BatchedHierarchy::level_counts(int l, int64_t* out) const {
    if (l < 0 || l >= lmax_) throw Error("batched level index out of range");
    const PLevel& D = L_[l];
    out[0] = D.nA;
    out[1] = D.nV;
    out[2] = D.A.nnz;
    out[3] = D.P.nnz;
    out[4] = D.Pt.nrow;
}

std::string BatchedHierarchy::tune_report() const {
    std::string out;
    char buf[200];
    const char* pn[2] = {"fp64", "fp32"};
    for (int p = 0; p < 2; ++p) {
        if (!tuned_[p]) continue;
        for (int l = 0; l < lmax_; ++l) {
            const PLevel& D = L_[l];
            auto m = [&](int k) {  // "4" or "4s" (LDS-staged)
                return std::to_string(std::abs(D.rpt[p][k])) + (D.rpt[p][k] < 0 ? "s" : "");
            };
            std::snprintf(buf, sizeof buf,
                          "%s level %d (%d rows x %d rhs): rhs/thread zero-step %s, restrict %s, prolong %s, "
                          "jacobi %s (%.1f / %.1f / %.1f / %.1f us)\n",
                          pn[p], l, D.nA, K_, m(0).c_str(), m(1).c_str(), m(2).c_str(), m(3).c_str(),
                          D.rpt_us[p][0], D.rpt_us[p][1], D.rpt_us[p][2], D.rpt_us[p][3]);
            out += buf;
        }
    }
    return out;
}

template <typename TX>
void BatchedHierarchy::precondition(TX* f0, const TX* b0, hipStream_t st, bool graphs) {
    ensure_vectors<TX>();
    tune<TX>();
    if (!graphs) {
        cycle_t<TX>(f0, b0, st);
        return;
    }
    const auto key = std::make_tuple((int)sizeof(TX), 0, (const void*)f0, (const void*)b0);
    auto it = graphs_.find(key);
    if (it == graphs_.end()) {
        if (graphs_.size() >= 32) {
            DRAUGR_HIP_CHECK(hipDeviceSynchronize());
            for (auto& kv : graphs_) (void)hipGraphExecDestroy(kv.second);
            graphs_.clear();
        }
        if (!cap_) DRAUGR_HIP_CHECK(hipStreamCreateWithFlags(&cap_, hipStreamNonBlocking));
        hipGraph_t g = nullptr;
        DRAUGR_HIP_CHECK(hipStreamBeginCapture(cap_, hipStreamCaptureModeThreadLocal));
        try {
            cycle_t<TX>(f0, b0, cap_);
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

void BatchedHierarchy::cycle(double* x, const double* b, const CycleOpts& o, hipStream_t st) {
    check_opts(o);
    if (o.prec == Precision::FP64) {
        precondition<double>(x, b, st, o.graphs);
        return;
    }
    const int64_t m = N0_ * K_;
    if (!w_r32_.p) { w_r32_.alloc(m); w_z32_.alloc(m); }
    float* r32 = w_r32_.p;
    float* z32 = w_z32_.p;
    kmap(m, [=] __device__(int64_t i) { r32[i] = (float)b[i]; }, st);
    precondition<float>(z32, r32, st, o.graphs);
    kmap(m, [=] __device__(int64_t i) { x[i] = (double)z32[i]; }, st);
}

// FP64 PCG (Hestenes-Stiefel) for every (system, rhs) at once, each stopping
// on its own recurrence residual ||r||/||b|| < tol (its step sizes are zeroed
// from then on, so its x and r stay frozen).  One iteration is two graph
// launches: [q = A p, alpha, x/r update, r.r, active count] and [z = V(0; r),
// beta]; the host reads the active count of iteration k while the cycle of
// iteration k runs.
int BatchedHierarchy::pcg(double* x, const double* b, double tol, int maxiter, const CycleOpts& o, hipStream_t st,
                          int32_t* iters, double* relres) {
    check_opts(o);
    const bool mixed = o.prec != Precision::FP64;
    const int S = nsys(), SK = S * K_;
    const int64_t m = N0_ * K_;
    if (!w_r_.p) {
        w_r_.alloc(m); w_p0_.alloc(m); w_p1_.alloc(m); w_q_.alloc(m);
        part_.alloc((size_t)nblk0_ * K_);
        st_bb_.alloc(SK); st_rr_.alloc(SK); st_rz_.alloc(SK); st_pq_.alloc(SK); st_alpha_.alloc(SK);
        st_beta_.alloc(SK); st_active_.alloc(SK); st_iters_.alloc(SK); st_count_.alloc(1); st_tol_.alloc(1);
        DRAUGR_HIP_CHECK(hipHostMalloc((void**)&pinned_, sizeof(int32_t) * 4, hipHostMallocDefault));
    }
    if (mixed && !w_r32_.p) { w_r32_.alloc(m); w_z32_.alloc(m); }
    if (!mixed && !w_z_.p) w_z_.alloc(m);
    const int ks = ks_;
    double* r = w_r_.p;
    double* q = w_q_.p;
    double* pa = w_p0_.p;
    double* pb = w_p1_.p;
    float* r32 = w_r32_.p;
    float* z32 = w_z32_.p;
    double* z = w_z_.p;
    double* part = part_.p;
    {
        double* tp = st_tol_.p;
        kmap(1, [=] __device__(int64_t) { tp[0] = tol; }, st);
    }
    KState S_{st_bb_.p, st_rr_.p, st_rz_.p, st_pq_.p, st_alpha_.p, st_beta_.p, st_active_.p, st_iters_.p, st_tol_.p};
    const PLevel& D0 = L_[0];
    const int nb = nblk0_;
    auto fin = [&](int mode, hipStream_t s) {
        hipLaunchKernelGGL(k_kfinal, dim3(S), dim3(kBlock), 0, s, mode, ks, sysb_.p, part, S_);
    };
    // bb, zero x where b = 0, r = b - A x, rr, active flags
    hipLaunchKernelGGL(k_kdot<double>, dim3(nb), dim3(kBlock), 0, st, ks, br0_.p, br1_.p, b, b, part);
    fin(F_BB, st);
    hipLaunchKernelGGL(k_kzero_bb, dim3(nb), dim3(kBlock), 0, st, ks, br0_.p, br1_.p, bs_.p, st_bb_.p, x);
    if (mixed)
        hipLaunchKernelGGL(k_kresid_rr<true>, dim3(nb), dim3(kBlock), 0, st, ks, br0_.p, br1_.p, D0.A.rp.p,
                           D0.A.cv.p, D0.A.v64.p, x, b, r, r32, part);
    else
        hipLaunchKernelGGL(k_kresid_rr<false>, dim3(nb), dim3(kBlock), 0, st, ks, br0_.p, br1_.p, D0.A.rp.p,
                           D0.A.cv.p, D0.A.v64.p, x, b, r, r32, part);
    fin(F_RR0, st);
    hipLaunchKernelGGL(k_kcount, dim3(1), dim3(kBlock), 0, st, SK, st_active_.p, st_count_.p);
    DRAUGR_HIP_CHECK(hipMemcpyAsync(pinned_, st_count_.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    DRAUGR_HIP_CHECK(hipStreamSynchronize(st));
    if (pinned_[0] > 0) {
        kmap(m, [=] __device__(int64_t i) { pa[i] = 0.0; }, st);
        auto apply = [&](hipStream_t s) {
            if (mixed) precondition<float>(z32, r32, s, false);
            else precondition<double>(z, r, s, false);
        };
        auto rz = [&](int mode, hipStream_t s) {
            if (mixed)
                hipLaunchKernelGGL(k_kdot<float>, dim3(nb), dim3(kBlock), 0, s, ks, br0_.p, br1_.p, r, z32, part);
            else
                hipLaunchKernelGGL(k_kdot<double>, dim3(nb), dim3(kBlock), 0, s, ks, br0_.p, br1_.p, r, z, part);
            fin(mode, s);
        };
        if (mixed) {  // the autotuner synchronizes: it runs before any capture
            ensure_vectors<float>();
            tune<float>();
        } else {
            ensure_vectors<double>();
            tune<double>();
        }
        apply(st);
        rz(F_RZ0, st);
        // the two graphs of an iteration (the first in two forms: p ping-pongs)
        auto step1 = [&](double* pold, double* pnew, hipStream_t s) {
            const size_t lds = pq_cap_ >= 0 ? ((size_t)pq_cap_ * 8 + 15) / 16 * 16 + (size_t)pq_cap_ * 4 : 0;
            const bool stage = pq_cap_ >= 0 && lds <= (size_t)kStageBytes;
            if (mixed && ks >= 2 && stage)
                hipLaunchKernelGGL((k_kpq4<float, true>), dim3(nb), dim3(kBlock), lds, s, ks, br0_.p, br1_.p, bs_.p,
                                   D0.A.rp.p, D0.A.cv.p, D0.A.v64.p, z32, pold, pnew, q, st_beta_.p, part, pq_cap_);
            else if (mixed && ks >= 2)
                hipLaunchKernelGGL((k_kpq4<float, false>), dim3(nb), dim3(kBlock), 0, s, ks, br0_.p, br1_.p, bs_.p,
                                   D0.A.rp.p, D0.A.cv.p, D0.A.v64.p, z32, pold, pnew, q, st_beta_.p, part, 0);
            else if (mixed)
                hipLaunchKernelGGL(k_kpq<float>, dim3(nb), dim3(kBlock), 0, s, ks, br0_.p, br1_.p, bs_.p, D0.A.rp.p,
                                   D0.A.cv.p, D0.A.v64.p, z32, pold, pnew, q, st_beta_.p, part);
            else if (ks >= 2 && stage)
                hipLaunchKernelGGL((k_kpq4<double, true>), dim3(nb), dim3(kBlock), lds, s, ks, br0_.p, br1_.p, bs_.p,
                                   D0.A.rp.p, D0.A.cv.p, D0.A.v64.p, z, pold, pnew, q, st_beta_.p, part, pq_cap_);
            else if (ks >= 2)
                hipLaunchKernelGGL((k_kpq4<double, false>), dim3(nb), dim3(kBlock), 0, s, ks, br0_.p, br1_.p, bs_.p,
                                   D0.A.rp.p, D0.A.cv.p, D0.A.v64.p, z, pold, pnew, q, st_beta_.p, part, 0);
            else
                hipLaunchKernelGGL(k_kpq<double>, dim3(nb), dim3(kBlock), 0, s, ks, br0_.p, br1_.p, bs_.p,
                                   D0.A.rp.p, D0.A.cv.p, D0.A.v64.p, z, pold, pnew, q, st_beta_.p, part);
            fin(F_PQ, s);
            if (mixed)
                hipLaunchKernelGGL(k_kupdate<true>, dim3(nb), dim3(kBlock), 0, s, ks, br0_.p, br1_.p, bs_.p,
                                   st_alpha_.p, x, pnew, r, q, r32, part);
            else
                hipLaunchKernelGGL(k_kupdate<false>, dim3(nb), dim3(kBlock), 0, s, ks, br0_.p, br1_.p, bs_.p,
                                   st_alpha_.p, x, pnew, r, q, r32, part);
            fin(F_RR, s);
            hipLaunchKernelGGL(k_kcount, dim3(1), dim3(kBlock), 0, s, SK, st_active_.p, st_count_.p);
        };
        auto step2 = [&](hipStream_t s) {
            apply(s);
            rz(F_RZ, s);
        };
        auto graph = [&](int tag, const void* a, const void* bp, auto&& body) -> hipGraphExec_t {
            const auto key = std::make_tuple(tag + (mixed ? 10 : 0), 1, a, bp);
            auto it = graphs_.find(key);
            if (it != graphs_.end()) return it->second;
            if (!cap_) DRAUGR_HIP_CHECK(hipStreamCreateWithFlags(&cap_, hipStreamNonBlocking));
            hipGraph_t g = nullptr;
            DRAUGR_HIP_CHECK(hipStreamBeginCapture(cap_, hipStreamCaptureModeThreadLocal));
            try {
                body(cap_);
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
            graphs_.emplace(key, ex);
            return ex;
        };
        hipGraphExec_t g1a = nullptr, g1b = nullptr, g2 = nullptr;
        if (o.graphs) {
            if (graphs_.size() >= 29) {
                DRAUGR_HIP_CHECK(hipDeviceSynchronize());
                for (auto& kv : graphs_) (void)hipGraphExecDestroy(kv.second);
                graphs_.clear();
            }
            g1a = graph(1, x, w_p0_.p, [&](hipStream_t s) { step1(w_p0_.p, w_p1_.p, s); });
            g1b = graph(2, x, w_p1_.p, [&](hipStream_t s) { step1(w_p1_.p, w_p0_.p, s); });
            g2 = graph(3, nullptr, nullptr, [&](hipStream_t s) { step2(s); });
        }
        hipEvent_t ev;
        DRAUGR_HIP_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        try {
            for (int it = 1; it <= maxiter; ++it) {
                const bool even = (it % 2) == 1;  // iteration 1 reads p0, writes p1
                if (o.graphs) DRAUGR_HIP_CHECK(hipGraphLaunch(even ? g1a : g1b, st));
                else step1(even ? w_p0_.p : w_p1_.p, even ? w_p1_.p : w_p0_.p, st);
                DRAUGR_HIP_CHECK(hipMemcpyAsync(pinned_, st_count_.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
                DRAUGR_HIP_CHECK(hipEventRecord(ev, st));
                if (o.graphs) DRAUGR_HIP_CHECK(hipGraphLaunch(g2, st));
                else step2(st);
                DRAUGR_HIP_CHECK(hipEventSynchronize(ev));
                if (pinned_[0] == 0) break;
            }
        } catch (...) {
            (void)hipStreamSynchronize(st);
            (void)hipEventDestroy(ev);
            throw;
        }
        DRAUGR_HIP_CHECK(hipEventDestroy(ev));
    }
    DRAUGR_HIP_CHECK(hipStreamSynchronize(st));
    std::vector<double> rr(SK), bbv(SK);
    std::vector<int32_t> itv(SK);
    DRAUGR_HIP_CHECK(hipMemcpy(rr.data(), st_rr_.p, SK * sizeof(double), hipMemcpyDeviceToHost));
    DRAUGR_HIP_CHECK(hipMemcpy(bbv.data(), st_bb_.p, SK * sizeof(double), hipMemcpyDeviceToHost));
    DRAUGR_HIP_CHECK(hipMemcpy(itv.data(), st_iters_.p, SK * sizeof(int32_t), hipMemcpyDeviceToHost));
    int mx = 0;
    for (int t = 0; t < SK; ++t) {
        if (iters) iters[t] = itv[t];
        if (relres) relres[t] = bbv[t] == 0.0 ? 0.0 : std::sqrt(rr[t] / bbv[t]);
        mx = std::max(mx, itv[t]);
    }
    return mx;
}

}  // namespace draugr
