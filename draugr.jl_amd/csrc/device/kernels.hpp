// HIP kernels for gfx950 (CDNA4) of the AMG cycle: every operator pass
// (SpMV, residual, smoother step, restriction, prolongation) is one
// CSR-stream launch with a fused per-row epilogue.
//
// CSR-stream: the rows of an operator are cut, once at upload, into row
// blocks of at most kStreamRows rows and kStreamNnz nonzeros.  A 256-thread
// workgroup streams its block's contiguous nonzero range: every thread forms
// products val[k] * x[col[k]] for k = tid, tid + 256, ... (coalesced loads of
// col/val, several independent gathers in flight per thread) into LDS; then
// one thread per row walks its row's products in row order and applies the
// epilogue.  The row sums are therefore formed in exactly the reference's
// sequential order (products rounded, then added left to right; device code is
// built with -ffp-contract=off), so FP64 results equal the reference's bit for
// bit.  A row longer than kStreamNnz gets a block of its own and is reduced by
// the whole workgroup (tree order).  Blocks are remapped so that each of the 8
// XCDs sweeps a contiguous slice of rows: neighbouring rows share x entries
// through that XCD's L2.
//
// Value storage TV is double, float or bf16 (uint16_t); vectors and
// accumulation TX are double (TV = double) or float (TV = float or bf16).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <vector>

namespace draugr {
namespace dev {

constexpr int kBlock = 256;
constexpr int kStreamNnz = 4096;    // products per row block at most (16 per thread)
constexpr int kStreamRows = kBlock; // one epilogue thread per row
// Column-sorted form: one 32-bit word per entry, the column offset from the
// block's base in the low kSortColBits bits and the entry's position in the
// block's row order (< kStreamNnz = 2^12) above them.
constexpr int kSortColBits = 20;
constexpr uint32_t kSortColMask = (1u << kSortColBits) - 1u;

typedef uint16_t bf16_t;

__device__ __forceinline__ float bf16_to_float(bf16_t v) { return __uint_as_float((uint32_t)v << 16); }
__host__ __device__ inline bf16_t float_to_bf16(float f) {
    // round to nearest even; NaN stays NaN
    uint32_t u;
#ifdef __HIP_DEVICE_COMPILE__
    u = __float_as_uint(f);
#else
    __builtin_memcpy(&u, &f, 4);
#endif
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (bf16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}

template <typename TX, typename TV> __device__ __forceinline__ TX ld(const TV* p, int64_t k);
template <> __device__ __forceinline__ double ld<double, double>(const double* p, int64_t k) { return __builtin_nontemporal_load(p + k); }
template <> __device__ __forceinline__ float ld<float, float>(const float* p, int64_t k) { return __builtin_nontemporal_load(p + k); }
template <> __device__ __forceinline__ float ld<float, bf16_t>(const bf16_t* p, int64_t k) {
    return bf16_to_float(__builtin_nontemporal_load(p + k));
}
__device__ __forceinline__ int ldi(const int* p, int64_t k) { return __builtin_nontemporal_load(p + k); }
__device__ __forceinline__ int ldc16(const uint16_t* p, int64_t k) { return (int)__builtin_nontemporal_load(p + k); }
__device__ __forceinline__ uint32_t ldu(const uint32_t* p, int64_t k) { return __builtin_nontemporal_load(p + k); }

// XCD-contiguous remap of the block index (bijective for any grid size).
__device__ __forceinline__ int xcd_block() {
    const int b = blockIdx.x, nb = gridDim.x;
    const int q = nb >> 3, r = nb & 7, x = b & 7, t = b >> 3;
    return x < r ? x * (q + 1) + t : r * (q + 1) + (x - r) * q + t;
}

// ── x functors: the vector entry a product multiplies ───────────────────────
template <typename TX> struct XVec {
    const TX* x;
    __device__ TX operator()(int j) const { return x[j]; }
};
// x_j = w * dinv_j * b_j: the first Jacobi step from a zero guess, evaluated
// on the fly so that its residual needs no stored x.
template <typename TX> struct XZeroJacobi {
    const TX* d; const TX* b; TX w;
    __device__ TX operator()(int j) const { return (w * d[j]) * b[j]; }
};

// ── epilogues.  A row's products p_k arrive in row order; every epilogue is
//    q = pre(i) (the row's own operands, loaded before the products are
//    staged so that their latency overlaps the gathers); a = init(q);
//    a = acc(a, p_k) for each k; fin(i, a, q) — the reference's own
//    accumulation for that operation. ─────────────────────────────────────────
struct NoPre {};
template <typename TX> struct EpiStore {          // y = A x   (src/csr_matrix.jl:150-159)
    TX* y;
    typedef NoPre Pre;
    __device__ Pre pre(int) const { return {}; }
    __device__ TX init(const Pre&) const { return 0; }
    __device__ TX acc(TX a, TX p) const { return a + p; }
    __device__ void fin(int i, TX a, const Pre&) const { y[i] = a; }
};
// Restriction that also forms the coarse level's first Jacobi iterate from a
// zero guess: bc = P^T r and x1 = (w dinv_c) bc, the value the coarse level's
// own first step would compute from bc.
template <typename TX> struct EpiStoreJac0 {
    TX* y; const TX* d; TX* x1; TX w;
    struct Pre { TX d; };
    __device__ Pre pre(int i) const { return {d[i]}; }
    __device__ TX init(const Pre&) const { return 0; }
    __device__ TX acc(TX a, TX p) const { return a + p; }
    __device__ void fin(int i, TX a, const Pre& q) const {
        y[i] = a;
        x1[i] = (w * q.d) * a;
    }
};
// Prolongation folded into the following smoothing pass: the prolonged
// iterate xp_k = x_k + sum_J P_kJ xc_J (EpiAdd's order) formed on the fly for
// every column k the pass gathers, and for the row itself in EpiJacobiP.
template <typename TP, typename TX> struct XProlong {
    const TX* x; const int* prp; const int* pcv; const TP* pv; const TX* xc;
    __device__ TX operator()(int k) const {
        TX a = x[k];
        const int e = prp[k + 1];
        for (int q = prp[k]; q < e; ++q) a = a + ld<TX, TP>(pv, q) * xc[ldi(pcv, q)];
        return a;
    }
};
// Jacobi step from the prolonged iterate: xo = xp + (w d)(b - A xp).
template <typename TP, typename TX> struct EpiJacobiP {
    XProlong<TP, TX> xp; const TX* b; const TX* d; TX* xo; TX w;
    struct Pre { TX x, b, d; };
    __device__ Pre pre(int i) const { return {xp(i), b[i], d[i]}; }
    __device__ TX init(const Pre& q) const { return q.b; }
    __device__ TX acc(TX a, TX p) const { return a - p; }
    __device__ void fin(int i, TX a, const Pre& q) const { xo[i] = q.x + (w * q.d) * a; }
};
template <typename TX> struct EpiResid {          // r = b - A x   (src/cycle.jl:32-43)
    const TX* b; TX* r;
    struct Pre { TX b; };
    __device__ Pre pre(int i) const { return {b[i]}; }
    __device__ TX init(const Pre&) const { return 0; }
    __device__ TX acc(TX a, TX p) const { return a + p; }
    __device__ void fin(int i, TX a, const Pre& q) const { r[i] = q.b - a; }
};
template <typename TX> struct EpiJacobi {         // xo = x + w d (b - A x)   (src/smoothers.jl:78-94)
    const TX* x; const TX* b; const TX* d; TX* xo; TX w;
    struct Pre { TX x, b, d; };
    __device__ Pre pre(int i) const { return {x[i], b[i], d[i]}; }
    __device__ TX init(const Pre& q) const { return q.b; }  // residual by successive subtraction
    __device__ TX acc(TX a, TX p) const { return a - p; }
    __device__ void fin(int i, TX a, const Pre& q) const { xo[i] = q.x + (w * q.d) * a; }
};
template <typename TX> struct EpiZeroJacobiResid {  // x1 = w d b, r = b - A x1
    const TX* b; const TX* d; TX* x1; TX* r; TX w;
    struct Pre { TX b, d; };
    __device__ Pre pre(int i) const { return {b[i], d[i]}; }
    __device__ TX init(const Pre&) const { return 0; }
    __device__ TX acc(TX a, TX p) const { return a + p; }
    __device__ void fin(int i, TX a, const Pre& q) const {
        x1[i] = (w * q.d) * q.b;
        r[i] = q.b - a;
    }
};
template <typename TX> struct EpiAdd {            // x += P xc, one product at a time (src/prolongation.jl:750-759)
    TX* x;
    struct Pre { TX x; };
    __device__ Pre pre(int i) const { return {x[i]}; }
    __device__ TX init(const Pre& q) const { return q.x; }
    __device__ TX acc(TX a, TX p) const { return a + p; }
    __device__ void fin(int i, TX a, const Pre&) const { x[i] = a; }
};

// Chebyshev smoother passes (src/smoothers.jl:925-990): with r = b - A x
// (row sum, then subtraction), the first pass of a step sets
// d = (dinv r) / theta, the others d = (c_r dinv) r + c_d d; both write
// x_out = x + d.  Out of place (the caller swaps x and x_out).
template <typename TX> struct EpiCheb0 {
    const TX* x; const TX* b; const TX* dv; TX* d; TX* xo; TX inv_theta;
    struct Pre { TX x, b, dv; };
    __device__ Pre pre(int i) const { return {x[i], b[i], dv[i]}; }
    __device__ TX init(const Pre&) const { return 0; }
    __device__ TX acc(TX a, TX p) const { return a + p; }
    __device__ void fin(int i, TX a, const Pre& q) const {
        const TX dn = q.dv * (q.b - a) * inv_theta;
        d[i] = dn;
        xo[i] = q.x + dn;
    }
};
template <typename TX> struct EpiChebK {
    const TX* x; const TX* b; const TX* dv; TX* d; TX* xo; TX cr, cd;
    struct Pre { TX x, b, dv, d; };
    __device__ Pre pre(int i) const { return {x[i], b[i], dv[i], d[i]}; }
    __device__ TX init(const Pre&) const { return 0; }
    __device__ TX acc(TX a, TX p) const { return a + p; }
    __device__ void fin(int i, TX a, const Pre& q) const {
        const TX dn = cr * q.dv * (q.b - a) + cd * q.d;
        d[i] = dn;
        xo[i] = q.x + dn;
    }
};
// w = (A v) .* dinv: one power-iteration step of src/smoothers.jl:866-885
template <typename TX> struct EpiStoreScaled {
    TX* y; const TX* dv;
    struct Pre { TX dv; };
    __device__ Pre pre(int i) const { return {dv[i]}; }
    __device__ TX init(const Pre&) const { return 0; }
    __device__ TX acc(TX a, TX p) const { return a + p; }
    __device__ void fin(int i, TX a, const Pre& q) const { y[i] = a * q.dv; }
};

// Row-block header of the CSR-stream kernels: first row, first nonzero,
// 16-bit column base (-1: the block reads the 32-bit columns) and end row of
// block b; entry nblk closes the table (its z0 is the end nonzero of the last
// block).  One 16-byte record per block, so a block's bounds and base arrive
// with one scalar load round trip (separate arrays cost a second, dependent
// one for the base).
struct alignas(16) BlkHdr {
    int32_t r0, z0, cb, r1;
};

// One row block: stage the products, then finish each row on its own thread
// with the epilogue's accumulation a = initf(q); a = accf(a, p_k) in row
// order; finf(row, a, q), q = pref(row).  The block header holds both row and
// nonzero offsets, so a block's range is known after one round of loads.  A
// block of at most PER * kBlock nonzeros (PER = the operator's budget /
// kBlock) is loaded without branches: every thread issues its PER column and
// value loads together (indices past the block's end are clamped to its last
// entry and the result discarded), then its PER gathers together, so each
// wave keeps 2 PER + PER loads in flight instead of waiting once per product.
// A row longer than the LDS stage (alone in its block) is staged kStreamNnz
// products at a time and summed by one thread in row order, so every row sum,
// long or short, is the reference's sequential one.
//
// C16: such blocks read their columns as 16-bit offsets from a per-block base
// (col = cb + c16[k]), 2 bytes per nonzero instead of 4, when the block spans
// fewer than 65536 columns (cb >= 0); wider and longer blocks read the 32-bit
// columns.
//
// SRT (launched with the column-sorted form's block headers, whose base cb is
// set for the branch-free blocks spanning fewer than 2^kSortColBits columns):
// the entries of each such block are stored sorted by column (av in that
// order) with one 32-bit word per entry, cs[k] = the column offset | the
// entry's position in the block's row order (where its product is staged)
// << kSortColBits, so column and position arrive with one load.  A wave's
// gathers then touch a few adjacent cache lines instead of the lines of every
// column cluster of its rows; the staged products, and so the row sums, are
// the same.  Other blocks keep row order (av in row order, columns from cv).
template <int PER, bool C16, typename TV, typename TX, class XF, class PREF, class INITF, class ACCF, class FINF,
          bool SRT = false>
__device__ __forceinline__ void stream_block(int b, const BlkHdr* __restrict__ hdr, const int* __restrict__ rp,
                                             const int* __restrict__ cv, const uint16_t* __restrict__ c16,
                                             const TV* __restrict__ av, XF xf, TX* prod, PREF pref, INITF initf,
                                             ACCF accf, FINF finf, const uint32_t* __restrict__ cs = nullptr) {
    const BlkHdr h0 = hdr[b];
    const int r0 = h0.r0, r1 = h0.r1;
    const int z0 = h0.z0, z1 = hdr[b + 1].z0;
    const int cb = C16 ? h0.cb : 0;
    const int nz = z1 - z0;
    const int t = (int)threadIdx.x;
    const int row = r0 + t;
    constexpr int S = PER * kBlock;  // the block's LDS stage
    if (nz <= S) {
        const bool own = row < r1;
        const int rs = own ? rp[row] : 0, re = own ? rp[row + 1] : 0;  // issued before the barrier
        decltype(pref(0)) q{};
        if (own) q = pref(row);
        {
            if (nz > 0) {
                int cj[PER], sl[PER];
                TX v[PER], g[PER];
                if (SRT && cb >= 0) {  // block-uniform choices, outside the unrolled loads
#pragma unroll
                    for (int u = 0; u < PER; ++u) {
                        const int k = min(t + u * kBlock, nz - 1);
                        const uint32_t e = ldu(cs, z0 + k);
                        cj[u] = cb + (int)(e & kSortColMask);
                        v[u] = ld<TX, TV>(av, z0 + k);
                        sl[u] = (int)(e >> kSortColBits);
                    }
                } else if (C16 && cb >= 0) {
#pragma unroll
                    for (int u = 0; u < PER; ++u) {
                        const int k = min(t + u * kBlock, nz - 1);
                        cj[u] = cb + ldc16(c16, z0 + k);
                        v[u] = ld<TX, TV>(av, z0 + k);
                        sl[u] = t + u * kBlock;
                    }
                } else {
#pragma unroll
                    for (int u = 0; u < PER; ++u) {
                        const int k = min(t + u * kBlock, nz - 1);
                        cj[u] = ldi(cv, z0 + k);
                        v[u] = ld<TX, TV>(av, z0 + k);
                        sl[u] = t + u * kBlock;
                    }
                }
#pragma unroll
                for (int u = 0; u < PER; ++u) g[u] = xf(cj[u]);
#pragma unroll
                for (int u = 0; u < PER; ++u)
                    if (t + u * kBlock < nz) prod[sl[u]] = v[u] * g[u];
            }
        }
        __syncthreads();
        if (own) {
            const TX* seg = prod + (rs - z0);
            auto a = initf(q);
            for (int k = 0; k < re - rs; ++k) a = accf(a, seg[k]);
            finf(row, a, q);
        }
    } else {  // one row longer than the LDS stage: chunk by chunk, summed in row order by thread 0
        decltype(pref(0)) q{};
        decltype(initf(q)) a{};
        if (t == 0) {
            q = pref(r0);
            a = initf(q);
        }
        for (int c0 = 0; c0 < nz; c0 += S) {
            const int cn = min(S, nz - c0);
            for (int k = t; k < cn; k += kBlock) prod[k] = ld<TX, TV>(av, z0 + c0 + k) * xf(ldi(cv, z0 + c0 + k));
            __syncthreads();
            if (t == 0)
                for (int k = 0; k < cn; ++k) a = accf(a, prod[k]);
            __syncthreads();
        }
        if (t == 0) finf(r0, a, q);
    }
}

template <class EPI, typename TX>
__device__ __forceinline__ void epi_segment(const EPI& epi, int row, const TX* seg, int len,
                                            const typename EPI::Pre& q) {
    TX a = epi.init(q);
    for (int k = 0; k < len; ++k) a = epi.acc(a, seg[k]);
    epi.fin(row, a, q);
}

template <bool XCD, int PER, bool C16, typename TV, typename TX, class XF, class EPI, bool SRT = false>
__global__ __launch_bounds__(kBlock) void k_csr_stream(const BlkHdr* __restrict__ hdr, const int* __restrict__ rp,
                                                       const int* __restrict__ cv, const uint16_t* __restrict__ c16,
                                                       const TV* __restrict__ av, XF xf, EPI epi,
                                                       const uint32_t* __restrict__ cs) {
    __shared__ TX prod[PER * kBlock];
    const int b = XCD ? xcd_block() : (int)blockIdx.x;
    auto pref = [&](int row) { return epi.pre(row); };
    auto initf = [&](const typename EPI::Pre& q) { return epi.init(q); };
    auto accf = [&](TX a, TX p) { return epi.acc(a, p); };
    auto finf = [&](int row, TX a, const typename EPI::Pre& q) { epi.fin(row, a, q); };
    stream_block<PER, C16, TV, TX, XF, decltype(pref), decltype(initf), decltype(accf), decltype(finf), SRT>(
        b, hdr, rp, cv, c16, av, xf, prod, pref, initf, accf, finf, cs);
}

// Four consecutive entries k0 .. k0 + 3 (k0 a multiple of 4) of a value array
// of the column-sorted form, in one or two 16-byte loads (those arrays are
// allocated to a multiple of 4 entries, so the load stays inside them).
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef uint16_t u16x4 __attribute__((ext_vector_type(4)));
template <typename TX, typename TV> struct Quad;
template <> struct Quad<double, double> {
    static __device__ __forceinline__ void load(const double* p, int k0, double (&o)[4]) {
        const f64x2 a = __builtin_nontemporal_load(reinterpret_cast<const f64x2*>(p + k0));
        const f64x2 b = __builtin_nontemporal_load(reinterpret_cast<const f64x2*>(p + k0) + 1);
        o[0] = a.x; o[1] = a.y; o[2] = b.x; o[3] = b.y;
    }
};
template <> struct Quad<float, float> {
    static __device__ __forceinline__ void load(const float* p, int k0, float (&o)[4]) {
        const f32x4 a = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(p + k0));
        o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w;
    }
};
template <> struct Quad<float, bf16_t> {
    static __device__ __forceinline__ void load(const bf16_t* p, int k0, float (&o)[4]) {
        const u16x4 a = __builtin_nontemporal_load(reinterpret_cast<const u16x4*>(p + k0));
        o[0] = bf16_to_float(a.x); o[1] = bf16_to_float(a.y); o[2] = bf16_to_float(a.z); o[3] = bf16_to_float(a.w);
    }
};

// One row block of the column-sorted form with four entries per load: the
// block's nonzero range [z0, z1) is covered by the aligned quads starting at
// z0 & ~3; thread t loads quads t, t + kBlock, ... (one 16-byte load of four
// packed column / position words, one or two of four values), gathers the
// four x entries and stages the products of the entries inside [z0, z1) at
// their positions; entries of the neighbouring blocks that share the first or
// last quad are loaded and dropped (their gathers read the block's base
// column instead).  Per product that is 1/4 column load, 1/4-1/2 value load
// and one gather, against one of each in stream_block.  The staged products
// and the row walk are stream_block's, so the values are the same bits.
// Blocks without a 16-bit base, and the rare block whose quads exceed the
// stage (nz within 3 of it and misaligned), take stream_block's path.
template <int PER, typename TV, typename TX, class XF, class PREF, class INITF, class ACCF, class FINF>
__device__ __forceinline__ void stream_block_quad(int b, const BlkHdr* __restrict__ hdr, const int* __restrict__ rp,
                                                  const int* __restrict__ cv, const uint32_t* __restrict__ cs,
                                                  const TV* __restrict__ av, XF xf, TX* prod, PREF pref, INITF initf,
                                                  ACCF accf, FINF finf) {
    static_assert(PER % 4 == 0, "four entries per load");
    constexpr int G = PER / 4;
    const BlkHdr h0 = hdr[b];
    const int z0 = h0.z0, z1 = hdr[b + 1].z0, cb = h0.cb;
    const int a0 = z0 & ~3;
    const int nq = (z1 - a0 + 3) >> 2;
    if (cb < 0 || nq > G * kBlock) {
        stream_block<PER, true, TV, TX, XF, PREF, INITF, ACCF, FINF, true>(b, hdr, rp, cv, nullptr, av, xf, prod,
                                                                          pref, initf, accf, finf, cs);
        return;
    }
    // cb >= 0: 0 < z1 - z0 <= the budget (16-bit bases are given to such blocks only)
    const int t = (int)threadIdx.x;
    const int row = h0.r0 + t;
    const bool own = row < h0.r1;
    const int rs = own ? rp[row] : 0, re = own ? rp[row + 1] : 0;  // issued before the barrier
    decltype(pref(0)) q{};
    if (own) q = pref(row);
    u32x4 c[G];
    TX v[G][4], g[G][4];
#pragma unroll
    for (int u = 0; u < G; ++u) {
        const int qi = min(t + u * kBlock, nq - 1);  // clamped: every load issued; the entries are masked below
        c[u] = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(cs + a0) + qi);
        Quad<TX, TV>::load(av, a0 + 4 * qi, v[u]);
    }
#pragma unroll
    for (int u = 0; u < G; ++u)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int k = a0 + 4 * (t + u * kBlock) + e;  // unclamped: k >= z1 for a clamped quad
            const bool in = k >= z0 && k < z1;
            g[u][e] = xf(cb + (in ? (int)(c[u][e] & kSortColMask) : 0));
        }
#pragma unroll
    for (int u = 0; u < G; ++u)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int k = a0 + 4 * (t + u * kBlock) + e;
            if (k >= z0 && k < z1) prod[c[u][e] >> kSortColBits] = v[u][e] * g[u][e];
        }
    __syncthreads();
    if (own) {
        const TX* seg = prod + (rs - z0);
        auto a = initf(q);
        for (int k = 0; k < re - rs; ++k) a = accf(a, seg[k]);
        finf(row, a, q);
    }
}

template <bool XCD, int PER, typename TV, typename TX, class XF, class EPI>
__global__ __launch_bounds__(kBlock) void k_csr_stream_quad(const BlkHdr* __restrict__ hdr, const int* __restrict__ rp,
                                                            const int* __restrict__ cv,
                                                            const uint32_t* __restrict__ cs,
                                                            const TV* __restrict__ av, XF xf, EPI epi) {
    __shared__ TX prod[PER * kBlock];
    const int b = XCD ? xcd_block() : (int)blockIdx.x;
    auto pref = [&](int row) { return epi.pre(row); };
    auto initf = [&](const typename EPI::Pre& q) { return epi.init(q); };
    auto accf = [&](TX a, TX p) { return epi.acc(a, p); };
    auto finf = [&](int row, TX a, const typename EPI::Pre& q) { epi.fin(row, a, q); };
    stream_block_quad<PER, TV, TX, XF, decltype(pref), decltype(initf), decltype(accf), decltype(finf)>(
        b, hdr, rp, cv, cs, av, xf, prod, pref, initf, accf, finf);
}

// Vector CSR (FP32 / BF16 cycles only): LPR consecutive lanes per row, lane l
// summing the row's products l, l + LPR, ... with the epilogue's own
// accumulation from zero, the LPR partial sums combined by xor shuffles, then
// a = init(q) + sum and fin -- LPR chosen by the operator's average row
// length (about four products per lane; 64 = a wavefront per row).  No LDS
// stage and no barrier: three dependent round trips (row bounds, columns and
// values, gathers).  The sum order differs from the reference's sequential
// one, so the FP64 cycle never uses this kernel (its kind falls back to the
// stream kernel there) and FP32 / BF16 results under it are within rounding,
// not bit-equal, to the other kinds'.  Every epilogue accumulates a +- p, so
// splitting the sum over lanes is valid.
template <int LPR, typename TV, typename TX, class XF, class EPI>
__global__ __launch_bounds__(kBlock) void k_csr_vector(int nrows, const int* __restrict__ rp,
                                                       const int* __restrict__ cv, const TV* __restrict__ av, XF xf,
                                                       EPI epi) {
    const int t = (int)threadIdx.x;
    const int lane = t & (LPR - 1);
    const int b = gridDim.x >= 64 ? xcd_block() : (int)blockIdx.x;
    const int row = b * (kBlock / LPR) + t / LPR;
    const bool own = row < nrows;
    typename EPI::Pre q{};
    if (own && lane == 0) q = epi.pre(row);
    TX sum = 0;
    if (own) {
        const int re = rp[row + 1];
#pragma unroll 4
        for (int k = rp[row] + lane; k < re; k += LPR) sum = epi.acc(sum, ld<TX, TV>(av, k) * xf(ldi(cv, k)));
    }
#pragma unroll
    for (int off = LPR / 2; off > 0; off >>= 1) sum += __shfl_xor(sum, off, LPR);  // every lane takes part
    if (own && lane == 0) {
        auto a = epi.init(q);
        a = a + sum;
        epi.fin(row, a, q);
    }
}

// Lanes per row of the vector CSR kernel: about four products per lane.
inline int vector_lanes(int64_t nnz, int64_t nrow) {
    const double avg = nrow > 0 ? (double)nnz / (double)nrow : 0.0;
    int l = 4;
    while (l < 64 && 4.0 * l < avg) l <<= 1;
    return l;
}

template <typename TV, typename TX, class XF, class EPI>
void launch_vector(int nrows, int64_t nnz, const int* rp, const int* cv, const TV* av, XF xf, EPI epi,
                   hipStream_t st) {
    if (nrows <= 0) return;
    const int lpr = vector_lanes(nnz, nrows);
    const int nb = (int)(((int64_t)nrows * lpr + kBlock - 1) / kBlock);
    switch (lpr) {
        case 4: hipLaunchKernelGGL((k_csr_vector<4, TV, TX, XF, EPI>), dim3(nb), dim3(kBlock), 0, st, nrows, rp, cv, av, xf, epi); break;
        case 8: hipLaunchKernelGGL((k_csr_vector<8, TV, TX, XF, EPI>), dim3(nb), dim3(kBlock), 0, st, nrows, rp, cv, av, xf, epi); break;
        case 16: hipLaunchKernelGGL((k_csr_vector<16, TV, TX, XF, EPI>), dim3(nb), dim3(kBlock), 0, st, nrows, rp, cv, av, xf, epi); break;
        case 32: hipLaunchKernelGGL((k_csr_vector<32, TV, TX, XF, EPI>), dim3(nb), dim3(kBlock), 0, st, nrows, rp, cv, av, xf, epi); break;
        default: hipLaunchKernelGGL((k_csr_vector<64, TV, TX, XF, EPI>), dim3(nb), dim3(kBlock), 0, st, nrows, rp, cv, av, xf, epi); break;
    }
}

// Pattern form, for operators whose rows repeat a few column patterns
// relative to the row index, as a stencil on a structured grid does (the
// 7-point operator of a 3-D grid has 27: the interior, 6 faces, 12 edges and
// 8 corners).  Per row a 16-bit pattern index; per pattern its length and its
// column offsets col - row in CSR order (npat rows of W + 1 ints, W = the
// longest row; at most kPatMax patterns of at most kPatSlots entries), staged
// in LDS by every workgroup.  The values sit in slices of 64 rows with entry
// j of the slice's rows contiguous (slot ((row >> 6) * W + j) * 64 +
// (row & 63), zero where a row is shorter), so a wavefront's value loads are
// coalesced and its gathers x[row + off] fall on one or two lines per
// offset (the BF16 copy pairs two entries per 32-bit word, pattern_slot2).
// No column index is read: 2 bytes per row instead of 4 per entry.
// Workgroup b takes the CSR-stream row block b, row r0 + t on thread t (the
// rows and threads of stream_block), and each row sums its products in CSR
// order, each rounded on its own: the same bits as every other kind.
constexpr int kPatMax = 64;
constexpr int kPatSlots = 16;
constexpr int kPatSlice = 64;

__host__ __device__ inline int64_t pattern_slot(int row, int j, int w) {
    return ((int64_t)(row >> 6) * w + j) * kPatSlice + (row & (kPatSlice - 1));
}

// Slots of the pattern form's value arrays for n rows of width w.
inline int64_t pattern_slots(int n, int w) { return (int64_t)((n + kPatSlice - 1) / kPatSlice) * kPatSlice * w; }

// The BF16 copy holds entries j and j + 1 (j even) of a row in one 32-bit
// word, low half first, the words in the same slice layout over ceil(w / 2)
// pairs: a wavefront loads two entries of its 64 rows per dword load (2-byte
// loads per lane run at a fraction of the rate).
__host__ __device__ inline int64_t pattern_slot2(int row, int j, int w) {
    return 2 * pattern_slot(row, j >> 1, (w + 1) >> 1) + (j & 1);
}
inline int64_t pattern_slots2(int n, int w) { return 2 * pattern_slots(n, (w + 1) >> 1); }

// The values of a row's slots j < w (KW >= w; the rest zero)
template <int KW, typename TX, typename TV> struct PatternLoad {
    static __device__ __forceinline__ void load(const TV* tv, int row, int w, TX (&v)[KW]) {
        const int64_t s0 = pattern_slot(row, 0, w);
#pragma unroll
        for (int j = 0; j < KW; ++j) v[j] = j < w ? ld<TX, TV>(tv, s0 + (int64_t)j * kPatSlice) : (TX)0;
    }
};
template <int KW> struct PatternLoad<KW, float, bf16_t> {
    static __device__ __forceinline__ void load(const bf16_t* tv, int row, int w, float (&v)[KW]) {
        const int wh = (w + 1) >> 1;
        const uint32_t* p = reinterpret_cast<const uint32_t*>(tv) + pattern_slot(row, 0, wh);
#pragma unroll
        for (int jj = 0; jj < KW / 2; ++jj) {
            const uint32_t u = jj < wh ? __builtin_nontemporal_load(p + (int64_t)jj * kPatSlice) : 0u;
            v[2 * jj] = bf16_to_float((bf16_t)(u & 0xFFFFu));
            v[2 * jj + 1] = bf16_to_float((bf16_t)(u >> 16));
        }
    }
};

template <int KW, typename TV, typename TX, class XF, class PREF, class INITF, class ACCF, class FINF>
__device__ __forceinline__ void pattern_block(int b, const BlkHdr* __restrict__ hdr, int npat, int w,
                                              const uint16_t* __restrict__ pid, const int* __restrict__ tab,
                                              const TV* __restrict__ tv, XF xf, int* stab, PREF pref, INITF initf,
                                              ACCF accf, FINF finf) {
    const int t = (int)threadIdx.x;
    const BlkHdr h0 = hdr[b];
    const int row = h0.r0 + t;
    const bool own = row < h0.r1;
    for (int k = t; k < npat * (w + 1); k += kBlock) stab[k] = tab[k];
    int p = 0;
    TX v[KW];
    decltype(pref(0)) q{};
    if (own) {  // issued before the barrier: the pattern index, every value slot, the epilogue's loads
        p = (int)pid[row];
        PatternLoad<KW, TX, TV>::load(tv, row, w, v);
        q = pref(row);
    }
    __syncthreads();
    if (own) {
        const int* pt = stab + p * (w + 1);
        const int len = pt[0];
        TX g[KW];
#pragma unroll
        for (int j = 0; j < KW; ++j) g[j] = j < len ? xf(row + pt[1 + j]) : (TX)0;
        auto a = initf(q);
#pragma unroll
        for (int j = 0; j < KW; ++j)
            if (j < len) a = accf(a, v[j] * g[j]);
        finf(row, a, q);
    }
}

template <bool XCD, int KW, typename TV, typename TX, class XF, class EPI>
__global__ __launch_bounds__(kBlock) void k_csr_pattern(const BlkHdr* __restrict__ hdr, int npat, int w,
                                                        const uint16_t* __restrict__ pid, const int* __restrict__ tab,
                                                        const TV* __restrict__ tv, XF xf, EPI epi) {
    __shared__ int stab[kPatMax * (kPatSlots + 1)];
    const int b = XCD ? xcd_block() : (int)blockIdx.x;
    auto pref = [&](int row) { return epi.pre(row); };
    auto initf = [&](const typename EPI::Pre& q) { return epi.init(q); };
    auto accf = [&](TX a, TX p) { return epi.acc(a, p); };
    auto finf = [&](int row, TX a, const typename EPI::Pre& q) { epi.fin(row, a, q); };
    pattern_block<KW, TV, TX, XF, decltype(pref), decltype(initf), decltype(accf), decltype(finf)>(
        b, hdr, npat, w, pid, tab, tv, xf, stab, pref, initf, accf, finf);
}

// w = the form's width (<= kPatSlots): 8 or 16 unrolled slots
template <typename TV, typename TX, class XF, class EPI>
void launch_pattern(int nblk, const BlkHdr* hdr, int npat, int w, const uint16_t* pid, const int* tab, const TV* tv,
                    XF xf, EPI epi, hipStream_t st) {
    if (nblk <= 0) return;
#define DRAUGR_PAT(X, KW)                                                                                    \
    hipLaunchKernelGGL((k_csr_pattern<X, KW, TV, TX, XF, EPI>), dim3(nblk), dim3(kBlock), 0, st, hdr, npat, w, \
                       pid, tab, tv, xf, epi)
    if (w <= 8) {
        if (nblk >= 64) DRAUGR_PAT(true, 8);
        else DRAUGR_PAT(false, 8);
    } else {
        if (nblk >= 64) DRAUGR_PAT(true, 16);
        else DRAUGR_PAT(false, 16);
    }
#undef DRAUGR_PAT
}

// Thread-per-row variant without LDS: fewer dependent round trips, for the
// small coarse levels where a kernel's latency, not bandwidth, is the cost.
// Same per-row operation order as the stream kernel.
template <typename TV, typename TX, class XF, class EPI>
__global__ __launch_bounds__(kBlock) void k_csr_rowseq(int nrows, const int* __restrict__ rp,
                                                       const int* __restrict__ cv, const TV* __restrict__ av,
                                                       XF xf, EPI epi) {
    const int row = blockIdx.x * kBlock + threadIdx.x;
    if (row >= nrows) return;
    const typename EPI::Pre q = epi.pre(row);
    TX a = epi.init(q);
    const int e = rp[row + 1];
    for (int k = rp[row]; k < e; ++k) a = epi.acc(a, ld<TX, TV>(av, k) * xf(ldi(cv, k)));
    epi.fin(row, a, q);
}

// Software-pipelined CSR-stream: workgroup w walks a contiguous run of row
// blocks; while block b's gathers, LDS stage and epilogue run, the column and
// value loads of block b+1 are already in flight, so each wave keeps two
// blocks' worth of memory requests outstanding (the plain kernel waits for
// one block's loads before it can issue its gathers).  Same per-row
// operation order as k_csr_stream.  Blocks longer than PER * kBlock nonzeros
// fall back to the plain kernel's loop / tree paths.
template <bool XCD, int PER, typename TV, typename TX, class XF, class EPI>
__global__ __launch_bounds__(kBlock) void k_csr_pipe(int nblk, const int* __restrict__ blk,
                                                     const int* __restrict__ bz, const int* __restrict__ rp,
                                                     const int* __restrict__ cv, const TV* __restrict__ av, XF xf,
                                                     EPI epi) {
    constexpr int S = PER * kBlock;  // the LDS stage
    __shared__ TX prod[S];
    const int w = XCD ? xcd_block() : (int)blockIdx.x;
    const int per_wg = (nblk + (int)gridDim.x - 1) / (int)gridDim.x;
    const int b0 = w * per_wg, b1 = min(nblk, b0 + per_wg);
    if (b0 >= b1) return;
    const int t = (int)threadIdx.x;
    int cj[PER];
    TX v[PER];
    auto load = [&](int z0, int nz) {
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int k = min(t + u * kBlock, nz - 1);
            cj[u] = ldi(cv, z0 + k);
            v[u] = ld<TX, TV>(av, z0 + k);
        }
    };
    int r0 = blk[b0], r1 = blk[b0 + 1], z0 = bz[b0], nz = bz[b0 + 1] - z0;
    if (nz > 0 && nz <= PER * kBlock) load(z0, nz);
    for (int b = b0; b < b1; ++b) {
        const int row = r0 + t;
        const bool own = row < r1;
        const int rs = own ? rp[row] : 0, re = own ? rp[row + 1] : 0;
        typename EPI::Pre q{};
        if (own) q = epi.pre(row);
        // next block's header (the loads themselves are issued after this block's gathers)
        int nr0 = 0, nr1 = 0, nz0 = 0, nnz = 0;
        if (b + 1 < b1) {
            nr0 = blk[b + 1];
            nr1 = blk[b + 2];
            nz0 = bz[b + 1];
            nnz = bz[b + 2] - nz0;
        }
        if (nz <= PER * kBlock) {
            if (nz > 0) {
                TX g[PER];
#pragma unroll
                for (int u = 0; u < PER; ++u) g[u] = xf(cj[u]);
                TX vv[PER];
#pragma unroll
                for (int u = 0; u < PER; ++u) vv[u] = v[u];
                if (nnz > 0 && nnz <= PER * kBlock) load(nz0, nnz);
#pragma unroll
                for (int u = 0; u < PER; ++u)
                    if (t + u * kBlock < nz) prod[t + u * kBlock] = vv[u] * g[u];
            } else if (nnz > 0 && nnz <= PER * kBlock) {
                load(nz0, nnz);
            }
            __syncthreads();
            if (own) epi_segment(epi, row, prod + (rs - z0), re - rs, q);
        } else {  // a row longer than the LDS stage: chunk by chunk, row order, on thread 0
            typename EPI::Pre q0{};
            TX a = 0;
            if (t == 0) {
                q0 = epi.pre(r0);
                a = epi.init(q0);
            }
            for (int c0 = 0; c0 < nz; c0 += S) {
                const int cn = min(S, nz - c0);
                for (int k = t; k < cn; k += kBlock) prod[k] = ld<TX, TV>(av, z0 + c0 + k) * xf(ldi(cv, z0 + c0 + k));
                __syncthreads();
                if (t == 0)
                    for (int k = 0; k < cn; ++k) a = epi.acc(a, prod[k]);
                __syncthreads();
            }
            if (t == 0) epi.fin(r0, a, q0);
            if (nnz > 0 && nnz <= PER * kBlock) load(nz0, nnz);
        }
        __syncthreads();  // the stage is reused by the next block
        r0 = nr0;
        r1 = nr1;
        z0 = nz0;
        nz = nnz;
    }
}

// Software-pipelined CSR-stream over the block-header table, for the 16-bit
// column form (C16) and its column-sorted copy (SRT): workgroup w walks a
// contiguous run of row blocks (the runs of one XCD are adjacent); block b+1's
// column, value and position loads are issued right after block b's gathers, so
// they are in flight while block b's gathers return, its products are staged
// and its rows are summed.  Block headers are read two blocks ahead.  Same
// products, same LDS slots and the same sequential row sums as k_csr_stream,
// so the results are the same bits.  Blocks longer than PER * kBlock nonzeros
// (a single long row) are staged kStreamNnz products at a time and summed in
// row order by one thread.
template <bool XCD, int PER, bool SRT, typename TV, typename TX, class XF, class EPI>
__global__ __launch_bounds__(kBlock) void k_stream_pipe(int nblk, const BlkHdr* __restrict__ hdr,
                                                        const int* __restrict__ rp, const int* __restrict__ cv,
                                                        const uint16_t* __restrict__ c16, const TV* __restrict__ av,
                                                        XF xf, EPI epi, const uint32_t* __restrict__ cs) {
    constexpr int S = PER * kBlock;  // the LDS stage
    __shared__ TX prod[S];
    const int w = XCD ? xcd_block() : (int)blockIdx.x;
    const int per_wg = (nblk + (int)gridDim.x - 1) / (int)gridDim.x;
    const int b0 = w * per_wg, b1 = min(nblk, b0 + per_wg);
    if (b0 >= b1) return;
    const int t = (int)threadIdx.x;
    int cj[PER], sl[PER];
    TX v[PER];
    // loads of a block of 0 < nz <= PER * kBlock nonzeros (indices past its end
    // clamped to its last entry; their products are discarded)
    auto load = [&](const BlkHdr& h, int nz) {
        if (SRT && h.cb >= 0) {
#pragma unroll
            for (int u = 0; u < PER; ++u) {
                const int k = min(t + u * kBlock, nz - 1);
                const uint32_t e = ldu(cs, h.z0 + k);
                cj[u] = h.cb + (int)(e & kSortColMask);
                v[u] = ld<TX, TV>(av, h.z0 + k);
                sl[u] = (int)(e >> kSortColBits);
            }
        } else if (h.cb >= 0) {
#pragma unroll
            for (int u = 0; u < PER; ++u) {
                const int k = min(t + u * kBlock, nz - 1);
                cj[u] = h.cb + ldc16(c16, h.z0 + k);
                v[u] = ld<TX, TV>(av, h.z0 + k);
                sl[u] = t + u * kBlock;
            }
        } else {
#pragma unroll
            for (int u = 0; u < PER; ++u) {
                const int k = min(t + u * kBlock, nz - 1);
                cj[u] = ldi(cv, h.z0 + k);
                v[u] = ld<TX, TV>(av, h.z0 + k);
                sl[u] = t + u * kBlock;
            }
        }
    };
    BlkHdr hc = hdr[b0], hn = hdr[b0 + 1];
    int nz = hn.z0 - hc.z0;
    if (nz > 0 && nz <= PER * kBlock) load(hc, nz);
    for (int b = b0; b < b1; ++b) {
        const bool more = b + 1 < b1;
        BlkHdr hn2{};
        if (more) hn2 = hdr[b + 2];  // entry nblk closes the table, so b + 2 <= nblk
        const int nz_next = more ? hn2.z0 - hn.z0 : 0;
        const bool pipe_next = nz_next > 0 && nz_next <= PER * kBlock;
        const int row = hc.r0 + t;
        const bool own = row < hc.r1;
        const int rs = own ? rp[row] : 0, re = own ? rp[row + 1] : 0;
        typename EPI::Pre q{};
        if (own) q = epi.pre(row);
        if (nz <= PER * kBlock) {
            if (nz > 0) {
                TX g[PER], vv[PER];
                int ss[PER];
#pragma unroll
                for (int u = 0; u < PER; ++u) g[u] = xf(cj[u]);
#pragma unroll
                for (int u = 0; u < PER; ++u) {
                    vv[u] = v[u];
                    ss[u] = sl[u];
                }
                if (pipe_next) load(hn, nz_next);
#pragma unroll
                for (int u = 0; u < PER; ++u)
                    if (t + u * kBlock < nz) prod[ss[u]] = vv[u] * g[u];
            } else if (pipe_next) {
                load(hn, nz_next);
            }
            __syncthreads();
            if (own) epi_segment(epi, row, prod + (rs - hc.z0), re - rs, q);
        } else {  // one long row: chunk by chunk, row order, on thread 0
            typename EPI::Pre q0{};
            TX a = 0;
            if (t == 0) {
                q0 = epi.pre(hc.r0);
                a = epi.init(q0);
            }
            for (int c0 = 0; c0 < nz; c0 += S) {
                const int cn = min(S, nz - c0);
                for (int k = t; k < cn; k += kBlock)
                    prod[k] = ld<TX, TV>(av, hc.z0 + c0 + k) * xf(ldi(cv, hc.z0 + c0 + k));
                __syncthreads();
                if (t == 0)
                    for (int k = 0; k < cn; ++k) a = epi.acc(a, prod[k]);
                __syncthreads();
            }
            if (t == 0) epi.fin(hc.r0, a, q0);
            if (pipe_next) load(hn, nz_next);
        }
        __syncthreads();  // the stage is reused by the next block
        hc = hn;
        hn = hn2;
        nz = nz_next;
    }
}

// Resident workgroups of the pipelined kernels: about 8 per CU on 256 CUs.
constexpr int kPipeGrid = 2048;

template <int PER, bool SRT, typename TV, typename TX, class XF, class EPI>
void launch_stream_pipe_per(int nblk, const BlkHdr* hdr, const int* rp, const int* cv, const uint16_t* c16,
                            const TV* av, XF xf, EPI epi, hipStream_t st, const uint32_t* cs) {
    const int grid = std::min(nblk, kPipeGrid);
    if (grid >= 64)
        hipLaunchKernelGGL((k_stream_pipe<true, PER, SRT, TV, TX, XF, EPI>), dim3(grid), dim3(kBlock), 0, st, nblk,
                           hdr, rp, cv, c16, av, xf, epi, cs);
    else
        hipLaunchKernelGGL((k_stream_pipe<false, PER, SRT, TV, TX, XF, EPI>), dim3(grid), dim3(kBlock), 0, st,
                           nblk, hdr, rp, cv, c16, av, xf, epi, cs);
}

// per = the operator's block budget / kBlock; cs non-null = the
// column-sorted form (cs and av sorted within each 16-bit block)
template <typename TV, typename TX, class XF, class EPI>
void launch_stream_pipe(int nblk, int per, const BlkHdr* hdr, const int* rp, const int* cv, const uint16_t* c16,
                        const TV* av, XF xf, EPI epi, hipStream_t st, const uint32_t* cs) {
    if (nblk <= 0) return;
    if (cs) {
        switch (per) {
            case 1: launch_stream_pipe_per<1, true, TV, TX>(nblk, hdr, rp, cv, c16, av, xf, epi, st, cs); break;
            case 2: launch_stream_pipe_per<2, true, TV, TX>(nblk, hdr, rp, cv, c16, av, xf, epi, st, cs); break;
            case 4: launch_stream_pipe_per<4, true, TV, TX>(nblk, hdr, rp, cv, c16, av, xf, epi, st, cs); break;
            case 8: launch_stream_pipe_per<8, true, TV, TX>(nblk, hdr, rp, cv, c16, av, xf, epi, st, cs); break;
            default: launch_stream_pipe_per<16, true, TV, TX>(nblk, hdr, rp, cv, c16, av, xf, epi, st, cs); break;
        }
    } else {
        switch (per) {
            case 1: launch_stream_pipe_per<1, false, TV, TX>(nblk, hdr, rp, cv, c16, av, xf, epi, st, cs); break;
            case 2: launch_stream_pipe_per<2, false, TV, TX>(nblk, hdr, rp, cv, c16, av, xf, epi, st, cs); break;
            case 4: launch_stream_pipe_per<4, false, TV, TX>(nblk, hdr, rp, cv, c16, av, xf, epi, st, cs); break;
            case 8: launch_stream_pipe_per<8, false, TV, TX>(nblk, hdr, rp, cv, c16, av, xf, epi, st, cs); break;
            default: launch_stream_pipe_per<16, false, TV, TX>(nblk, hdr, rp, cv, c16, av, xf, epi, st, cs); break;
        }
    }
}

template <int PER, typename TV, typename TX, class XF, class EPI>
void launch_pipe_per(int nblk, const int* blk, const int* bz, const int* rp, const int* cv, const TV* av, XF xf,
                     EPI epi, hipStream_t st) {
    const int grid = std::min(nblk, kPipeGrid);
    if (grid >= 64)
        hipLaunchKernelGGL((k_csr_pipe<true, PER, TV, TX, XF, EPI>), dim3(grid), dim3(kBlock), 0, st, nblk, blk, bz,
                           rp, cv, av, xf, epi);
    else
        hipLaunchKernelGGL((k_csr_pipe<false, PER, TV, TX, XF, EPI>), dim3(grid), dim3(kBlock), 0, st, nblk, blk,
                           bz, rp, cv, av, xf, epi);
}

template <typename TV, typename TX, class XF, class EPI>
void launch_pipe(int nblk, int per, const int* blk, const int* bz, const int* rp, const int* cv, const TV* av, XF xf,
                 EPI epi, hipStream_t st) {
    if (nblk <= 0) return;
    switch (per) {
        case 1: launch_pipe_per<1, TV, TX>(nblk, blk, bz, rp, cv, av, xf, epi, st); break;
        case 2: launch_pipe_per<2, TV, TX>(nblk, blk, bz, rp, cv, av, xf, epi, st); break;
        case 4: launch_pipe_per<4, TV, TX>(nblk, blk, bz, rp, cv, av, xf, epi, st); break;
        case 8: launch_pipe_per<8, TV, TX>(nblk, blk, bz, rp, cv, av, xf, epi, st); break;
        default: launch_pipe_per<16, TV, TX>(nblk, blk, bz, rp, cv, av, xf, epi, st); break;
    }
}

enum OpKernel : int {
    kOpStream = 0, kOpRowSeq = 1, kOpPipe = 2, kOpStream16 = 3, kOpSorted = 4, kOpStream16Pipe = 5, kOpSortedPipe = 6,
    kOpSortedQuad = 7, kOpVector = 8, kOpPattern = 9
};
constexpr int kOpKinds = 10;

template <int PER, bool C16, bool SRT, typename TV, typename TX, class XF, class EPI>
void launch_stream_per(int nblk, const BlkHdr* hdr, const int* rp, const int* cv, const uint16_t* c16, const TV* av,
                       XF xf, EPI epi, hipStream_t st, const uint32_t* cs) {
    if (nblk >= 64)
        hipLaunchKernelGGL((k_csr_stream<true, PER, C16, TV, TX, XF, EPI, SRT>), dim3(nblk), dim3(kBlock), 0, st, hdr,
                           rp, cv, c16, av, xf, epi, cs);
    else
        hipLaunchKernelGGL((k_csr_stream<false, PER, C16, TV, TX, XF, EPI, SRT>), dim3(nblk), dim3(kBlock), 0, st,
                           hdr, rp, cv, c16, av, xf, epi, cs);
}

template <bool C16, bool SRT, typename TV, typename TX, class XF, class EPI>
void launch_stream_c(int nblk, int per, const BlkHdr* hdr, const int* rp, const int* cv, const uint16_t* c16,
                     const TV* av, XF xf, EPI epi, hipStream_t st, const uint32_t* cs) {
    switch (per) {
        case 1: launch_stream_per<1, C16, SRT, TV, TX>(nblk, hdr, rp, cv, c16, av, xf, epi, st, cs); break;
        case 2: launch_stream_per<2, C16, SRT, TV, TX>(nblk, hdr, rp, cv, c16, av, xf, epi, st, cs); break;
        case 4: launch_stream_per<4, C16, SRT, TV, TX>(nblk, hdr, rp, cv, c16, av, xf, epi, st, cs); break;
        case 8: launch_stream_per<8, C16, SRT, TV, TX>(nblk, hdr, rp, cv, c16, av, xf, epi, st, cs); break;
        default: launch_stream_per<16, C16, SRT, TV, TX>(nblk, hdr, rp, cv, c16, av, xf, epi, st, cs); break;
    }
}

// per = the operator's block budget / kBlock (1, 2, 4, 8 or 16); c16 non-null =
// the 16-bit column form (bases in the block headers); cs non-null = its
// column-sorted form (cs and av sorted within each 16-bit block)
template <typename TV, typename TX, class XF, class EPI>
void launch_stream(int nblk, int per, const BlkHdr* hdr, const int* rp, const int* cv, const TV* av, XF xf, EPI epi,
                   hipStream_t st, const uint16_t* c16 = nullptr, const uint32_t* cs = nullptr) {
    if (nblk <= 0) return;
    if (cs) launch_stream_c<true, true, TV, TX>(nblk, per, hdr, rp, cv, nullptr, av, xf, epi, st, cs);
    else if (c16) launch_stream_c<true, false, TV, TX>(nblk, per, hdr, rp, cv, c16, av, xf, epi, st, nullptr);
    else launch_stream_c<false, false, TV, TX>(nblk, per, hdr, rp, cv, nullptr, av, xf, epi, st, nullptr);
}

// The column-sorted form with four entries per load (budgets of 4, 8 or 16
// per thread; smaller budgets take the entry-by-entry kernel).
template <int PER, typename TV, typename TX, class XF, class EPI>
void launch_stream_quad_per(int nblk, const BlkHdr* hdr, const int* rp, const int* cv, const uint32_t* cs,
                            const TV* av, XF xf, EPI epi, hipStream_t st) {
    if (nblk >= 64)
        hipLaunchKernelGGL((k_csr_stream_quad<true, PER, TV, TX, XF, EPI>), dim3(nblk), dim3(kBlock), 0, st, hdr, rp,
                           cv, cs, av, xf, epi);
    else
        hipLaunchKernelGGL((k_csr_stream_quad<false, PER, TV, TX, XF, EPI>), dim3(nblk), dim3(kBlock), 0, st, hdr,
                           rp, cv, cs, av, xf, epi);
}

template <typename TV, typename TX, class XF, class EPI>
void launch_stream_quad(int nblk, int per, const BlkHdr* hdr, const int* rp, const int* cv, const uint32_t* cs,
                        const TV* av, XF xf, EPI epi, hipStream_t st) {
    if (nblk <= 0) return;
    switch (per) {
        case 4: launch_stream_quad_per<4, TV, TX>(nblk, hdr, rp, cv, cs, av, xf, epi, st); break;
        case 8: launch_stream_quad_per<8, TV, TX>(nblk, hdr, rp, cv, cs, av, xf, epi, st); break;
        case 16: launch_stream_quad_per<16, TV, TX>(nblk, hdr, rp, cv, cs, av, xf, epi, st); break;
        default: launch_stream_c<true, true, TV, TX>(nblk, per, hdr, rp, cv, nullptr, av, xf, epi, st, cs); break;
    }
}

template <typename TV, typename TX, class XF, class EPI>
void launch_rowseq(int nrows, const int* rp, const int* cv, const TV* av, XF xf, EPI epi, hipStream_t st) {
    if (nrows <= 0) return;
    hipLaunchKernelGGL((k_csr_rowseq<TV, TX, XF, EPI>), dim3((nrows + kBlock - 1) / kBlock), dim3(kBlock), 0, st,
                       nrows, rp, cv, av, xf, epi);
}

// Nonzeros per row block: kStreamNnz for large operators; halved (down to
// 256) while the operator would give fewer than ~8 blocks per CU, so that
// small coarse levels still spread over the chip and each block's serial
// phases stay short, and while a block of kStreamRows average rows would fill
// less than half of it (its threads' clamped loads would mostly repeat).
inline int stream_budget(int64_t nnz, int64_t nrow) {
    int budget = kStreamNnz;
    const double avg = nrow > 0 ? (double)nnz / (double)nrow : 0.0;
    while (budget > 256 && (nnz / budget < 2048 || budget > 2.0 * avg * kStreamRows)) budget >>= 1;
    return budget;
}

// Row blocks of at most kStreamRows rows and `budget` nonzeros (a longer row
// alone in its block).  Returns offsets: block b = rows [o[b], o[b+1]).
inline std::vector<int32_t> build_row_blocks(const std::vector<int32_t>& rp, int budget = kStreamNnz) {
    const int n = (int)rp.size() - 1;
    std::vector<int32_t> o;
    o.push_back(0);
    int64_t acc = 0;
    int rows = 0;
    for (int i = 0; i < n; ++i) {
        const int len = rp[i + 1] - rp[i];
        if (rows > 0 && (acc + len > budget || rows == kStreamRows)) {
            o.push_back(i);
            acc = 0;
            rows = 0;
        }
        acc += len;
        ++rows;
    }
    if (n > 0) o.push_back(n);
    return o;
}

// Coloured Gauss-Seidel update of the rows in one colour list, in place
// (src/smoothers.jl:203-217): one thread per row, residual by successive
// subtraction in row order.  Rows of one colour do not couple.
template <typename TV, typename TX>
__global__ __launch_bounds__(kBlock) void k_gs_color(int cnt, const int* __restrict__ list, const int* __restrict__ rp,
                                                     const int* __restrict__ cv, const TV* __restrict__ av, TX* x,
                                                     const TX* __restrict__ b, const TX* __restrict__ d) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= cnt) return;
    const int i = list[t];
    TX r = b[i];
    for (int k = rp[i]; k < rp[i + 1]; ++k) r -= ld<TX, TV>(av, k) * x[cv[k]];
    x[i] += d[i] * r;
}

}  // namespace dev
}  // namespace draugr
