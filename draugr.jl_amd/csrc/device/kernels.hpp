// HIP kernels for gfx950 (CDNA4): every hot operation of the AMG cycle is a
// CSR row-group pass with a fused epilogue.
//
// A "row group" is G consecutive lanes of a 64-wide wavefront (G in
// {1,2,4,8,16,32,64}); the group strides over the nonzeros of one row, so a
// wavefront covers 64/G rows with coalesced loads of colval/nzval, and the
// partial sums are combined with cross-lane shuffles.  G is picked per operator
// from its mean row length (see pick_group()).  Blocks are remapped so that
// each of the 8 XCDs sweeps a contiguous slice of rows: neighbouring rows share
// x entries through that XCD's L2.
//
// Value storage TV is double, float or bf16 (uint16_t); vectors / accumulation
// TX is double (TV = double) or float (TV = float or bf16).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace draugr {
namespace dev {

constexpr int kBlock = 256;

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

// XCD-contiguous remap of the block index (bijective for any grid size).
__device__ __forceinline__ int xcd_block() {
    const int b = blockIdx.x, nb = gridDim.x;
    const int q = nb >> 3, r = nb & 7, x = b & 7, t = b >> 3;
    return x < r ? x * (q + 1) + t : r * (q + 1) + (x - r) * q + t;
}

template <int G, typename T> __device__ __forceinline__ T group_sum(T v) {
#pragma unroll
    for (int off = G >> 1; off > 0; off >>= 1) v += __shfl_xor(v, off, G);
    return v;
}

// ── row functors: list position -> row index ────────────────────────────────
struct RowId { __device__ int operator()(int i) const { return i; } };
struct RowList {
    const int* list;
    __device__ int operator()(int i) const { return list[i]; }
};

// ── x functors: value of the vector the row multiplies ──────────────────────
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

// ── epilogues ────────────────────────────────────────────────────────────────
template <typename TX> struct EpiStore {          // y = A x
    TX* y;
    __device__ void operator()(int i, TX s) const { y[i] = s; }
};
template <typename TX> struct EpiResid {          // r = b - A x
    const TX* b; TX* r;
    __device__ void operator()(int i, TX s) const { r[i] = b[i] - s; }
};
template <typename TX> struct EpiJacobi {         // xo = x + w d (b - A x)
    const TX* x; const TX* b; const TX* d; TX* xo; TX w;
    __device__ void operator()(int i, TX s) const { xo[i] = x[i] + (w * d[i]) * (b[i] - s); }
};
template <typename TX> struct EpiZeroJacobiResid {  // x1 = w d b, r = b - A x1
    const TX* b; const TX* d; TX* x1; TX* r; TX w;
    __device__ void operator()(int i, TX s) const {
        x1[i] = (w * d[i]) * b[i];
        r[i] = b[i] - s;
    }
};
template <typename TX> struct EpiGS {             // x_i += d_i (b_i - A_i x), in place
    TX* x; const TX* b; const TX* d;
    __device__ void operator()(int i, TX s) const { x[i] += d[i] * (b[i] - s); }
};
template <typename TX> struct EpiAdd {            // x += P xc
    TX* x;
    __device__ void operator()(int i, TX s) const { x[i] += s; }
};

template <int G, bool XCD, typename TV, typename TX, class RowF, class XF, class EPI>
__global__ __launch_bounds__(kBlock) void k_csr_rows(int nitems, const int* __restrict__ rp,
                                                     const int* __restrict__ cv, const TV* __restrict__ av,
                                                     RowF rowf, XF xf, EPI epi) {
    const int b = XCD ? xcd_block() : (int)blockIdx.x;
    const int t = b * kBlock + (int)threadIdx.x;
    const int item = t / G, lane = t % G;
    if (item >= nitems) return;
    const int row = rowf(item);
    const int s = rp[row], e = rp[row + 1];
    TX acc = 0;
    for (int k = s + lane; k < e; k += G) acc += ld<TX, TV>(av, k) * xf(ldi(cv, k));
    acc = group_sum<G>(acc);
    if (lane == 0) epi(row, acc);
}

template <int G, typename TV, typename TX, class RowF, class XF, class EPI>
void launch_rows_g(int nitems, const int* rp, const int* cv, const TV* av, RowF rowf, XF xf, EPI epi,
                   hipStream_t st) {
    if (nitems <= 0) return;
    const int64_t threads = (int64_t)nitems * G;
    const int nb = (int)((threads + kBlock - 1) / kBlock);
    if (nb >= 64)
        hipLaunchKernelGGL((k_csr_rows<G, true, TV, TX, RowF, XF, EPI>), dim3(nb), dim3(kBlock), 0, st,
                           nitems, rp, cv, av, rowf, xf, epi);
    else
        hipLaunchKernelGGL((k_csr_rows<G, false, TV, TX, RowF, XF, EPI>), dim3(nb), dim3(kBlock), 0, st,
                           nitems, rp, cv, av, rowf, xf, epi);
}

template <typename TV, typename TX, class RowF, class XF, class EPI>
void launch_rows(int G, int nitems, const int* rp, const int* cv, const TV* av, RowF rowf, XF xf, EPI epi,
                 hipStream_t st) {
    switch (G) {
        case 1: launch_rows_g<1, TV, TX>(nitems, rp, cv, av, rowf, xf, epi, st); break;
        case 2: launch_rows_g<2, TV, TX>(nitems, rp, cv, av, rowf, xf, epi, st); break;
        case 4: launch_rows_g<4, TV, TX>(nitems, rp, cv, av, rowf, xf, epi, st); break;
        case 8: launch_rows_g<8, TV, TX>(nitems, rp, cv, av, rowf, xf, epi, st); break;
        case 16: launch_rows_g<16, TV, TX>(nitems, rp, cv, av, rowf, xf, epi, st); break;
        case 32: launch_rows_g<32, TV, TX>(nitems, rp, cv, av, rowf, xf, epi, st); break;
        default: launch_rows_g<64, TV, TX>(nitems, rp, cv, av, rowf, xf, epi, st); break;
    }
}

// Lanes per row from the mean row length: the smallest power of two that
// covers the mean, capped at a wavefront.
inline int pick_group(int64_t nnz, int64_t nrows) {
    if (nrows <= 0) return 1;
    const double mean = (double)nnz / (double)nrows;
    int g = 1;
    while (g < 64 && g < mean) g <<= 1;
    return g;
}

}  // namespace dev
}  // namespace draugr
