// Differentiation of x = A(k)^-1 b (no counterpart in the reference; the
// reference's operator convention is kept: CSR, columns sorted in rows).
//
//  * assemble_7pt: values of the cell-centred 7-point diffusion operator on an
//    m^3 grid from the cell permeabilities k (harmonic-mean face
//    transmissibility T = 2 ka kb / (ka + kb), Dirichlet faces through a
//    half-cell term 2 k on the diagonal), in the row layout and summation
//    order of the host generator, so that equal k give bit-equal values.  Used
//    to resetup the hierarchy on the device when k changes.
//  * matrix_grad: dL/dA_ij = -lambda_i x_j on the pattern of A, for a loss L
//    with adjoint lambda = A^-T dL/dx.
//  * grad_logk_7pt: chain rule from the matrix entries to log k:
//    dL/dT_f = -(l_a - l_b)(x_a - x_b) for a face (a, b);
//    dT/dlog k_a = 2 k_a k_b^2 / (k_a + k_b)^2;
//    a Dirichlet face of cell c adds -l_c x_c * 2 k_c.
//    One thread per cell gathers its own faces (no atomics).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "dbuf.hpp"
#include "kernels.hpp"

namespace draugr {

namespace {

using dev::kBlock;

__device__ __forceinline__ double trans(double ka, double kb) { return 2.0 * ka * kb / (ka + kb); }

// cell c = i + m (j + m kz); rows in the order z-1, y-1, x-1, diag, x+1, y+1, z+1
__global__ __launch_bounds__(kBlock) void k_assemble_7pt(int m, const double* __restrict__ k,
                                                          const int* __restrict__ rp, double* __restrict__ nz) {
    const int64_t n = (int64_t)m * m * m;
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const int i = (int)(c % m), j = (int)((c / m) % m), z = (int)(c / ((int64_t)m * m));
    const int64_t mm = (int64_t)m * m;
    const double kc = k[c];
    // diagonal: per axis (z, y, x): face to +1, face to -1, then the boundary terms
    double d = 0.0;
    double tzm = 0, tzp = 0, tym = 0, typ = 0, txm = 0, txp = 0;
    if (z < m - 1) { tzp = trans(kc, k[c + mm]); d += tzp; }
    if (z > 0) { tzm = trans(k[c - mm], kc); d += tzm; }
    if (z == 0) d += 2.0 * kc;
    if (z == m - 1) d += 2.0 * kc;
    if (j < m - 1) { typ = trans(kc, k[c + m]); d += typ; }
    if (j > 0) { tym = trans(k[c - m], kc); d += tym; }
    if (j == 0) d += 2.0 * kc;
    if (j == m - 1) d += 2.0 * kc;
    if (i < m - 1) { txp = trans(kc, k[c + 1]); d += txp; }
    if (i > 0) { txm = trans(k[c - 1], kc); d += txm; }
    if (i == 0) d += 2.0 * kc;
    if (i == m - 1) d += 2.0 * kc;
    int p = rp[c];
    if (z > 0) nz[p++] = -tzm;
    if (j > 0) nz[p++] = -tym;
    if (i > 0) nz[p++] = -txm;
    nz[p++] = d;
    if (i < m - 1) nz[p++] = -txp;
    if (j < m - 1) nz[p++] = -typ;
    if (z < m - 1) nz[p++] = -tzp;
}

__global__ __launch_bounds__(kBlock) void k_matrix_grad(int n, const int* __restrict__ rp, const int* __restrict__ cv,
                                                         const double* __restrict__ x, const double* __restrict__ lam,
                                                         double* __restrict__ g) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const double li = lam[i];
    for (int q = rp[i]; q < rp[i + 1]; ++q) g[q] = -li * x[cv[q]];
}

__device__ __forceinline__ double dtrans_dlog(double ka, double kb) {  // k_a dT/dk_a
    const double s = ka + kb;
    return 2.0 * ka * kb * kb / (s * s);
}

__global__ __launch_bounds__(kBlock) void k_grad_logk_7pt(int m, const double* __restrict__ k,
                                                           const double* __restrict__ x,
                                                           const double* __restrict__ lam, double* __restrict__ g) {
    const int64_t n = (int64_t)m * m * m;
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const int i = (int)(c % m), j = (int)((c / m) % m), z = (int)(c / ((int64_t)m * m));
    const int64_t mm = (int64_t)m * m;
    const double kc = k[c], xc = x[c], lc = lam[c];
    double acc = 0.0;
    auto face = [&](int64_t o) {
        const double dT = -(lc - lam[o]) * (xc - x[o]);
        acc += dT * dtrans_dlog(kc, k[o]);
    };
    if (z > 0) face(c - mm);
    if (j > 0) face(c - m);
    if (i > 0) face(c - 1);
    if (i < m - 1) face(c + 1);
    if (j < m - 1) face(c + m);
    if (z < m - 1) face(c + mm);
    const int nb = (z == 0) + (z == m - 1) + (j == 0) + (j == m - 1) + (i == 0) + (i == m - 1);
    acc += -lc * xc * (2.0 * kc) * nb;
    g[c] = acc;
}

}  // namespace

void assemble_7pt(int m, const double* k, const int* rp, double* nz, hipStream_t st) {
    const int64_t n = (int64_t)m * m * m;
    if (n == 0) return;
    hipLaunchKernelGGL(k_assemble_7pt, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, m, k, rp,
                       nz);
}

void matrix_grad(int n, const int* rp, const int* cv, const double* x, const double* lam, double* g,
                 hipStream_t st) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_matrix_grad, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, n, rp, cv, x, lam, g);
}

void grad_logk_7pt(int m, const double* k, const double* x, const double* lam, double* g, hipStream_t st) {
    const int64_t n = (int64_t)m * m * m;
    if (n == 0) return;
    hipLaunchKernelGGL(k_grad_logk_7pt, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, m, k, x,
                       lam, g);
}

}  // namespace draugr
