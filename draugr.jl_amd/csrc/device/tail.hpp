// The last AMG level of a V-cycle and the coarsest direct solve in one
// launch, resident in LDS.  On that level every pass is tiny and a separate
// launch costs its launch plus a chain of dependent memory round trips
// (block bounds, columns and values, gathers) -- about five launches and
// 25-30 us per cycle at 128^3.  Here one workgroup copies a packed image of the
// level (A, P, P^T with 16-bit columns, the smoother's inverse diagonal, the
// coarsest LU factors and permutation; 16-byte loads, all in flight) and the
// level's right-hand side into LDS, then runs every pass from LDS with
// workgroup barriers between them -- pre-smoothing from a zero guess,
// residual, restriction, the coarsest solve (one wave, getrs order),
// prolongation, post-smoothing -- and writes the level's iterate back.  Every
// row is summed on its own thread in row order with the products rounded one
// by one and the stream kernels' epilogue arithmetic, so the values are the
// same bits as the separate launches'.  The parent level's restriction (into
// this level's right-hand side) and prolongation (of this level's iterate)
// can be folded into the same launch, read from global memory (TailRun).
//
// TV: values of A (double, float or bf16); TP: values of P and P^T; TX:
// vectors.  Eligible (checked on the host): a Jacobi-type smoother, a
// coarsest level of at most 64 rows, and the image plus the level's vectors
// within kTailLds bytes.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace draugr {
namespace dev {

constexpr int kTailThreads = 512;
constexpr int kTailLds = 156 * 1024;

__device__ __forceinline__ double bcast_lane(double v, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

// Forward and back substitution in one wavefront for n <= 64 unknowns, lane i
// holding row i of the LU factors and x_i: the operations of the reference's
// getrs loops -- x_i -= L_ij x_j for i > j with j ascending, then x_j /= U_jj
// and x_i -= U_ij x_j for i < j with j descending -- x_j broadcast from lane
// j with readlane.  Every lane forms a step's candidate and keeps it by a
// select (no divergent branch); the pivot division is formed on the broadcast
// operands x_j and U_jj, the same IEEE quotient in every lane.  Lanes i >= n
// hold the identity row and x_i = 0 (wave_lu_row), so that the steps j >= n
// inside the last group of 8 leave every value unchanged (x - 0 * 0); one
// uniform branch per group of 8 steps, not per step, skips the rest.
template <class AT>
__device__ __forceinline__ void wave_lu_row(double (&row)[64], int n, int i, AT at) {  // at(j) = LU(i, j)
#pragma unroll
    for (int j = 0; j < 64; ++j) row[j] = (j < n && i < n) ? at(j) : (i == j ? 1.0 : 0.0);
}

__device__ __forceinline__ double wave_lu_solve(double xi, const double (&row)[64], int n, int i) {
    const bool in = i < n;
#pragma unroll
    for (int j0 = 0; j0 < 64; j0 += 8) {
        if (j0 < n) {
#pragma unroll
            for (int j = j0; j < j0 + 8; ++j) {
                const double xj = bcast_lane(xi, j);
                const double t = xi - row[j] * xj;
                xi = (in && i > j) ? t : xi;
            }
        }
    }
#pragma unroll
    for (int j0 = 56; j0 >= 0; j0 -= 8) {
        if (j0 < n) {
#pragma unroll
            for (int j = j0 + 7; j >= j0; --j) {
                const double xd = bcast_lane(xi, j) / bcast_lane(row[j], j);
                const double t = xi - row[j] * xd;
                xi = i == j ? xd : (i < j ? t : xi);
            }
        }
    }
    return xi;
}

// Byte offsets of the image's segments (16-byte aligned) and of the LDS
// vectors that follow the image.
struct TailLayout {
    int n, nc, nnz_a, nnz_p;
    int a_rp, a_cv, a_v, p_rp, p_cv, p_v, t_rp, t_cv, t_v, d, lu, perm;
    int img_bytes;
    int v_b, v_x, v_alt, v_r, v_bc, v_xc, v_prod;
    int lds_bytes;
};

inline int tail_align(int b) { return (b + 15) & ~15; }

// The layout of a level with n rows, nc coarse rows, nnz(A) and nnz(P)
// entries for value types of sizes sv (A), sp (P) and sx (vectors).
inline TailLayout tail_layout(int n, int nc, int nnz_a, int nnz_p, int sv, int sp, int sx) {
    TailLayout L{};
    L.n = n;
    L.nc = nc;
    L.nnz_a = nnz_a;
    L.nnz_p = nnz_p;
    int o = 0;
    auto seg = [&](int bytes) { const int at = o; o = tail_align(o + bytes); return at; };
    L.a_rp = seg(4 * (n + 1));
    L.a_cv = seg(2 * nnz_a);
    L.a_v = seg(sv * nnz_a);
    L.p_rp = seg(4 * (n + 1));
    L.p_cv = seg(2 * nnz_p);
    L.p_v = seg(sp * nnz_p);
    L.t_rp = seg(4 * (nc + 1));
    L.t_cv = seg(2 * nnz_p);
    L.t_v = seg(sp * nnz_p);
    L.d = seg(sx * n);
    L.lu = seg(8 * nc * nc);
    L.perm = seg(4 * nc);
    L.img_bytes = o;
    L.v_b = seg(sx * n);
    L.v_x = seg(sx * n);
    L.v_alt = seg(sx * n);
    L.v_r = seg(sx * n);
    L.v_bc = seg(sx * nc);
    L.v_xc = seg(sx * nc);
    L.v_prod = seg(sx * (nnz_a > nnz_p ? nnz_a : nnz_p));
    L.lds_bytes = o;
    return L;
}

template <typename TX> struct TailRun {
    const uint8_t* img;  // the packed image (device memory)
    const TX* b;         // the level's right-hand side (its parent's restricted residual)
    TX* x;               // the level's iterate out (its parent's coarse vector)
    TX w;                // smoother weight
    int pre, post;
    TailLayout lay;
    // The parent level's restriction and prolongation folded into the launch
    // (values of P^T and P in the transfer type TP of the kernel):
    // trp set: b = P^T rin formed here (rows of P^T in global memory, each
    // summed in row order as the restriction pass sums it) instead of read;
    // prp set: xpar += P x at the end (each of the parent's npar rows in row
    // order, as the prolongation pass adds it).
    const int* trp = nullptr;
    const int* tcv = nullptr;
    const void* tvv = nullptr;
    const TX* rin = nullptr;
    const int* prp = nullptr;
    const int* pcv = nullptr;
    const void* pvv = nullptr;
    TX* xpar = nullptr;
    int npar = 0;
};

template <typename TV, typename TP, typename TX>
__global__ __launch_bounds__(kTailThreads) void k_tail_lds(TailRun<TX> R) {
    __shared__ __align__(16) uint8_t sm[kTailLds];
    const TailLayout Y = R.lay;  // by value: a reference to the kernel argument would copy it to scratch
    const int t = (int)threadIdx.x;
    // stage the image: every thread's 16-byte loads in flight together
    {
        const int nv = Y.img_bytes / 16;
        const uint4* src = reinterpret_cast<const uint4*>(R.img);
        uint4* dst = reinterpret_cast<uint4*>(sm);
        constexpr int kU = 16;
        for (int base = 0; base < nv; base += kTailThreads * kU) {
            uint4 tmp[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) tmp[u] = src[min(base + t + u * kTailThreads, nv - 1)];  // clamped: every
            // load is issued (the array stays in registers); stores past the end are skipped
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const int i = base + t + u * kTailThreads;
                if (i < nv) dst[i] = tmp[u];
            }
        }
    }
    const int n = Y.n, nc = Y.nc;
    const int* arp = reinterpret_cast<const int*>(sm + Y.a_rp);
    const uint16_t* acv = reinterpret_cast<const uint16_t*>(sm + Y.a_cv);
    const TV* av = reinterpret_cast<const TV*>(sm + Y.a_v);
    const int* prp = reinterpret_cast<const int*>(sm + Y.p_rp);
    const uint16_t* pcv = reinterpret_cast<const uint16_t*>(sm + Y.p_cv);
    const TP* pv = reinterpret_cast<const TP*>(sm + Y.p_v);
    const int* trp = reinterpret_cast<const int*>(sm + Y.t_rp);
    const uint16_t* tcv = reinterpret_cast<const uint16_t*>(sm + Y.t_cv);
    const TP* tv = reinterpret_cast<const TP*>(sm + Y.t_v);
    const TX* d = reinterpret_cast<const TX*>(sm + Y.d);
    const double* lu = reinterpret_cast<const double*>(sm + Y.lu);
    const int* perm = reinterpret_cast<const int*>(sm + Y.perm);
    TX* b = reinterpret_cast<TX*>(sm + Y.v_b);
    TX* xa = reinterpret_cast<TX*>(sm + Y.v_x);
    TX* xb = reinterpret_cast<TX*>(sm + Y.v_alt);
    TX* r = reinterpret_cast<TX*>(sm + Y.v_r);
    TX* bc = reinterpret_cast<TX*>(sm + Y.v_bc);
    TX* xc = reinterpret_cast<TX*>(sm + Y.v_xc);
    if (R.trp) {  // the parent's restriction: b_I = sum_k P^T_Ik rin_k, products rounded one by one, row order
        const TP* ptv = static_cast<const TP*>(R.tvv);
        for (int I = t; I < n; I += kTailThreads) {
            TX a = (TX)0;
            const int e = R.trp[I + 1];
            for (int k = R.trp[I]; k < e; ++k) a = a + ld<TX, TP>(ptv, k) * R.rin[R.tcv[k]];
            b[I] = a;
        }
    } else {
        for (int i = t; i < n; i += kTailThreads) b[i] = R.b[i];
    }
    TX* prod = reinterpret_cast<TX*>(sm + Y.v_prod);
    __syncthreads();
    const TX w = R.w;
    auto aval = [&](int k) -> TX {
        if constexpr (sizeof(TV) == 2) return bf16_to_float(av[k]);
        else return (TX)av[k];
    };
    // one sparse pass: every product of the operator formed in parallel (each
    // rounded on its own; eight per thread formed before any is stored, so
    // their LDS reads overlap), then every row summed on its own thread in
    // row order: a = init(i); a = acc(a, p_k); fin(i, a)
    auto pass = [&](const int* rp, int nrow, int nnz, auto prodf, auto initf, auto accf, auto finf) {
        constexpr int kU = 8;
        for (int k0 = t; k0 < nnz; k0 += kU * kTailThreads) {
            TX v[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) v[u] = prodf(min(k0 + u * kTailThreads, nnz - 1));
#pragma unroll
            for (int u = 0; u < kU; ++u)
                if (k0 + u * kTailThreads < nnz) prod[k0 + u * kTailThreads] = v[u];
        }
        __syncthreads();
        for (int i = t; i < nrow; i += kTailThreads) {
            auto a = initf(i);
            const int e = rp[i + 1];
#pragma unroll 8
            for (int k = rp[i]; k < e; ++k) a = accf(a, prod[k]);
            finf(i, a);
        }
        __syncthreads();
    };
    const int nnz_a = Y.nnz_a, nnz_p = Y.nnz_p;
    // one Jacobi step out of place: xo = x + (w d) (b - A x), residual by successive subtraction
    auto jacobi = [&](const TX* x, TX* xo) {
        pass(arp, n, nnz_a, [&](int k) { return aval(k) * x[acv[k]]; }, [&](int i) { return b[i]; },
             [](TX a, TX p) { return a - p; }, [&](int i, TX a) { xo[i] = x[i] + (w * d[i]) * a; });
    };
    // pre-smoothing from a zero guess: x1 = (w d) b, then Jacobi steps
    TX* x = xa;
    TX* alt = xb;
    if (R.pre == 0) {
        for (int i = t; i < n; i += kTailThreads) x[i] = (TX)0;
    } else {
        for (int i = t; i < n; i += kTailThreads) x[i] = (w * d[i]) * b[i];
    }
    __syncthreads();
    for (int s = 1; s < R.pre; ++s) {
        jacobi(x, alt);
        TX* q = x; x = alt; alt = q;
    }
    // residual r = b - A x (the row's sum, then the subtraction)
    pass(arp, n, nnz_a, [&](int k) { return aval(k) * x[acv[k]]; }, [](int) { return (TX)0; },
         [](TX a, TX p) { return a + p; }, [&](int i, TX a) { r[i] = b[i] - a; });
    // restriction bc = P^T r (fine rows ascending within each coarse row)
    pass(trp, nc, nnz_p, [&](int k) { return (TX)tv[k] * r[tcv[k]]; }, [](int) { return (TX)0; },
         [](TX a, TX p) { return a + p; }, [&](int I, TX a) { bc[I] = a; });
    // coarsest solve by wave 0: the operations of k_coarse_lu_solve_wave
    if (t < 64) {
        const int i = t;
        double row[64];
        wave_lu_row(row, nc, i, [&](int j) { return lu[j * nc + i]; });
        const double xi = wave_lu_solve(i < nc ? (double)bc[perm[i]] : 0.0, row, nc, i);
        if (i < nc) xc[i] = (TX)xi;
    }
    __syncthreads();
    // prolongation x += P xc, one product at a time
    pass(prp, n, nnz_p, [&](int k) { return (TX)pv[k] * xc[pcv[k]]; }, [&](int i) { return x[i]; },
         [](TX a, TX p) { return a + p; }, [&](int i, TX a) { x[i] = a; });
    for (int s = 0; s < R.post; ++s) {
        jacobi(x, alt);
        TX* q = x; x = alt; alt = q;
    }
    for (int i = t; i < n; i += kTailThreads) R.x[i] = x[i];
    if (R.prp) {  // the parent's prolongation: xpar_i += sum_k P_ik x_k, one product at a time
        const TP* ppv = static_cast<const TP*>(R.pvv);
        for (int i = t; i < R.npar; i += kTailThreads) {
            TX a = R.xpar[i];
            const int e = R.prp[i + 1];
            for (int k = R.prp[i]; k < e; ++k) a = a + ld<TX, TP>(ppv, k) * x[R.pcv[k]];
            R.xpar[i] = a;
        }
    }
}

}  // namespace dev
}  // namespace draugr
