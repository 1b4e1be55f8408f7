// PMIS sweeps on the GPU at setup; src/coarsening.jl:202-283 (PMIS) and
// 575-624 (HMIS's second phase), and the sweeps of aggressive coarsening's
// second stage over the graph of C points.
//
// A sweep is two passes over the undecided nodes (cf == 0), each reading the
// state the previous pass left (a snapshot) and writing a second array:
//   select: a node whose strong neighbours that are not F all have a measure
//           no larger than its own becomes C (HMIS: a node of measure < 1
//           becomes F first);
//   fine:   a node with a strong C neighbour becomes F.
// Marking C only turns 0 into 1 and the select test asks "not F", marking F
// only turns 0 into -1 and the fine test asks "is C", so the result does not
// depend on the order the nodes are visited in: it is the host's, node for
// node.  Sweeps repeat until a select pass finds no undecided node (one host
// read of a flag per sweep; a few sweeps on the hierarchies here).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "coarsen.hpp"
#include "dbuf.hpp"

namespace draugr {

namespace {

constexpr int kCB = 256;

inline unsigned grid_of(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kCB - 1) / kCB); }

// nxt = cf on entry; writes nxt[i] for undecided i, sets *any when one exists
__global__ __launch_bounds__(kCB) void k_pmis_select(int32_t n, const int64_t* __restrict__ off,
                                                     const int32_t* __restrict__ adj, const uint8_t* __restrict__ s,
                                                     const double* __restrict__ meas, const int8_t* __restrict__ cf,
                                                     int8_t* __restrict__ nxt, int low_fine, int* __restrict__ any) {
    const int32_t i = blockIdx.x * kCB + threadIdx.x;
    if (i >= n || cf[i] != 0) return;
    *any = 1;
    const double mi = meas[i];
    if (low_fine && mi < 1.0) {
        nxt[i] = -1;
        return;
    }
    for (int64_t k = off[i]; k < off[i + 1]; ++k) {
        const int32_t j = adj[k];
        if (j != i && (s == nullptr || s[k]) && cf[j] != -1 && meas[j] > mi) return;
    }
    nxt[i] = 1;
}

__global__ __launch_bounds__(kCB) void k_pmis_fine(int32_t n, const int64_t* __restrict__ off,
                                                   const int32_t* __restrict__ adj, const uint8_t* __restrict__ s,
                                                   const int8_t* __restrict__ cf, int8_t* __restrict__ nxt) {
    const int32_t i = blockIdx.x * kCB + threadIdx.x;
    if (i >= n || cf[i] != 0) return;
    for (int64_t k = off[i]; k < off[i + 1]; ++k)
        if ((s == nullptr || s[k]) && cf[adj[k]] == 1) {
            nxt[i] = -1;
            return;
        }
}

__global__ __launch_bounds__(kCB) void k_copy_i8(int32_t n, const int8_t* __restrict__ a, int8_t* __restrict__ b) {
    const int32_t i = blockIdx.x * kCB + threadIdx.x;
    if (i < n) b[i] = a[i];
}

__global__ void k_clear_flag(int* __restrict__ f) {
    if (threadIdx.x == 0) *f = 0;
}

}  // namespace

void pmis_sweeps_device(std::vector<int8_t>& cf, int32_t n, const int64_t* off, const int32_t* adj,
                        const uint8_t* s, const std::vector<double>& measure, bool low_measure_fine) {
    if (n <= 0) return;
    if ((int64_t)cf.size() != n || (int64_t)measure.size() != n) throw Error("pmis_sweeps_device: size mismatch");
    const int64_t nnz = off[n];
    hipStream_t st;
    DRAUGR_HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct StreamOwner {
        hipStream_t s;
        ~StreamOwner() { (void)hipStreamDestroy(s); }
    } owner{st};
    DBuf<int64_t> d_off;
    DBuf<int32_t> d_adj, d_any;
    DBuf<uint8_t> d_s;
    DBuf<double> d_meas;
    DBuf<int8_t> d_cf, d_nxt;
    int* pinned = nullptr;  // the sweep's flag, read by the host each pass
    DRAUGR_HIP_CHECK(hipHostMalloc((void**)&pinned, sizeof(int), hipHostMallocDefault));
    struct PinnedOwner {
        int* p;
        ~PinnedOwner() { (void)hipHostFree(p); }
    } pinned_owner{pinned};
    struct SyncOnExit {  // runs before the buffers above are freed, also when a step throws
        hipStream_t s;
        ~SyncOnExit() { (void)hipStreamSynchronize(s); }
    } sync_on_exit{st};
    d_off.upload(off, (size_t)n + 1, st);
    d_adj.upload(adj, (size_t)nnz, st);
    if (s) d_s.upload(s, (size_t)nnz, st);
    d_meas.upload(measure, st);
    d_cf.upload(cf, st);
    d_nxt.upload(cf, st);
    d_any.alloc(1);
    const uint8_t* sp = s ? d_s.p : nullptr;
    for (int64_t iter = 0; iter <= (int64_t)n; ++iter) {
        hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(64), 0, st, d_any.p);
        hipLaunchKernelGGL(k_pmis_select, dim3(grid_of(n)), dim3(kCB), 0, st, n, d_off.p, d_adj.p, sp, d_meas.p,
                           d_cf.p, d_nxt.p, (int)low_measure_fine, d_any.p);
        hipLaunchKernelGGL(k_copy_i8, dim3(grid_of(n)), dim3(kCB), 0, st, n, d_nxt.p, d_cf.p);
        hipLaunchKernelGGL(k_pmis_fine, dim3(grid_of(n)), dim3(kCB), 0, st, n, d_off.p, d_adj.p, sp, d_cf.p,
                           d_nxt.p);
        hipLaunchKernelGGL(k_copy_i8, dim3(grid_of(n)), dim3(kCB), 0, st, n, d_nxt.p, d_cf.p);
        DRAUGR_HIP_CHECK(hipMemcpyAsync(pinned, d_any.p, sizeof(int), hipMemcpyDeviceToHost, st));
        DRAUGR_HIP_CHECK(hipStreamSynchronize(st));
        if (*pinned == 0) break;
    }
    DRAUGR_HIP_CHECK(hipMemcpyAsync(cf.data(), d_cf.p, (size_t)n, hipMemcpyDeviceToHost, st));
    DRAUGR_HIP_CHECK(hipStreamSynchronize(st));
}

}  // namespace draugr
