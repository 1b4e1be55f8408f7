// Batched path: many independent systems, each with its own AMG hierarchy
// (from the reference's setup, one per system) and K right-hand sides, solved
// together.  No counterpart in the reference, which solves one system and one
// right-hand side per call (src/cycle.jl:12, :110).
//
// Layout ("packed level table"): level l of the batch is the block-diagonal
// concatenation of level l of every system that has one, so every operator
// pass of the V-cycle is ONE launch over all systems.  Level 0 keeps the
// caller's system order; from level 1 on the systems are ordered by
// decreasing depth, so at every level the systems that still smooth are a
// prefix and the systems whose coarsest (direct) level this is form the
// suffix.  Vectors are row-major n x K (entry (i, j) at i*K + j), so a product
// a_ij * x_j gathers K contiguous values: the cycle's sparse passes are SpMM.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "../host/types.hpp"
#include "dbuf.hpp"
#include "hierarchy.hpp"

namespace draugr {

struct CoarseDesc {         // one system's coarsest level inside a packed level vector
    int32_t row0, n;
    const double* lu;       // column-major LU factors (n x n)
    const int32_t* perm;    // row permutation of the pivoting
};

class BatchedHierarchy {
public:
    BatchedHierarchy(const std::vector<const HostHierarchy*>& hs, const Config& cfg, int nrhs);
    ~BatchedHierarchy();
    BatchedHierarchy(const BatchedHierarchy&) = delete;
    BatchedHierarchy& operator=(const BatchedHierarchy&) = delete;

    int nsys() const { return (int)n0_.size(); }
    int nrhs() const { return K_; }
    int nlevels_max() const { return lmax_; }
    int64_t rows() const { return N0_; }
    int64_t row_offset(int s) const { return off0_[s]; }
    int sys_levels(int s) const { return sys_levels_[s]; }
    int sys_coarse(int s) const { return sys_coarse_[s]; }
    int64_t bytes_resident() const;
    // the autotuner's right-hand sides per thread per level and pass, as text
    std::string tune_report() const;
    // one K-wide Jacobi step on packed level l, out = x + w d (b - A x) (hook)
    void jacobi_pass(int l, Precision p, const void* x, const void* b, void* out, hipStream_t st);
    // out[0..5): smoothing rows nA, vector rows nV, nnz(A), nnz(P), rows of P^T of level l
    void level_counts(int l, int64_t* out) const;

    // x = V(0; b) for every system and right-hand side (packed n x K, FP64).
    void cycle(double* x, const double* b, const CycleOpts& o, hipStream_t st);
    // FP64 PCG with the batched cycle as preconditioner, every (system, rhs)
    // stopping on its own recurrence residual; iters / relres: nsys * K.
    int pcg(double* x, const double* b, double tol, int maxiter, const CycleOpts& o, hipStream_t st,
            int32_t* iters, double* relres);

private:
    struct PCSR {
        int32_t nrow = 0, ncol = 0;
        int64_t nnz = 0;
        DBuf<int32_t> rp, cv;
        DBuf<double> v64;
        DBuf<float> v32;
        // the most nonzeros of one workgroup's rows under the one (cap[0]) and
        // four (cap[1]; -1 for fewer than 4 rhs) rhs-per-thread mappings: the
        // LDS stage of the staged pass kernels
        int32_t cap[2] = {-1, -1};
    };
    struct PLevel {
        int32_t nA = 0;   // rows of the systems that smooth on this level (a prefix)
        int32_t nV = 0;   // rows of the level's vectors (also the suffix of coarsest levels)
        PCSR A, P, Pt;    // P: nA x nV(l+1); Pt: nV(l+1) x nA
        DBuf<double> d64;
        DBuf<float> d32;
        std::vector<CoarseDesc> coarse_wave, coarse_seq;  // coarsest levels on this level (n <= 64 / larger)
        DBuf<CoarseDesc> cw, cs;
        DBuf<double> x64, r64, b64, f64;  // K-wide cycle vectors (level 0: only x, r)
        DBuf<float> x32, r32, b32, f32;
        // right-hand sides per thread of the zero-guess step, restriction,
        // prolongation and Jacobi passes, per precision (the autotuner's
        // choice; negative: the LDS-staged kernel of that mapping) and the
        // time it measured for each
        int rpt[2][4] = {{4, 4, 4, 4}, {4, 4, 4, 4}};
        float rpt_us[2][4] = {};
    };
    template <typename TX> void ensure_vectors();
    template <typename TX> void tune();
    template <typename TX> void cycle_t(TX* f0, const TX* b0, hipStream_t st);
    template <typename TX> void precondition(TX* f0, const TX* b0, hipStream_t st, bool graphs);
    void check_opts(const CycleOpts& o) const;

    int K_ = 1, ks_ = 0;
    int forced_rpt_ = 0;                         // config batch_rhs_per_thread (0: autotuned)
    bool tuned_[2] = {false, false};
    int lmax_ = 0;
    int64_t N0_ = 0;
    double omega_ = 2.0 / 3.0;
    bool l1_ = true;
    std::vector<int32_t> n0_, sys_levels_, sys_coarse_;
    std::vector<int64_t> off0_;
    std::vector<PLevel> L_;                       // lmax_ + 1 levels
    std::vector<DBuf<double>> lu_;                // per system
    std::vector<DBuf<int32_t>> piv_;
    bool have_vec_[2] = {false, false};
    // level-0 blocks for the PCG reductions: never straddle two systems
    int nblk0_ = 0;
    int32_t pq_cap_ = -1;  // the most nonzeros of one reduction block's rows of level 0 (-1: not staged)
    DBuf<int32_t> br0_, br1_, bs_, sysb_;
    // PCG state (per system and rhs) and workspaces
    DBuf<double> st_bb_, st_rr_, st_rz_, st_pq_, st_alpha_, st_beta_, st_tol_, part_;
    DBuf<int32_t> st_active_, st_iters_, st_count_;
    DBuf<double> w_r_, w_p0_, w_p1_, w_q_, w_z_;
    DBuf<float> w_r32_, w_z32_;
    int32_t* pinned_ = nullptr;
    hipStream_t cap_ = nullptr;
    std::map<std::tuple<int, int, const void*, const void*>, hipGraphExec_t> graphs_;
};

}  // namespace draugr
