// Device-resident AMG hierarchy: the level table uploaded once to HBM, and the
// stream-ordered cycle / solve / PCG / resetup drivers that run on it.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <tuple>
#include <string>
#include <vector>

#include "../host/types.hpp"
#include "dbuf.hpp"
#include "galerkin.hpp"
#include "kernels.hpp"
#include "tail.hpp"

namespace draugr {

// Value type of the transfer operators for a value type of A: FP32 when A is
// stored in BF16, else the same.
template <typename TV> struct TransferOf { typedef TV type; };
template <> struct TransferOf<dev::bf16_t> { typedef float type; };

struct DevCSR {
    int32_t nrow = 0, ncol = 0;
    int64_t nnz = 0;
    DBuf<int32_t> rp, cv;
    DBuf<double> v64;
    DBuf<float> v32;
    DBuf<dev::bf16_t> v16;
    DBuf<int32_t> blk;  // CSR-stream row blocks (offsets into rows)
    DBuf<int32_t> bz;   // nonzero offset of each block boundary (rp[blk[b]])
    int nblk = 0;
    int per = 8;        // block budget / kBlock: products per thread
    int kind[3] = {0, 0, 0};  // kernel per cycle precision (dev::OpKernel), set by the autotuner
    bool has16 = false;       // 16-bit column form present (some branch-free block spans < 65536 columns)
    int64_t nnz16 = 0;        // nonzeros of the blocks read through 16-bit columns
    DBuf<uint16_t> c16;       // column - (the block's base) per nonzero
    DBuf<dev::BlkHdr> hdr;    // per block: first row, first nonzero, 16-bit column base (-1: 32-bit columns)
    // Column-sorted form (kernel kinds sorted, sortedpipe, sortedquad) of the
    // branch-free blocks spanning fewer than 2^20 columns: its own block
    // headers (base = the block's first column, -1 for the other blocks, which
    // keep row order and read the 32-bit columns), the values in column order
    // and per entry one word, column offset | the entry's position in the
    // block's row order << 20.  Allocated to a multiple of 4 entries
    // (four-entry loads).
    bool hasS = false;
    int64_t nnzS = 0;          // nonzeros of the blocks that have the sorted form
    DBuf<dev::BlkHdr> shdr;
    DBuf<uint32_t> scs;
    DBuf<double> sv64;
    DBuf<float> sv32;
    DBuf<dev::bf16_t> sv16;
    // Pattern form (kernel kind pattern), present when the rows use at most
    // dev::kPatMax column patterns relative to the row, none longer than
    // dev::kPatSlots entries: per row its pattern index, per pattern
    // [length, col - row ...] in tw + 1 ints, the values in 64-row slices
    // (dev::pattern_slot), refreshed with every new set of values.
    bool hasT = false;
    int tw = 0, npat = 0;
    DBuf<uint16_t> tpid;
    DBuf<int32_t> ttab;
    DBuf<double> tv64;
    DBuf<float> tv32;
    DBuf<dev::bf16_t> tv16;
};

struct DevLevel {
    DevCSR A;                 // level operator
    DevCSR P;                 // prolongation n_l x n_{l+1}
    DevCSR Pt;                // P^T as CSR (values gathered from P)
    DBuf<int32_t> pt_pnz;     // Pt entry -> P entry
    // Chebyshev smoother: eigenvalue bounds of D^-1 A, power-iteration vectors,
    // the direction vector d of the recurrence per precision
    double cheb_lmin = 0.0, cheb_lmax = 0.0;
    DBuf<double> pv0, pv, pw, cd64;
    DBuf<float> cd32;
    // ILU(0) smoother: level schedules (host offsets, device orders), diagonal
    // positions, factors on A's pattern, the host pattern for refactorization
    IluStructure ilu;
    DBuf<int32_t> ilu_diag, ilu_fwd, ilu_bwd;
    DBuf<double> ilu_L64, ilu_U64;
    DBuf<float> ilu_L32, ilu_U32;
    std::vector<int32_t> h_rp, h_cv;
    bool has_rmap = false;
    std::shared_ptr<DeviceRMap> rmap;  // Galerkin triple map of A_{l+1} (shared with the host level when built on the GPU)
    DBuf<double> d64;         // smoother inverse diagonal (plain or l1)
    DBuf<float> d32;
    int ncolors = 0;
    std::vector<int32_t> color_off;  // host copy (loop control)
    DBuf<int32_t> color_order;
    bool fuse_zero[3] = {false, false, false};  // zero-guess step fused with its residual (autotuned)
    bool fuse_down[3] = {false, false, false};  // zero-guess step + residual + restriction in one pass (autotuned)
    int down_state = 0;       // blocks of the one-pass down sweep: 0 not built, 1 built, -1 not possible
    int ndblk = 0;
    DBuf<int32_t> dblk;       // coarse-row offsets of its blocks
    bool fuse_up[3] = {false, false, false};    // prolongation + first post-smoothing step in one pass (autotuned)
    // row chunks of A, P and P^T for the one-launch tail (tail_ok: every row
    // of the three fits one LDS stage and the level is small enough)
    float tA_us[3][dev::kOpKinds] = {};  // the autotuner's time per kernel kind of the Jacobi pass over A
    // workspaces: residual, ping-pong smoother buffer, coarse x and b
    DBuf<double> r64, t64, xc64, bc64;
    DBuf<float> r32, t32, xc32, bc32;
};

// Device scalars used by the fused reductions of PCG / stationary solve.
enum Scalar : int {
    S_RZ = 0, S_PQ, S_ALPHA, S_BETA, S_RR, S_BB, S_TMP, S_RHO, S_COUNT_ = 16
};

// Device memory for the PCG's scalars (S_COUNT_ doubles) and reduction
// partials (partial_cap doubles) supplied by the caller instead of the
// hierarchy's own buffers (the allocation-free entry point).
struct PcgScratch {
    double* scal;
    double* partial;
    int64_t partial_cap;
};

struct CycleOpts {
    bool w_cycle = false;
    int pre = 1, post = 1;
    Precision prec = Precision::FP64;
    bool graphs = true;  // zero-guess cycles on fixed buffers replayed from a hipGraph
    bool x1_ready0 = false;  // the finest level's first step from zero is already in its work vector
    bool gs_sym = false;     // coloured GS post-smoothing sweeps the colours in reverse order
};

// Kernels shared with the batched path (batch.hip):
// dense LU with partial pivoting in place (column-major n x n; piv[0..n) the
// LAPACK pivots, piv[n..2n) the row permutation; *flag = 1 if singular), and
// the plain / l1 inverse diagonal of a CSR operator.
// stepwise: one launch pair per column over the whole chip (large n); the
// factors are the same bits either way.
void dense_lu_device(int n, double* a, int* piv, int* flag, hipStream_t st, bool stepwise = false);
// mode 0: 1 / a_ii (src/smoothers.jl:47-63), 1: 1 / sum |a_ij| (:816-827),
// 2: SPAI(0) a_ii / sum a_ij^2 (:502-532).
void invdiag_device(int n, const int* rp, const int* cv, const double* av, double* d, int mode, hipStream_t st);

// Differentiation kernels (adjoint.hip)
void assemble_7pt(int m, const double* k, const int* rp, double* nz, hipStream_t st);
void matrix_grad(int n, const int* rp, const int* cv, const double* x, const double* lam, double* g, hipStream_t st);
void grad_logk_7pt(int m, const double* k, const double* x, const double* lam, double* g, hipStream_t st);

class DeviceHierarchy {
public:
    DeviceHierarchy(const HostHierarchy& h, const Config& cfg);
    ~DeviceHierarchy();

    int32_t n_fine() const { return n0_; }
    // finest operator (device): row pointer, columns, nnz
    const int32_t* finest_rp() const { return L_.empty() ? Ac_.rp.p : L_[0].A.rp.p; }
    const int32_t* finest_cv() const { return L_.empty() ? Ac_.cv.p : L_[0].A.cv.p; }
    int64_t finest_nnz() const { return L_.empty() ? Ac_.nnz : L_[0].A.nnz; }
    int nlevels() const { return (int)L_.size(); }
    int level_blocks(int l) const { return l < nlevels() ? L_[l].A.nblk : Ac_.nblk; }
    bool level_has16(int l) const { return l < nlevels() ? L_[l].A.has16 : Ac_.has16; }
    int64_t level_nnz_sorted(int l) const { return l < nlevels() ? L_[l].A.nnzS : Ac_.nnzS; }
    hipStream_t stream() const { return stream_; }

    // One V/W cycle on device vectors x (in/out) and b; zero_guess asserts x == 0
    // on entry (its content is ignored then).
    void cycle(double* x, const double* b, bool zero_guess, const CycleOpts& o, hipStream_t st);
    // Stationary AMG iteration (src/cycle.jl:110-160); returns the iteration count.
    int solve(double* x, const double* b, double tol, int maxiter, const CycleOpts& o, hipStream_t st,
              double* relres, int verbose);
    // FP64 PCG preconditioned by one cycle; returns the iteration count.
    int pcg(double* x, const double* b, double tol, int maxiter, const CycleOpts& o, hipStream_t st,
            double* relres, int verbose, const PcgScratch* ws = nullptr);
    // The same PCG with its scalars and partials in the library's static
    // device scratch: allocates nothing (see pcg); calls that overlap on one
    // GPU race on that scratch.
    int pcg_no_alloc(double* x, const double* b, double tol, int maxiter, const CycleOpts& o, hipStream_t st,
                     double* relres, int verbose);
    int device() const { return device_; }
    // Partial numeric resetup from new fine-level values (device pointer, FP64).
    void resetup(const double* nzval_dev, hipStream_t st);

    // ── test / measurement hooks ─────────────────────────────────────────────
    // Autotuner report: per level and precision the kernel picked for A, P^T, P
    // and the zero-guess fusion, as text.
    std::string tune_report() const;
    // op: 0 y=A x, 1 r=b-Ax, 2 jacobi step xo=x+w d(b-Ax), 3 zero-guess jacobi
    // + residual (x1, r), 4 restrict bc=P^T r, 5 prolong x+=P xc, 6 GS sweep
    // (in place, all colours), 7 coarse solve, 8 galerkin (A_{l+1} values),
    // 9 inverse diagonal (returned in out)
    void level_op(int lvl, int op, int prec, const void* in1, const void* in2, void* out, void* out2,
                  hipStream_t st);
    void export_level_values(int lvl, int which, double* host_dst);  // 0 A vals, 1 dinv, 2 coarse LU, 3 coarsest vals
    int64_t bytes_resident() const;
    double cycle_bytes(const CycleOpts& o) const;  // algorithmic bytes of one V/W cycle (zero guess)
    // The PCG's preconditioner application (zero-guess cycle, replayed from its
    // graph when o.graphs) `reps` times back to back on the PCG's own residual
    // and correction vectors, the residual set from b first (measurement hook)
    void precondition_repeat(const double* b, const CycleOpts& o, int reps, hipStream_t st);
    // bytes of one pass over operator which (0 A, 1 P, 2 P^T) of level l at
    // precision p: values, column indices, row pointers (vectors excluded);
    // as_stored: indices in the form the tuned kernel reads (16-bit where it
    // chose that form), else int32 (the algorithmic count)
    double op_bytes(int l, int which, Precision p, bool as_stored = true) const;
    void ensure_precision(Precision p);
    // Stream ordering between calls: work a call enqueues on stream st waits
    // for everything the previous call enqueued (possibly on another stream),
    // since all calls share the hierarchy's workspaces.
    void begin(hipStream_t st);
    void end(hipStream_t st);

private:
    template <typename TV, typename TX>
    void cycle_level(int l, TX* x, const TX* b, bool zero_guess, const CycleOpts& o, hipStream_t st,
                     bool x1_ready = false);
    // reverse: coloured GS sweeps the colours in reverse order (the
    // post-smoothing of the symmetric option)
    template <typename TV, typename TX>
    void smooth(int l, TX*& x, TX*& alt, const TX* b, int steps, bool& zero, bool fuse_resid, bool& resid_done,
                hipStream_t st, bool x1_ready = false, bool reverse = false);
    template <typename TV> bool zero_step_is_map(int l, int pre) const;
    bool down_blocks(int l);
    template <typename TV, typename TX>
    void launch_down(int l, const TX* b, TX* x1, TX* bc, const TX* dc, TX* x1c, hipStream_t st);
    // global_x: keep the vector in a global workspace (the form for a coarsest
    // level whose vector does not fit in LDS)
    template <typename TX> void coarse_solve(TX* x, const TX* b, hipStream_t st, bool global_x = false);
    template <typename TV, typename TX> void cycle_top(TX* x, const TX* b, bool zero_guess, const CycleOpts& o,
                                                       hipStream_t st);
    // zero-guess cycle x = V(0; b) on buffers that stay fixed for the
    // hierarchy's life: captured once into a hipGraph, then replayed
    // capture = false: a cycle with no graph yet is launched eagerly instead
    template <typename TV, typename TX>
    void precondition(TX* x, const TX* b, const CycleOpts& o, hipStream_t st, bool capture = true);
    void build(const HostHierarchy& h);
    void release_handles();
    void refresh_level_mirrors(int l, hipStream_t st);
    template <typename TV, typename TX> void tune();
    // the LDS-resident last level (with the coarse solve) in one launch: the
    // level it covers per cycle precision (-1: none), its packed image and
    // layout per precision, and the tuner's measurements
    int tail_lt_[3] = {-1, -1, -1};
    int forced_tail_ = -1;  // config "tail": -1 tuned, 0 never, 1 whenever the last level qualifies
    float tail_us_[3] = {0.f, 0.f, 0.f}, notail_us_[3] = {0.f, 0.f, 0.f};
    DBuf<uint8_t> tail_img_[3];
    dev::TailLayout tail_lay_[3] = {};
    template <typename TV, typename TX> bool tail_fits(dev::TailLayout* lay) const;
    template <typename TV, typename TX> void pack_tail(hipStream_t st);
    void refresh_tails(hipStream_t st);
    template <typename TV, typename TX> void tune_tail();
    template <typename TV, typename TX>
    void launch_tail(const CycleOpts& o, hipStream_t st, const TX* rin = nullptr, TX* xpar = nullptr);
    void compute_dinv(int l, hipStream_t st);
    double estimate_rho(int l, hipStream_t st);  // power iteration on D^-1 A (FP64)
    void chebyshev_bounds(hipStream_t st);
    void ilu_factor(int l, const std::vector<double>& values, hipStream_t st);  // host factorization, upload
    template <typename TV, typename TX>
    void ilu_smooth(int l, TX* x, TX* tmp, const TX* b, int steps, bool& zero, hipStream_t st);
    template <typename TV, typename TX>
    void chebyshev(int l, TX*& x, TX*& alt, const TX* b, int steps, bool& zero, hipStream_t st);
    double weight() const { return smoother_ == Smoother::SPAI0 ? 1.0 : omega_; }
    void galerkin(int l, hipStream_t st);
    void coarse_factor(hipStream_t st, bool stepwise = false);
    double reduce_to_host(hipStream_t st, int slot);
    double reduce_to_host(hipStream_t st, const double* scal, int slot);

    std::vector<DevLevel> L_;
    int32_t n0_ = 0;
    int32_t nc_ = 0;
    DevCSR Ac_;                        // coarsest operator (sparse; values by Galerkin)
    DBuf<double> clu_;                 // dense LU factors of the coarsest operator (column-major)
    DBuf<int32_t> flag_;               // [0] singular flag, [1..n] pivots, [n+1..2n] permutation
    DBuf<double> cx_;                  // coarse solve vector when it does not fit in LDS
    Smoother smoother_;
    double omega_;
    uint64_t seed_ = 0;
    int forced_kind_ = -1;  // config "kernel": every sparse pass uses this kernel (no tuning)
    int forced_fuse_ = -1;  // config "fuse": -1 tuned, 0 never, 1 always (Jacobi-type smoothers)
    int pcg_form_ = -1;     // form of the PCG's q = A p pass: -1 not timed yet, else stream / sorted / pattern
    int pcg_form(const DevCSR& A, hipStream_t st);
    bool have32_ = false, have16_ = false;
    bool tuned_[3] = {false, false, false};
    // solver workspaces at the finest size
    DBuf<double> w_r_, w_p0_, w_p1_, w_q_, w_z_, w_x_;
    DBuf<float> w_r32_, w_z32_;
    DBuf<double> scal_, partial_;
    // mapped, coherent host memory: [0, 16) scalars read back by the host,
    // [kPinPart, kPinPart + kRedBlocks) the per-block partials of the PCG's
    // r.r, which the host sums itself
    double* pinned_ = nullptr;
    double* pinned_dev_ = nullptr;  // its device address
    static constexpr int kPinPart = 16;
    double host_finalize(int nblocks) const;
    hipStream_t stream_ = nullptr;
    hipEvent_t last_ev_ = nullptr;
    hipEvent_t pcg_ev_ = nullptr;  // marks each PCG iteration's r.r (the host waits on it)
    int device_ = 0;               // the HIP device the hierarchy lives on
    hipStream_t cap_ = nullptr;  // capture stream for the graphs
    std::map<std::tuple<int, int, int, int, int, int, const void*, const void*>, hipGraphExec_t> graphs_;
    hipStream_t last_st_ = nullptr;
    bool last_valid_ = false;
};

}  // namespace draugr
