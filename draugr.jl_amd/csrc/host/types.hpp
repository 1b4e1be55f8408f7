// Host-side data layout of the AMG hierarchy and its configuration.
//
// Everything here is 0-based, int32 indices, FP64 values.  The layout mirrors
// the reference's level contract (src/types.jl:368-467): per level the operator
// A (CSR, columns sorted within rows), the prolongation P (CSR, n_fine x n_coarse),
// the transpose map of P (coarse row -> fine rows, P nz index), the Galerkin
// triple map grouped by destination coarse nonzero, and the smoother colouring.
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace draugr {

struct Error : std::runtime_error {
    explicit Error(const std::string& m) : std::runtime_error(m) {}
};

// Compressed sparse row matrix, 0-based.
struct CSR {
    int32_t nrow = 0, ncol = 0;
    std::vector<int32_t> rp;   // nrow + 1
    std::vector<int32_t> cv;   // nnz
    std::vector<double> nz;    // nnz
    int64_t nnz() const { return (int64_t)cv.size(); }
};

// ── configuration (field set and defaults follow src/types.jl:497-536) ──────
enum class Interp : int { Direct = 0, Standard = 1, ExtendedI = 2 };
enum class Coarsen : int {
    Aggregation = 0, PMIS = 1, HMIS = 2, RS = 3,
    AggressivePMIS = 4, AggressiveHMIS = 5, SmoothedAggregation = 6,
    AggressiveAggregation = 7  // AggressiveCoarsening with a base other than :pmis/:hmis
};
enum class Smoother : int {
    Jacobi = 0, ColoredGS = 1, SerialGS = 2, SPAI0 = 3, SPAI1 = 4, L1Jacobi = 5,
    L1ColoredGS = 6, L1SerialGS = 7, Chebyshev = 8, ILU0 = 9, SerialILU0 = 10
};
enum class Precision : int { FP64 = 0, FP32 = 1, BF16 = 2 };

struct InterpSpec {
    Interp type = Interp::ExtendedI;
    double trunc_factor = 0.0;
    int max_elements = 0;
    int norm_p = 1;
    bool rescale = false;
};

struct CoarsenSpec {
    Coarsen type = Coarsen::HMIS;
    double theta = 0.5;
    InterpSpec interp;
    bool filtering = false;     // aggregation / smoothed aggregation
    double filter_tol = 0.1;
    double sa_omega = 2.0 / 3.0;  // smoothed aggregation damping
};

struct Config {
    CoarsenSpec coarsening;
    Smoother smoother = Smoother::L1ColoredGS;
    int max_levels = 20;
    int max_coarse_size = 50;
    int pre_smoothing_steps = 1;
    int post_smoothing_steps = 1;
    double jacobi_omega = 2.0 / 3.0;
    int verbose = 0;
    CoarsenSpec initial_coarsening;
    int initial_coarsening_levels = 0;
    double max_row_sum = 1.0;
    bool w_cycle = false;
    bool signed_strength = false;  // stored only: coarsening always uses absolute strength
    bool coarse_solve_on_cpu = false;
    // ── extensions of this backend (no counterpart in the reference config) ──
    Precision precision = Precision::FP64;  // arithmetic of the V/W cycle
    uint64_t seed = 0x5EEDull;              // seed of the setup RNG (PMIS/HMIS measures)
    int graphs = -1;                        // -1 auto (batched PCG: replayed graphs; one system: eager launches),
                                            // 1: replay the preconditioner cycle / PCG iteration from a captured
                                            // hipGraph, 0: launch eagerly
    int fuse = -1;                          // zero-guess step + residual, one-pass down / up sweeps: -1 autotuned, 0 never, 1 always
    int kernel = -1;                        // sparse-pass kernel: -1 autotuned, 0 stream, 1 row-sequential, 2 pipelined,
                                            // 3 stream with 16-bit columns, 4 column-sorted, 5 pipelined 16-bit,
                                            // 6 pipelined column-sorted, 7 column-sorted with four-entry loads,
                                            // 8 vector CSR (FP32 / BF16 only)
    bool gpu_setup = true;                  // interpolation, coarse operators and triple maps built on the GPU
    int gpu_setup_min_rows = 32768;         // levels below this size interpolate on the host (faster there)
    bool gs_symmetric = false;              // coloured GS post-smoothing sweeps the colours in reverse order
    int batch_rpt = 0;                      // batched passes: right-hand sides per thread (0: autotuned; 1, 4, 8, 16;
                                            // -1, -4: the LDS-staged kernels of 1 and 4)
    int tail = -1;                          // small levels + coarse solve in one launch: -1 autotuned, 0 never, 1 deepest
};

// ── hierarchy (host copy; uploaded once to HBM) ──────────────────────────────
struct TransposeMap {            // src/types.jl:394-398
    std::vector<int32_t> offsets;    // n_coarse + 1
    std::vector<int32_t> fine_rows;  // nnz(P)
    std::vector<int32_t> p_nz_idx;   // nnz(P)
};

struct RestrictionMap {          // src/types.jl:413-418
    std::vector<int64_t> nz_offsets; // nnz(A_coarse) + 1
    std::vector<int32_t> pi, ai, pj; // ntriples each
    // built on the GPU: the map stays in HBM (an opaque device object shared
    // with the device hierarchy) and the host vectors above stay empty
    std::shared_ptr<void> device;
    int64_t device_ntriples = 0;
    int64_t ntriples() const { return device ? device_ntriples : (int64_t)pi.size(); }
};

struct Coloring {                // src/types.jl:206-228
    int num_colors = 0;
    std::vector<int32_t> colors;        // 1-based colour per row
    std::vector<int32_t> color_offsets; // num_colors + 1 (0-based offsets into order)
    std::vector<int32_t> color_order;   // rows sorted by colour
};

struct HostLevel {               // src/types.jl:431-440
    CSR A;
    CSR P;
    TransposeMap Pt;
    bool has_rmap = false;
    RestrictionMap R;
    Coloring coloring;           // only for the coloured Gauss-Seidel smoothers
};

struct HostHierarchy {           // src/types.jl:457-467
    std::vector<HostLevel> levels;
    CSR A_coarsest;                   // sparse coarsest operator (n_c x n_c)
    int32_t n_coarse = 0;
    std::vector<double> coarse_dense; // column-major n_c x n_c
    double setup_seconds = 0.0;
};

// Deterministic generator for the setup's random measures: SplitMix64 with
// 53-bit uniforms in [0, 1).
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

// ── setup entry points (host) ────────────────────────────────────────────────
std::vector<uint8_t> strength_absolute(const CSR& A, double theta);
CSR apply_max_row_sum(const CSR& A, double threshold);

// PMIS sweeps (src/coarsening.jl:202-283, 575-624) over a graph: the nodes
// with cf == 0 are decided by snapshot sweeps (select C, then mark F) over row
// i's entries off[i]..off[i+1] of adj that are strong (s[k] != 0; every entry
// when s is null), a row's own index never counting; with low_measure_fine an
// undecided node of measure < 1 becomes F first (HMIS's second phase).  Nodes
// still undecided after n sweeps keep 0.  pmis_sweeps is the host's.
using PmisFn = void (*)(std::vector<int8_t>& cf, int32_t n, const int64_t* off, const int32_t* adj, const uint8_t* s,
                        const std::vector<double>& measure, bool low_measure_fine);
void pmis_sweeps(std::vector<int8_t>& cf, int32_t n, const int64_t* off, const int32_t* adj, const uint8_t* s,
                 const std::vector<double>& measure, bool low_measure_fine);
// Sweeps of graphs of at least min_rows nodes by fn (null: the host's).
struct PmisHook {
    PmisFn fn = nullptr;
    int32_t min_rows = 0;
};

void coarsen_aggregation(const CSR& A, double theta, std::vector<int32_t>& agg, int32_t& nc);
void coarsen_aggressive_agg(const CSR& A, double theta, Rng& rng, std::vector<int32_t>& agg, int32_t& nc,
                            const PmisHook& ph = PmisHook());
void coarsen_pmis(const CSR& A, double theta, Rng& rng, std::vector<int8_t>& cf, int32_t& nc,
                  const PmisHook& ph = PmisHook());
void coarsen_hmis(const CSR& A, double theta, Rng& rng, std::vector<int8_t>& cf, int32_t& nc,
                  const PmisHook& ph = PmisHook());
void coarsen_rs(const CSR& A, double theta, std::vector<int8_t>& cf, int32_t& nc);
void coarsen_aggressive_cf(const CSR& A, double theta, bool hmis_base, Rng& rng,
                           std::vector<int8_t>& cf, int32_t& nc, const PmisHook& ph = PmisHook());

CSR prolongation_from_aggregates(const std::vector<int32_t>& agg, int32_t nc);
CSR smooth_prolongation(const CSR& A, const CSR& Ptent, double omega);
CSR filter_prolongation(const CSR& P, double tol);
// Work arrays of ext_i_row (pmark / fmark: n entries, -1 where unused).
struct ExtIWork {
    std::vector<int32_t> pmark, fmark, chat, keep;
    std::vector<double> pdata;
    explicit ExtIWork(int32_t n) : pmark(n, -1), fmark(n, -1) {}
};
void ext_i_row(const CSR& A, const std::vector<int8_t>& cf, const std::vector<int32_t>& cmap,
               const std::vector<uint8_t>& s, const InterpSpec& ip, int32_t i, ExtIWork& w,
               std::vector<std::pair<int32_t, double>>& out);
CSR interpolation_cf(const CSR& A, const std::vector<int8_t>& cf, int32_t nc,
                     const InterpSpec& ip, double theta);

void coarse_sparsity(const CSR& A, const CSR& P, int32_t nc, CSR& Ac, RestrictionMap* rmap);
void build_transpose_map(const CSR& P, TransposeMap& t);
void greedy_coloring(const CSR& A, Coloring& c);

// ILU(0) (src/smoothers.jl:1016-1098, 1228-1307): diagonal positions and
// the forward / backward level schedules of a CSR operator, and its factors
// L (strict lower, unit diagonal implied) and U (upper with diagonal) on A's
// pattern.
struct IluStructure {
    std::vector<int32_t> diag_idx;
    std::vector<int32_t> fwd_order, fwd_offsets, bwd_order, bwd_offsets;
};
void ilu0_structure(const CSR& A, IluStructure& s);
void ilu0_factorize(const CSR& A, const std::vector<int32_t>& diag_idx, std::vector<double>& L,
                    std::vector<double>& U);

// Builder of a level's coarse operator Ac = P^T A P, its triple map (when
// rmap is not null) and the transpose map of P.  Null = the host's
// coarse_sparsity + build_transpose_map.
using GalerkinFn = void (*)(const CSR& A, const CSR& P, int32_t nc, CSR& Ac, RestrictionMap* rmap, TransposeMap& T);
// Builder of P from a C/F splitting, followed by filter_prolongation(P,
// ip.trunc_factor) when truncate is set; levels of fewer than min_rows rows
// may use the host's routines.  Null = the host's routines.
using InterpFn = CSR (*)(const CSR& A, const std::vector<int8_t>& cf, int32_t nc, const InterpSpec& ip, double theta,
                         bool truncate, int32_t min_rows);
struct SetupHooks {
    GalerkinFn galerkin = nullptr;
    InterpFn interp = nullptr;
    PmisFn pmis = nullptr;  // PMIS sweeps of C/F splittings of >= cfg.gpu_setup_min_rows rows
};
HostHierarchy amg_setup_host(const CSR& A, const Config& cfg, bool allow_partial_resetup,
                             const SetupHooks* hooks = nullptr);
// Hierarchies of independent systems, built concurrently by a few host
// threads (each with its share of the OpenMP threads); the same hierarchies
// as one amg_setup_host call per system.  The first failure (by system
// index) is rethrown after every worker has finished.
std::vector<HostHierarchy> amg_setup_host_many(const std::vector<CSR>& As, const Config& cfg,
                                               bool allow_partial_resetup, const SetupHooks* hooks = nullptr);
std::string hierarchy_summary(const HostHierarchy& h, const Config& cfg, int64_t n_finest);

// Dense LU with partial pivoting (column-major), and its explicit inverse.
// Returns false when the matrix is singular to working precision.
bool dense_lu_inverse(int n, const double* A_colmajor, double* inv_colmajor);

// JSON configuration (keys of c_api/draugr_amg.h:58-89 plus this backend's).
Config config_from_json(const char* json);

}  // namespace draugr
