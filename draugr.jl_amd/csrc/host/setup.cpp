// Host-side hierarchy construction (setup phase, run once; the result is
// uploaded to HBM and stays resident).
#include "types.hpp"

#include <omp.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <exception>
#include <sstream>
#include <thread>

namespace draugr {

namespace {

bool uses_cf(Coarsen c) {
    return c == Coarsen::PMIS || c == Coarsen::HMIS || c == Coarsen::RS ||
           c == Coarsen::AggressivePMIS || c == Coarsen::AggressiveHMIS;
}

// Coarsen once and build P; src/setup.jl:256-327.
CSR coarsen_and_build_p(const CSR& A, const CoarsenSpec& alg, const Config& cfg, Rng& rng, int32_t& nc,
                        const SetupHooks& hooks) {
    const bool weaken = cfg.max_row_sum < 1.0;
    CSR Aw;
    if (weaken) Aw = apply_max_row_sum(A, cfg.max_row_sum);
    const CSR& As = weaken ? Aw : A;  // matrix the splitting / aggregation sees
    const PmisHook ph{hooks.pmis, cfg.gpu_setup_min_rows};
    if (uses_cf(alg.type)) {
        std::vector<int8_t> cf;
        const auto t0 = std::chrono::steady_clock::now();
        switch (alg.type) {
            case Coarsen::PMIS: coarsen_pmis(As, alg.theta, rng, cf, nc, ph); break;
            case Coarsen::HMIS: coarsen_hmis(As, alg.theta, rng, cf, nc, ph); break;
            case Coarsen::RS: coarsen_rs(As, alg.theta, cf, nc); break;
            case Coarsen::AggressivePMIS: coarsen_aggressive_cf(As, alg.theta, false, rng, cf, nc, ph); break;
            case Coarsen::AggressiveHMIS: coarsen_aggressive_cf(As, alg.theta, true, rng, cf, nc, ph); break;
            default: break;
        }
        const auto t1 = std::chrono::steady_clock::now();
        const bool truncate = alg.interp.trunc_factor > 0;
        CSR P;
        if (hooks.interp) P = hooks.interp(A, cf, nc, alg.interp, alg.theta, truncate, cfg.gpu_setup_min_rows);
        else P = interpolation_cf(A, cf, nc, alg.interp, alg.theta);
        const auto t2 = std::chrono::steady_clock::now();
        if (truncate && !hooks.interp) P = filter_prolongation(P, alg.interp.trunc_factor);
        if (cfg.verbose >= 3) {
            const auto sec = [](auto a, auto b) { return std::chrono::duration<double>(b - a).count(); };
            std::printf("    C/F splitting %.3fs, interpolation %.3fs, truncation %.3fs\n", sec(t0, t1), sec(t1, t2),
                        sec(t2, std::chrono::steady_clock::now()));
        }
        return P;
    }
    std::vector<int32_t> agg;
    if (alg.type == Coarsen::AggressiveAggregation) {
        coarsen_aggressive_agg(As, alg.theta, rng, agg, nc, ph);
        return prolongation_from_aggregates(agg, nc);
    }
    coarsen_aggregation(As, alg.theta, agg, nc);
    CSR P = prolongation_from_aggregates(agg, nc);
    if (alg.type == Coarsen::SmoothedAggregation) P = smooth_prolongation(A, P, alg.sa_omega);
    if (alg.filtering) P = filter_prolongation(P, alg.filter_tol);
    return P;
}

// Retry with theta halved (up to 3 times) when n_coarse >= 0.8 n;
// src/setup.jl:219-237.
CSR coarsen_with_fallback(const CSR& A, const CoarsenSpec& alg, const Config& cfg, Rng& rng, int32_t& nc,
                          const SetupHooks& hooks) {
    const int32_t n = A.nrow;
    CSR P = coarsen_and_build_p(A, alg, cfg, rng, nc, hooks);
    if (nc < 0.8 * n || n <= cfg.max_coarse_size) return P;
    for (int attempt = 1; attempt <= 3; ++attempt) {
        CoarsenSpec red = alg;
        red.theta = alg.theta * std::pow(0.5, attempt);
        int32_t nc2 = 0;
        CSR P2 = coarsen_and_build_p(A, red, cfg, rng, nc2, hooks);
        if (nc2 < nc) {
            P = std::move(P2);
            nc = nc2;
        }
        if (nc < 0.8 * n) break;
    }
    return P;
}

bool needs_coloring(Smoother s) { return s == Smoother::ColoredGS || s == Smoother::L1ColoredGS; }

}  // namespace

// Level loop of src/setup.jl:123-166 and the dense coarse operator of
// src/setup.jl:177-209.
HostHierarchy amg_setup_host(const CSR& A_in, const Config& cfg, bool allow_partial_resetup, const SetupHooks* hooks_in) {
    const SetupHooks hooks = hooks_in ? *hooks_in : SetupHooks();
    auto t0 = std::chrono::steady_clock::now();
    HostHierarchy h;
    Rng rng(cfg.seed);
    CSR A = A_in;
    for (int lvl = 1; lvl <= cfg.max_levels - 1; ++lvl) {
        const int32_t n = A.nrow;
        if (n <= cfg.max_coarse_size) break;
        const CoarsenSpec& alg = lvl <= cfg.initial_coarsening_levels ? cfg.initial_coarsening : cfg.coarsening;
        int32_t nc = 0;
        auto ta = std::chrono::steady_clock::now();
        CSR P = coarsen_with_fallback(A, alg, cfg, rng, nc, hooks);
        if (nc >= n || nc == 0) break;
        HostLevel L;
        CSR Ac;
        auto tb = std::chrono::steady_clock::now();
        auto tc = tb;
        if (hooks.galerkin) {
            hooks.galerkin(A, P, nc, Ac, allow_partial_resetup ? &L.R : nullptr, L.Pt);
            tc = std::chrono::steady_clock::now();
        } else {
            coarse_sparsity(A, P, nc, Ac, allow_partial_resetup ? &L.R : nullptr);
            tc = std::chrono::steady_clock::now();
            build_transpose_map(P, L.Pt);
        }
        L.has_rmap = allow_partial_resetup;
        if (needs_coloring(cfg.smoother)) greedy_coloring(A, L.coloring);
        if (cfg.verbose >= 2) {
            const auto sec = [](auto a, auto b) { return std::chrono::duration<double>(b - a).count(); };
            std::printf("  setup level %d: n=%d coarsening+P %.3fs, coarse operator %.3fs, maps %.3fs\n", lvl, n,
                        sec(ta, tb), sec(tb, tc), sec(tc, std::chrono::steady_clock::now()));
        }
        L.A = std::move(A);
        L.P = std::move(P);
        h.levels.push_back(std::move(L));
        A = std::move(Ac);
    }
    h.n_coarse = A.nrow;
    h.coarse_dense.assign((size_t)A.nrow * A.nrow, 0.0);
    for (int32_t i = 0; i < A.nrow; ++i)
        for (int32_t k = A.rp[i]; k < A.rp[i + 1]; ++k)
            h.coarse_dense[(size_t)A.cv[k] * A.nrow + i] = A.nz[k];
    h.A_coarsest = std::move(A);
    h.setup_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return h;
}

std::vector<HostHierarchy> amg_setup_host_many(const std::vector<CSR>& As, const Config& cfg,
                                               bool allow_partial_resetup, const SetupHooks* hooks) {
    const int nsys = (int)As.size();
    std::vector<HostHierarchy> out(nsys);
    if (nsys == 0) return out;
    // the thread budget is the calling thread's (OMP_NUM_THREADS or the caller's
    // omp_set_num_threads), shared out: up to 8 systems at once, each worker
    // with its share of the threads for its own parallel regions
    const int budget = std::max(1, omp_get_max_threads());
    const int workers = std::min(nsys, std::max(1, std::min(8, budget / 2)));
    const int per_worker = std::max(1, budget / workers);
    std::atomic<int> next{0};
    std::vector<std::exception_ptr> errs(nsys);
    auto work = [&]() {
        omp_set_num_threads(per_worker);  // this thread's parallel regions
        for (int s = next++; s < nsys; s = next++) {
            try {
                out[s] = amg_setup_host(As[s], cfg, allow_partial_resetup, hooks);
            } catch (...) {
                errs[s] = std::current_exception();
            }
        }
    };
    std::vector<std::thread> pool;
    for (int w = 0; w < workers; ++w) pool.emplace_back(work);
    for (auto& t : pool) t.join();
    for (auto& e : errs)
        if (e) std::rethrow_exception(e);
    return out;
}

// Hierarchy table in the shape of src/setup.jl:405-473.
std::string hierarchy_summary(const HostHierarchy& h, const Config& cfg, int64_t n_finest) {
    std::ostringstream o;
    char buf[256];
    const size_t L = h.levels.size();
    o << "+=============================================================================+\n";
    o << "                         AMG Hierarchy Summary (HIP, gfx950)\n";
    o << "+=============================================================================+\n";
    std::snprintf(buf, sizeof buf, "  Setup time:      %.4f s\n", h.setup_seconds);
    o << buf;
    o << "  Levels:          " << L + 1 << " (" << L << " AMG + 1 coarse direct)\n";
    o << "  Cycle type:      " << (cfg.w_cycle ? "W" : "V") << "-cycle\n";
    o << "  Level     Rows        NNZ      Ratio\n";
    int64_t tot_nnz = 0, tot_rows = 0;
    for (size_t i = 0; i < L; ++i) {
        const CSR& A = h.levels[i].A;
        tot_nnz += A.nnz();
        tot_rows += A.nrow;
        if (i == 0)
            std::snprintf(buf, sizeof buf, "  %5zu   %9d  %10lld        -\n", i + 1, A.nrow, (long long)A.nnz());
        else
            std::snprintf(buf, sizeof buf, "  %5zu   %9d  %10lld    %5.3f\n", i + 1, A.nrow, (long long)A.nnz(),
                          (double)A.nrow / h.levels[i - 1].A.nrow);
        o << buf;
    }
    int64_t cnnz = 0;
    for (double v : h.coarse_dense) cnnz += (v != 0.0);
    tot_nnz += cnnz;
    tot_rows += h.n_coarse;
    std::snprintf(buf, sizeof buf, "  %5zu   %9d  %10lld    %5.3f   Direct (LU)\n", L + 1, h.n_coarse,
                  (long long)cnnz, L ? (double)h.n_coarse / h.levels[L - 1].A.nrow : 1.0);
    o << buf;
    if (L) {
        std::snprintf(buf, sizeof buf, "  Operator complexity: %.3f\n  Grid complexity:     %.3f\n",
                      (double)tot_nnz / h.levels[0].A.nnz(), (double)tot_rows / n_finest);
        o << buf;
    }
    return o.str();
}

}  // namespace draugr
