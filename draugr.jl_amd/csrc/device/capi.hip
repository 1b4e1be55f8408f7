// C-ABI (include/draugr_hip.h): handle tables, error convention and the
// entry points.  Conventions follow src/cfunction_interface.jl: int32 handles
// from one counter shared by configs and hierarchies, -1 plus a message on
// failure, setup/resetup copy their inputs, solve/cycle work on the caller's
// x in place.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>

#include "../../../include/draugr_hip.h"
#include "batch.hpp"
#include "coarsen.hpp"
#include "galerkin.hpp"
#include "interp.hpp"
#include "hierarchy.hpp"

namespace draugr {
namespace {

thread_local std::string g_last_error;

struct Hier {
    Config cfg;                  // configuration the hierarchy was built with
    bool allow_partial = true;
    HostHierarchy host;
    std::unique_ptr<DeviceHierarchy> dev;
    DBuf<double> hx, hb, hv;     // staging for host-pointer entry points
    std::mutex mu;
};

struct Batch {
    Config cfg;
    std::unique_ptr<BatchedHierarchy> dev;
    double setup_seconds = 0.0;
    hipEvent_t last = nullptr;   // end of the previous call's work (calls share workspaces)
    bool last_valid = false;
    std::mutex mu;
    ~Batch() {
        if (last_valid) (void)hipEventSynchronize(last);
        if (last) (void)hipEventDestroy(last);
    }
};

std::mutex g_lock;
std::map<int32_t, std::shared_ptr<Hier>> g_hier;
std::map<int32_t, std::shared_ptr<Batch>> g_batch;
std::map<int32_t, Config> g_cfg;
int32_t g_next = 1;

int32_t new_handle() { return g_next++; }

std::shared_ptr<Hier> get_hier(int32_t h) {
    std::lock_guard<std::mutex> g(g_lock);
    auto it = g_hier.find(h);
    return it == g_hier.end() ? nullptr : it->second;
}
bool get_cfg(int32_t h, Config& c) {
    std::lock_guard<std::mutex> g(g_lock);
    auto it = g_cfg.find(h);
    if (it == g_cfg.end()) return false;
    c = it->second;
    return true;
}

template <class F>
int32_t guarded(F&& f) {
    try {
        g_last_error.clear();
        return f();
    } catch (const std::exception& e) {
        g_last_error = e.what();
    } catch (...) {
        g_last_error = "unknown error";
    }
    return -1;
}

[[noreturn]] void invalid(int32_t h, int32_t c) {
    throw Error("Invalid handle (hierarchy=" + std::to_string(h) + ", config=" + std::to_string(c) + ")");
}

// Raw CSR -> 0-based CSR copy, with the checks the reference implies
// (index_base in {0,1}: src/csr_matrix.jl:133-143).
CSR csr_from_raw(int32_t n, int32_t nnz, const int32_t* rp, const int32_t* cv, const double* nz, int32_t base) {
    if (base != 0 && base != 1)
        throw Error("ArgumentError: index_base must be 0 or 1, got " + std::to_string(base));
    if (n < 0 || nnz < 0) throw Error("ArgumentError: n and nnz must be non-negative");
    if (!rp || (nnz > 0 && (!cv || !nz))) throw Error("ArgumentError: null CSR array");
    CSR A;
    A.nrow = A.ncol = n;
    A.rp.resize(n + 1);
    for (int32_t i = 0; i <= n; ++i) A.rp[i] = rp[i] - base;
    if (A.rp[0] != 0 || A.rp[n] != nnz)
        throw Error("ArgumentError: rowptr inconsistent with nnz (rowptr[end] - index_base must equal nnz)");
    for (int32_t i = 0; i < n; ++i)
        if (A.rp[i + 1] < A.rp[i]) throw Error("ArgumentError: rowptr must be non-decreasing");
    A.cv.resize(nnz);
    A.nz.assign(nz, nz + nnz);
    for (int32_t k = 0; k < nnz; ++k) {
        const int32_t j = cv[k] - base;
        if (j < 0 || j >= n) throw Error("ArgumentError: column index out of range at position " + std::to_string(k));
        A.cv[k] = j;
    }
    return A;
}

// graphs "auto": replayed graphs for the batched PCG (its iterations are
// many launches over all systems), eager launches for one system (the
// calling thread keeps ahead of the GPU, and a switch between graph and
// eager submissions on a stream costs more than the launches it saves)
CycleOpts opts_of(const Config& c, bool batched = false) {
    CycleOpts o;
    o.w_cycle = c.w_cycle;
    o.pre = c.pre_smoothing_steps;
    o.post = c.post_smoothing_steps;
    o.prec = c.precision;
    o.graphs = c.graphs < 0 ? batched : c.graphs != 0;
    o.gs_sym = c.gs_symmetric;
    return o;
}

void require_device(Hier& H) {
    if (!H.dev) throw Error("hierarchy is not uploaded to a GPU (draugr_hip_upload)");
}

void check_n(Hier& H, int32_t n) {
    if (n != H.dev->n_fine())
        throw Error("DimensionMismatch: vector length " + std::to_string(n) + " != matrix size " +
                    std::to_string(H.dev->n_fine()));
}

// Setup phases run on the GPU (config gpu_setup): interpolation and the
// Galerkin product; null = everything on the host.
const SetupHooks* gpu_hooks(const Config& cfg) {
    static const SetupHooks hooks{galerkin_device, interpolation_device, pmis_sweeps_device};
    return cfg.gpu_setup ? &hooks : nullptr;
}

int32_t setup_impl(int32_t n, int32_t nnz, const int32_t* rp, const int32_t* cv, const double* nz, int32_t ch,
                   int32_t base, int32_t allow, bool upload) {
    Config cfg;
    if (!get_cfg(ch, cfg)) throw Error("Invalid config handle: " + std::to_string(ch));
    CSR A = csr_from_raw(n, nnz, rp, cv, nz, base);
    auto H = std::make_shared<Hier>();
    H->cfg = cfg;
    H->allow_partial = allow != 0;
    H->host = amg_setup_host(A, cfg, H->allow_partial, upload ? gpu_hooks(cfg) : nullptr);
    if (upload) H->dev = std::make_unique<DeviceHierarchy>(H->host, cfg);
    if (cfg.verbose >= 1) std::printf("%s", hierarchy_summary(H->host, cfg, n).c_str());
    std::lock_guard<std::mutex> g(g_lock);
    const int32_t h = new_handle();
    g_hier[h] = H;
    return h;
}

std::shared_ptr<Batch> get_batch(int32_t h) {
    std::lock_guard<std::mutex> g(g_lock);
    auto it = g_batch.find(h);
    if (it == g_batch.end()) throw Error("Invalid batch handle: " + std::to_string(h));
    return it->second;
}

void batch_begin(Batch& B, hipStream_t st) {
    if (B.last_valid) DRAUGR_HIP_CHECK(hipStreamWaitEvent(st, B.last, 0));
}
void batch_end(Batch& B, hipStream_t st) {
    if (!B.last) DRAUGR_HIP_CHECK(hipEventCreateWithFlags(&B.last, hipEventDisableTiming));
    DRAUGR_HIP_CHECK(hipEventRecord(B.last, st));
    B.last_valid = true;
}

// The caller's stream as given: NULL is HIP's default (null) stream.
hipStream_t stream_of(Hier&, void* s) { return (hipStream_t)s; }

}  // namespace
}  // namespace draugr

using namespace draugr;

extern "C" {

void init_julia(int, char**) {}

void shutdown_julia(int) {
    std::map<int32_t, std::shared_ptr<Hier>> hier;
    std::map<int32_t, std::shared_ptr<Batch>> batch;
    {
        std::lock_guard<std::mutex> g(g_lock);
        hier.swap(g_hier);
        batch.swap(g_batch);
        g_cfg.clear();
    }
    for (auto& kv : hier) std::lock_guard<std::mutex> g(kv.second->mu);  // wait for calls in flight
    for (auto& kv : batch) std::lock_guard<std::mutex> g(kv.second->mu);
}

const char* draugr_amg_last_error(void) { return g_last_error.c_str(); }

int32_t draugr_amg_config_from_json(const char* json_str) {
    return guarded([&]() -> int32_t {
        Config c = config_from_json(json_str);
        std::lock_guard<std::mutex> g(g_lock);
        const int32_t h = new_handle();
        g_cfg[h] = c;
        return h;
    });
}

int32_t draugr_amg_setup(int32_t n, int32_t nnz, const int32_t* rowptr, const int32_t* colval, const double* nzval,
                         int32_t config_handle, int32_t index_base, int32_t allow_partial_resetup) {
    return guarded([&]() {
        return setup_impl(n, nnz, rowptr, colval, nzval, config_handle, index_base, allow_partial_resetup, true);
    });
}

int32_t draugr_hip_setup_host(int32_t n, int32_t nnz, const int32_t* rowptr, const int32_t* colval,
                              const double* nzval, int32_t config_handle, int32_t index_base,
                              int32_t allow_partial_resetup) {
    return guarded([&]() {
        return setup_impl(n, nnz, rowptr, colval, nzval, config_handle, index_base, allow_partial_resetup, false);
    });
}

int32_t draugr_hip_upload(int32_t handle) {
    return guarded([&]() -> int32_t {
        auto H = get_hier(handle);
        if (!H) throw Error("Invalid hierarchy handle: " + std::to_string(handle));
        std::lock_guard<std::mutex> g(H->mu);
        if (!H->dev) H->dev = std::make_unique<DeviceHierarchy>(H->host, H->cfg);
        return 0;
    });
}

int32_t draugr_amg_resetup(int32_t handle, int32_t n, int32_t nnz, const int32_t* rowptr, const int32_t* colval,
                           const double* nzval, int32_t config_handle, int32_t index_base, int32_t partial,
                           int32_t allow_partial_resetup) {
    return guarded([&]() -> int32_t {
        CSR A = csr_from_raw(n, nnz, rowptr, colval, nzval, index_base);
        auto H = get_hier(handle);
        Config cfg;
        if (!H || !get_cfg(config_handle, cfg)) invalid(handle, config_handle);
        std::lock_guard<std::mutex> g(H->mu);
        if (partial) {
            const CSR& A0 = H->host.levels.empty() ? H->host.A_coarsest : H->host.levels[0].A;
            if (A.nrow != A0.nrow || A.nnz() != A0.nnz())
                throw Error("DimensionMismatch: partial resetup needs the sparsity pattern of the setup matrix");
            if (!H->allow_partial && !H->host.levels.empty())
                throw Error("partial resetup needs restriction maps: setup with allow_partial_resetup=1");
            require_device(*H);
            hipStream_t st = H->dev->stream();
            H->dev->begin(st);
            H->hv.alloc(A.nnz());
            DRAUGR_HIP_CHECK(hipMemcpyAsync(H->hv.p, A.nz.data(), A.nnz() * sizeof(double), hipMemcpyHostToDevice, st));
            H->dev->resetup(H->hv.p, st);
            H->dev->end(st);
            if (H->host.levels.empty()) H->host.A_coarsest.nz = A.nz;
            else H->host.levels[0].A.nz = A.nz;
        } else {
            H->cfg = cfg;
            H->allow_partial = allow_partial_resetup != 0;
            H->dev.reset();
            H->host = amg_setup_host(A, cfg, H->allow_partial, gpu_hooks(cfg));
            H->dev = std::make_unique<DeviceHierarchy>(H->host, cfg);
            if (cfg.verbose >= 1) std::printf("%s", hierarchy_summary(H->host, cfg, n).c_str());
        }
        return 0;
    });
}

int32_t draugr_hip_resetup_device(int32_t handle, int64_t nnz, const double* nzval, void* stream) {
    return guarded([&]() -> int32_t {
        auto H = get_hier(handle);
        if (!H) throw Error("Invalid hierarchy handle: " + std::to_string(handle));
        std::lock_guard<std::mutex> g(H->mu);
        require_device(*H);
        const CSR& A0 = H->host.levels.empty() ? H->host.A_coarsest : H->host.levels[0].A;
        if (nnz != A0.nnz()) throw Error("DimensionMismatch: nnz differs from the setup matrix");
        if (!H->allow_partial && !H->host.levels.empty())
            throw Error("partial resetup needs restriction maps: setup with allow_partial_resetup=1");
        hipStream_t st = stream_of(*H, stream);
        H->dev->begin(st);
        H->dev->resetup(nzval, st);
        H->dev->end(st);
        return 0;
    });
}

static int32_t host_vector_call(int32_t handle, int32_t n, double* x, const double* b, int32_t ch, int kind,
                                double tol, int32_t maxiter, double* relres) {
    return guarded([&]() -> int32_t {
        auto H = get_hier(handle);
        Config cfg;
        if (!H || !get_cfg(ch, cfg)) invalid(handle, ch);
        if (n > 0 && (!x || !b)) throw Error("ArgumentError: null vector");
        std::lock_guard<std::mutex> g(H->mu);
        require_device(*H);
        check_n(*H, n);
        hipStream_t st = H->dev->stream();
        H->dev->begin(st);
        if (H->hx.n != (size_t)n) { H->hx.alloc(n); H->hb.alloc(n); }
        DRAUGR_HIP_CHECK(hipMemcpyAsync(H->hx.p, x, n * sizeof(double), hipMemcpyHostToDevice, st));
        DRAUGR_HIP_CHECK(hipMemcpyAsync(H->hb.p, b, n * sizeof(double), hipMemcpyHostToDevice, st));
        int32_t ret = 0;
        double rr = 0.0;
        const CycleOpts o = opts_of(cfg);
        if (kind == 0) H->dev->cycle(H->hx.p, H->hb.p, false, o, st);
        else if (kind == 1) ret = H->dev->solve(H->hx.p, H->hb.p, tol, maxiter, o, st, &rr, cfg.verbose);
        else ret = H->dev->pcg(H->hx.p, H->hb.p, tol, maxiter, o, st, &rr, cfg.verbose);
        DRAUGR_HIP_CHECK(hipMemcpyAsync(x, H->hx.p, n * sizeof(double), hipMemcpyDeviceToHost, st));
        H->dev->end(st);
        DRAUGR_HIP_CHECK(hipStreamSynchronize(st));
        if (relres) *relres = rr;
        if (kind != 0 && cfg.verbose >= 1)
            std::printf("AMG %s %s: %d iterations, final residual %.2e\n", kind == 1 ? "solve" : "PCG",
                        rr < tol ? "converged" : "did NOT converge", ret, rr);
        return ret;
    });
}

int32_t draugr_amg_solve(int32_t handle, int32_t n, double* x, const double* b, int32_t config_handle, double tol,
                         int32_t maxiter) {
    return host_vector_call(handle, n, x, b, config_handle, 1, tol, maxiter, nullptr);
}

int32_t draugr_amg_cycle(int32_t handle, int32_t n, double* x, const double* b, int32_t config_handle) {
    return host_vector_call(handle, n, x, b, config_handle, 0, 0.0, 0, nullptr);
}

int32_t draugr_hip_pcg(int32_t handle, int32_t n, double* x, const double* b, int32_t config_handle, double tol,
                       int32_t maxiter, double* relres) {
    return host_vector_call(handle, n, x, b, config_handle, 2, tol, maxiter, relres);
}

static int32_t device_vector_call(int32_t handle, int32_t n, double* x, const double* b, int32_t ch, int kind,
                                  int32_t zero_guess, double tol, int32_t maxiter, void* stream, double* relres) {
    return guarded([&]() -> int32_t {
        auto H = get_hier(handle);
        Config cfg;
        if (!H || !get_cfg(ch, cfg)) invalid(handle, ch);
        if (n > 0 && (!x || !b)) throw Error("ArgumentError: null vector");
        std::lock_guard<std::mutex> g(H->mu);
        require_device(*H);
        check_n(*H, n);
        hipStream_t st = stream_of(*H, stream);
        const CycleOpts o = opts_of(cfg);
        double rr = 0.0;
        int32_t ret = 0;
        H->dev->begin(st);
        if (kind == 0) H->dev->cycle(x, b, zero_guess != 0, o, st);
        else if (kind == 1) ret = H->dev->solve(x, b, tol, maxiter, o, st, &rr, cfg.verbose);
        else if (kind == 2) ret = H->dev->pcg(x, b, tol, maxiter, o, st, &rr, cfg.verbose);
        else ret = H->dev->pcg_no_alloc(x, b, tol, maxiter, o, st, &rr, cfg.verbose);
        H->dev->end(st);
        if (relres) *relres = rr;
        return ret;
    });
}

int32_t draugr_hip_cycle_device(int32_t handle, int32_t n, double* x, const double* b, int32_t config_handle,
                                int32_t zero_guess, void* stream) {
    return device_vector_call(handle, n, x, b, config_handle, 0, zero_guess, 0.0, 0, stream, nullptr);
}

int32_t draugr_hip_solve_device(int32_t handle, int32_t n, double* x, const double* b, int32_t config_handle,
                                double tol, int32_t maxiter, void* stream, double* relres) {
    return device_vector_call(handle, n, x, b, config_handle, 1, 0, tol, maxiter, stream, relres);
}

int32_t draugr_hip_pcg_device(int32_t handle, int32_t n, double* x, const double* b, int32_t config_handle,
                              double tol, int32_t maxiter, void* stream, double* relres) {
    return device_vector_call(handle, n, x, b, config_handle, 2, 0, tol, maxiter, stream, relres);
}

int32_t draugr_hip_pcg_device_noalloc(int32_t handle, int32_t n, double* x, const double* b, int32_t config_handle,
                                      double tol, int32_t maxiter, void* stream, double* relres) {
    return device_vector_call(handle, n, x, b, config_handle, 3, 0, tol, maxiter, stream, relres);
}

int32_t draugr_amg_free(int32_t handle) {
    return guarded([&]() -> int32_t {
        std::shared_ptr<Hier> H;
        {
            std::lock_guard<std::mutex> g(g_lock);
            auto it = g_hier.find(handle);
            if (it == g_hier.end()) throw Error("Invalid hierarchy handle: " + std::to_string(handle));
            H = it->second;
            g_hier.erase(it);
        }
        std::lock_guard<std::mutex> g(H->mu);  // wait for a call in flight
        return 0;
    });
}

int32_t draugr_amg_config_free(int32_t handle) {
    return guarded([&]() -> int32_t {
        std::lock_guard<std::mutex> g(g_lock);
        if (!g_cfg.erase(handle)) throw Error("Invalid config handle: " + std::to_string(handle));
        return 0;
    });
}

int32_t draugr_hip_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

int32_t draugr_hip_hierarchy_info(int32_t handle, int64_t* out, int32_t nout) {
    return guarded([&]() -> int32_t {
        auto H = get_hier(handle);
        if (!H) throw Error("Invalid hierarchy handle: " + std::to_string(handle));
        const HostHierarchy& h = H->host;
        int64_t tot_nnz = 0, tot_rows = 0;
        for (auto& l : h.levels) { tot_nnz += l.A.nnz(); tot_rows += l.A.nrow; }
        tot_nnz += h.A_coarsest.nnz();
        tot_rows += h.n_coarse;
        const int64_t v[] = {(int64_t)h.levels.size(), h.n_coarse,
                             h.levels.empty() ? h.A_coarsest.nrow : h.levels[0].A.nrow,
                             (int64_t)(h.setup_seconds * 1e6), H->dev ? 1 : 0, tot_nnz, tot_rows};
        const int32_t m = std::min<int32_t>(nout, (int32_t)(sizeof v / sizeof v[0]));
        for (int32_t i = 0; i < m; ++i) out[i] = v[i];
        return m;
    });
}

int32_t draugr_hip_level_info(int32_t handle, int32_t level, int64_t* out, int32_t nout) {
    return guarded([&]() -> int32_t {
        auto H = get_hier(handle);
        if (!H) throw Error("Invalid hierarchy handle: " + std::to_string(handle));
        if (level < 0 || level >= (int32_t)H->host.levels.size()) throw Error("level index out of range");
        const HostLevel& L = H->host.levels[level];
        const int64_t v[] = {L.A.nrow, L.A.nnz(), L.P.ncol, L.P.nnz(), L.R.ntriples(), L.coloring.num_colors,
                             L.has_rmap ? 1 : 0, H->dev ? H->dev->level_blocks(level) : 0,
                             H->dev && H->dev->level_has16(level) ? 1 : 0,
                             H->dev ? H->dev->level_nnz_sorted(level) : 0};
        const int32_t m = std::min<int32_t>(nout, (int32_t)(sizeof v / sizeof v[0]));
        for (int32_t i = 0; i < m; ++i) out[i] = v[i];
        return m;
    });
}

int32_t draugr_hip_level_export(int32_t handle, int32_t level, int32_t which, void* dst) {
    return guarded([&]() -> int32_t {
        auto H = get_hier(handle);
        if (!H) throw Error("Invalid hierarchy handle: " + std::to_string(handle));
        auto put = [&](const auto& v) {
            std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
            return 0;
        };
        if (level == -1) {
            const CSR& C = H->host.A_coarsest;
            switch (which) {
                case 0: return put(C.rp);
                case 1: return put(C.cv);
                case 2: return put(C.nz);
                case 15: return put(H->host.coarse_dense);
                default: throw Error("unknown export selector");
            }
        }
        if (level < 0 || level >= (int32_t)H->host.levels.size()) throw Error("level index out of range");
        const HostLevel& L = H->host.levels[level];
        // triple map left in HBM by the GPU setup: copied out
        auto rm = [&]() { return std::static_pointer_cast<DeviceRMap>(L.R.device); };
        auto get_dev = [&](const auto* src, size_t count) {
            if (count) DRAUGR_HIP_CHECK(hipMemcpy(dst, src, count * sizeof(src[0]), hipMemcpyDeviceToHost));
            return 0;
        };
        switch (which) {
            case 0: return put(L.A.rp);
            case 1: return put(L.A.cv);
            case 2: return put(L.A.nz);
            case 3: return put(L.P.rp);
            case 4: return put(L.P.cv);
            case 5: return put(L.P.nz);
            case 6: return put(L.Pt.offsets);
            case 7: return put(L.Pt.fine_rows);
            case 8: return put(L.Pt.p_nz_idx);
            case 9: return L.R.device ? get_dev(rm()->off.p, rm()->off.n) : put(L.R.nz_offsets);
            case 10: return L.R.device ? get_dev(rm()->pi.p, (size_t)rm()->ntriples) : put(L.R.pi);
            case 11: return L.R.device ? get_dev(rm()->ai.p, (size_t)rm()->ntriples) : put(L.R.ai);
            case 12: return L.R.device ? get_dev(rm()->pj.p, (size_t)rm()->ntriples) : put(L.R.pj);
            case 13: return put(L.coloring.color_offsets);
            case 14: return put(L.coloring.color_order);
            default: throw Error("unknown export selector");
        }
    });
}

int32_t draugr_hip_summary(int32_t handle, char* buf, int32_t buflen) {
    return guarded([&]() -> int32_t {
        auto H = get_hier(handle);
        if (!H) throw Error("Invalid hierarchy handle: " + std::to_string(handle));
        const int64_t n = H->host.levels.empty() ? H->host.A_coarsest.nrow : H->host.levels[0].A.nrow;
        const std::string s = hierarchy_summary(H->host, H->cfg, n);
        if (buf && buflen > 0) {
            const size_t m = std::min<size_t>(s.size(), (size_t)buflen - 1);
            std::memcpy(buf, s.data(), m);
            buf[m] = 0;
        }
        return (int32_t)s.size();
    });
}

int32_t draugr_hip_level_op(int32_t handle, int32_t level, int32_t op, int32_t prec, const void* in1,
                            const void* in2, void* out, void* out2, void* stream) {
    return guarded([&]() -> int32_t {
        auto H = get_hier(handle);
        if (!H) throw Error("Invalid hierarchy handle: " + std::to_string(handle));
        std::lock_guard<std::mutex> g(H->mu);
        require_device(*H);
        if (prec < 0 || prec > 2) throw Error("prec must be 0, 1 or 2");
        hipStream_t st = stream_of(*H, stream);
        H->dev->begin(st);
        H->dev->level_op(level, op, prec, in1, in2, out, out2, st);
        H->dev->end(st);
        return 0;
    });
}

int32_t draugr_hip_level_op_repeat(int32_t handle, int32_t level, int32_t op, int32_t prec, const void* in1,
                                   const void* in2, void* out, void* out2, int32_t reps, void* stream) {
    return guarded([&]() -> int32_t {
        auto H = get_hier(handle);
        if (!H) throw Error("Invalid hierarchy handle: " + std::to_string(handle));
        if (reps < 1) throw Error("ArgumentError: reps must be positive");
        std::lock_guard<std::mutex> g(H->mu);
        require_device(*H);
        if (prec < 0 || prec > 2) throw Error("prec must be 0, 1 or 2");
        hipStream_t st = stream_of(*H, stream);
        H->dev->begin(st);
        for (int32_t r = 0; r < reps; ++r) H->dev->level_op(level, op, prec, in1, in2, out, out2, st);
        H->dev->end(st);
        return 0;
    });
}

int32_t draugr_hip_export_device_values(int32_t handle, int32_t level, int32_t which, double* dst) {
    return guarded([&]() -> int32_t {
        auto H = get_hier(handle);
        if (!H) throw Error("Invalid hierarchy handle: " + std::to_string(handle));
        std::lock_guard<std::mutex> g(H->mu);
        require_device(*H);
        H->dev->export_level_values(level, which, dst);
        return 0;
    });
}

int32_t draugr_hip_cycle_bytes(int32_t handle, int32_t config_handle, double* bytes) {
    return guarded([&]() -> int32_t {
        auto H = get_hier(handle);
        Config cfg;
        if (!H || !get_cfg(config_handle, cfg)) invalid(handle, config_handle);
        require_device(*H);
        *bytes = H->dev->cycle_bytes(opts_of(cfg));
        return 0;
    });
}

int32_t draugr_hip_precondition_repeat(int32_t handle, int32_t config_handle, const double* b, int32_t reps,
                                       void* stream) {
    return guarded([&]() -> int32_t {
        auto H = get_hier(handle);
        Config cfg;
        if (!H || !get_cfg(config_handle, cfg)) invalid(handle, config_handle);
        if (reps < 1 || !b) throw Error("ArgumentError: reps must be positive and b non-null");
        std::lock_guard<std::mutex> g(H->mu);
        require_device(*H);
        hipStream_t st = stream_of(*H, stream);
        H->dev->begin(st);
        H->dev->precondition_repeat(b, opts_of(cfg), reps, st);
        H->dev->end(st);
        return 0;
    });
}

int32_t draugr_hip_op_bytes(int32_t handle, int32_t level, int32_t which, int32_t precision, int32_t as_stored,
                            double* bytes) {
    return guarded([&]() -> int32_t {
        auto H = get_hier(handle);
        if (!H) throw Error("Invalid hierarchy handle: " + std::to_string(handle));
        require_device(*H);
        if (which < 0 || which > 2 || precision < 0 || precision > 2 || !bytes)
            throw Error("ArgumentError: which in 0..2, precision in 0..2, bytes non-null");
        *bytes = H->dev->op_bytes(level, which, (Precision)precision, as_stored != 0);
        return 0;
    });
}

int32_t draugr_hip_assemble_7pt(int32_t handle, int32_t m, const double* k, double* nzval, void* stream) {
    return guarded([&]() -> int32_t {
        auto H = get_hier(handle);
        if (!H) throw Error("Invalid hierarchy handle: " + std::to_string(handle));
        std::lock_guard<std::mutex> g(H->mu);
        require_device(*H);
        const int64_t n = (int64_t)m * m * m;
        if (m < 1 || n != H->dev->n_fine() || H->dev->finest_nnz() != 7 * n - 6 * (int64_t)m * m)
            throw Error("DimensionMismatch: the hierarchy's finest operator is not the 7-point operator on m^3 cells");
        if (!k || !nzval) throw Error("ArgumentError: null pointer");
        hipStream_t st = stream_of(*H, stream);
        H->dev->begin(st);
        assemble_7pt(m, k, H->dev->finest_rp(), nzval, st);
        H->dev->end(st);
        return 0;
    });
}

int32_t draugr_hip_matrix_grad(int32_t handle, const double* x, const double* lambda, double* grad, void* stream) {
    return guarded([&]() -> int32_t {
        auto H = get_hier(handle);
        if (!H) throw Error("Invalid hierarchy handle: " + std::to_string(handle));
        std::lock_guard<std::mutex> g(H->mu);
        require_device(*H);
        if (!x || !lambda || !grad) throw Error("ArgumentError: null pointer");
        hipStream_t st = stream_of(*H, stream);
        H->dev->begin(st);
        matrix_grad(H->dev->n_fine(), H->dev->finest_rp(), H->dev->finest_cv(), x, lambda, grad, st);
        H->dev->end(st);
        return 0;
    });
}

int32_t draugr_hip_grad_logk_7pt(int32_t m, const double* k, const double* x, const double* lambda, double* grad,
                                 void* stream) {
    return guarded([&]() -> int32_t {
        if (m < 0) throw Error("ArgumentError: m must be non-negative");
        if (m > 0 && (!k || !x || !lambda || !grad)) throw Error("ArgumentError: null pointer");
        grad_logk_7pt(m, k, x, lambda, grad, (hipStream_t)stream);
        return 0;
    });
}

int32_t draugr_hip_tune_report(int32_t handle, char* buf, int32_t buflen) {
    return guarded([&]() -> int32_t {
        auto H = get_hier(handle);
        if (!H) throw Error("Invalid hierarchy handle: " + std::to_string(handle));
        require_device(*H);
        const std::string s = H->dev->tune_report();
        if (buf && buflen > 0) {
            const size_t m = std::min<size_t>(s.size(), (size_t)buflen - 1);
            std::memcpy(buf, s.data(), m);
            buf[m] = 0;
        }
        return (int32_t)s.size();
    });
}

int32_t draugr_hip_resident_bytes(int32_t handle, int64_t* bytes) {
    return guarded([&]() -> int32_t {
        auto H = get_hier(handle);
        if (!H) throw Error("Invalid hierarchy handle: " + std::to_string(handle));
        require_device(*H);
        *bytes = H->dev->bytes_resident();
        return 0;
    });
}

// ── batches of independent systems ──────────────────────────────────────────
int32_t draugr_hip_batch_setup(int32_t nsys, const int32_t* n, const int32_t* nnz, const int32_t* const* rowptr,
                               const int32_t* const* colval, const double* const* nzval, int32_t config_handle,
                               int32_t index_base, int32_t nrhs) {
    return guarded([&]() -> int32_t {
        Config cfg;
        if (!get_cfg(config_handle, cfg)) throw Error("Invalid config handle: " + std::to_string(config_handle));
        if (nsys < 1 || !n || !nnz || !rowptr || !colval || !nzval)
            throw Error("ArgumentError: a batch needs nsys >= 1 and non-null per-system arrays");
        const auto t0 = std::chrono::steady_clock::now();
        std::vector<CSR> As(nsys);
        for (int32_t s = 0; s < nsys; ++s) As[s] = csr_from_raw(n[s], nnz[s], rowptr[s], colval[s], nzval[s], index_base);
        // independent systems: built concurrently, each hierarchy as a lone setup would build it
        std::vector<HostHierarchy> host = amg_setup_host_many(As, cfg, false, gpu_hooks(cfg));
        std::vector<const HostHierarchy*> hp(nsys);
        for (int32_t s = 0; s < nsys; ++s) hp[s] = &host[s];
        auto B = std::make_shared<Batch>();
        B->cfg = cfg;
        B->dev = std::make_unique<BatchedHierarchy>(hp, cfg, nrhs);
        B->setup_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::lock_guard<std::mutex> g(g_lock);
        const int32_t h = new_handle();
        g_batch[h] = B;
        return h;
    });
}

int32_t draugr_hip_batch_info(int32_t batch, int64_t* out, int32_t nout) {
    return guarded([&]() -> int32_t {
        auto B = get_batch(batch);
        std::lock_guard<std::mutex> g(B->mu);
        const BatchedHierarchy& D = *B->dev;
        const int64_t v[] = {D.nsys(), D.nrhs(), D.rows(), D.nlevels_max(), (int64_t)(B->setup_seconds * 1e6),
                             D.bytes_resident()};
        const int32_t m = std::min<int32_t>(nout, (int32_t)(sizeof v / sizeof v[0]));
        for (int32_t i = 0; i < m; ++i) out[i] = v[i];
        return m;
    });
}

int32_t draugr_hip_batch_system_info(int32_t batch, int32_t sys, int64_t* out, int32_t nout) {
    return guarded([&]() -> int32_t {
        auto B = get_batch(batch);
        const BatchedHierarchy& D = *B->dev;
        if (sys < 0 || sys >= D.nsys()) throw Error("system index out of range");
        const int64_t v[] = {D.row_offset(sys + 1) - D.row_offset(sys), D.row_offset(sys), D.sys_levels(sys),
                             D.sys_coarse(sys)};
        const int32_t m = std::min<int32_t>(nout, (int32_t)(sizeof v / sizeof v[0]));
        for (int32_t i = 0; i < m; ++i) out[i] = v[i];
        return m;
    });
}

int32_t draugr_hip_batch_cycle_device(int32_t batch, double* x, const double* b, int32_t config_handle,
                                      void* stream) {
    return guarded([&]() -> int32_t {
        auto B = get_batch(batch);
        Config cfg;
        if (!get_cfg(config_handle, cfg)) throw Error("Invalid config handle: " + std::to_string(config_handle));
        if (!x || !b) throw Error("ArgumentError: null vector");
        std::lock_guard<std::mutex> g(B->mu);
        hipStream_t st = (hipStream_t)stream;
        batch_begin(*B, st);
        B->dev->cycle(x, b, opts_of(cfg, true), st);
        batch_end(*B, st);
        return 0;
    });
}

int32_t draugr_hip_batch_pcg_device(int32_t batch, double* x, const double* b, int32_t config_handle, double tol,
                                    int32_t maxiter, void* stream, int32_t* iters, double* relres) {
    return guarded([&]() -> int32_t {
        auto B = get_batch(batch);
        Config cfg;
        if (!get_cfg(config_handle, cfg)) throw Error("Invalid config handle: " + std::to_string(config_handle));
        if (!x || !b) throw Error("ArgumentError: null vector");
        if (maxiter < 0) throw Error("ArgumentError: maxiter must be non-negative");
        std::lock_guard<std::mutex> g(B->mu);
        hipStream_t st = (hipStream_t)stream;
        batch_begin(*B, st);
        const int it = B->dev->pcg(x, b, tol, maxiter, opts_of(cfg, true), st, iters, relres);
        batch_end(*B, st);
        return it;
    });
}

int32_t draugr_hip_batch_tune_report(int32_t batch, char* buf, int32_t buflen) {
    return guarded([&]() -> int32_t {
        auto B = get_batch(batch);
        std::lock_guard<std::mutex> g(B->mu);
        const std::string s = B->dev->tune_report();
        if (buf && buflen > 0) {
            const size_t m = std::min<size_t>(s.size(), (size_t)buflen - 1);
            std::memcpy(buf, s.data(), m);
            buf[m] = 0;
        }
        return (int32_t)s.size();
    });
}

int32_t draugr_hip_batch_level_info(int32_t batch, int32_t level, int64_t* out, int32_t nout) {
    return guarded([&]() -> int32_t {
        auto B = get_batch(batch);
        int64_t v[5];
        B->dev->level_counts(level, v);
        const int32_t m = std::min<int32_t>(nout, 5);
        for (int32_t i = 0; i < m; ++i) out[i] = v[i];
        return m;
    });
}

int32_t draugr_hip_batch_jacobi_pass(int32_t batch, int32_t level, int32_t prec, const void* x, const void* b,
                                     void* out, void* stream) {
    return guarded([&]() -> int32_t {
        auto B = get_batch(batch);
        if (!x || !b || !out) throw Error("ArgumentError: null vector");
        if (prec < 0 || prec > 1) throw Error("prec must be 0 (fp64) or 1 (fp32)");
        std::lock_guard<std::mutex> g(B->mu);
        hipStream_t st = (hipStream_t)stream;
        batch_begin(*B, st);
        B->dev->jacobi_pass(level, prec == 0 ? Precision::FP64 : Precision::FP32, x, b, out, st);
        batch_end(*B, st);
        return 0;
    });
}

int32_t draugr_hip_batch_free(int32_t batch) {
    return guarded([&]() -> int32_t {
        std::shared_ptr<Batch> B;
        {
            std::lock_guard<std::mutex> g(g_lock);
            auto it = g_batch.find(batch);
            if (it == g_batch.end()) throw Error("Invalid batch handle: " + std::to_string(batch));
            B = it->second;
            g_batch.erase(it);
        }
        std::lock_guard<std::mutex> g(B->mu);
        return 0;
    });
}

int32_t draugr_hip_ilu0_factor(int32_t n, const int32_t* rowptr, const int32_t* colval, const double* nzval,
                               int32_t* diag, double* L, double* U) {
    return guarded([&]() -> int32_t {
        if (!diag || !L || !U) throw Error("ArgumentError: null output array");
        const int32_t nnz = rowptr ? rowptr[n] : 0;
        CSR A = csr_from_raw(n, nnz, rowptr, colval, nzval, 0);
        IluStructure s;
        ilu0_structure(A, s);
        std::vector<double> Lv, Uv;
        ilu0_factorize(A, s.diag_idx, Lv, Uv);
        std::copy(s.diag_idx.begin(), s.diag_idx.end(), diag);
        std::copy(Lv.begin(), Lv.end(), L);
        std::copy(Uv.begin(), Uv.end(), U);
        return 0;
    });
}

}  // extern "C"
