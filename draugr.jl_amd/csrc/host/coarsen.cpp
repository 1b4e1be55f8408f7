// Host-side coarsening (setup phase, runs once; its output is uploaded to HBM).
//
// Behaviour follows the reference algorithms; each function cites the file:line
// of the Draugr.jl routine it restates.  All arrays are 0-based here; the
// C/F state uses the reference's codes: 1 = C, -1 = F, 0 = undecided,
// -2 = tentative fine (Z_PT, HMIS first pass only).
#include "types.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>

namespace draugr {

// Parallel loops here are the order-independent ones: per-row strength,
// column counts (integer atomics), per-node graph construction (each node's
// list in the sequential order), and the PMIS sweeps.  A PMIS sweep reads
// the C/F state of the previous sweep (snapshot): marking C only turns 0 into
// 1 and the C test asks "not F", marking F only turns 0 into -1 and the F test
// asks "is C", so the sequential in-place sweep of the reference decides
// every node exactly as the snapshot sweep does.

namespace {

constexpr double kEps = std::numeric_limits<double>::epsilon();
// Loops over fewer nodes run on one thread: waking the team costs more than
// the loop on the small levels (every parallel loop here is order-independent).
constexpr int32_t kOmpMin = 32768;

inline bool strong_offdiag(const std::vector<uint8_t>& s, int64_t nz, int32_t j, int32_t i) {
    return j != i && s[nz];
}

int32_t uf_find(std::vector<int32_t>& parent, int32_t i) {
    while (parent[i] != i) {
        parent[i] = parent[parent[i]];
        i = parent[i];
    }
    return i;
}

void uf_union(std::vector<int32_t>& parent, int32_t a, int32_t b) {
    int32_t ra = uf_find(parent, a), rb = uf_find(parent, b);
    if (ra != rb) parent[rb] = ra;
}

// Number of rows that strongly depend on each node (column count of S).
// src/coarsening.jl:179-192
std::vector<int32_t> strong_transpose_count(const CSR& A, const std::vector<uint8_t>& s) {
    std::vector<int32_t> cnt(A.nrow, 0);
    int32_t* c = cnt.data();
#pragma omp parallel for schedule(static, 4096) if(A.nrow >= kOmpMin)
    for (int32_t i = 0; i < A.nrow; ++i)
        for (int32_t k = A.rp[i]; k < A.rp[i + 1]; ++k)
            if (strong_offdiag(s, k, A.cv[k], i)) {
#pragma omp atomic
                c[A.cv[k]]++;
            }
    return cnt;
}

// Transpose adjacency of S: for node j, the rows i (ascending) with i -> j strong.
// src/coarsening.jl:827-860.  Rows claim slots of their columns' lists in any
// order, then each list is sorted: a row appears at most once in a column's
// list, so the sorted list is the one the sequential row-order build makes.
void strong_transpose_adj(const CSR& A, const std::vector<uint8_t>& s,
                          std::vector<int64_t>& off, std::vector<int32_t>& src) {
    const int32_t n = A.nrow;
    const std::vector<int32_t> cnt = strong_transpose_count(A, s);
    off.assign(n + 1, 0);
    for (int32_t j = 0; j < n; ++j) off[j + 1] = off[j] + cnt[j];
    src.resize(off[n]);
    std::vector<int64_t> pos(off.begin(), off.end() - 1);
    int64_t* ps = pos.data();
    int32_t* sr = src.data();
#pragma omp parallel for schedule(static, 4096) if(n >= kOmpMin)
    for (int32_t i = 0; i < n; ++i)
        for (int32_t k = A.rp[i]; k < A.rp[i + 1]; ++k)
            if (strong_offdiag(s, k, A.cv[k], i)) {
                int64_t p;
#pragma omp atomic capture
                p = ps[A.cv[k]]++;
                sr[p] = i;
            }
    const int64_t* o = off.data();
#pragma omp parallel for schedule(static, 4096) if(n >= kOmpMin)
    for (int32_t j = 0; j < n; ++j) std::sort(sr + o[j], sr + o[j + 1]);
}

bool has_strong_offdiag(const CSR& A, const std::vector<uint8_t>& s, int32_t i) {
    for (int32_t k = A.rp[i]; k < A.rp[i + 1]; ++k)
        if (strong_offdiag(s, k, A.cv[k], i)) return true;
    return false;
}

// Doubly linked buckets keyed by the measure lambda, head insertion, with the
// same selection semantics as the reference's bucket lists
// (src/coarsening.jl:450-507, 706-776).  A node's links, measure and C/F state
// share one 16-byte record, so a bucket move touches three records (the node
// and its two neighbours) instead of a line per array.
struct Node {
    int32_t next, prev, lam, cf;  // node ids, -1 = none
};
struct Buckets {
    std::vector<int32_t> head;
    Node* nd;
    Buckets(int32_t nb, Node* n) : head(std::max(nb, 1), -1), nd(n) {}
    void push(int32_t i) {
        Node& a = nd[i];
        const int32_t k = a.lam;
        if (k >= (int32_t)head.size()) head.resize(k + 1, -1);
        const int32_t h = head[k];
        head[k] = i;
        a.next = h;
        a.prev = -1;
        if (h >= 0) nd[h].prev = i;
    }
    void remove(int32_t i) {
        Node& a = nd[i];
        const int32_t p = a.prev, nx = a.next;
        if (p >= 0) nd[p].next = nx; else head[a.lam] = nx;
        if (nx >= 0) nd[nx].prev = p;
        a.next = a.prev = -1;
    }
    void update(int32_t i, int32_t v) {
        remove(i);
        nd[i].lam = v;
        push(i);
    }
};

// Ruge-Stueben greedy first pass.  With zpt = true (HMIS), nodes found to have
// zero measure become Z_PT (-2) instead of F; src/coarsening.jl:406-567.
// With classical = true the reference's coarsen_rs variant is run instead
// (no zero-measure preprocessing, plain decrement): src/coarsening.jl:671-805.
std::vector<int8_t> rs_first_pass(const CSR& A, const std::vector<uint8_t>& s, bool zpt, bool classical) {
    const int32_t n = A.nrow;
    const int32_t fpnt = zpt ? -2 : -1;
    std::vector<Node> nd(n);
    {
        const std::vector<int32_t> lam0 = strong_transpose_count(A, s);
        Node* q = nd.data();
#pragma omp parallel for schedule(static, 4096) if(n >= kOmpMin)
        for (int32_t i = 0; i < n; ++i)
            q[i] = Node{-1, -1, lam0[i], (!has_strong_offdiag(A, s, i) && lam0[i] == 0) ? 1 : 0};
    }
    std::vector<int64_t> st_off;
    std::vector<int32_t> st_src;
    strong_transpose_adj(A, s, st_off, st_src);
    const int32_t* rp = A.rp.data();
    const int32_t* cv = A.cv.data();
    const uint8_t* sk = s.data();
    if (!classical) {
        for (int32_t i = 0; i < n; ++i) {
            if (nd[i].cf != 0 || nd[i].lam != 0) continue;
            nd[i].cf = fpnt;
            for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
                const int32_t j = cv[k];
                if (sk[k] && j != i && nd[j].cf == 0) nd[j].lam++;
            }
        }
    }
    int32_t maxlam = 0;
    for (int32_t i = 0; i < n; ++i)
        if (nd[i].cf == 0) maxlam = std::max(maxlam, nd[i].lam);
    Buckets B(classical ? std::max(maxlam + 1, n) : maxlam + 1, nd.data());
    for (int32_t i = 0; i < n; ++i)
        if (nd[i].cf == 0) B.push(i);
    // lambda + 1 for the undecided strong neighbours of a new F point j
    auto raise_neighbours = [&](int32_t j, int32_t& top) {
        for (int32_t k2 = rp[j]; k2 < rp[j + 1]; ++k2) {
            const int32_t k = cv[k2];
            if (sk[k2] && k != j && nd[k].cf == 0) {
                const int32_t v = nd[k].lam + 1;
                B.update(k, v);
                if (v > top) top = v;
            }
        }
    };
    int32_t top = maxlam;
    for (;;) {
        while (top >= 0 && B.head[top] < 0) --top;
        if (top < 0) break;
        const int32_t best = B.head[top];
        B.remove(best);
        nd[best].cf = 1;
        for (int64_t t = st_off[best]; t < st_off[best + 1]; ++t) {
            const int32_t j = st_src[t];
            if (nd[j].cf != 0) continue;
            B.remove(j);
            nd[j].cf = -1;
            raise_neighbours(j, top);
        }
        for (int32_t kk = rp[best]; kk < rp[best + 1]; ++kk) {
            const int32_t j = cv[kk];
            if (!(sk[kk] && j != best && nd[j].cf == 0)) continue;
            const int32_t v = std::max(0, nd[j].lam - 1);
            B.update(j, v);
            if (!classical && v == 0) {
                B.remove(j);
                nd[j].cf = fpnt;
                raise_neighbours(j, top);
            }
        }
    }
    std::vector<int8_t> cf(n);
    for (int32_t i = 0; i < n; ++i) cf[i] = (int8_t)nd[i].cf;
    return cf;
}

// Snapshot PMIS sweeps (see the note at the top); src/coarsening.jl:227-283,
// 575-624.
template <typename OFF>
void sweeps_host(std::vector<int8_t>& cf, int32_t n, const OFF* off, const int32_t* adj, const uint8_t* s,
                 const std::vector<double>& measure, bool low_measure_fine) {
    std::vector<int8_t> nxt(cf);
    for (int64_t iter = 0; iter <= (int64_t)n; ++iter) {
        int any_undecided = 0;
#pragma omp parallel for schedule(static, 4096) reduction(| : any_undecided) if(n >= kOmpMin)
        for (int32_t i = 0; i < n; ++i) {
            if (cf[i] != 0) continue;
            any_undecided = 1;
            if (low_measure_fine && measure[i] < 1.0) { nxt[i] = -1; continue; }
            bool is_max = true;
            for (OFF k = off[i]; k < off[i + 1]; ++k) {
                const int32_t j = adj[k];
                if (j != i && (!s || s[k]) && cf[j] != -1 && measure[j] > measure[i]) {
                    is_max = false;
                    break;
                }
            }
            if (is_max) nxt[i] = 1;
        }
        cf = nxt;
#pragma omp parallel for schedule(static, 4096) if(n >= kOmpMin)
        for (int32_t i = 0; i < n; ++i) {
            if (cf[i] != 0) continue;
            for (OFF k = off[i]; k < off[i + 1]; ++k)
                if ((!s || s[k]) && cf[adj[k]] == 1) { nxt[i] = -1; break; }
        }
        cf = nxt;
        if (!any_undecided) break;
    }
}

// The sweeps over the strong graph of A, on the device when the hook asks.
void pmis_on_matrix(std::vector<int8_t>& cf, const CSR& A, const std::vector<uint8_t>& s,
                    const std::vector<double>& measure, bool low_measure_fine, const PmisHook& ph) {
    if (ph.fn && A.nrow >= ph.min_rows) {
        const std::vector<int64_t> off(A.rp.begin(), A.rp.end());
        ph.fn(cf, A.nrow, off.data(), A.cv.data(), s.data(), measure, low_measure_fine);
    } else {
        sweeps_host<int32_t>(cf, A.nrow, A.rp.data(), A.cv.data(), s.data(), measure, low_measure_fine);
    }
}

// Promote F points without a strong C neighbour; src/coarsening.jl:291-313.
void ensure_fine_have_coarse_neighbor(std::vector<int8_t>& cf, const CSR& A, const std::vector<uint8_t>& s) {
    bool changed = true;
    while (changed) {
        changed = false;
        for (int32_t i = 0; i < A.nrow; ++i) {
            if (cf[i] != -1) continue;
            bool has_c = false;
            for (int32_t k = A.rp[i]; k < A.rp[i + 1]; ++k)
                if (strong_offdiag(s, k, A.cv[k], i) && cf[A.cv[k]] == 1) { has_c = true; break; }
            if (!has_c) { cf[i] = 1; changed = true; }
        }
    }
}

int32_t count_coarse(const std::vector<int8_t>& cf) {
    int32_t nc = 0;
    for (auto c : cf) nc += (c == 1);
    return nc;
}

}  // namespace

void pmis_sweeps(std::vector<int8_t>& cf, int32_t n, const int64_t* off, const int32_t* adj, const uint8_t* s,
                 const std::vector<double>& measure, bool low_measure_fine) {
    sweeps_host<int64_t>(cf, n, off, adj, s, measure, low_measure_fine);
}

// Absolute strength of connection; src/strength.jl:24-61.
std::vector<uint8_t> strength_absolute(const CSR& A, double theta) {
    std::vector<uint8_t> s(A.nnz(), 0);
#pragma omp parallel for schedule(static, 4096) if(A.nrow >= kOmpMin)
    for (int32_t i = 0; i < A.nrow; ++i) {
        double mx = 0.0;
        for (int32_t k = A.rp[i]; k < A.rp[i + 1]; ++k)
            if (A.cv[k] != i) mx = std::max(mx, std::fabs(A.nz[k]));
        const double thr = theta * mx;
        for (int32_t k = A.rp[i]; k < A.rp[i + 1]; ++k)
            s[k] = (A.cv[k] != i) && (std::fabs(A.nz[k]) >= thr);
    }
    return s;
}

// Max-row-sum dependency weakening, used for strength only; src/strength.jl:183-213.
CSR apply_max_row_sum(const CSR& A, double threshold) {
    CSR W = A;
    for (int32_t i = 0; i < A.nrow; ++i) {
        double diag = 0.0, rsum = 0.0;
        for (int32_t k = A.rp[i]; k < A.rp[i + 1]; ++k) {
            rsum += A.nz[k];
            if (A.cv[k] == i) diag = A.nz[k];
        }
        const double ad = std::fabs(diag);
        if (ad < kEps) continue;
        if (std::fabs(rsum) > ad * threshold && threshold < 1.0)
            for (int32_t k = A.rp[i]; k < A.rp[i + 1]; ++k)
                if (A.cv[k] != i) W.nz[k] = 0.0;
    }
    return W;
}

// Greedy aggregation with distance-2 seeds, neighbour assignment and singleton
// merging; src/coarsening.jl:14-168.  agg[i] is 0-based.
void coarsen_aggregation(const CSR& A, double theta, std::vector<int32_t>& agg, int32_t& nc) {
    const int32_t n = A.nrow;
    if (n <= 1) {
        agg.assign(n, 0);
        nc = n;
        return;
    }
    const std::vector<uint8_t> s = strength_absolute(A, theta);
    agg.assign(n, -1);
    nc = 0;
    std::vector<int32_t> scount(n, 0);
    for (int32_t i = 0; i < n; ++i)
        for (int32_t k = A.rp[i]; k < A.rp[i + 1]; ++k)
            if (strong_offdiag(s, k, A.cv[k], i)) scount[i]++;
    // phase 1: seeds in decreasing strong-count order (stable)
    std::vector<int8_t> state(n, 0);
    std::vector<int32_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t a, int32_t b) { return scount[a] > scount[b]; });
    for (int32_t idx = 0; idx < n; ++idx) {
        const int32_t i = order[idx];
        if (state[i] != 0) continue;
        bool can_seed = true;
        for (int32_t k = A.rp[i]; k < A.rp[i + 1]; ++k) {
            const int32_t j = A.cv[k];
            if (strong_offdiag(s, k, j, i) && (state[j] == 1 || state[j] == -1)) { can_seed = false; break; }
        }
        if (!can_seed) continue;
        state[i] = 1;
        agg[i] = nc++;
        for (int32_t k = A.rp[i]; k < A.rp[i + 1]; ++k) {
            const int32_t j = A.cv[k];
            if (!(strong_offdiag(s, k, j, i) && state[j] == 0)) continue;
            state[j] = -1;
            agg[j] = agg[i];
            for (int32_t k2 = A.rp[j]; k2 < A.rp[j + 1]; ++k2) {
                const int32_t m = A.cv[k2];
                if (strong_offdiag(s, k2, m, j) && state[m] == 0) state[m] = -2;
            }
        }
    }
    // phase 2: attach the rest to the strongest aggregated neighbour
    for (int32_t i = 0; i < n; ++i) {
        if (agg[i] >= 0) continue;
        int32_t best = -1;
        double bv = 0.0;
        for (int32_t k = A.rp[i]; k < A.rp[i + 1]; ++k) {
            const int32_t j = A.cv[k];
            if (strong_offdiag(s, k, j, i) && agg[j] >= 0 && std::fabs(A.nz[k]) > bv) {
                bv = std::fabs(A.nz[k]);
                best = agg[j];
            }
        }
        if (best < 0)
            for (int32_t k = A.rp[i]; k < A.rp[i + 1]; ++k) {
                const int32_t j = A.cv[k];
                if (j != i && agg[j] >= 0 && std::fabs(A.nz[k]) > bv) {
                    bv = std::fabs(A.nz[k]);
                    best = agg[j];
                }
            }
        agg[i] = best >= 0 ? best : nc++;
    }
    // phase 3: merge aggregates of size <= 1 into the strongest neighbour
    std::vector<int32_t> size(nc, 0);
    for (int32_t i = 0; i < n; ++i) size[agg[i]]++;
    std::vector<int32_t> parent(nc);
    std::iota(parent.begin(), parent.end(), 0);
    for (int32_t i = 0; i < n; ++i) {
        const int32_t mine = uf_find(parent, agg[i]);
        if (size[mine] > 1) continue;
        int32_t target = -1;
        double bv = 0.0;
        for (int32_t k = A.rp[i]; k < A.rp[i + 1]; ++k) {
            const int32_t j = A.cv[k];
            if (j == i) continue;
            const int32_t rj = uf_find(parent, agg[j]);
            if (rj != mine && std::fabs(A.nz[k]) > bv) {
                bv = std::fabs(A.nz[k]);
                target = rj;
            }
        }
        if (target >= 0) {
            const int32_t sm = size[mine], st = size[target];
            uf_union(parent, target, mine);
            size[uf_find(parent, target)] = sm + st;
        }
    }
    std::vector<int32_t> new_id(nc, -1);
    int32_t cnt = 0;
    for (int32_t a = 0; a < nc; ++a) {
        const int32_t r = uf_find(parent, a);
        if (new_id[r] < 0) new_id[r] = cnt++;
        new_id[a] = new_id[r];
    }
    for (int32_t i = 0; i < n; ++i) agg[i] = new_id[uf_find(parent, agg[i])];
    nc = cnt;
}

// PMIS with the column-count measure plus a uniform random; src/coarsening.jl:202-283.
void coarsen_pmis(const CSR& A, double theta, Rng& rng, std::vector<int8_t>& cf, int32_t& nc, const PmisHook& ph) {
    const int32_t n = A.nrow;
    if (n <= 1) {
        cf.assign(n, 1);
        nc = n;
        return;
    }
    const std::vector<uint8_t> s = strength_absolute(A, theta);
    const std::vector<int32_t> st = strong_transpose_count(A, s);
    std::vector<double> measure(n);
    for (int32_t i = 0; i < n; ++i) measure[i] = (double)st[i] + rng.uniform();
    cf.assign(n, 0);
    for (int32_t i = 0; i < n; ++i)
        if (!has_strong_offdiag(A, s, i) && st[i] == 0) cf[i] = 1;
    pmis_on_matrix(cf, A, s, measure, false, ph);
    for (auto& c : cf)
        if (c == 0) c = 1;
    nc = count_coarse(cf);
}

// HMIS = RS first pass (Z_PT variant) + PMIS on the rest; src/coarsening.jl:330-394.
void coarsen_hmis(const CSR& A, double theta, Rng& rng, std::vector<int8_t>& cf, int32_t& nc, const PmisHook& ph) {
    const int32_t n = A.nrow;
    if (n <= 1) {
        cf.assign(n, 1);
        nc = n;
        return;
    }
    const std::vector<uint8_t> s = strength_absolute(A, theta);
    cf = rs_first_pass(A, s, /*zpt=*/true, /*classical=*/false);
    const std::vector<int32_t> st = strong_transpose_count(A, s);
    std::vector<double> measure(n);
    for (int32_t i = 0; i < n; ++i) measure[i] = (double)st[i] + rng.uniform();
    for (int32_t i = 0; i < n; ++i) {
        if (cf[i] != -2) continue;
        cf[i] = (measure[i] >= 1.0 || has_strong_offdiag(A, s, i)) ? 0 : -1;
    }
    pmis_on_matrix(cf, A, s, measure, true, ph);
    for (auto& c : cf)
        if (c == 0) c = -1;
    nc = count_coarse(cf);
}

// Classical Ruge-Stueben (first + second pass); src/coarsening.jl:671-818.
void coarsen_rs(const CSR& A, double theta, std::vector<int8_t>& cf, int32_t& nc) {
    const int32_t n = A.nrow;
    if (n <= 1) {
        cf.assign(n, 1);
        nc = n;
        return;
    }
    const std::vector<uint8_t> s = strength_absolute(A, theta);
    cf = rs_first_pass(A, s, /*zpt=*/false, /*classical=*/true);
    ensure_fine_have_coarse_neighbor(cf, A, s);
    nc = count_coarse(cf);
}

// Two-pass aggressive C/F splitting: base splitting, then PMIS on the
// distance-2 C graph; src/coarsening.jl:990-1116.
void coarsen_aggressive_cf(const CSR& A_in, double theta, bool hmis_base, Rng& rng,
                           std::vector<int8_t>& cf, int32_t& nc, const PmisHook& ph) {
    const int32_t n = A_in.nrow;
    if (n <= 1) {
        cf.assign(n, 1);
        nc = n;
        return;
    }
    std::vector<int8_t> cf1;
    int32_t nc1 = 0;
    if (hmis_base) coarsen_hmis(A_in, theta, rng, cf1, nc1, ph);
    else coarsen_pmis(A_in, theta, rng, cf1, nc1, ph);
    const CSR& A = A_in;
    const std::vector<uint8_t> s = strength_absolute(A, theta);

    std::vector<int32_t> c_idx, c_loc(n, -1);
    for (int32_t i = 0; i < n; ++i)
        if (cf1[i] == 1) { c_loc[i] = (int32_t)c_idx.size(); c_idx.push_back(i); }
    const int32_t m = (int32_t)c_idx.size();

    // Strong C neighbours of every row (CSR), and its transpose restricted to C.
    // Per-row lists in parallel (count, scan, fill); the transpose claims slots
    // in any order and sorts each list, which gives the ascending row order of
    // the sequential build (a row appears at most once in a list).
    std::vector<int64_t> sc_off(n + 1, 0);
    int64_t* so = sc_off.data();
#pragma omp parallel for schedule(static, 4096) if(n >= kOmpMin)
    for (int32_t i = 0; i < n; ++i) {
        int64_t c = 0;
        for (int32_t k = A.rp[i]; k < A.rp[i + 1]; ++k)
            if (strong_offdiag(s, k, A.cv[k], i) && cf1[A.cv[k]] == 1) ++c;
        so[i + 1] = c;
    }
    for (int32_t i = 0; i < n; ++i) sc_off[i + 1] += sc_off[i];
    std::vector<int32_t> sc(sc_off[n]);
    int32_t* scp = sc.data();
#pragma omp parallel for schedule(static, 4096) if(n >= kOmpMin)
    for (int32_t i = 0; i < n; ++i) {
        int64_t p = so[i];
        for (int32_t k = A.rp[i]; k < A.rp[i + 1]; ++k)
            if (strong_offdiag(s, k, A.cv[k], i) && cf1[A.cv[k]] == 1) scp[p++] = c_loc[A.cv[k]];
    }
    std::vector<int64_t> tr_off(m + 1, 0);
    {
        std::vector<int32_t> tcnt(m, 0);
        int32_t* tc = tcnt.data();
#pragma omp parallel for schedule(static, 65536) if(n >= kOmpMin)
        for (int64_t t = 0; t < sc_off[n]; ++t) {
#pragma omp atomic
            tc[scp[t]]++;
        }
        for (int32_t a = 0; a < m; ++a) tr_off[a + 1] = tr_off[a] + tcnt[a];
    }
    std::vector<int32_t> tr(tr_off[m]);
    {
        std::vector<int64_t> pos(tr_off.begin(), tr_off.end() - 1);
        int64_t* ps = pos.data();
        int32_t* trp = tr.data();
#pragma omp parallel for schedule(static, 4096) if(n >= kOmpMin)
        for (int32_t i = 0; i < n; ++i)
            for (int64_t t = so[i]; t < so[i + 1]; ++t) {
                int64_t p;
#pragma omp atomic capture
                p = ps[scp[t]]++;
                trp[p] = i;
            }
        const int64_t* to = tr_off.data();
#pragma omp parallel for schedule(static, 4096) if(m >= kOmpMin)
        for (int32_t a = 0; a < m; ++a) std::sort(trp + to[a], trp + to[a + 1]);
    }
    // Neighbour set of C point a (local ids): pairs of strong C neighbours of a
    // common row, plus direct strong C-C connections in both directions.
    // Built per node in parallel (count, then fill), each list in the order
    // the sequential construction appends it.
    std::vector<int64_t> g_off(m + 1, 0);
    std::vector<int32_t> g;
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1) {
            for (int32_t a = 0; a < m; ++a) g_off[a + 1] += g_off[a];
            g.resize(g_off[m]);
        }
#pragma omp parallel if(m >= kOmpMin)
        {
            std::vector<int32_t> mark(m, -1);
#pragma omp for schedule(dynamic, 1024)
            for (int32_t a = 0; a < m; ++a) {
                int64_t cnt = 0;
                const int64_t base = pass == 1 ? g_off[a] : 0;
                auto add = [&](int32_t b) {
                    if (b != a && mark[b] != a) {
                        mark[b] = a;
                        if (pass == 1) g[base + cnt] = b;
                        ++cnt;
                    }
                };
                const int32_t ia = c_idx[a];
                for (int64_t t = sc_off[ia]; t < sc_off[ia + 1]; ++t) add(sc[t]);
                for (int64_t t = tr_off[a]; t < tr_off[a + 1]; ++t) {
                    const int32_t i = tr[t];
                    for (int64_t u = sc_off[i]; u < sc_off[i + 1]; ++u) add(sc[u]);
                    if (cf1[i] == 1) add(c_loc[i]);
                }
                if (pass == 0) g_off[a + 1] = cnt;
            }
        }
    }
    std::vector<double> measure(m);
    for (int32_t a = 0; a < m; ++a) measure[a] = (double)(g_off[a + 1] - g_off[a]) + rng.uniform();
    std::vector<int8_t> cf2(m, 0);  // every edge of the C graph is strong, no node lists itself
    if (ph.fn && m >= ph.min_rows) ph.fn(cf2, m, g_off.data(), g.data(), nullptr, measure, false);
    else sweeps_host<int64_t>(cf2, m, g_off.data(), g.data(), nullptr, measure, false);
    for (auto& c : cf2)
        if (c == 0) c = 1;
    cf = cf1;
    for (int32_t a = 0; a < m; ++a)
        if (cf2[a] == -1) cf[c_idx[a]] = -1;
    ensure_fine_have_coarse_neighbor(cf, A, s);
    nc = count_coarse(cf);
}

// Aggregation-style aggressive coarsening built on PMIS (AggressiveCoarsening
// with a base other than :pmis/:hmis); src/coarsening.jl:869-957.
void coarsen_aggressive_agg(const CSR& A, double theta, Rng& rng, std::vector<int32_t>& agg, int32_t& nc,
                            const PmisHook& ph) {
    const int32_t n = A.nrow;
    std::vector<int8_t> cf;
    int32_t nc1 = 0;
    coarsen_pmis(A, theta, rng, cf, nc1, ph);
    const std::vector<uint8_t> s = strength_absolute(A, theta);
    agg.assign(n, -1);
    int32_t cnt = 0;
    for (int32_t i = 0; i < n; ++i)
        if (cf[i] == 1) agg[i] = cnt++;
    for (int32_t i = 0; i < n; ++i) {
        if (cf[i] == 1) continue;
        int32_t best = -1;
        double bv = 0.0;
        for (int32_t k = A.rp[i]; k < A.rp[i + 1]; ++k) {
            const int32_t j = A.cv[k];
            if (j != i && cf[j] == 1 && std::fabs(A.nz[k]) > bv) { bv = std::fabs(A.nz[k]); best = agg[j]; }
        }
        if (best >= 0) agg[i] = best;
    }
    std::vector<int32_t> parent(cnt);
    std::iota(parent.begin(), parent.end(), 0);
    std::vector<int32_t> roots;
    for (int32_t i = 0; i < n; ++i) {
        if (cf[i] != -1) continue;
        roots.clear();
        for (int32_t k = A.rp[i]; k < A.rp[i + 1]; ++k)
            if (s[k] && cf[A.cv[k]] == 1) roots.push_back(uf_find(parent, agg[A.cv[k]]));
        for (size_t r = 1; r < roots.size(); ++r) uf_union(parent, roots[0], roots[r]);
    }
    std::vector<int32_t> new_id(cnt, -1);
    nc = 0;
    for (int32_t a = 0; a < cnt; ++a) {
        const int32_t r = uf_find(parent, a);
        if (new_id[r] < 0) new_id[r] = nc++;
        new_id[a] = new_id[r];
    }
    for (int32_t i = 0; i < n; ++i)
        if (agg[i] >= 0) agg[i] = new_id[uf_find(parent, agg[i])];
    for (int32_t i = 0; i < n; ++i) {
        if (agg[i] >= 0) continue;
        int32_t best = -1;
        double bv = 0.0;
        for (int32_t k = A.rp[i]; k < A.rp[i + 1]; ++k) {
            const int32_t j = A.cv[k];
            if (j != i && agg[j] >= 0 && std::fabs(A.nz[k]) > bv) { bv = std::fabs(A.nz[k]); best = agg[j]; }
        }
        agg[i] = best >= 0 ? best : nc++;
    }
}

}  // namespace draugr
