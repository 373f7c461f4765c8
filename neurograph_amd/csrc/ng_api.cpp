// C-ABI layer of libng_hip.so: contexts, device-resident graphs, the batched path driver.
// See include/neurograph_hip.h for the contract of every entry point.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <numeric>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "ng_internal.h"

using namespace ng;

namespace {

thread_local std::string g_error;

ng_status fail(ng_status st, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_error = buf;
    return st;
}

#define NG_HIP(call)                                                                                  \
    do {                                                                                              \
        hipError_t e_ = (call);                                                                       \
        if (e_ != hipSuccess) {                                                                       \
            return fail(e_ == hipErrorOutOfMemory ? NG_ENOMEM : NG_EDEVICE, "%s failed: %s (%s:%d)", #call, \
                        hipGetErrorString(e_), __FILE__, __LINE__);                                   \
        }                                                                                             \
    } while (0)

template <class T>
void dfree(T*& p) {
    if (p) (void)hipFree(static_cast<void*>(p));
    p = nullptr;
}

size_t dk_size(DKind dk) { return (dk == DKind::U32 || dk == DKind::F32) ? 4 : 8; }

int dist_kind_of(DKind dk) {
    switch (dk) {
    case DKind::U32:
    case DKind::U64: return NG_DIST_U64;
    case DKind::F32: return NG_DIST_F32;
    default: return NG_DIST_F64;
    }
}

// Runs fn(begin, end) over [0, count) on up to hardware_concurrency threads.
template <class Fn>
void parallel_for(uint64_t count, Fn&& fn) {
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (count < 4096) nt = 1;
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nt; ++t) {
        const uint64_t b = count * t / nt, e = count * (t + 1) / nt;
        if (nt == 1) {
            fn(b, e);
        } else {
            pool.emplace_back([b, e, &fn] { fn(b, e); });
        }
    }
    for (auto& th : pool) th.join();
}

}  // namespace

struct ng_context {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int cus = 0;
    Workspace ws;
    hipError_t scratch_status = hipSuccess;  // result of resolving the device's scratch array at creation
    // pinned staging of a call's betweenness vector on its way to the caller's (calls on a context are sequential)
    double* host_bc = nullptr;
    uint32_t host_bc_cap = 0;
    // the creator's handle plus one per live graph and cohort: ng_context_destroy releases the creator's share, and
    // the context is torn down when the last share goes, so graphs and cohorts may be destroyed after it
    std::atomic<uint32_t> refs{1};
};

namespace {

void free_ws(Workspace& w);

void context_release(ng_context* ctx) {
    if (ctx->refs.fetch_sub(1, std::memory_order_acq_rel) != 1) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->host_bc) (void)hipHostFree(ctx->host_bc);
    free_ws(ctx->ws);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

}  // namespace

struct ng_graph {
    ng_context* ctx = nullptr;
    uint32_t n = 0;
    uint64_t m = 0;
    ng_weight_kind wk = NG_W_I32;
    std::vector<uint64_t> off;
    std::vector<uint32_t> tgt;
    std::vector<unsigned char> w;  // raw weights
    bool has_negative = false;
    View views[2];                 // [0] |w|, [1] 1/|w|
    ng_call_stats last{};
    // launch options (ng_graph_set_option): 0 / -1 = automatic
    int opt_block = 0;
    uint32_t opt_max_slots = 0;
    int opt_lds_maps = -1;
    uint32_t opt_pile_capacity = 0;
    uint32_t opt_log_capacity = 0;
    uint64_t weights_version = 0;  // bumped by ng_graph_set_weights (cohorts re-capture)
};

// Cohort: all sources of many graphs in one persistent launch, captured once in a hipGraph and replayed.
struct ng_cohort {
    ng_context* ctx = nullptr;
    std::vector<ng_graph*> graphs;
    DKind dk = DKind::U32;
    uint32_t n_max = 0;
    uint64_t m_max = 0;
    uint32_t total = 0;                // sum of member sizes = number of (graph, source) items
    std::vector<uint32_t> prefix;      // [G+1] item / node prefix per graph
    std::vector<double> delta;         // per graph (0: the graph's tuned or heuristic width)
    int block = kBlock;                // lanes per workgroup
    int bpc = 0;                       // resident workgroups per CU (0: occupancy maximum)
    uint32_t slots = 0;
    uint64_t ws_generation = 0;        // context workspace generation the capture was made against
    CohortGraph* d_gtab = nullptr;
    RowRange* d_rows = nullptr;        // row bounds and light-prefix ends of every member for its own width
    uint32_t* d_prefix = nullptr;
    uint64_t* d_fbase = nullptr;
    uint32_t* d_seg_ptr = nullptr;
    uint64_t* d_seg_off = nullptr;
    double* d_flush = nullptr;
    double* d_bc = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool built = false;
    std::vector<uint64_t> versions;    // members' weights_version the capture was made against
    ng_call_stats last{};
};

namespace {

double weight_at(const ng_graph* g, uint64_t e) {
    switch (g->wk) {
    case NG_W_I32: return static_cast<double>(reinterpret_cast<const int32_t*>(g->w.data())[e]);
    case NG_W_F32: return static_cast<double>(reinterpret_cast<const float*>(g->w.data())[e]);
    default: return reinterpret_cast<const double*>(g->w.data())[e];
    }
}

// Lengths of the caller's arcs for |w| (inverse = false) or 1/|w| (inverse = true), following
// DijkstraCommon::transform_weight (DijkstraCommon.h:36-54).
template <class L>
void arc_lengths(const ng_graph* g, bool inverse, std::vector<L>& len) {
    len.resize(g->m);
    parallel_for(g->m, [&](uint64_t b, uint64_t e) {
        for (uint64_t a = b; a < e; ++a) {
            if (g->wk == NG_W_I32) {
                const long long aw = std::llabs(static_cast<long long>(reinterpret_cast<const int32_t*>(g->w.data())[a]));
                if (inverse) {
                    len[a] = aw == 0 ? std::numeric_limits<L>::infinity() : static_cast<L>(1.0 / static_cast<double>(aw));
                } else {
                    len[a] = static_cast<L>(aw);
                }
            } else if (g->wk == NG_W_F32) {
                const float aw = std::fabs(reinterpret_cast<const float*>(g->w.data())[a]);
                if (inverse) {
                    len[a] = aw == 0.0f ? std::numeric_limits<L>::infinity() : static_cast<L>(1.0 / static_cast<double>(aw));
                } else {
                    len[a] = static_cast<L>(aw);
                }
            } else {
                const double aw = std::fabs(reinterpret_cast<const double*>(g->w.data())[a]);
                if (inverse) {
                    len[a] = aw == 0.0 ? std::numeric_limits<L>::infinity() : static_cast<L>(1.0 / aw);
                } else {
                    len[a] = static_cast<L>(aw);
                }
            }
        }
    });
}

// Sorts row u of the view by (length, target), ties in the caller's arc order (a stable sort of the row's arcs).
template <class L>
void sort_row(const ng_graph* g, uint32_t u, const std::vector<L>& len, std::vector<uint32_t>& idx, uint32_t* perm) {
    const uint64_t r0 = g->off[u], r1 = g->off[u + 1];
    idx.resize(r1 - r0);
    std::iota(idx.begin(), idx.end(), static_cast<uint32_t>(r0));
    std::stable_sort(idx.begin(), idx.end(), [&](uint32_t x, uint32_t y) {
        if (len[x] != len[y]) return len[x] < len[y];
        return g->tgt[x] < g->tgt[y];
    });
    std::copy(idx.begin(), idx.end(), perm + r0);
}

// Whether row u of perm is still the stable (length, target) order of its arcs under `len`.
template <class L>
bool row_in_order(const ng_graph* g, uint32_t u, const std::vector<L>& len, const uint32_t* perm) {
    for (uint64_t k = g->off[u]; k + 1 < g->off[u + 1]; ++k) {
        const uint32_t x = perm[k], y = perm[k + 1];
        if (len[x] < len[y]) continue;
        if (len[x] > len[y]) return false;
        if (g->tgt[x] < g->tgt[y]) continue;
        if (g->tgt[x] > g->tgt[y] || x > y) return false;
    }
    return true;
}

// Whether the arcs of finite length (lengths in the caller's arc order) connect every node to every other: a
// search from node 0 along them and one against them both reach every node.
template <class L>
bool strongly_connected(const ng_graph* g, const std::vector<L>& len) {
    const uint32_t n = g->n;
    if (n <= 1) return true;
    auto finite = [&](uint64_t a) -> bool {
        if constexpr (std::is_floating_point_v<L>) return std::isfinite(len[a]);
        else return true;
    };
    std::vector<uint8_t> seen(n, 0);
    std::vector<uint32_t> stack;
    stack.reserve(n);
    auto search = [&](const uint64_t* off, const uint32_t* tgt, bool all) -> bool {
        std::fill(seen.begin(), seen.end(), uint8_t(0));
        seen[0] = 1;
        stack.assign(1, 0u);
        uint32_t reached = 1;
        while (!stack.empty()) {
            const uint32_t u = stack.back();
            stack.pop_back();
            for (uint64_t a = off[u]; a < off[u + 1]; ++a) {
                if (all && !finite(a)) continue;
                const uint32_t t = tgt[a];
                if (!seen[t]) {
                    seen[t] = 1;
                    ++reached;
                    stack.push_back(t);
                }
            }
        }
        return reached == n;
    };
    if (!search(g->off.data(), g->tgt.data(), true)) return false;
    // the reversed arcs of finite length
    std::vector<uint64_t> roff(static_cast<size_t>(n) + 1, 0);
    for (uint32_t u = 0; u < n; ++u)
        for (uint64_t a = g->off[u]; a < g->off[u + 1]; ++a)
            if (finite(a)) ++roff[g->tgt[a] + 1];
    for (uint32_t v = 0; v < n; ++v) roff[v + 1] += roff[v];
    std::vector<uint32_t> rsrc(roff[n]);
    std::vector<uint64_t> cur(roff.begin(), roff.end() - 1);
    for (uint32_t u = 0; u < n; ++u)
        for (uint64_t a = g->off[u]; a < g->off[u + 1]; ++a)
            if (finite(a)) rsrc[cur[g->tgt[a]]++] = u;
    return search(roff.data(), rsrc.data(), false);
}

// Host-side construction of the device CSR for |w| (vi = 0) or 1/|w| (vi = 1). Rows are sorted by (length,
// target) so that the arcs lighter than the bucket width form a prefix. With the view already built in kind dk
// (a weight update), its device arrays are reused: rows whose order the new lengths keep are not re-sorted, and
// when no row changes order only the lengths are uploaded.
template <class L>
ng_status build_view_typed(ng_graph* g, int vi, DKind dk) {
    View& v = g->views[vi];
    const uint32_t n = g->n;
    const uint64_t m = g->m;
    const bool update = v.built && v.dk == dk;
    std::vector<L> len;
    arc_lengths<L>(g, vi == 1, len);
    bool reordered = !update;
    if (!update) {
        v.perm_host.assign(m, 0u);
        parallel_for(n, [&](uint64_t b, uint64_t e) {
            std::vector<uint32_t> idx;
            for (uint64_t u = b; u < e; ++u) sort_row<L>(g, static_cast<uint32_t>(u), len, idx, v.perm_host.data());
        });
    } else {
        std::atomic<bool> any{false};
        parallel_for(n, [&](uint64_t b, uint64_t e) {
            std::vector<uint32_t> idx;
            for (uint64_t u = b; u < e; ++u) {
                if (row_in_order<L>(g, static_cast<uint32_t>(u), len, v.perm_host.data())) continue;
                sort_row<L>(g, static_cast<uint32_t>(u), len, idx, v.perm_host.data());
                any.store(true, std::memory_order_relaxed);
            }
        });
        reordered = any.load();
    }
    std::vector<uint32_t> stgt(reordered ? m : 0);
    std::vector<L> slen(m);
    parallel_for(m, [&](uint64_t b, uint64_t e) {
        for (uint64_t k = b; k < e; ++k) {
            slen[k] = len[v.perm_host[k]];
            if (reordered) stgt[k] = g->tgt[v.perm_host[k]];
        }
    });
    // statistics over finite lengths
    double mn = std::numeric_limits<double>::infinity(), mx = 0.0, sum = 0.0;
    uint64_t fin = 0;
    bool zero = false;
    for (uint64_t a = 0; a < m; ++a) {
        const double x = static_cast<double>(slen[a]);
        if (!std::isfinite(x)) continue;
        ++fin;
        sum += x;
        mn = std::min(mn, x);
        mx = std::max(mx, x);
        if (x == 0.0) zero = true;
    }
    double q05 = 1.0;
    if (fin) {
        std::vector<double> sample;
        const uint64_t step = std::max<uint64_t>(1, m / 65536);
        for (uint64_t a = 0; a < m; a += step) {
            const double x = static_cast<double>(slen[a]);
            if (std::isfinite(x)) sample.push_back(x);
        }
        if (!sample.empty()) {
            const size_t k = sample.size() / 20;
            std::nth_element(sample.begin(), sample.begin() + k, sample.end());
            q05 = sample[k];
        }
    }
    NG_HIP(hipSetDevice(g->ctx->device));
    if (!update) {
        std::vector<uint32_t> off32(static_cast<size_t>(n) + 1);
        for (uint32_t u = 0; u <= n; ++u) off32[u] = static_cast<uint32_t>(g->off[u]);
        NG_HIP(hipMalloc(&v.off, sizeof(uint32_t) * (static_cast<size_t>(n) + 1)));
        NG_HIP(hipMalloc(&v.tgt, sizeof(uint32_t) * std::max<uint64_t>(m, 1)));
        NG_HIP(hipMalloc(&v.len, sizeof(L) * std::max<uint64_t>(m, 1)));
        NG_HIP(hipMalloc(&v.perm, sizeof(uint32_t) * std::max<uint64_t>(m, 1)));
        NG_HIP(hipMalloc(&v.rows, sizeof(RowRange) * std::max<uint32_t>(n, 1)));
        NG_HIP(hipMemcpy(v.off, off32.data(), sizeof(uint32_t) * off32.size(), hipMemcpyHostToDevice));
    } else {
        // earlier calls on the context stream may still read the arrays
        NG_HIP(hipStreamSynchronize(g->ctx->stream));
    }
    if (m) {
        if (reordered) {
            NG_HIP(hipMemcpy(v.tgt, stgt.data(), sizeof(uint32_t) * m, hipMemcpyHostToDevice));
            NG_HIP(hipMemcpy(v.perm, v.perm_host.data(), sizeof(uint32_t) * m, hipMemcpyHostToDevice));
        }
        NG_HIP(hipMemcpy(v.len, slen.data(), sizeof(L) * m, hipMemcpyHostToDevice));
    }
    if constexpr (sizeof(L) == 4) {
        if (v.arcs == nullptr) NG_HIP(hipMalloc(&v.arcs, 8 * std::max<uint64_t>(m, 1)));
        if (m) {
            NG_HIP(launch_interleave_arcs(v.tgt, static_cast<const uint32_t*>(v.len), m, v.arcs, g->ctx->stream));
            NG_HIP(hipStreamSynchronize(g->ctx->stream));
        }
    } else {
        dfree(v.arcs);  // 8-byte lengths are read from tgt and len
    }
    {
        // the arcs of finite length are those of before when every length was and is finite
        const bool all_finite = fin == m;
        if (!(update && all_finite && v.all_finite)) v.strong = strongly_connected<L>(g, len);
        v.all_finite = all_finite;
    }
    v.dk = dk;
    v.min_len = fin ? mn : 0.0;
    v.max_len = mx;
    v.mean_len = fin ? sum / static_cast<double>(fin) : 0.0;
    v.q05_len = q05;
    v.has_zero = zero;
    v.lend_delta = -1.0;   // light-prefix ends follow the lengths
    v.bound_inv_q = 0.0;   // the 4-bit code scale is re-estimated from the new distances
    v.built = true;
    v.reordered_last = reordered;
    return NG_OK;
}

// Distance kind of view vi: 1/|w| in f64; |w| as u32 for integer weights whose longest simple path ((n-1) max|w|)
// stays below the sentinel, else u64; f32 and f64 weights keep their type.
DKind wanted_kind(const ng_graph* g, int vi) {
    if (vi == 1) return DKind::F64;
    switch (g->wk) {
    case NG_W_I32: {
        uint64_t maxw = 0;
        const int32_t* w = reinterpret_cast<const int32_t*>(g->w.data());
        for (uint64_t a = 0; a < g->m; ++a) maxw = std::max<uint64_t>(maxw, static_cast<uint64_t>(std::llabs(static_cast<long long>(w[a]))));
        const long double bound = static_cast<long double>(g->n) * static_cast<long double>(maxw);
        return bound < static_cast<long double>(0xFFFFFFF0u) ? DKind::U32 : DKind::U64;
    }
    case NG_W_F32: return DKind::F32;
    default: return DKind::F64;
    }
}

ng_status build_view(ng_graph* g, int vi, DKind dk) {
    switch (dk) {
    case DKind::U32: return build_view_typed<uint32_t>(g, vi, dk);
    case DKind::U64: return build_view_typed<uint64_t>(g, vi, dk);
    case DKind::F32: return build_view_typed<float>(g, vi, dk);
    default: return build_view_typed<double>(g, vi, dk);
    }
}

ng_status ensure_view(ng_graph* g, int vi) {
    if (g->views[vi].built) return NG_OK;
    return build_view(g, vi, wanted_kind(g, vi));
}

void free_view(View& v) {
    dfree(v.off);
    dfree(v.tgt);
    if (v.len) (void)hipFree(v.len);
    v.len = nullptr;
    dfree(v.perm);
    dfree(v.rows);
    dfree(v.arcs);
    std::vector<uint32_t>().swap(v.perm_host);
    v.built = false;
}

void free_ws(Workspace& w) {
    if (w.dist) (void)hipFree(w.dist);
    w.dist = nullptr;
    dfree(w.fstamp);
    dfree(w.sstamp);
    dfree(w.rec);
    dfree(w.delta2);
    dfree(w.lists);
    if (w.pile) (void)hipFree(w.pile);
    w.pile = nullptr;
    dfree(w.counters);
    dfree(w.order);
    dfree(w.layers);
    dfree(w.bits);
    dfree(w.log);
    dfree(w.log_hi);
    if (w.summaries) (void)hipFree(w.summaries);
    w.summaries = nullptr;
    dfree(w.arcs_reached);
    dfree(w.err);
    dfree(w.sources);
    dfree(w.bc_out);
    dfree(w.grad1);
    dfree(w.grad2);
    const uint64_t gen = w.generation;
    w = Workspace{};
    w.generation = gen + 1;
}

// Extra LDS maps, a set used only when all of it fits and it costs no resident workgroups, the distance
// bounds first: MAP_BOUND skips the distance probes of candidates above their target's bound (74 % of the
// probes on a 50k-node spatial graph), with 4-bit codes (MAP_BOUND4, when allowed) where 16-bit bounds do
// not fit; MAP_SORTED makes settle passes expand each bucket's settled nodes in increasing id; MAP_FRONTIER
// replaces the global frontier stamps. forced >= 0 (the graph's NG_OPT_LDS_MAPS) overrides the choice; LDS
// distances that do not fit are dropped from it.
int lds_maps_for(DKind dk, int mode, uint32_t n, int block, bool allow4, int forced = -1) {
    const int b = allow4 ? (MAP_BOUND | MAP_BOUND4) : static_cast<int>(MAP_BOUND);
    if (forced >= 0) {
        int maps = forced & (allow4 ? 31 : 23);
        if ((maps & MAP_LDIST) != 0 && (lds_plan_maps(dk, n, maps, block) & MAP_LDIST) == 0) maps &= ~MAP_LDIST;
        return maps;
    }
    const int base = batch_blocks_per_cu(dk, mode, n, 0, block);
    // graphs small enough: the distances themselves in LDS (256- and 1024-lane workgroups)
    if (block == 256 || block == 1024)
        for (int maps : {MAP_LDIST | MAP_SORTED | MAP_FRONTIER, MAP_LDIST | MAP_FRONTIER, MAP_LDIST | MAP_SORTED, static_cast<int>(MAP_LDIST)})
            if (lds_plan_maps(dk, n, maps, block) == maps && batch_blocks_per_cu(dk, mode, n, maps, block) >= base) return maps;
    // 16-bit bounds where they fit, else 4-bit codes (also where 16-bit bounds would cost workgroups: on a
    // 50k-node graph at 512 lanes, 4-bit codes keep two workgroups per CU, 12 % faster than no bounds)
    const int b4 = allow4 ? static_cast<int>(MAP_BOUND4) : 0;
    for (int maps : {b | MAP_SORTED | MAP_FRONTIER, b | MAP_FRONTIER, b | MAP_SORTED, b, b4 | MAP_SORTED | MAP_FRONTIER,
                     b4 | MAP_FRONTIER, b4 | MAP_SORTED, b4, MAP_SORTED | MAP_FRONTIER, static_cast<int>(MAP_FRONTIER),
                     static_cast<int>(MAP_SORTED)})
        if (lds_plan_maps(dk, n, maps, block) == maps && batch_blocks_per_cu(dk, mode, n, maps, block) >= base) return maps;
    return 0;
}

bool valid_block(int b) { return b == 64 || b == 256 || b == 512 || b == 1024; }

// Lanes per workgroup for a call: requested by the caller (the autotuner's trials), else the graph's
// NG_OPT_BLOCK, else tuned, else the default.
int pick_block(const View& v, int mode_idx, int requested, int option) {
    if (valid_block(requested)) return requested;
    if (valid_block(option)) return option;
    if (v.tuned_block[mode_idx] > 0) return v.tuned_block[mode_idx];
    return kBlock;
}

// Pile capacity per list: n first parkings plus 2n re-parkings with improved values.
uint32_t pile_capacity(uint32_t n) { return static_cast<uint32_t>(std::min<uint64_t>(3ull * n + 1024, 0xFFFFFF00u)); }

// Per-slot device bytes of the batched kernel for (n, m, mode): distances, frontier and settled stamps, the
// three lists and the pile; in Brandes modes also the records, second dependency, Kahn order, layer bounds,
// tight-arc bitmap and log.
uint64_t slot_bytes(uint32_t n, uint64_t m, int mode, uint32_t logcap) {
    uint64_t b = static_cast<uint64_t>(n) * (8 + 4 + 4 + 12) + 16 + 32ull * pile_capacity(n);
    if (mode != MODE_SSSP) {
        b += static_cast<uint64_t>(n) * (sizeof(BcRec) + 8 + 4 + 4) + 4 + ((m + 31) / 32) * 4 +
             static_cast<uint64_t>(logcap) * (sizeof(LogEntry) + 4);
    }
    return b;
}

// Candidate tight-arc log entries per slot; `option` > 0 (NG_OPT_LOG_CAPACITY) replaces the default.
uint32_t log_capacity(uint32_t n, uint64_t m, uint32_t option = 0) {
    if (option > 0) return option;
    const uint64_t want = std::max<uint64_t>(static_cast<uint64_t>(n) * 8, 4096);
    return static_cast<uint32_t>(std::min<uint64_t>(std::max<uint64_t>(m, 1), want));
}

// Pile entries per list for a slot stride of n nodes; `option` > 0 (NG_OPT_PILE_CAPACITY) shrinks it, down to
// n + 2 x the largest workgroup (tests use it to force the exact bucket-advance path).
uint32_t wanted_pcap(uint32_t n, uint32_t option = 0) {
    uint32_t p = pile_capacity(n);
    if (option > 0) p = static_cast<uint32_t>(std::min<uint64_t>(std::max<uint64_t>(option, static_cast<uint64_t>(n) + 2 * kMaxBlock), p));
    return p;
}

// Capacities a call asks of the workspace (the graph's options; the cohort takes the defaults).
struct WsCaps {
    uint32_t pile = 0;
    uint32_t log = 0;
};

// Makes sure workspace w holds at least `slots` slots for graphs of up to n nodes and m arcs, with every
// array `mode` needs and dist in kind dk. Growing keeps the larger of the old and new dimensions, except that
// old slots beyond `slot_cap` (the caller's memory budget at the new dimensions) are dropped: many small
// slots of a small graph must not be carried over to a large one.
ng_status ensure_ws(Workspace& w, hipStream_t st, uint32_t n, uint64_t m, uint32_t slots, int mode, DKind dk, uint32_t n_sources,
                    uint32_t slot_cap, WsCaps caps = {}) {
    int need_bits = (mode == MODE_SSSP) ? 0 : (MODE_BC | MODE_ARC1 | MODE_ARC2);
    const bool has_log = ((need_bits | w.mode_bits) & MODE_BC) != 0;
    const bool realloc = w.slots < slots || w.n < n || w.m < m || dk_size(dk) > w.dist_bytes || (need_bits & ~w.mode_bits) != 0 ||
                         w.pcap != wanted_pcap(std::max(w.n, n), caps.pile) ||
                         (has_log && w.logcap != log_capacity(std::max(w.n, n), std::max(w.m, m), caps.log));
    if (realloc) {
        if (w.dist != nullptr) {
            n = std::max(n, w.n);
            m = std::max(m, w.m);
            slots = std::max(slots, std::min(w.slots, slot_cap));
            n_sources = std::max(n_sources, w.summaries_cap);
            need_bits |= w.mode_bits;
        }
        free_ws(w);
        const size_t ns = static_cast<size_t>(slots) * n;
        w.slots = slots;
        w.n = n;
        w.m = m;
        w.mode_bits = need_bits;
        w.dist_bytes = 8;  // room for either element size
        NG_HIP(hipMalloc(&w.dist, 8 * std::max<size_t>(ns, 1)));
        NG_HIP(hipMalloc(&w.fstamp, 4 * std::max<size_t>(ns, 1)));
        NG_HIP(hipMalloc(&w.sstamp, 4 * std::max<size_t>(ns, 1)));
        NG_HIP(hipMalloc(&w.lists, 12 * std::max<size_t>(ns, 1)));
        w.pcap = wanted_pcap(n, caps.pile);
        NG_HIP(hipMalloc(&w.pile, 16ull * 2 * w.pcap * slots));
        NG_HIP(hipMalloc(&w.counters, 16 * static_cast<size_t>(slots)));
        NG_HIP(launch_fill_u32(w.fstamp, ns, 0u, st));
        NG_HIP(launch_fill_u32(w.sstamp, ns, 0u, st));
        NG_HIP(launch_fill_u32(w.counters, 4ull * slots, 1u, st));
        if (need_bits) {
            w.logcap = log_capacity(n, m, caps.log);
            w.words = (m + 31) / 32;
            NG_HIP(hipMalloc(&w.rec, sizeof(BcRec) * std::max<size_t>(ns, 1)));
            NG_HIP(hipMalloc(&w.delta2, 8 * std::max<size_t>(ns, 1)));
            NG_HIP(launch_rec_init(w.rec, ns, st));
            w.bc_zero = true;
            NG_HIP(launch_fill_f64(w.delta2, ns, 0.0, st));
            NG_HIP(hipMalloc(&w.order, 4 * std::max<size_t>(ns, 1)));
            NG_HIP(hipMalloc(&w.layers, 4 * static_cast<size_t>(slots) * (n + 1)));
            NG_HIP(hipMalloc(&w.bits, 4 * std::max<uint64_t>(w.words * slots, 1)));
            NG_HIP(hipMalloc(&w.log, sizeof(LogEntry) * static_cast<size_t>(w.logcap) * slots));
            NG_HIP(hipMalloc(&w.log_hi, 4 * std::max<size_t>(static_cast<size_t>(w.logcap) * slots, 1)));
            NG_HIP(hipMalloc(&w.bc_out, 8 * std::max<uint32_t>(n, 1)));
            NG_HIP(hipMalloc(&w.grad1, 8 * std::max<uint64_t>(m, 1)));
            NG_HIP(hipMalloc(&w.grad2, 8 * std::max<uint64_t>(m, 1)));
            NG_HIP(launch_fill_u32(w.bits, w.words * slots, 0u, st));
        }
        NG_HIP(hipMalloc(&w.arcs_reached, 8 * 32));
        NG_HIP(hipMalloc(&w.err, 4));
        w.dist_valid = false;
    }
    if (!w.dist_valid || w.dist_kind != dk) {
        NG_HIP(launch_fill_dist(dk, w.dist, static_cast<uint64_t>(w.slots) * w.n, st));
        w.dist_kind = dk;
        w.dist_valid = true;
    }
    if (w.summaries_cap < n_sources) {
        if (w.summaries) (void)hipFree(w.summaries);
        w.summaries = nullptr;
        NG_HIP(hipMalloc(&w.summaries, sizeof(ng_sssp_summary) * std::max<uint32_t>(n_sources, 1)));
        w.summaries_cap = n_sources;
        ++w.generation;
    }
    if (w.sources_cap < n_sources) {
        dfree(w.sources);
        NG_HIP(hipMalloc(&w.sources, sizeof(uint32_t) * std::max<uint32_t>(n_sources, 1)));
        w.sources_cap = n_sources;
        ++w.generation;
    }
    return NG_OK;
}

double pick_delta(const ng_graph* g, const View& v, int vi, int mode_idx, double delta) {
    if (delta > 0.0) return delta;
    (void)g;
    (void)vi;
    if (v.tuned_delta[mode_idx] > 0.0) return v.tuned_delta[mode_idx];
    double d = v.q05_len;
    if (!(d > 0.0) || !std::isfinite(d)) d = v.min_len > 0.0 ? v.min_len : 1.0;
    if (v.dk == DKind::U32 || v.dk == DKind::U64) d = std::max(1.0, std::round(d));
    return d;
}

struct RunSpec {
    int vi;           // view index
    int mode;         // MODE_*
    double delta;
    const uint32_t* sources;
    uint32_t n_sources;
    void* dist_out;   // host, optional
    void* dist_dev = nullptr;  // device, optional (n_sources * n elements, kept on the device)
    int block = 0;             // lanes per workgroup (0: tuned or default)
    bool no_bound4 = false;    // no 4-bit distance codes (the run that estimates their scale)
};

ng_status estimate_bound_scale(ng_graph* g, int vi, double delta);

// Uploads sources, sizes the workspace, launches the batched kernel once, and synchronizes.
ng_status run_batch(ng_graph* g, const RunSpec& rs, uint32_t* slots_used, double* used_delta) {
    ng_context* ctx = g->ctx;
    NG_HIP(hipSetDevice(ctx->device));
    for (uint32_t i = 0; i < rs.n_sources; ++i) {
        if (rs.sources[i] >= g->n) {
            return fail(NG_EINVAL, "source %u is out of range, the graph has %u nodes (Dijkstra.h:66-67)", rs.sources[i], g->n);
        }
    }
    ng_status st = ensure_view(g, rs.vi);
    if (st != NG_OK) return st;
    View& v = g->views[rs.vi];
    const int mode_idx = rs.mode == MODE_SSSP ? 0 : 1;
    const double delta = pick_delta(g, v, rs.vi, mode_idx, rs.delta);
    if (used_delta) *used_delta = delta;
    if (v.lend_delta != delta) {
        NG_HIP(launch_light_end(v.dk, g->n, v.off, v.len, delta, v.rows, ctx->stream));
        v.lend_delta = delta;
    }
    // slots: resident workgroups of this kernel, bounded by the sources and by device memory
    const int block = pick_block(v, mode_idx, rs.block, g->opt_block);
    const int maps = lds_maps_for(v.dk, rs.mode, g->n, block, !rs.no_bound4, g->opt_lds_maps);
    if (lds_plan_bound_kind(v.dk, g->n, maps, block) == 2 && !(v.bound_inv_q > 0.0)) {
        st = estimate_bound_scale(g, rs.vi, delta);
        if (st != NG_OK) return st;
    }
    const int bpc = batch_blocks_per_cu(v.dk, rs.mode, g->n, maps, block);
    uint64_t slots = static_cast<uint64_t>(bpc) * static_cast<uint64_t>(ctx->cus);
    // NG_OPT_MAX_SLOTS caps the concurrent source slots (tuning sweeps, tests of slot reuse)
    if (g->opt_max_slots > 0) slots = std::min<uint64_t>(slots, g->opt_max_slots);
    slots = std::min<uint64_t>(slots, std::max<uint32_t>(rs.n_sources, 1));
    size_t free_b = 0, total_b = 0;
    NG_HIP(hipMemGetInfo(&free_b, &total_b));
    Workspace& w = ctx->ws;
    const uint32_t wn = std::max(w.n, g->n);
    const uint64_t wm = std::max(w.m, g->m);
    const WsCaps caps{g->opt_pile_capacity, g->opt_log_capacity};
    const uint64_t per = slot_bytes(wn, wm, rs.mode, log_capacity(wn, wm, caps.log));
    // memory already held by the context's workspace counts as available
    const uint64_t held = static_cast<uint64_t>(w.slots) * slot_bytes(w.n, w.m, rs.mode, w.logcap);
    const uint64_t budget = static_cast<uint64_t>(static_cast<double>(free_b + held) * 0.6);
    const uint64_t slot_cap = per > 0 ? std::max<uint64_t>(1, budget / per) : slots;
    slots = std::min<uint64_t>(slots, slot_cap);
    st = ensure_ws(w, ctx->stream, g->n, g->m, static_cast<uint32_t>(slots), rs.mode, v.dk, rs.n_sources,
                   static_cast<uint32_t>(std::min<uint64_t>(slot_cap, 0xFFFFFFFFu)), caps);
    if (st != NG_OK) return st;
    const hipStream_t s = ctx->stream;
    if (rs.n_sources) NG_HIP(hipMemcpyAsync(w.sources, rs.sources, sizeof(uint32_t) * rs.n_sources, hipMemcpyHostToDevice, s));
    NG_HIP(hipMemsetAsync(w.arcs_reached, 0, 8 * 32, s));
    NG_HIP(hipMemsetAsync(w.err, 0, 4, s));
    if (rs.mode != MODE_SSSP) {
        // every slot's betweenness partials start at zero: the last Brandes call's slot reduction cleared them, else
        // they are cleared here (all slots, so that the flag covers them all); until this call's reduction they are not
        if ((rs.mode & MODE_BC) && !w.bc_zero) NG_HIP(launch_rec_zero_bc(w.rec, static_cast<uint64_t>(w.slots) * w.n, s));
        w.bc_zero = false;
    }
    if (rs.mode & MODE_ARC1) NG_HIP(launch_fill_f64(w.grad1, g->m, 0.0, s));
    if (rs.mode & MODE_ARC2) NG_HIP(launch_fill_f64(w.grad2, g->m, 0.0, s));
    void* d_dist_out = rs.dist_dev;
    const int dok = dist_kind_of(v.dk);
    const size_t dob = (dok == NG_DIST_F32 ? 4 : 8);
    if (rs.dist_out) NG_HIP(hipMalloc(&d_dist_out, dob * std::max<size_t>(static_cast<size_t>(rs.n_sources) * g->n, 1)));
    LaunchArgs la{};
    la.dk = v.dk;
    la.mode = rs.mode;
    la.n = g->n;
    la.view = &v;
    la.sources = w.sources;
    la.n_sources = rs.n_sources;
    la.delta = delta;
    la.grid = static_cast<uint32_t>(slots);
    la.ws = &w;
    la.dist_out = d_dist_out;
    la.dist_out_kind = dok;
    la.maps = maps;
    la.bd_inv_q = v.bound_inv_q;
    la.block = block;
    float ms = 0.0f;
    if (rs.n_sources) {
        NG_HIP(hipEventRecord(ctx->ev0, s));
        hipError_t le = launch_batch(la, s);
        if (le != hipSuccess) {
            if (rs.dist_out) (void)hipFree(d_dist_out);
            return fail(NG_EDEVICE, "batched kernel launch failed: %s", hipGetErrorString(le));
        }
        NG_HIP(hipEventRecord(ctx->ev1, s));
    }
    hipError_t se = hipStreamSynchronize(s);
    if (se != hipSuccess) {
        if (rs.dist_out) (void)hipFree(d_dist_out);
        return fail(NG_EDEVICE, "batched kernel failed: %s", hipGetErrorString(se));
    }
    if (rs.n_sources) NG_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    uint32_t err = 0;
    unsigned long long ctrs[32] = {};
    NG_HIP(hipMemcpy(&err, w.err, 4, hipMemcpyDeviceToHost));
    NG_HIP(hipMemcpy(ctrs, w.arcs_reached, 8 * 32, hipMemcpyDeviceToHost));
    for (int sct = 0; sct < 8; ++sct) g->last.section_ticks[sct] = ctrs[9 + sct];
    g->last.tree_sources = ctrs[17];
    g->last.near_tree_sources = ctrs[18];
    const unsigned long long arcs = ctrs[0];
    if (rs.dist_out) {
        NG_HIP(hipMemcpy(rs.dist_out, d_dist_out, dob * static_cast<size_t>(rs.n_sources) * g->n, hipMemcpyDeviceToHost));
        NG_HIP(hipFree(d_dist_out));
    }
    if (err & ERR_KAHN) return fail(NG_EDEVICE, "tight-arc DAG has a cycle (zero-length arcs?)");
    if (err & ERR_GUARD) return fail(NG_EDEVICE, "iteration guard hit in the batched kernel");
    if (err & ERR_OVERFLOW) return fail(NG_EDEVICE, "queue overflow in the batched kernel");
    g->last.kernel_ms = ms;
    g->last.launches = rs.n_sources ? 1 : 0;
    g->last.slots = static_cast<uint32_t>(slots);
    g->last.block = static_cast<uint32_t>(block);
    g->last.delta = delta;
    g->last.arcs_reached = arcs;
    g->last.flags = err;
    g->last.phases = ctrs[1];
    g->last.buckets = ctrs[2];
    g->last.light_arcs = ctrs[3];
    g->last.settle_arcs = ctrs[4];
    g->last.logged_arcs = ctrs[5];
    g->last.layers = ctrs[6];
    g->last.probes = ctrs[7];
    g->last.improvements = ctrs[8];
    if (slots_used) *slots_used = static_cast<uint32_t>(slots);
    return NG_OK;
}

ng_status check_graph(const ng_graph* g) {
    if (!g || !g->ctx) return fail(NG_EINVAL, "null graph");
    return NG_OK;
}

ng_status copy_summaries(ng_graph* g, uint32_t n_sources, ng_sssp_summary* out) {
    if (!out || !n_sources) return NG_OK;
    NG_HIP(hipMemcpy(out, g->ctx->ws.summaries, sizeof(ng_sssp_summary) * n_sources, hipMemcpyDeviceToHost));
    return NG_OK;
}

// Scale of the 4-bit distance codes of a view: 15 codes over 1.1 x the largest distance from 32 evenly spaced
// sources (one small SSSP launch, once per view). Farther nodes keep code 15 (no bound), so the estimate only
// affects how many probes are skipped, never the results.
ng_status estimate_bound_scale(ng_graph* g, int vi, double delta) {
    const uint32_t k = std::min<uint32_t>(g->n, 32);
    std::vector<uint32_t> src(k);
    for (uint32_t i = 0; i < k; ++i) src[i] = static_cast<uint32_t>(static_cast<uint64_t>(i) * g->n / k);
    RunSpec rs{vi, MODE_SSSP, delta, src.data(), k, nullptr};
    rs.no_bound4 = true;
    ng_status st = run_batch(g, rs, nullptr, nullptr);
    if (st != NG_OK) return st;
    std::vector<ng_sssp_summary> sm(k);
    st = copy_summaries(g, k, sm.data());
    if (st != NG_OK) return st;
    double mx = 0.0;
    for (const ng_sssp_summary& r : sm) mx = std::max(mx, r.max_d);
    g->views[vi].bound_inv_q = mx > 0.0 && std::isfinite(mx) ? 15.0 / (1.1 * mx) : 1.0;
    return NG_OK;
}

// Brandes-family run on the |w| view; requires positive lengths.
ng_status run_dependency(ng_graph* g, const uint32_t* sources, uint32_t n_sources, double delta, int mode, ng_sssp_summary* out_opt) {
    ng_status st = ensure_view(g, 0);
    if (st != NG_OK) return st;
    if (g->views[0].has_zero) {
        return fail(NG_EUNSUPPORTED, "zero-length arcs: Brandes needs non-zero arc weights (BrandesVariant.h:30-31)");
    }
    RunSpec rs{0, mode, delta, sources, n_sources, nullptr};
    st = run_batch(g, rs, nullptr, nullptr);
    if (st != NG_OK) return st;
    return copy_summaries(g, n_sources, out_opt);
}

}  // namespace

extern "C" {

const char* ng_last_error(void) { return g_error.c_str(); }

// Internal: lets the I/O translation unit report through ng_last_error().
void ng_set_error_message(const char* msg) { g_error = msg ? msg : ""; }

const char* ng_version(void) { return "ng_hip 0.1 gfx950"; }

ng_status ng_context_create(int device, ng_context** out) {
    if (!out) return fail(NG_EINVAL, "null output pointer");
    *out = nullptr;
    int count = 0;
    NG_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return fail(NG_EINVAL, "device %d out of range (%d devices)", device, count);
    NG_HIP(hipSetDevice(device));
    auto ctx = std::make_unique<ng_context>();
    ctx->device = device;
    hipDeviceProp_t prop;
    NG_HIP(hipGetDeviceProperties(&prop, device));
    ctx->cus = prop.multiProcessorCount;
    NG_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    NG_HIP(hipEventCreate(&ctx->ev0));
    NG_HIP(hipEventCreate(&ctx->ev1));
    // the allocation-free entry point's scratch array on this device (device current): its address is resolved
    // here, once per device; a failure is reported by that entry point, the others do not use the array
    ctx->scratch_status = scratch_setup(device);
    *out = ctx.release();
    return NG_OK;
}

ng_status ng_context_synchronize(ng_context* ctx) {
    if (!ctx) return fail(NG_EINVAL, "null context");
    NG_HIP(hipSetDevice(ctx->device));
    NG_HIP(hipStreamSynchronize(ctx->stream));
    return NG_OK;
}

ng_status ng_copy_calibration(ng_context* ctx, uint64_t bytes, double* ms_out) {
    if (!ctx) return fail(NG_EINVAL, "null context");
    bytes &= ~static_cast<uint64_t>(15);
    if (bytes == 0) return fail(NG_EINVAL, "copy size must be at least 16 bytes");
    NG_HIP(hipSetDevice(ctx->device));
    void* src = nullptr;
    void* dst = nullptr;
    NG_HIP(hipMalloc(&src, bytes));
    hipError_t e = hipMalloc(&dst, bytes);
    if (e != hipSuccess) {
        (void)hipFree(src);
        return fail(NG_ENOMEM, "hipMalloc failed: %s", hipGetErrorString(e));
    }
    float ms = 0.0f;
    e = launch_fill_u32(static_cast<uint32_t*>(src), bytes / 4, 1u, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(ctx->ev0, ctx->stream);
    if (e == hipSuccess) e = launch_copy(src, dst, bytes, ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(ctx->ev1, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    (void)hipFree(src);
    (void)hipFree(dst);
    if (e != hipSuccess) return fail(NG_EDEVICE, "copy calibration failed: %s", hipGetErrorString(e));
    if (ms_out) *ms_out = ms;
    return NG_OK;
}

void ng_context_destroy(ng_context* ctx) {
    if (!ctx) return;
    context_release(ctx);
}

ng_status ng_graph_create(ng_context* ctx, uint32_t n, uint64_t m, const uint64_t* out_offsets, const uint32_t* out_targets,
                          const void* out_weights, ng_weight_kind wkind, ng_graph** out) {
    if (!out) return fail(NG_EINVAL, "null output pointer");
    *out = nullptr;
    if (!ctx) return fail(NG_EINVAL, "null context");
    if (!out_offsets) return fail(NG_EINVAL, "null offsets");
    if (m && (!out_targets || !out_weights)) return fail(NG_EINVAL, "null targets or weights");
    if (wkind != NG_W_I32 && wkind != NG_W_F32 && wkind != NG_W_F64) return fail(NG_EINVAL, "unknown weight kind %d", wkind);
    if (m >= 0xFFFFFFFFull) return fail(NG_EUNSUPPORTED, "%llu arcs: arc ids are 32-bit (arc_id_type, Types.h:27)", (unsigned long long)m);
    if (n == 0xFFFFFFFFu) return fail(NG_EUNSUPPORTED, "too many nodes");
    if (out_offsets[0] != 0 || out_offsets[n] != m) return fail(NG_EINVAL, "offsets must start at 0 and end at m");
    for (uint32_t u = 0; u < n; ++u) {
        if (out_offsets[u] > out_offsets[u + 1]) return fail(NG_EINVAL, "offsets decrease at node %u", u);
    }
    for (uint64_t a = 0; a < m; ++a) {
        if (out_targets[a] >= n) return fail(NG_EINVAL, "arc %llu targets node %u, the graph has %u nodes", (unsigned long long)a, out_targets[a], n);
    }
    auto g = std::make_unique<ng_graph>();
    g->ctx = ctx;
    g->n = n;
    g->m = m;
    g->wk = wkind;
    g->off.assign(out_offsets, out_offsets + static_cast<size_t>(n) + 1);
    g->tgt.assign(out_targets, out_targets + m);
    const size_t wb = wkind == NG_W_F64 ? 8 : 4;
    g->w.resize(wb * m);
    if (m) std::memcpy(g->w.data(), out_weights, wb * m);
    for (uint64_t a = 0; a < m; ++a) {
        if (!std::isfinite(weight_at(g.get(), a))) {
            return fail(NG_EINVAL, "the weight of arc %llu is not a finite number", (unsigned long long)a);
        }
    }
    for (uint64_t a = 0; a < m; ++a) {
        if (weight_at(g.get(), a) < 0.0) {
            g->has_negative = true;
            break;
        }
    }
    ng_status st = ensure_view(g.get(), 0);
    if (st != NG_OK) {
        free_view(g->views[0]);
        return st;
    }
    ctx->refs.fetch_add(1, std::memory_order_relaxed);
    *out = g.release();
    return NG_OK;
}

void ng_graph_destroy(ng_graph* g) {
    if (!g) return;
    ng_context* ctx = g->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    free_view(g->views[0]);
    free_view(g->views[1]);
    delete g;
    context_release(ctx);
}

ng_status ng_graph_dist_kind(const ng_graph* g, ng_length_kind lkind, ng_dist_kind* out) {
    if (check_graph(g) != NG_OK) return NG_EINVAL;
    if (!out) return fail(NG_EINVAL, "null output pointer");
    if (lkind == NG_LEN_INVERSE) {
        *out = NG_DIST_F64;
        return NG_OK;
    }
    ng_status st = ensure_view(const_cast<ng_graph*>(g), 0);
    if (st != NG_OK) return st;
    *out = static_cast<ng_dist_kind>(dist_kind_of(g->views[0].dk));
    return NG_OK;
}

ng_status ng_graph_set_option(ng_graph* g, ng_graph_option opt, int64_t value) {
    if (check_graph(g) != NG_OK) return NG_EINVAL;
    switch (opt) {
    case NG_OPT_BLOCK:
        if (value != 0 && !valid_block(static_cast<int>(value))) return fail(NG_EINVAL, "lanes per workgroup must be 0, 64, 256, 512 or 1024, not %lld", (long long)value);
        g->opt_block = static_cast<int>(value);
        return NG_OK;
    case NG_OPT_MAX_SLOTS:
        if (value < 0 || value > 0xFFFFFFFFll) return fail(NG_EINVAL, "slot cap %lld out of range", (long long)value);
        g->opt_max_slots = static_cast<uint32_t>(value);
        return NG_OK;
    case NG_OPT_LDS_MAPS:
        if (value < -1 || value > 31) return fail(NG_EINVAL, "LDS maps must be -1 (automatic) or 0..31, not %lld", (long long)value);
        g->opt_lds_maps = static_cast<int>(value);
        return NG_OK;
    case NG_OPT_PILE_CAPACITY:
        if (value < 0 || value > 0xFFFFFF00ll) return fail(NG_EINVAL, "pile capacity %lld out of range", (long long)value);
        g->opt_pile_capacity = static_cast<uint32_t>(value);
        return NG_OK;
    case NG_OPT_LOG_CAPACITY:
        if (value < 0 || value > 0xFFFFFFFFll) return fail(NG_EINVAL, "log capacity %lld out of range", (long long)value);
        g->opt_log_capacity = static_cast<uint32_t>(value);
        return NG_OK;
    }
    return fail(NG_EINVAL, "unknown graph option %d", static_cast<int>(opt));
}

ng_status ng_graph_get_option(const ng_graph* g, ng_graph_option opt, int64_t* value) {
    if (check_graph(g) != NG_OK) return NG_EINVAL;
    if (!value) return fail(NG_EINVAL, "null output pointer");
    switch (opt) {
    case NG_OPT_BLOCK: *value = g->opt_block; return NG_OK;
    case NG_OPT_MAX_SLOTS: *value = g->opt_max_slots; return NG_OK;
    case NG_OPT_LDS_MAPS: *value = g->opt_lds_maps; return NG_OK;
    case NG_OPT_PILE_CAPACITY: *value = g->opt_pile_capacity; return NG_OK;
    case NG_OPT_LOG_CAPACITY: *value = g->opt_log_capacity; return NG_OK;
    }
    return fail(NG_EINVAL, "unknown graph option %d", static_cast<int>(opt));
}

ng_status ng_graph_set_weights(ng_graph* g, const void* weights) {
    if (check_graph(g) != NG_OK) return NG_EINVAL;
    if (g->m && !weights) return fail(NG_EINVAL, "null weights");
    const size_t wb = g->wk == NG_W_F64 ? 8 : 4;
    std::vector<unsigned char> old(g->w);
    if (g->m) std::memcpy(g->w.data(), weights, wb * g->m);
    for (uint64_t a = 0; a < g->m; ++a) {
        if (!std::isfinite(weight_at(g, a))) {
            g->w.swap(old);
            return fail(NG_EINVAL, "the weight of arc %llu is not a finite number", (unsigned long long)a);
        }
    }
    g->has_negative = false;
    for (uint64_t a = 0; a < g->m && !g->has_negative; ++a) g->has_negative = weight_at(g, a) < 0.0;
    for (int vi = 0; vi < 2; ++vi) {
        View& v = g->views[vi];
        if (!v.built) continue;
        const DKind dk = wanted_kind(g, vi);
        if (dk != v.dk) {  // integer distances outgrew (or fit again) 32 bits: new arrays
            NG_HIP(hipSetDevice(g->ctx->device));
            NG_HIP(hipStreamSynchronize(g->ctx->stream));
            free_view(v);
        }
        const double tuned_d[2] = {v.tuned_delta[0], v.tuned_delta[1]};
        const int tuned_b[2] = {v.tuned_block[0], v.tuned_block[1]};
        ng_status st = build_view(g, vi, dk);
        if (st != NG_OK) return st;
        for (int k = 0; k < 2; ++k) {  // tuned shapes are speed hints; they stay valid for any lengths
            v.tuned_delta[k] = tuned_d[k];
            v.tuned_block[k] = tuned_b[k];
        }
    }
    ++g->weights_version;
    return NG_OK;
}

ng_status ng_graph_weights_reordered(const ng_graph* g, int* reordered) {
    if (check_graph(g) != NG_OK) return NG_EINVAL;
    if (!reordered) return fail(NG_EINVAL, "null output pointer");
    *reordered = g->views[0].reordered_last ? 1 : 0;
    return NG_OK;
}

ng_status ng_sssp_summaries(ng_graph* g, const uint32_t* sources, uint32_t n_sources, ng_length_kind lkind, double delta,
                            ng_sssp_summary* out, void* dist_out) {
    if (check_graph(g) != NG_OK) return NG_EINVAL;
    if (n_sources && (!sources || !out)) return fail(NG_EINVAL, "null sources or output");
    int vi = 0;
    switch (lkind) {
    case NG_LEN_ABS: vi = 0; break;
    case NG_LEN_INVERSE: vi = 1; break;
    case NG_LEN_RAW:
        if (g->has_negative) return fail(NG_EINVAL, "negative arc weight: delta-stepping casts weights to distance_type (PushCheckVariant.h:82)");
        vi = 0;
        break;
    default: return fail(NG_EINVAL, "unknown length kind %d", lkind);
    }
    RunSpec rs{vi, MODE_SSSP, delta, sources, n_sources, dist_out};
    ng_status st = run_batch(g, rs, nullptr, nullptr);
    if (st != NG_OK) return st;
    return copy_summaries(g, n_sources, out);
}

ng_status ng_distance_histogram(ng_graph* g, const uint32_t* sources, uint32_t n_sources, ng_length_kind lkind, double delta,
                                double bin_width, uint32_t n_bins, uint64_t* counts, ng_sssp_summary* out_opt) {
    if (check_graph(g) != NG_OK) return NG_EINVAL;
    if (n_sources && !sources) return fail(NG_EINVAL, "null sources");
    if (!counts || n_bins == 0) return fail(NG_EINVAL, "empty histogram");
    if (!(bin_width > 0.0) || !std::isfinite(bin_width)) return fail(NG_EINVAL, "bin width must be positive and finite");
    int vi = 0;
    switch (lkind) {
    case NG_LEN_ABS: vi = 0; break;
    case NG_LEN_INVERSE: vi = 1; break;
    case NG_LEN_RAW:
        if (g->has_negative) return fail(NG_EINVAL, "negative arc weight: delta-stepping casts weights to distance_type (PushCheckVariant.h:82)");
        vi = 0;
        break;
    default: return fail(NG_EINVAL, "unknown length kind %d", lkind);
    }
    ng_status st = ensure_view(g, vi);
    if (st != NG_OK) return st;
    NG_HIP(hipSetDevice(g->ctx->device));
    const int dok = dist_kind_of(g->views[vi].dk);
    const size_t dob = dok == NG_DIST_F32 ? 4 : 8;
    // sources in chunks whose distance rows take at most 1 GiB of device memory
    const uint64_t row = std::max<uint64_t>(static_cast<uint64_t>(g->n) * dob, 1);
    const uint32_t chunk = static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(n_sources, (1ull << 30) / row)));
    void* d_dist = nullptr;
    unsigned long long* d_counts = nullptr;
    uint32_t* d_src = nullptr;
    auto cleanup = [&]() {
        if (d_dist) (void)hipFree(d_dist);
        if (d_counts) (void)hipFree(d_counts);
        if (d_src) (void)hipFree(d_src);
    };
    if (hipMalloc(&d_dist, row * std::max<uint32_t>(chunk, 1)) != hipSuccess || hipMalloc(&d_counts, 8ull * n_bins) != hipSuccess ||
        hipMalloc(&d_src, 4ull * std::max<uint32_t>(chunk, 1)) != hipSuccess) {
        cleanup();
        return fail(NG_ENOMEM, "histogram buffers");
    }
    hipError_t he = launch_fill_u64(reinterpret_cast<uint64_t*>(d_counts), n_bins, 0, g->ctx->stream);
    for (uint32_t b = 0; he == hipSuccess && b < n_sources; b += chunk) {
        const uint32_t k = std::min(chunk, n_sources - b);
        RunSpec rs{vi, MODE_SSSP, delta, sources + b, k, nullptr};
        rs.dist_dev = d_dist;
        st = run_batch(g, rs, nullptr, nullptr);
        if (st != NG_OK) {
            cleanup();
            return st;
        }
        if (out_opt) {
            st = copy_summaries(g, k, out_opt + b);
            if (st != NG_OK) {
                cleanup();
                return st;
            }
        }
        he = hipMemcpyAsync(d_src, sources + b, 4ull * k, hipMemcpyHostToDevice, g->ctx->stream);
        if (he == hipSuccess) he = launch_distance_histogram(dok, d_dist, d_src, k, g->n, bin_width, n_bins, d_counts, g->ctx->stream);
    }
    if (he == hipSuccess) he = hipStreamSynchronize(g->ctx->stream);
    if (he == hipSuccess) he = hipMemcpy(counts, d_counts, 8ull * n_bins, hipMemcpyDeviceToHost);
    cleanup();
    if (he != hipSuccess) return fail(NG_EDEVICE, "distance histogram failed: %s", hipGetErrorString(he));
    return NG_OK;
}

ng_status ng_brandes_accumulate(ng_graph* g, const uint32_t* sources, uint32_t n_sources, double delta, double* bc_inout,
                                ng_sssp_summary* out_opt) {
    if (check_graph(g) != NG_OK) return NG_EINVAL;
    if (n_sources && !sources) return fail(NG_EINVAL, "null sources");
    if (!bc_inout) return fail(NG_EINVAL, "null betweenness buffer");
    ng_status st = run_dependency(g, sources, n_sources, delta, MODE_BC, out_opt);
    if (st != NG_OK) return st;
    if (!n_sources) return NG_OK;
    Workspace& w = g->ctx->ws;
    NG_HIP(launch_reduce_slots(w.rec, g->last.slots, g->n, w.n, w.bc_out, true, g->ctx->stream));
    w.bc_zero = true;
    ng_context* const ctx = g->ctx;
    if (ctx->host_bc_cap < g->n) {
        if (ctx->host_bc) NG_HIP(hipHostFree(ctx->host_bc));
        ctx->host_bc = nullptr;
        ctx->host_bc_cap = 0;
        NG_HIP(hipHostMalloc(reinterpret_cast<void**>(&ctx->host_bc), 8 * static_cast<size_t>(g->n), hipHostMallocDefault));
        ctx->host_bc_cap = g->n;
    }
    NG_HIP(hipMemcpyAsync(ctx->host_bc, w.bc_out, 8 * static_cast<size_t>(g->n), hipMemcpyDeviceToHost, ctx->stream));
    NG_HIP(hipStreamSynchronize(ctx->stream));
    for (uint32_t v = 0; v < g->n; ++v) bc_inout[v] += ctx->host_bc[v];
    return NG_OK;
}

}  // extern "C"

namespace {

size_t align256(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }

// Layout of the allocation-free Brandes entry point in the scratch array: the fixed part (work counters, error
// word, one chunk of sources and summaries, the betweenness output) and `slots` source slots of the batched
// kernel's Brandes arrays, each array 256-byte aligned. Returns the bytes used; with base != nullptr also points
// w's arrays into it (second dependencies and gradients stay null: Brandes mode does not use them).
size_t scratch_layout(char* base, Workspace& w, uint32_t n, uint64_t m, DKind dk, uint32_t slots, uint32_t chunk,
                      uint32_t pcap, uint32_t logcap) {
    const size_t ns = static_cast<size_t>(slots) * n;
    const uint64_t words = (m + 31) / 32;
    size_t at = 0;
    auto take = [&](size_t bytes) -> char* {
        char* p = base ? base + at : nullptr;
        at += align256(bytes);
        return p;
    };
    w = Workspace{};
    w.arcs_reached = reinterpret_cast<uint64_t*>(take(8 * 32));
    w.err = reinterpret_cast<uint32_t*>(take(4));
    w.sources = reinterpret_cast<uint32_t*>(take(4ull * chunk));
    w.summaries = take(sizeof(ng_sssp_summary) * chunk);
    w.bc_out = reinterpret_cast<double*>(take(8ull * n));
    w.dist = take(dk_size(dk) * ns);
    w.fstamp = reinterpret_cast<uint32_t*>(take(4 * ns));
    w.sstamp = reinterpret_cast<uint32_t*>(take(4 * ns));
    w.lists = reinterpret_cast<uint32_t*>(take(12 * ns));
    w.pile = take(16ull * 2 * pcap * slots);
    w.counters = reinterpret_cast<uint32_t*>(take(16ull * slots));
    w.rec = reinterpret_cast<BcRec*>(take(sizeof(BcRec) * ns));
    w.order = reinterpret_cast<uint32_t*>(take(4 * ns));
    w.layers = reinterpret_cast<uint32_t*>(take(4ull * slots * (static_cast<size_t>(n) + 1)));
    w.bits = reinterpret_cast<uint32_t*>(take(4ull * words * slots));
    w.log = reinterpret_cast<LogEntry*>(take(sizeof(LogEntry) * static_cast<size_t>(logcap) * slots));
    if (dk_size(dk) == 8) w.log_hi = reinterpret_cast<uint32_t*>(take(4ull * logcap * slots));  // 8-byte distances only
    w.slots = slots;
    w.n = n;
    w.m = m;
    w.mode_bits = MODE_BC;
    w.logcap = logcap;
    w.pcap = pcap;
    w.words = words;
    w.summaries_cap = chunk;
    w.sources_cap = chunk;
    w.dist_bytes = dk_size(dk);
    w.dist_kind = dk;
    w.dist_valid = true;
    return at;
}

}  // namespace

extern "C" {

ng_status ng_brandes_accumulate_noalloc(ng_graph* g, const uint32_t* sources, uint32_t n_sources, double delta,
                                        double* bc_inout, ng_sssp_summary* out_opt) {
    if (check_graph(g) != NG_OK) return NG_EINVAL;
    if (n_sources && !sources) return fail(NG_EINVAL, "null sources");
    if (!bc_inout) return fail(NG_EINVAL, "null betweenness buffer");
    for (uint32_t i = 0; i < n_sources; ++i) {
        if (sources[i] >= g->n) {
            return fail(NG_EINVAL, "source %u is out of range, the graph has %u nodes (Dijkstra.h:66-67)", sources[i], g->n);
        }
    }
    ng_context* ctx = g->ctx;
    View& v = g->views[0];
    if (!v.built) return fail(NG_EDEVICE, "the graph's device arrays are missing");
    if (v.has_zero) return fail(NG_EUNSUPPORTED, "zero-length arcs: Brandes needs non-zero arc weights (BrandesVariant.h:30-31)");
    if (ctx->scratch_status != hipSuccess) {
        return fail(NG_EDEVICE, "the scratch array of device %d is unavailable: %s", ctx->device, hipGetErrorString(ctx->scratch_status));
    }
    NG_HIP(hipSetDevice(ctx->device));
    char* const base = static_cast<char*>(scratch_device_ptr(ctx->device));
    if (base == nullptr) return fail(NG_EDEVICE, "the scratch array of device %d was not resolved", ctx->device);
    if (n_sources == 0) return NG_OK;
    const uint32_t n = g->n;
    const uint64_t m = g->m;
    const hipStream_t s = ctx->stream;
    const double d = pick_delta(g, v, 0, 1, delta);
    if (v.lend_delta != d) {
        NG_HIP(launch_light_end(v.dk, n, v.off, v.len, d, v.rows, s));
        v.lend_delta = d;
    }
    const int block = pick_block(v, 1, 0, g->opt_block);
    // 4-bit distance codes only with a scale estimated earlier (estimating it would need workspace memory)
    const int maps = lds_maps_for(v.dk, MODE_BC, n, block, v.bound_inv_q > 0.0, g->opt_lds_maps) &
                     (v.bound_inv_q > 0.0 ? ~0 : ~static_cast<int>(MAP_BOUND4));
    uint64_t slots = static_cast<uint64_t>(batch_blocks_per_cu(v.dk, MODE_BC, n, maps, block)) * static_cast<uint64_t>(ctx->cus);
    if (g->opt_max_slots > 0) slots = std::min<uint64_t>(slots, g->opt_max_slots);
    slots = std::min<uint64_t>(slots, n_sources);
    const uint32_t chunk = std::min(n_sources, kScratchChunk);
    const uint32_t pcap = wanted_pcap(n, g->opt_pile_capacity);
    const uint32_t logcap = log_capacity(n, m, g->opt_log_capacity);
    Workspace w;
    const size_t fixed = scratch_layout(nullptr, w, n, m, v.dk, 0, chunk, pcap, logcap);
    const size_t per = scratch_layout(nullptr, w, n, m, v.dk, 1, chunk, pcap, logcap) - fixed;
    if (fixed + per > kScratchBytes) {
        return fail(NG_EINVAL,
                    "graph too large for the allocation-free entry point: %u nodes and %llu arcs need %zu bytes for one "
                    "source slot, the scratch array holds %zu bytes (use ng_brandes_accumulate)",
                    n, (unsigned long long)m, fixed + per, static_cast<size_t>(kScratchBytes));
    }
    slots = std::min<uint64_t>(slots, (kScratchBytes - fixed) / per);
    const size_t used = scratch_layout(base, w, n, m, v.dk, static_cast<uint32_t>(slots), chunk, pcap, logcap);
    if (used > kScratchBytes) return fail(NG_EDEVICE, "scratch layout exceeds the array (%zu bytes)", used);
    // every call starts from a clean scratch: earlier calls on this GPU (any graph) left their own contents
    const uint64_t ns = slots * static_cast<uint64_t>(n);
    NG_HIP(launch_fill_u64(w.arcs_reached, 32, 0, s));
    NG_HIP(launch_fill_u32(w.err, 1, 0u, s));
    NG_HIP(launch_fill_dist(v.dk, w.dist, ns, s));
    NG_HIP(launch_fill_u32(w.fstamp, ns, 0u, s));
    NG_HIP(launch_fill_u32(w.sstamp, ns, 0u, s));
    NG_HIP(launch_fill_u32(w.counters, 4 * slots, 1u, s));
    NG_HIP(launch_rec_init(w.rec, ns, s));
    NG_HIP(launch_fill_u32(w.bits, w.words * slots, 0u, s));
    double total_ms = 0.0;
    uint32_t launches = 0;
    for (uint32_t b = 0; b < n_sources; b += chunk) {
        const uint32_t k = std::min(chunk, n_sources - b);
        NG_HIP(hipMemcpyAsync(w.sources, sources + b, sizeof(uint32_t) * k, hipMemcpyHostToDevice, s));
        LaunchArgs la{};
        la.dk = v.dk;
        la.mode = MODE_BC;
        la.n = n;
        la.view = &v;
        la.sources = w.sources;
        la.n_sources = k;
        la.delta = d;
        la.grid = static_cast<uint32_t>(std::min<uint64_t>(slots, k));
        la.ws = &w;
        la.dist_out = nullptr;
        la.dist_out_kind = dist_kind_of(v.dk);
        la.maps = maps;
        la.bd_inv_q = v.bound_inv_q;
        la.block = block;
        NG_HIP(hipEventRecord(ctx->ev0, s));
        const hipError_t le = launch_batch(la, s);
        if (le != hipSuccess) return fail(NG_EDEVICE, "batched kernel launch failed: %s", hipGetErrorString(le));
        NG_HIP(hipEventRecord(ctx->ev1, s));
        const hipError_t se = hipStreamSynchronize(s);
        if (se != hipSuccess) return fail(NG_EDEVICE, "batched kernel failed: %s", hipGetErrorString(se));
        float ms = 0.0f;
        NG_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        total_ms += ms;
        ++launches;
        uint32_t err = 0;
        NG_HIP(hipMemcpy(&err, w.err, 4, hipMemcpyDeviceToHost));
        if (err & ERR_KAHN) return fail(NG_EDEVICE, "tight-arc DAG has a cycle (zero-length arcs?)");
        if (err & ERR_GUARD) return fail(NG_EDEVICE, "iteration guard hit in the batched kernel");
        if (err & ERR_OVERFLOW) return fail(NG_EDEVICE, "queue overflow in the batched kernel");
        if (out_opt) NG_HIP(hipMemcpy(out_opt + b, w.summaries, sizeof(ng_sssp_summary) * k, hipMemcpyDeviceToHost));
    }
    NG_HIP(launch_reduce_slots(w.rec, static_cast<uint32_t>(slots), n, n, w.bc_out, false, s));
    NG_HIP(hipStreamSynchronize(s));
    // accumulate into the caller's vector through a fixed host buffer (no host allocation either)
    double part[4096];
    for (uint32_t v0 = 0; v0 < n; v0 += 4096) {
        const uint32_t c = std::min<uint32_t>(4096, n - v0);
        NG_HIP(hipMemcpy(part, w.bc_out + v0, 8ull * c, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < c; ++i) bc_inout[v0 + i] += part[i];
    }
    uint32_t err = 0;
    unsigned long long ctrs[32] = {};
    NG_HIP(hipMemcpy(&err, w.err, 4, hipMemcpyDeviceToHost));
    NG_HIP(hipMemcpy(ctrs, w.arcs_reached, 8 * 32, hipMemcpyDeviceToHost));
    ng_call_stats& l = g->last;
    l = ng_call_stats{};
    l.kernel_ms = total_ms;
    l.launches = launches;
    l.slots = static_cast<uint32_t>(slots);
    l.block = static_cast<uint32_t>(block);
    l.delta = d;
    l.arcs_reached = ctrs[0];
    l.flags = err;
    l.phases = ctrs[1];
    l.buckets = ctrs[2];
    l.light_arcs = ctrs[3];
    l.settle_arcs = ctrs[4];
    l.logged_arcs = ctrs[5];
    l.layers = ctrs[6];
    l.probes = ctrs[7];
    l.improvements = ctrs[8];
    for (int sct = 0; sct < 8; ++sct) l.section_ticks[sct] = ctrs[9 + sct];
    l.tree_sources = ctrs[17];
    l.near_tree_sources = ctrs[18];
    return NG_OK;
}

ng_status ng_edge_dependency(ng_graph* g, const uint32_t* sources, uint32_t n_sources, double delta, double* edge_inout) {
    if (check_graph(g) != NG_OK) return NG_EINVAL;
    if (n_sources && !sources) return fail(NG_EINVAL, "null sources");
    if (!edge_inout) return fail(NG_EINVAL, "null edge buffer");
    ng_status st = run_dependency(g, sources, n_sources, delta, MODE_ARC1, nullptr);
    if (st != NG_OK) return st;
    if (!n_sources) return NG_OK;
    std::vector<double> tmp(g->m);
    if (g->m) NG_HIP(hipMemcpy(tmp.data(), g->ctx->ws.grad1, 8 * g->m, hipMemcpyDeviceToHost));
    for (uint64_t a = 0; a < g->m; ++a) edge_inout[a] += tmp[a];
    return NG_OK;
}

ng_status ng_apsp_adjoint(ng_graph* g, const uint32_t* sources, uint32_t n_sources, double delta, double* grad_sum_d,
                          double* grad_sum_eff, ng_sssp_summary* out_opt) {
    if (check_graph(g) != NG_OK) return NG_EINVAL;
    if (n_sources && !sources) return fail(NG_EINVAL, "null sources");
    if (!grad_sum_d || !grad_sum_eff) return fail(NG_EINVAL, "null gradient buffer");
    ng_status st = run_dependency(g, sources, n_sources, delta, MODE_ARC1 | MODE_ARC2, out_opt);
    if (st != NG_OK) return st;
    if (!n_sources) return NG_OK;
    std::vector<double> t1(g->m), t2(g->m);
    if (g->m) {
        NG_HIP(hipMemcpy(t1.data(), g->ctx->ws.grad1, 8 * g->m, hipMemcpyDeviceToHost));
        NG_HIP(hipMemcpy(t2.data(), g->ctx->ws.grad2, 8 * g->m, hipMemcpyDeviceToHost));
    }
    // d|w|/dw = sign(w); the efficiency sum decreases with every length on a shortest path
    for (uint64_t a = 0; a < g->m; ++a) {
        const double w = weight_at(g, a);
        const double sg = w > 0.0 ? 1.0 : (w < 0.0 ? -1.0 : 0.0);
        grad_sum_d[a] += sg * t1[a];
        grad_sum_eff[a] -= sg * t2[a];
    }
    return NG_OK;
}

ng_status ng_autotune_delta(ng_graph* g, ng_length_kind lkind, int mode, const uint32_t* sample, uint32_t n_sample,
                            double* delta_out) {
    if (check_graph(g) != NG_OK) return NG_EINVAL;
    if (!sample || !n_sample) return fail(NG_EINVAL, "empty sample");
    const int vi = lkind == NG_LEN_INVERSE ? 1 : 0;
    ng_status st = ensure_view(g, vi);
    if (st != NG_OK) return st;
    View& v = g->views[vi];
    if (mode != 0 && v.has_zero) return fail(NG_EUNSUPPORTED, "zero-length arcs");
    const bool integer = v.dk == DKind::U32 || v.dk == DKind::U64;
    // candidates: multiples of the 5% length quantile between the minimum length and the mean length
    std::vector<double> cand;
    const double base = v.q05_len > 0.0 ? v.q05_len : (v.min_len > 0.0 ? v.min_len : 1.0);
    for (double f : {0.25, 0.5, 1.0, 2.0, 4.0, 8.0}) {
        double d = base * f;
        if (integer) d = std::max(1.0, std::round(d));
        if (std::find(cand.begin(), cand.end(), d) == cand.end()) cand.push_back(d);
    }
    const int run_mode = mode == 0 ? MODE_SSSP : MODE_BC;
    double best = cand.front(), best_ms = std::numeric_limits<double>::infinity();
    for (double d : cand) {
        double total = 0.0;
        for (int rep = 0; rep < 2; ++rep) {
            RunSpec rs{vi, run_mode, d, sample, n_sample, nullptr};
            st = run_batch(g, rs, nullptr, nullptr);
            if (st != NG_OK) return st;
            if (rep == 1) total += g->last.kernel_ms;
        }
        if (total < best_ms) {
            best_ms = total;
            best = d;
        }
    }
    v.tuned_delta[mode == 0 ? 0 : 1] = best;
    if (delta_out) *delta_out = best;
    return NG_OK;
}

ng_status ng_autotune(ng_graph* g, ng_length_kind lkind, int mode, const uint32_t* sample, uint32_t n_sample, double* delta_out,
                      uint32_t* block_out) {
    if (check_graph(g) != NG_OK) return NG_EINVAL;
    if (!sample || !n_sample) return fail(NG_EINVAL, "empty sample");
    const int vi = lkind == NG_LEN_INVERSE ? 1 : 0;
    ng_status st = ensure_view(g, vi);
    if (st != NG_OK) return st;
    View& v = g->views[vi];
    if (mode != 0 && v.has_zero) return fail(NG_EUNSUPPORTED, "zero-length arcs");
    const int mi = mode == 0 ? 0 : 1;
    const int run_mode = mode == 0 ? MODE_SSSP : MODE_BC;
    const bool integer = v.dk == DKind::U32 || v.dk == DKind::U64;
    std::vector<double> cand;
    const double base = v.q05_len > 0.0 ? v.q05_len : (v.min_len > 0.0 ? v.min_len : 1.0);
    for (double f : {0.25, 0.5, 1.0, 2.0, 4.0, 8.0}) {
        double d = base * f;
        if (integer) d = std::max(1.0, std::round(d));
        if (std::find(cand.begin(), cand.end(), d) == cand.end()) cand.push_back(d);
    }
    double best_d = cand.front(), best_ms = std::numeric_limits<double>::infinity();
    int best_b = kBlock;
    for (int b : {64, 256, 512, 1024}) {
        if (b == 64 && g->n > kWaveSourceMaxNodes) continue;
        bool first = true;
        for (double d : cand) {
            for (int rep = first ? 0 : 1; rep < 2; ++rep) {  // one untimed run per block size (workspace, caches)
                RunSpec rs{vi, run_mode, d, sample, n_sample, nullptr};
                rs.block = b;
                st = run_batch(g, rs, nullptr, nullptr);
                if (st != NG_OK) return st;
            }
            first = false;
            if (g->last.kernel_ms < best_ms) {
                best_ms = g->last.kernel_ms;
                best_d = d;
                best_b = b;
            }
        }
    }
    // refinement between the powers of two around the best width, at the best workgroup size
    const double grid_best = best_d;
    for (double f : {0.7071067811865476, 1.4142135623730951}) {
        double d = grid_best * f;
        if (integer) d = std::max(1.0, std::round(d));
        if (std::find(cand.begin(), cand.end(), d) != cand.end()) continue;
        cand.push_back(d);
        RunSpec rs{vi, run_mode, d, sample, n_sample, nullptr};
        rs.block = best_b;
        st = run_batch(g, rs, nullptr, nullptr);
        if (st != NG_OK) return st;
        if (g->last.kernel_ms < best_ms) {
            best_ms = g->last.kernel_ms;
            best_d = d;
        }
    }
    v.tuned_delta[mi] = best_d;
    v.tuned_block[mi] = best_b;
    if (delta_out) *delta_out = best_d;
    if (block_out) *block_out = static_cast<uint32_t>(best_b);
    return NG_OK;
}

ng_status ng_graph_last_stats(const ng_graph* g, ng_call_stats* out) {
    if (check_graph(g) != NG_OK) return NG_EINVAL;
    if (!out) return fail(NG_EINVAL, "null output pointer");
    *out = g->last;
    return NG_OK;
}

}  // extern "C"

namespace {

void cohort_free_device(ng_cohort* c) {
    if (c->exec) (void)hipGraphExecDestroy(c->exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->exec = nullptr;
    c->graph = nullptr;
    dfree(c->d_gtab);
    dfree(c->d_rows);
    dfree(c->d_prefix);
    dfree(c->d_fbase);
    dfree(c->d_seg_ptr);
    dfree(c->d_seg_off);
    dfree(c->d_flush);
    dfree(c->d_bc);
    c->built = false;
}

template <class T>
ng_status upload(T** dst, const std::vector<T>& src, hipStream_t st) {
    dfree(*dst);
    NG_HIP(hipMalloc(reinterpret_cast<void**>(dst), sizeof(T) * std::max<size_t>(src.size(), 1)));
    if (!src.empty()) NG_HIP(hipMemcpyAsync(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice, st));
    return NG_OK;
}

// Sizes and captures the cohort launch:
//   * items = (graph g, source s) in graph order; slot b runs items b, b + S, b + 2S, ... (S slots), so all
//     slots work on the same member graph at a time;
//   * each slot flushes its betweenness partials once per graph it touches, into its own flush region;
//     the reduction sums a graph's regions in slot order, so results are deterministic;
//   * the captured hipGraph holds fill kernels (no memset nodes), the batched kernel and the reduction.
ng_status cohort_build(ng_cohort* c) {
    ng_context* ctx = c->ctx;
    const hipStream_t st = ctx->stream;
    NG_HIP(hipSetDevice(ctx->device));
    cohort_free_device(c);
    const uint32_t G = static_cast<uint32_t>(c->graphs.size());
    std::vector<CohortGraph> gtab(G);
    std::vector<double> eff_delta(G);
    c->prefix.assign(G + 1, 0);
    c->n_max = 0;
    c->m_max = 0;
    for (uint32_t gi = 0; gi < G; ++gi) {
        ng_graph* g = c->graphs[gi];
        ng_status s0 = ensure_view(g, 0);
        if (s0 != NG_OK) return s0;
        const View& v = g->views[0];
        if (v.has_zero) return fail(NG_EUNSUPPORTED, "cohort graph %u has zero-length arcs (BrandesVariant.h:30-31)", gi);
        if (gi == 0) c->dk = v.dk;
        if (v.dk != c->dk) return fail(NG_EUNSUPPORTED, "cohort graphs must share one distance type (graph %u differs)", gi);
        const uint64_t np = static_cast<uint64_t>(c->prefix[gi]) + g->n;
        if (np > 0xFFFFFFF0ull) return fail(NG_EINVAL, "cohort too large: more than 2^32 sources in total");
        c->prefix[gi + 1] = static_cast<uint32_t>(np);
        c->n_max = std::max(c->n_max, g->n);
        c->m_max = std::max(c->m_max, g->m);
        double d = pick_delta(g, v, 0, 1, c->delta[gi]);
        if (c->dk == DKind::U32 || c->dk == DKind::U64) d = std::max(1.0, std::round(d));
        eff_delta[gi] = d;
        gtab[gi] = CohortGraph{v.off, v.tgt, v.len, v.arcs, nullptr, v.perm, g->n, v.strong ? 1u : 0u, d};
    }
    c->total = c->prefix[G];
    c->versions.resize(G);
    for (uint32_t gi = 0; gi < G; ++gi) c->versions[gi] = c->graphs[gi]->weights_version;
    // per-graph light-prefix ends for the cohort's own widths
    NG_HIP(hipMalloc(&c->d_rows, sizeof(RowRange) * std::max<size_t>(c->total, 1)));
    for (uint32_t gi = 0; gi < G; ++gi) {
        const View& v = c->graphs[gi]->views[0];
        RowRange* rows = c->d_rows + c->prefix[gi];
        NG_HIP(launch_light_end(v.dk, c->graphs[gi]->n, v.off, v.len, eff_delta[gi], rows, st));
        gtab[gi].rows = rows;
    }
    // slots
    const int maps = lds_maps_for(c->dk, MODE_BC, c->n_max, c->block, false);
    const int bpc_max = batch_blocks_per_cu(c->dk, MODE_BC, c->n_max, maps, c->block);
    const int bpc = c->bpc > 0 ? std::min(c->bpc, bpc_max) : bpc_max;
    uint64_t slots = static_cast<uint64_t>(bpc) * static_cast<uint64_t>(ctx->cus);
    slots = std::min<uint64_t>(slots, std::max<uint32_t>(c->total, 1));
    size_t free_b = 0, total_b = 0;
    NG_HIP(hipMemGetInfo(&free_b, &total_b));
    Workspace& w = ctx->ws;
    const uint32_t wn = std::max(w.n, c->n_max);
    const uint64_t wm = std::max(w.m, c->m_max);
    const uint64_t per = slot_bytes(wn, wm, MODE_BC, log_capacity(wn, wm)) + 16ull * wn;
    const uint64_t held = static_cast<uint64_t>(w.slots) * slot_bytes(w.n, w.m, MODE_BC, w.logcap);
    const uint64_t budget = static_cast<uint64_t>(static_cast<double>(free_b + held) * 0.5);
    slots = std::min<uint64_t>(slots, std::max<uint64_t>(1, budget / per));
    // flush layout: slot b touches graph g when one of its items b + kS falls into g's item range
    auto flush_layout = [&](uint64_t S, std::vector<uint64_t>* fbase, std::vector<std::vector<uint64_t>>* segs) {
        uint64_t F = 0;
        for (uint64_t b = 0; b < S; ++b) {
            if (fbase) (*fbase)[b] = F;
            for (uint32_t gi = 0; gi < G; ++gi) {
                const uint64_t p0 = c->prefix[gi], p1 = c->prefix[gi + 1];
                const uint64_t first = p0 + (b + S - p0 % S) % S;
                if (first < p1) {
                    if (segs) (*segs)[gi].push_back(F);
                    F += p1 - p0;
                }
            }
        }
        return F;
    };
    // the flush regions take up to slots * total doubles: bound them by the memory budget too
    while (slots > 1 && (slots * per + 8 * flush_layout(slots, nullptr, nullptr)) > budget) slots = slots * 7 / 8;
    c->slots = static_cast<uint32_t>(slots);
    std::vector<uint64_t> fbase(slots, 0);
    std::vector<std::vector<uint64_t>> segs(G);
    const uint64_t F = flush_layout(slots, &fbase, &segs);
    std::vector<uint32_t> seg_ptr(G + 1, 0);
    std::vector<uint64_t> seg_off;
    for (uint32_t gi = 0; gi < G; ++gi) {
        seg_ptr[gi + 1] = seg_ptr[gi] + static_cast<uint32_t>(segs[gi].size());
        seg_off.insert(seg_off.end(), segs[gi].begin(), segs[gi].end());
    }
    ng_status s1 = upload(&c->d_gtab, gtab, st);
    if (s1 == NG_OK) s1 = upload(&c->d_prefix, c->prefix, st);
    if (s1 == NG_OK) s1 = upload(&c->d_fbase, fbase, st);
    if (s1 == NG_OK) s1 = upload(&c->d_seg_ptr, seg_ptr, st);
    if (s1 == NG_OK) s1 = upload(&c->d_seg_off, seg_off, st);
    if (s1 != NG_OK) return s1;
    NG_HIP(hipMalloc(&c->d_flush, 8 * std::max<uint64_t>(F, 1)));
    NG_HIP(hipMalloc(&c->d_bc, 8 * std::max<uint32_t>(c->total, 1)));
    ng_status s2 = ensure_ws(w, st, c->n_max, c->m_max, c->slots, MODE_BC, c->dk, c->total, c->slots);
    if (s2 != NG_OK) return s2;
    NG_HIP(hipStreamSynchronize(st));
    c->ws_generation = w.generation;

    LaunchArgs la{};
    la.dk = c->dk;
    la.mode = MODE_BC;
    la.n = c->n_max;
    la.view = nullptr;
    la.sources = nullptr;
    la.n_sources = c->total;
    la.delta = 0.0;
    la.grid = c->slots;
    la.ws = &w;
    la.dist_out = nullptr;
    la.dist_out_kind = 0;
    la.maps = maps;
    la.block = c->block;
    la.gtab = c->d_gtab;
    la.n_graphs = G;
    la.node_prefix = c->d_prefix;
    la.flush = c->d_flush;
    la.flush_base = c->d_fbase;
    NG_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    hipError_t e = launch_rec_zero_bc(w.rec, static_cast<uint64_t>(c->slots) * w.n, st);
    if (e == hipSuccess) e = launch_fill_u64(w.arcs_reached, 32, 0, st);
    if (e == hipSuccess) e = launch_fill_u32(w.err, 1, 0u, st);
    if (e == hipSuccess) e = launch_batch(la, st);
    if (e == hipSuccess) e = launch_cohort_reduce(c->d_flush, c->d_prefix, G, c->d_seg_ptr, c->d_seg_off, c->total, c->d_bc, st);
    hipGraph_t graph = nullptr;
    const hipError_t ee = hipStreamEndCapture(st, &graph);
    if (e != hipSuccess || ee != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return fail(NG_EDEVICE, "cohort graph capture failed: %s", hipGetErrorString(e != hipSuccess ? e : ee));
    }
    c->graph = graph;
    NG_HIP(hipGraphInstantiate(&c->exec, c->graph, nullptr, nullptr, 0));
    c->built = true;
    return NG_OK;
}

// One replay of the captured cohort launch; fills c->last.
ng_status cohort_replay(ng_cohort* c) {
    ng_context* ctx = c->ctx;
    const hipStream_t st = ctx->stream;
    NG_HIP(hipSetDevice(ctx->device));
    // the context workspace is shared: plain calls in between may have grown it (new buffers: re-capture) or
    // left +infinity of another distance kind in it (refilled here, outside the capture)
    if (c->built) {
        ng_status s0 = ensure_ws(ctx->ws, st, c->n_max, c->m_max, c->slots, MODE_BC, c->dk, c->total, c->slots);
        if (s0 != NG_OK) return s0;
        if (ctx->ws.generation != c->ws_generation) c->built = false;
        for (size_t gi = 0; gi < c->graphs.size() && c->built; ++gi)
            if (c->graphs[gi]->weights_version != c->versions[gi]) c->built = false;  // lengths changed: new light prefixes
    }
    if (!c->built) {
        ng_status s0 = cohort_build(c);
        if (s0 != NG_OK) return s0;
    }
    NG_HIP(hipEventRecord(ctx->ev0, st));
    ctx->ws.bc_zero = false;  // the replay adds betweenness partials of its own
    NG_HIP(hipGraphLaunch(c->exec, st));
    NG_HIP(hipEventRecord(ctx->ev1, st));
    const hipError_t se = hipStreamSynchronize(st);
    if (se != hipSuccess) return fail(NG_EDEVICE, "cohort launch failed: %s", hipGetErrorString(se));
    float ms = 0.0f;
    NG_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    uint32_t err = 0;
    unsigned long long ctrs[32] = {};
    NG_HIP(hipMemcpy(&err, ctx->ws.err, 4, hipMemcpyDeviceToHost));
    NG_HIP(hipMemcpy(ctrs, ctx->ws.arcs_reached, 8 * 32, hipMemcpyDeviceToHost));
    if (err & ERR_KAHN) return fail(NG_EDEVICE, "tight-arc DAG has a cycle (zero-length arcs?)");
    if (err & ERR_GUARD) return fail(NG_EDEVICE, "iteration guard hit in the batched kernel");
    if (err & ERR_OVERFLOW) return fail(NG_EDEVICE, "queue overflow in the batched kernel");
    ng_call_stats& l = c->last;
    l = ng_call_stats{};
    l.kernel_ms = ms;
    l.launches = 1;
    l.slots = c->slots;
    l.block = static_cast<uint32_t>(c->block);
    l.delta = 0.0;
    l.arcs_reached = ctrs[0];
    l.flags = err;
    l.phases = ctrs[1];
    l.buckets = ctrs[2];
    l.light_arcs = ctrs[3];
    l.settle_arcs = ctrs[4];
    l.logged_arcs = ctrs[5];
    l.layers = ctrs[6];
    l.probes = ctrs[7];
    l.improvements = ctrs[8];
    for (int sct = 0; sct < 8; ++sct) l.section_ticks[sct] = ctrs[9 + sct];
    l.tree_sources = ctrs[17];
    l.near_tree_sources = ctrs[18];
    return NG_OK;
}

}  // namespace

extern "C" {

ng_status ng_cohort_create(ng_context* ctx, ng_graph* const* graphs, uint32_t n_graphs, ng_cohort** out) {
    if (!out) return fail(NG_EINVAL, "null output pointer");
    *out = nullptr;
    if (!ctx || !graphs || !n_graphs) return fail(NG_EINVAL, "empty cohort");
    auto c = std::make_unique<ng_cohort>();
    c->ctx = ctx;
    for (uint32_t i = 0; i < n_graphs; ++i) {
        if (!graphs[i] || graphs[i]->ctx != ctx) return fail(NG_EINVAL, "graph %u is null or belongs to another context", i);
        c->graphs.push_back(graphs[i]);
        c->delta.push_back(0.0);
    }
    ctx->refs.fetch_add(1, std::memory_order_relaxed);
    *out = c.release();
    return NG_OK;
}

void ng_cohort_destroy(ng_cohort* c) {
    if (!c) return;
    ng_context* ctx = c->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    cohort_free_device(c);
    delete c;
    context_release(ctx);
}

ng_status ng_cohort_brandes(ng_cohort* c, double* bc_out, ng_sssp_summary* summaries_out, double* kernel_ms) {
    if (!c) return fail(NG_EINVAL, "null cohort");
    if (!bc_out) return fail(NG_EINVAL, "null betweenness buffer");
    ng_status st = cohort_replay(c);
    if (st != NG_OK) return st;
    NG_HIP(hipMemcpy(bc_out, c->d_bc, 8 * static_cast<size_t>(c->total), hipMemcpyDeviceToHost));
    if (summaries_out && c->total) {
        NG_HIP(hipMemcpy(summaries_out, c->ctx->ws.summaries, sizeof(ng_sssp_summary) * c->total, hipMemcpyDeviceToHost));
    }
    if (kernel_ms) *kernel_ms = c->last.kernel_ms;
    return NG_OK;
}

ng_status ng_cohort_autotune(ng_cohort* c, uint32_t n_sample) {
    if (!c) return fail(NG_EINVAL, "null cohort");
    if (n_sample == 0) return fail(NG_EINVAL, "empty sample");
    // per graph: the bucket width
    for (size_t gi = 0; gi < c->graphs.size(); ++gi) {
        ng_graph* g = c->graphs[gi];
        const uint32_t k = std::min(n_sample, g->n);
        std::vector<uint32_t> sample(k);
        for (uint32_t i = 0; i < k; ++i) sample[i] = static_cast<uint32_t>((static_cast<uint64_t>(i) * g->n) / k);
        double d = 0.0;
        ng_status st = ng_autotune_delta(g, NG_LEN_ABS, 1, sample.data(), k, &d);
        if (st != NG_OK) return st;
        c->delta[gi] = d;
    }
    // launch shape, one whole replay each: 256-lane workgroups at the occupancy maximum and one below it,
    // 512- and 1024-lane workgroups at theirs
    int best_block = kBlock, best_bpc = 0;
    double best_ms = std::numeric_limits<double>::infinity();
    uint32_t slots_at_max = 0;
    for (int cand = 0; cand < 4; ++cand) {
        c->block = cand < 2 ? 256 : (cand == 2 ? 512 : 1024);
        c->bpc = 0;
        if (cand == 1) {
            const int at_max = static_cast<int>(slots_at_max / std::max(1u, static_cast<uint32_t>(c->ctx->cus)));
            if (at_max <= 1) continue;
            c->bpc = at_max - 1;
        }
        c->built = false;
        ng_status st = cohort_replay(c);
        if (st != NG_OK) return st;
        if (cand == 0) slots_at_max = c->slots;
        if (c->last.kernel_ms < best_ms) {
            best_ms = c->last.kernel_ms;
            best_block = c->block;
            best_bpc = c->bpc;
        }
    }
    c->block = best_block;
    c->bpc = best_bpc;
    c->built = false;
    return NG_OK;
}

ng_status ng_cohort_last_stats(const ng_cohort* c, ng_call_stats* out) {
    if (!c) return fail(NG_EINVAL, "null cohort");
    if (!out) return fail(NG_EINVAL, "null output pointer");
    *out = c->last;
    return NG_OK;
}

}  // extern "C"
