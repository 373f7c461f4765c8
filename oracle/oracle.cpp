// ORACLE — test infrastructure only. Nothing in the product links, loads or calls this file.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, and only as the reference
// check (or the timed CPU baseline), never as the thing that is measured or shipped.
//
// A sequential CPU restatement of neurograph's APSP / Brandes hot path, generic in the weight type
// (the reference is integer-only: weight_type = int32, distance_type = uint64, Types.h:31-33).
// Every function cites the reference file:line it restates. Paths are relative to the reference
// checkout (tuda-hpclab/neurograph).
//
// Parity status of each function (how it is pinned) is listed in the header comment of oracle.py.
//
// Build: g++ -O2 -std=c++20 -ffp-contract=off -fno-fast-math -fPIC -shared -pthread (see oracle/Makefile).
// -ffp-contract=off matters: distances are single IEEE adds, which must not fuse with anything.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <limits>
#include <queue>
#include <thread>
#include <type_traits>
#include <utility>
#include <vector>

namespace {

enum WeightKind : int { W_I32 = 0, W_F32 = 1, W_F64 = 2 };
enum LengthKind : int { L_ABS = 0, L_INVERSE = 1, L_RAW = 2 };
enum Status : int { OK = 0, EINVAL_ = 1, ENOMEM_ = 2, EUNSUPPORTED_ = 4 };

// Per-source summary, SsspResult.h:19-25 plus the exact integer channels.
struct Summary {
    double sum_d;        // sum of finite distances (exact for integer distances up to 2^53)
    uint64_t sum_d_u64;  // sum of finite distances, integer distances only (else 0)
    double max_d;        // maximum finite distance
    uint64_t max_d_u64;  // maximum finite distance, integer distances only (else 0)
    double sum_eff;      // sum of 1/d over finite, non-zero d, accumulated in global-id order
    uint64_t n_unreached;
};

struct ApspResult {  // ApspGlobalResult, Types.h:237-244, plus the raw sums
    double avg_spl;
    double avg_eff;
    double diameter;
    uint64_t diameter_u64;
    uint64_t n_disconnected;
    uint64_t sum_d_u64;
    double sum_d;
    double sum_eff;
};

template <class W>
struct Csr {
    uint32_t n;
    const uint64_t* off;
    const uint32_t* tgt;
    const W* w;
};

// DijkstraCommon::transform_weight, DijkstraCommon.h:36-54.
//   |w| (widened before abs, :41-43) for the plain metric; 1/|w| in double with 0 -> max for the inverse metric.
// The real-valued weight kinds are an extension of this library: |w| keeps the weight's own precision.
template <class D, class W>
inline D transform_weight(W w, int lkind) {
    if (lkind == L_INVERSE) {
        double aw;
        if constexpr (std::is_integral_v<W>) {
            aw = static_cast<double>(std::llabs(static_cast<long long>(w)));
        } else {
            aw = static_cast<double>(std::fabs(w));
        }
        if (aw == 0.0) return std::numeric_limits<D>::max();
        return static_cast<D>(1.0 / aw);
    }
    if constexpr (std::is_integral_v<W>) {
        return static_cast<D>(std::llabs(static_cast<long long>(w)));
    } else {
        return static_cast<D>(std::fabs(w));
    }
}

// summarize_distances, SsspResult.h:43-64: sum of finite d, max finite d, sum of 1/d over d != 0
// (in global-id order), count of d == max. Sums of real-valued distances are kept in double.
template <class D>
Summary summarize(const std::vector<D>& dist) {
    constexpr D maxd = std::numeric_limits<D>::max();
    Summary s{};
    double sum_eff = 0.0;
    using Acc = std::conditional_t<std::is_integral_v<D>, uint64_t, double>;
    Acc sum = 0;
    D mx = 0;
    uint64_t unreached = 0;
    for (const D d : dist) {
        if (d == maxd) {
            ++unreached;
        } else {
            sum += static_cast<Acc>(d);
            if (d != 0) sum_eff += 1.0 / static_cast<double>(d);
            mx = std::max(mx, d);
        }
    }
    s.sum_d = static_cast<double>(sum);
    s.sum_d_u64 = std::is_integral_v<D> ? static_cast<uint64_t>(sum) : 0;
    s.max_d = static_cast<double>(mx);
    s.max_d_u64 = std::is_integral_v<D> ? static_cast<uint64_t>(mx) : 0;
    s.sum_eff = sum_eff;
    s.n_unreached = unreached;
    return s;
}

template <class D>
struct VertexDist {  // VertexDistance.h:17-24; ordered by distance only (GreaterVertexDistance, :31-36)
    D d;
    uint32_t id;
};
template <class D>
struct Greater {
    bool operator()(const VertexDist<D>& a, const VertexDist<D>& b) const { return a.d > b.d; }
};

// PriorityQueueVariant::compute_distances, PriorityQueueVariant.h:47-93, with the set-up of
// Dijkstra::compute_sssp, Dijkstra.h:69-76: fill with max, root = 0, lazy-deletion binary heap,
// stale entries skipped (:66-69), infinitely long arcs skipped (:75-79), strict improvement only (:85-87).
template <class D, class W>
void pq_dijkstra(const Csr<W>& g, int lkind, uint32_t root, std::vector<D>& dist) {
    constexpr D maxd = std::numeric_limits<D>::max();
    std::fill(dist.begin(), dist.end(), maxd);
    dist[root] = 0;
    std::priority_queue<VertexDist<D>, std::vector<VertexDist<D>>, Greater<D>> q;
    q.push({D(0), root});
    while (!q.empty()) {
        const auto [du, u] = q.top();
        q.pop();
        if (dist[u] < du) continue;
        for (uint64_t e = g.off[u]; e < g.off[u + 1]; ++e) {
            const D len = transform_weight<D>(g.w[e], lkind);
            if (len == maxd) continue;
            const D nd = du + len;
            const uint32_t t = g.tgt[e];
            if (dist[t] <= nd) continue;
            dist[t] = nd;
            q.push({nd, t});
        }
    }
}

// Sequential delta-stepping with the reference's bucket semantics: bucket = d / delta (EpochCommon.h:46-49),
// improved nodes become active and move to a smaller bucket (PushCheckVariant.h:93-109), phases repeat
// until no node is active (:57-65), epochs run while a non-empty bucket is left (:123-125, Buckets.h:90-104).
// Weights are used raw and must be non-negative (safe_cast<distance_type>(weight) throws, PushCheckVariant.h:82).
// Only the integer weight kind has this path in the reference.
Status delta_stepping(const Csr<int32_t>& g, uint64_t delta, uint32_t root, std::vector<uint64_t>& dist) {
    constexpr uint64_t maxd = std::numeric_limits<uint64_t>::max();
    if (delta == 0) return EINVAL_;
    std::fill(dist.begin(), dist.end(), maxd);
    dist[root] = 0;
    std::vector<std::vector<uint32_t>> buckets(1);
    std::vector<uint64_t> node_bucket(g.n, std::numeric_limits<uint64_t>::max());
    std::vector<char> active(g.n, 0);
    std::vector<uint32_t> active_list{root};
    active[root] = 1;
    node_bucket[root] = 0;
    buckets[0].push_back(root);
    auto has_nodes_left = [&](size_t from) {
        for (size_t b = from; b < buckets.size(); ++b) {
            for (uint32_t v : buckets[b]) {
                if (node_bucket[v] == b) return true;
            }
        }
        return false;
    };
    for (size_t epoch = 0; has_nodes_left(epoch); ++epoch) {
        while (!active_list.empty()) {
            std::vector<std::pair<uint32_t, uint64_t>> questions;
            for (uint32_t u : active_list) active[u] = 0;
            for (uint32_t u : active_list) {
                const uint64_t du = dist[u];
                for (uint64_t e = g.off[u]; e < g.off[u + 1]; ++e) {
                    if (g.w[e] < 0) return EINVAL_;
                    const uint64_t nd = du + static_cast<uint64_t>(g.w[e]);
                    if (nd < dist[g.tgt[e]]) questions.emplace_back(g.tgt[e], nd);
                }
            }
            active_list.clear();
            for (const auto& [t, nd] : questions) {
                const uint64_t old = dist[t];
                if (old <= nd) continue;
                dist[t] = nd;
                if (!active[t]) {
                    active[t] = 1;
                    active_list.push_back(t);
                }
                const uint64_t nb = nd / delta;
                const uint64_t ob = old == maxd ? std::numeric_limits<uint64_t>::max() : old / delta;
                if (nb < ob) {
                    node_bucket[t] = nb;
                    if (nb >= buckets.size()) buckets.resize(nb * 2 + 1);
                    buckets[nb].push_back(t);
                }
            }
        }
    }
    return OK;
}

// BrandesVariant::accumulate_from_source, BrandesVariant.h:79-159, generalised in two ways that leave the
// reference arithmetic untouched when seeds are 1:
//   * predecessors are recorded as arc ids instead of node ids (the reference keeps one entry per parallel
//     arc anyway, :133-139), so the per-arc term sigma(v) * portion(w) (:151-154) can be accumulated per arc;
//   * the backward sweep takes a per-node seed g(w) in place of the constant 1 in (1 + delta(w)) (:149):
//     g = 1 is Brandes; g(w) = 1/d(w)^2 gives the dependency used by the efficiency adjoint.
// Returns the settled count. `bc` accumulates like local_betweenness (:157); `edge` (optional) per arc.
template <class D, class W>
struct BrandesState {
    std::vector<D> dist;
    std::vector<double> sigma, delta;
    std::vector<std::vector<uint64_t>> pred_arcs;
    std::vector<uint32_t> settled;
    std::vector<uint32_t> arc_src;
    explicit BrandesState(uint32_t n) : dist(n, std::numeric_limits<D>::max()), sigma(n, 0.0), delta(n, 0.0), pred_arcs(n) {
        settled.reserve(n);
    }
};

enum SeedKind : int { SEED_ONE = 0, SEED_INV_D2 = 1 };

template <class D, class W>
void brandes_source(const Csr<W>& g, uint32_t root, BrandesState<D, W>& st, int seed_kind, double* bc, double* edge,
                    std::vector<D>* dist_copy, std::vector<double>* sigma_copy) {
    constexpr D maxd = std::numeric_limits<D>::max();
    for (uint32_t v : st.settled) {  // reset exactly the previously settled nodes, :90-96
        st.dist[v] = maxd;
        st.sigma[v] = 0.0;
        st.delta[v] = 0.0;
        st.pred_arcs[v].clear();
    }
    st.settled.clear();
    st.dist[root] = 0;
    st.sigma[root] = 1.0;
    std::priority_queue<VertexDist<D>, std::vector<VertexDist<D>>, Greater<D>> q;
    q.push({D(0), root});
    while (!q.empty()) {  // :104-141
        const auto [du, u] = q.top();
        q.pop();
        if (st.dist[u] < du) continue;
        st.settled.push_back(u);
        for (uint64_t e = g.off[u]; e < g.off[u + 1]; ++e) {
            const D nd = du + transform_weight<D>(g.w[e], L_ABS);
            const uint32_t t = g.tgt[e];
            if (nd < st.dist[t]) {
                st.dist[t] = nd;
                st.sigma[t] = st.sigma[u];
                st.pred_arcs[t].clear();
                st.pred_arcs[t].push_back(e);
                q.push({nd, t});
            } else if (nd == st.dist[t]) {
                st.sigma[t] += st.sigma[u];
                st.pred_arcs[t].push_back(e);
            }
        }
    }
    if (dist_copy) *dist_copy = st.dist;
    if (sigma_copy) *sigma_copy = st.sigma;
    // Backward sweep in reverse settle order, skipping the root (index 0), :146-158.
    for (size_t idx = st.settled.size(); idx > 1;) {
        --idx;
        const uint32_t w = st.settled[idx];
        double seed = 1.0;
        if (seed_kind == SEED_INV_D2) {
            const double dw = static_cast<double>(st.dist[w]);
            seed = 1.0 / (dw * dw);
        }
        const double portion = (seed + st.delta[w]) / st.sigma[w];
        for (const uint64_t e : st.pred_arcs[w]) {
            const uint32_t v = st.arc_src[e];
            const double term = st.sigma[v] * portion;
            st.delta[v] += term;
            if (edge) edge[e] += term;
        }
        if (bc) bc[w] += st.delta[w];
    }
}

template <class W>
std::vector<uint32_t> make_arc_src(const Csr<W>& g) {
    std::vector<uint32_t> src(g.off[g.n]);
    for (uint32_t u = 0; u < g.n; ++u)
        for (uint64_t e = g.off[u]; e < g.off[u + 1]; ++e) src[e] = u;
    return src;
}

// Splits [0, ns) into `threads` contiguous blocks, like rank-owned local sources (AllPairsShortestPath.h:130),
// runs fn(block_begin, block_end, worker) on each, and joins. Workers own private buffers.
template <class Fn>
void run_blocks(uint32_t ns, int threads, Fn&& fn) {
    if (threads < 1) threads = 1;
    if (static_cast<uint32_t>(threads) > ns) threads = static_cast<int>(std::max<uint32_t>(ns, 1));
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t) {
        const uint32_t b = static_cast<uint32_t>((static_cast<uint64_t>(ns) * t) / threads);
        const uint32_t e = static_cast<uint32_t>((static_cast<uint64_t>(ns) * (t + 1)) / threads);
        if (threads == 1) {
            fn(b, e, t);
        } else {
            pool.emplace_back([=, &fn] { fn(b, e, t); });
        }
    }
    for (auto& th : pool) th.join();
}

// compute_apsp_impl, AllPairsShortestPath.h:111-163: per source Dijkstra + summarize; sums, max, unreachable
// count and ++cluster_sizes[N - unreachable] (:130-141); then averages over reached ordered pairs (:154-158).
// The worker partial sums are reduced in worker order (the MPI reductions of :149-152 on P ranks).
template <class D, class W>
Status apsp_impl(const Csr<W>& g, int lkind, const uint32_t* sources, uint32_t ns, int threads, Summary* per_source,
                 ApspResult* res, uint64_t* cluster_sizes) {
    for (uint32_t i = 0; i < ns; ++i)
        if (sources[i] >= g.n) return EINVAL_;
    using Acc = std::conditional_t<std::is_integral_v<D>, uint64_t, double>;
    struct Part {
        Acc sum = 0;
        double eff = 0.0;
        D mx = 0;
        uint64_t unreach = 0;
        std::vector<uint64_t> clusters;
    };
    const int nt = std::max(1, std::min<int>(threads, static_cast<int>(std::max<uint32_t>(ns, 1))));
    std::vector<Part> parts(nt);
    run_blocks(ns, nt, [&](uint32_t b, uint32_t e, int t) {
        Part& p = parts[t];
        p.clusters.assign(static_cast<size_t>(g.n) + 1, 0);
        std::vector<D> dist(g.n);
        for (uint32_t i = b; i < e; ++i) {
            pq_dijkstra<D>(g, lkind, sources[i], dist);
            const Summary s = summarize(dist);
            if (per_source) per_source[i] = s;
            if constexpr (std::is_integral_v<D>) {
                p.sum += s.sum_d_u64;
                p.mx = std::max<D>(p.mx, static_cast<D>(s.max_d_u64));
            } else {
                p.sum += s.sum_d;
                p.mx = std::max<D>(p.mx, static_cast<D>(s.max_d));
            }
            p.eff += s.sum_eff;
            p.unreach += s.n_unreached;
            ++p.clusters[g.n - s.n_unreached];
        }
    });
    Acc sum = 0;
    double eff = 0.0;
    D mx = 0;
    uint64_t unreach = 0;
    if (cluster_sizes) std::fill(cluster_sizes, cluster_sizes + g.n + 1, 0);
    for (const Part& p : parts) {
        sum += p.sum;
        eff += p.eff;
        mx = std::max(mx, p.mx);
        unreach += p.unreach;
        if (cluster_sizes && !p.clusters.empty())
            for (size_t k = 0; k <= g.n; ++k) cluster_sizes[k] += p.clusters[k];
    }
    if (res) {
        const uint64_t N = g.n;
        const uint64_t pairs = N * N - N;
        const uint64_t reached = pairs - unreach;
        res->avg_spl = static_cast<double>(sum) / static_cast<double>(reached);
        res->avg_eff = eff / static_cast<double>(reached);
        res->diameter = static_cast<double>(mx);
        res->diameter_u64 = std::is_integral_v<D> ? static_cast<uint64_t>(mx) : 0;
        res->n_disconnected = unreach;
        res->sum_d_u64 = std::is_integral_v<D> ? static_cast<uint64_t>(sum) : 0;
        res->sum_d = static_cast<double>(sum);
        res->sum_eff = eff;
    }
    return OK;
}

// BrandesVariant::accumulate_betweenness, BrandesVariant.h:50-74, over the given sources; worker-private
// betweenness vectors are reduced componentwise in worker order (BetweennessCentrality.h:74).
template <class D, class W>
Status brandes_impl(const Csr<W>& g, const uint32_t* sources, uint32_t ns, int threads, int seed_kind, double* bc,
                    double* edge, Summary* per_source, double* sigma_out, void* dist_out) {
    for (uint32_t i = 0; i < ns; ++i)
        if (sources[i] >= g.n) return EINVAL_;
    const std::vector<uint32_t> arc_src = make_arc_src(g);
    const uint64_t m = g.off[g.n];
    const int nt = std::max(1, std::min<int>(threads, static_cast<int>(std::max<uint32_t>(ns, 1))));
    std::vector<std::vector<double>> bcs(nt), edges(nt);
    run_blocks(ns, nt, [&](uint32_t b, uint32_t e, int t) {
        BrandesState<D, W> st(g.n);
        st.arc_src = arc_src;
        bcs[t].assign(g.n, 0.0);
        if (edge) edges[t].assign(m, 0.0);
        std::vector<D> dcopy;
        std::vector<double> scopy;
        const bool want_copy = per_source || sigma_out || dist_out;
        for (uint32_t i = b; i < e; ++i) {
            brandes_source<D, W>(g, sources[i], st, seed_kind, bcs[t].data(), edge ? edges[t].data() : nullptr,
                                 want_copy ? &dcopy : nullptr, sigma_out ? &scopy : nullptr);
            if (per_source) per_source[i] = summarize(dcopy);
            if (sigma_out) std::memcpy(sigma_out + static_cast<size_t>(i) * g.n, scopy.data(), sizeof(double) * g.n);
            if (dist_out) std::memcpy(static_cast<D*>(dist_out) + static_cast<size_t>(i) * g.n, dcopy.data(), sizeof(D) * g.n);
        }
    });
    for (int t = 0; t < nt; ++t) {
        for (uint32_t v = 0; v < g.n; ++v) bc[v] += bcs[t][v];
        if (edge)
            for (uint64_t a = 0; a < m; ++a) edge[a] += edges[t][a];
    }
    return OK;
}

template <class Fn>
int dispatch_weights(int wkind, const void* w, Fn&& fn) {
    switch (wkind) {
    case W_I32: return fn(static_cast<const int32_t*>(w));
    case W_F32: return fn(static_cast<const float*>(w));
    case W_F64: return fn(static_cast<const double*>(w));
    default: return EUNSUPPORTED_;
    }
}

}  // namespace

extern "C" {

// Distance type selected by (weight kind, length kind): int32 + |w| -> uint64 (distance_type),
// anything + 1/|w| -> double (inverse_distance_type), f32 + |w| -> float, f64 + |w| -> double.
int orc_dist_kind(int wkind, int lkind) {
    if (lkind == L_INVERSE) return 2;
    if (wkind == W_I32) return 0;
    if (wkind == W_F32) return 1;
    return 2;  // 0 = u64, 1 = f32, 2 = f64
}

// One SSSP from `root`: Dijkstra::compute_sssp<d> with the priority-queue variant (Dijkstra.h:56-98).
int orc_sssp(uint32_t n, const uint64_t* off, const uint32_t* tgt, const void* w, int wkind, int lkind, uint32_t root,
             void* dist_out, Summary* summary_out) {
    if (root >= n) return EINVAL_;  // Dijkstra.h:66-67
    return dispatch_weights(wkind, w, [&](auto* wp) -> int {
        using W = std::remove_const_t<std::remove_pointer_t<decltype(wp)>>;
        const Csr<W> g{n, off, tgt, wp};
        auto run = [&](auto tag) -> int {
            using D = decltype(tag);
            std::vector<D> dist(n);
            pq_dijkstra<D>(g, lkind, root, dist);
            if (dist_out) std::memcpy(dist_out, dist.data(), sizeof(D) * n);
            if (summary_out) *summary_out = summarize(dist);
            return OK;
        };
        switch (orc_dist_kind(wkind, lkind)) {
        case 0: return run(uint64_t{});
        case 1: return run(float{});
        default: return run(double{});
        }
    });
}

// AllPairsShortestPath::compute_apsp / compute_apsp_inverse over `sources` (all nodes for the reference result).
int orc_apsp(uint32_t n, const uint64_t* off, const uint32_t* tgt, const void* w, int wkind, int lkind,
             const uint32_t* sources, uint32_t ns, int threads, Summary* per_source, ApspResult* res, uint64_t* cluster_sizes) {
    return dispatch_weights(wkind, w, [&](auto* wp) -> int {
        using W = std::remove_const_t<std::remove_pointer_t<decltype(wp)>>;
        const Csr<W> g{n, off, tgt, wp};
        switch (orc_dist_kind(wkind, lkind)) {
        case 0: return apsp_impl<uint64_t>(g, lkind, sources, ns, threads, per_source, res, cluster_sizes);
        case 1: return apsp_impl<float>(g, lkind, sources, ns, threads, per_source, res, cluster_sizes);
        default: return apsp_impl<double>(g, lkind, sources, ns, threads, per_source, res, cluster_sizes);
        }
    });
}

// Brandes over `sources`, lengths |w| (BrandesVariant.h:127 uses transform_weight<distance_type>).
// seed_kind 0: betweenness; 1: the 1/d^2-seeded dependency (efficiency adjoint). bc (n) and edge (m) accumulate.
int orc_brandes(uint32_t n, const uint64_t* off, const uint32_t* tgt, const void* w, int wkind, const uint32_t* sources,
                uint32_t ns, int threads, int seed_kind, double* bc, double* edge, Summary* per_source, double* sigma_out,
                void* dist_out) {
    return dispatch_weights(wkind, w, [&](auto* wp) -> int {
        using W = std::remove_const_t<std::remove_pointer_t<decltype(wp)>>;
        const Csr<W> g{n, off, tgt, wp};
        switch (orc_dist_kind(wkind, L_ABS)) {
        case 0: return brandes_impl<uint64_t>(g, sources, ns, threads, seed_kind, bc, edge, per_source, sigma_out, dist_out);
        case 1: return brandes_impl<float>(g, sources, ns, threads, seed_kind, bc, edge, per_source, sigma_out, dist_out);
        default: return brandes_impl<double>(g, sources, ns, threads, seed_kind, bc, edge, per_source, sigma_out, dist_out);
        }
    });
}

// DeltaStepping::compute_shortest_distances semantics (DeltaStepping.h:36-82) for integer weights.
int orc_delta_stepping(uint32_t n, const uint64_t* off, const uint32_t* tgt, const int32_t* w, uint64_t delta, uint32_t root,
                       uint64_t* dist_out, Summary* summary_out) {
    if (root >= n) return EINVAL_;
    const Csr<int32_t> g{n, off, tgt, w};
    std::vector<uint64_t> dist(n);
    const Status st = delta_stepping(g, delta, root, dist);
    if (st != OK) return st;
    if (dist_out) std::memcpy(dist_out, dist.data(), sizeof(uint64_t) * n);
    if (summary_out) *summary_out = summarize(dist);
    return OK;
}

int orc_hardware_threads(void) { return static_cast<int>(std::thread::hardware_concurrency()); }

}  // extern "C"
