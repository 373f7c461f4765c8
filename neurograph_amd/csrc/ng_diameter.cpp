// Diameter approximation on one rank: the reference's ball clustering and contraction on the host, the exact
// diameter of the contracted graph on the device. See ng_contract_balls / ng_diameter_approximation in
// include/neurograph_io.h.

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <numeric>
#include <random>
#include <utility>
#include <vector>

#include "../../include/neurograph_hip.h"
#include "../../include/neurograph_io.h"

extern "C" void ng_set_error_message(const char* msg);

namespace {

ng_status dia_fail(ng_status st, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    ng_set_error_message(buf);
    return st;
}

constexpr uint32_t kNoCenter = std::numeric_limits<uint32_t>::max();
constexpr uint64_t kFar = std::numeric_limits<uint64_t>::max();

// Per-node clustering state (the reference keeps it in four RMA windows, DiameterApproximation.h:129-152).
struct Balls {
    std::vector<uint32_t> center;
    std::vector<uint64_t> dist;
    std::vector<uint64_t> generation;
    std::vector<uint8_t> stable;
    explicit Balls(uint32_t n) : center(n, kNoCenter), dist(n, kFar), generation(n, kFar), stable(n, 0) {}
};

// A claim of one growing round: target t joins `center` at distance d (generation of the claiming ball).
struct Claim {
    uint32_t t;
    uint32_t center;
    uint64_t d;
    uint64_t generation;
};

// Draws of the reference's DefaultRandomNumberGenerator (RandomNumberGenerator.h): std::mt19937 behind
// std::uniform_real_distribution<double>(0, 1).
struct DefaultDraws {
    std::mt19937 prng;
    std::uniform_real_distribution<double> dist{0.0, 1.0};
    explicit DefaultDraws(uint32_t seed) : prng(seed) {}
    static double draw(void* self) {
        auto* d = static_cast<DefaultDraws*>(self);
        return d->dist(d->prng);
    }
};

// One growing step (DiameterApproximation.h:204-284): synchronous rounds over every node with a center until no
// claim succeeds. Claims are built from the state at the start of a round, then applied in claiming order.
void grow(const ng_host_csr& g, const int32_t* w, uint64_t r, uint64_t iteration, Balls& b) {
    std::vector<Claim> claims;
    const uint64_t two_r = 2 * r;
    for (;;) {
        claims.clear();
        for (uint32_t u = 0; u < g.n; ++u) {
            if (b.center[u] == kNoCenter) continue;
            const uint64_t du = b.dist[u];
            const uint64_t gu = b.generation[u];
            const uint64_t reach = (iteration - gu + 1) * two_r;
            for (uint64_t a = g.offsets[u]; a < g.offsets[u + 1]; ++a) {
                const uint64_t len = static_cast<uint64_t>(w[a]);
                if (len == 0 || len > two_r) continue;  // weight-0 arcs are not stored (DistributedGraph.cpp:222-225)
                const uint32_t t = g.targets[a];
                if (b.stable[t] != 0) continue;
                const uint64_t d = du + len;
                if (reach < d || d >= b.dist[t]) continue;
                claims.push_back(Claim{t, b.center[u], d, gu});
            }
        }
        bool changed = false;
        for (const Claim& c : claims) {
            if (b.dist[c.t] <= c.d) continue;
            b.center[c.t] = c.center;
            b.dist[c.t] = c.d;
            b.generation[c.t] = c.generation;
            b.stable[c.t] = 0;
            changed = true;
        }
        if (!changed) return;
    }
}

ng_status new_csr(uint32_t n, uint64_t m, ng_host_csr* out) {
    std::memset(out, 0, sizeof *out);
    out->n = n;
    out->m = m;
    out->wkind = NG_W_I32;
    out->offsets = static_cast<uint64_t*>(std::calloc(static_cast<size_t>(n) + 1, sizeof(uint64_t)));
    out->targets = static_cast<uint32_t*>(std::malloc(std::max<size_t>(m, 1) * sizeof(uint32_t)));
    out->weights = std::malloc(std::max<size_t>(m, 1) * sizeof(int32_t));
    if (!out->offsets || !out->targets || !out->weights) {
        ng_host_csr_free(out);
        return dia_fail(NG_ENOMEM, "host allocation failed");
    }
    return NG_OK;
}

}  // namespace

extern "C" {

ng_status ng_contract_balls(const ng_host_csr* g, uint64_t r, ng_draw_fn draw, void* user, uint32_t seed, ng_host_csr* out,
                            uint32_t* centers_out) {
    if (!g || !out) return dia_fail(NG_EINVAL, "null argument");
    std::memset(out, 0, sizeof *out);
    if (g->wkind != NG_W_I32) return dia_fail(NG_EUNSUPPORTED, "the diameter approximation takes integer weights (distance_type, Types.h:32)");
    if (g->n == 0) return dia_fail(NG_EINVAL, "empty graph: no node can become a center");
    if (g->m && (!g->targets || !g->weights)) return dia_fail(NG_EINVAL, "null targets or weights");
    if (r > (std::numeric_limits<uint64_t>::max() >> 8)) return dia_fail(NG_EINVAL, "radius %llu too large", (unsigned long long)r);
    const uint32_t n = g->n;
    const int32_t* w = static_cast<const int32_t*>(g->weights);
    if (g->offsets[0] != 0 || g->offsets[n] != g->m) return dia_fail(NG_EINVAL, "offsets must start at 0 and end at m");
    for (uint32_t u = 0; u < n; ++u)
        if (g->offsets[u] > g->offsets[u + 1]) return dia_fail(NG_EINVAL, "offsets decrease at node %u", u);
    for (uint64_t a = 0; a < g->m; ++a) {
        if (g->targets[a] >= n) return dia_fail(NG_EINVAL, "arc %llu targets node %u, the graph has %u nodes", (unsigned long long)a, g->targets[a], n);
        // safe_cast<distance_type>(weight) throws on a negative weight (DiameterApproximation.h:231)
        if (w[a] < 0) return dia_fail(NG_EINVAL, "negative weight %d on arc %llu: the diameter approximation casts weights to distance_type", w[a], (unsigned long long)a);
    }
    DefaultDraws defaults(seed);
    if (draw == nullptr) {
        draw = &DefaultDraws::draw;
        user = &defaults;
    }
    Balls b(n);
    // DiameterApproximation.h:79: ceil(log2(number of nodes)) iterations at most
    const uint64_t last_iteration = static_cast<uint64_t>(std::ceil(std::log2(static_cast<double>(n))));
    for (uint64_t it = 0; it < last_iteration; ++it) {
        // mark_as_centers (:154-179): nodes without a center draw, in id order
        const double threshold = std::pow(2.0, static_cast<double>(it)) / static_cast<double>(n);
        bool has_center = false;
        for (uint32_t v = 0; v < n; ++v) {
            if (b.center[v] != kNoCenter) {
                has_center = true;
                continue;
            }
            if (draw(user) >= threshold) continue;
            b.center[v] = v;
            b.dist[v] = 0;
            b.generation[v] = it;
            b.stable[v] = 1;
            has_center = true;
        }
        if (!has_center) continue;
        grow(*g, w, r, it, b);
        // stabilize_nodes (:286-295), check_all_stable (:191-202)
        bool all_stable = true;
        for (uint32_t v = 0; v < n; ++v) {
            if (b.center[v] != kNoCenter) b.stable[v] = 1;
            all_stable = all_stable && b.stable[v] != 0;
        }
        if (all_stable) break;
    }
    // contract (:297-561): centers in increasing id become the new nodes
    std::vector<uint32_t> new_id(n, kNoCenter);
    uint32_t k = 0;
    for (uint32_t v = 0; v < n; ++v) {
        if (b.center[v] == kNoCenter) return dia_fail(NG_EINVAL, "node %u was reached by no ball (DiameterApproximation.h:329)", v);
        if (b.center[v] == v) new_id[v] = k++;
    }
    // one arc per (center(u), center(v)); a later arc between the same centers replaces the earlier one
    std::vector<std::map<uint32_t, int32_t>> rows(k);
    for (uint32_t u = 0; u < n; ++u) {
        const uint32_t cu = new_id[b.center[u]];
        for (uint64_t a = g->offsets[u]; a < g->offsets[u + 1]; ++a) {
            if (w[a] == 0) continue;  // not stored by the reference's upload (DistributedGraph.cpp:222-225)
            const uint32_t t = g->targets[a];
            const uint64_t nw = static_cast<uint64_t>(w[a]) + b.dist[u] + b.dist[t];
            if (nw > static_cast<uint64_t>(std::numeric_limits<int32_t>::max())) {
                return dia_fail(NG_EINVAL, "contracted weight %llu does not fit weight_type (DiameterApproximation.h:474)", (unsigned long long)nw);
            }
            rows[cu][new_id[b.center[t]]] = static_cast<int32_t>(nw);
        }
    }
    uint64_t m = 0;
    for (const auto& row : rows) m += row.size();
    ng_status st = new_csr(k, m, out);
    if (st != NG_OK) return st;
    uint64_t at = 0;
    for (uint32_t c = 0; c < k; ++c) {
        for (const auto& [t, wt] : rows[c]) {
            out->targets[at] = t;
            static_cast<int32_t*>(out->weights)[at] = wt;
            ++at;
        }
        out->offsets[c + 1] = at;
    }
    if (centers_out)
        for (uint32_t v = 0; v < n; ++v) centers_out[v] = b.center[v];
    return NG_OK;
}

ng_status ng_diameter_approximation(ng_context* ctx, const ng_host_csr* g, uint64_t r, ng_draw_fn draw, void* user, uint32_t seed,
                                    uint64_t* diameter_out, uint32_t* contracted_nodes_out) {
    if (!ctx || !diameter_out) return dia_fail(NG_EINVAL, "null argument");
    ng_host_csr small{};
    ng_status st = ng_contract_balls(g, r, draw, user, seed, &small, nullptr);
    if (st != NG_OK) return st;
    ng_graph* dg = nullptr;
    st = ng_graph_create(ctx, small.n, small.m, small.offsets, small.targets, small.weights, small.wkind, &dg);
    if (st != NG_OK) {
        ng_host_csr_free(&small);
        return st;
    }
    // AllPairsShortestPath::compute_apsp on the contracted graph (DiameterApproximation.h:118): its diameter is the
    // largest finite distance over all sources
    std::vector<uint32_t> sources(small.n);
    std::iota(sources.begin(), sources.end(), 0u);
    std::vector<ng_sssp_summary> per(small.n);
    st = ng_sssp_summaries(dg, sources.data(), small.n, NG_LEN_ABS, 0.0, per.data(), nullptr);
    ng_graph_destroy(dg);
    const uint32_t k = small.n;
    ng_host_csr_free(&small);
    if (st != NG_OK) return st;
    uint64_t diameter = 0;
    for (const ng_sssp_summary& s : per) diameter = std::max<uint64_t>(diameter, s.max_d_u64);
    *diameter_out = diameter + 2 * r;
    if (contracted_nodes_out) *contracted_nodes_out = k;
    return NG_OK;
}

}  // extern "C"
