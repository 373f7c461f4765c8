// C++20 host mirror of neurograph's APSP / betweenness drivers over the MI355X path (libng_hip.so).
//
// Same names, argument meaning and error behaviour as the reference's metric classes, with the GPU
// variant added to each selector enum:
//   AllPairsShortestPath::compute_apsp / compute_apsp_inverse / compute_apsp_delta_stepping
//       (source/metrics/paths/AllPairsShortestPath.h:54-98)
//   BetweennessCentrality::compute_average_betweenness_centrality (BetweennessCentrality.h:46-80)
//   DiameterApproximation::compute_approximation (DiameterApproximation.h:54-126)
//   DijkstraQueueType / DeltaSteppingEpochType / BetweennessCentralityAlgorithm (source/Types.h:36-60)
// The rank shape of the reference is kept: a RankSlice (prefix distribution + own rank, AllPairsShortestPath.h:
// 113-117) selects the sources a rank runs. Dijkstra APSP and Brandes run the rank's own nodes and leave a partial
// that the MPI reductions combine (reduce_sum / reduce_max / reduce_map, AllPairsShortestPath.h:145-162;
// reduce_componentwise_sum, BetweennessCentrality.h:74), here ApspPartial::reduce; delta-stepping APSP runs all
// N global sources on every rank and needs no reduction (AllPairsShortestPath.h:197-247).
// Errors throw ng_host::Exception, as the reference throws utility::Exception (Exception.hpp:76).
// The graph is a device-resident ng_graph built once from host CSR (the loader of neurograph_io.h or
// caller arrays), replacing DistributedGraph's RMA windows on one GPU.
#pragma once

#include <algorithm>
#include <cstdint>
#include <functional>
#include <memory>
#include <numeric>
#include <span>
#include <stdexcept>
#include <type_traits>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/neurograph_hip.h"
#include "../../include/neurograph_io.h"

namespace ng_host {

class Exception : public std::runtime_error {
public:
    using std::runtime_error::runtime_error;
};

inline void check(ng_status st, const char* what) {
    if (st != NG_OK) throw Exception(std::string(what) + ": " + ng_last_error());
}

// source/Types.h:47-51 plus the GPU variant. Every value yields identical distances; the CPU queue
// kinds are accepted and served by the GPU path, except that BucketQueue keeps the reference's
// refusal of the inverse metric (Dijkstra.h:89-91).
enum class DijkstraQueueType : std::uint8_t { PriorityQueue, FibonacciHeap, BucketQueue, Gpu };

// source/Types.h:36-42 plus Gpu. All epoch strategies compute the same distances
// (tests/apsp/test_apsp_delta_stepping.cpp), so each is served by the batched delta-stepping kernel.
enum class DeltaSteppingEpochType : std::uint8_t { PushBruteForce, PushCheck, PushShortLong, PushInnerShort, PullModel, Gpu };

// source/Types.h:57-60 plus Gpu.
enum class BetweennessCentralityAlgorithm : std::uint8_t { Brandes, PathEnumeration, Gpu };

// ApspGlobalResult<d_type>, source/Types.h:237-244.
template <typename d_type>
struct ApspGlobalResult {
    double average_shortest_path_length{};
    double average_efficiency{};
    d_type diameter{};
    std::uint64_t number_disconnected_pairs{};
    std::unordered_map<std::uint64_t, std::uint64_t> cluster_sizes{};
};

class Context {
public:
    explicit Context(int device = 0) { check(ng_context_create(device, &ctx_), "ng_context_create"); }
    ~Context() { ng_context_destroy(ctx_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    ng_context* get() const { return ctx_; }

private:
    ng_context* ctx_ = nullptr;
};

// Device-resident graph (the one-GPU replacement of DistributedGraph).
class Graph {
public:
    Graph(const Context& ctx, const ng_host_csr& csr) : n_(csr.n), m_(csr.m) {
        check(ng_graph_create(ctx.get(), csr.n, csr.m, csr.offsets, csr.targets, csr.weights, csr.wkind, &g_), "ng_graph_create");
    }
    ~Graph() { ng_graph_destroy(g_); }
    Graph(const Graph&) = delete;
    Graph& operator=(const Graph&) = delete;
    ng_graph* get() const { return g_; }
    std::uint32_t number_nodes() const { return n_; }
    std::uint64_t number_arcs() const { return m_; }
    std::vector<std::uint32_t> all_sources() const {
        std::vector<std::uint32_t> s(n_);
        std::iota(s.begin(), s.end(), 0u);
        return s;
    }

private:
    ng_graph* g_ = nullptr;
    std::uint32_t n_ = 0;
    std::uint64_t m_ = 0;
};

// Which global sources one rank runs: prefix[P+1] is the exclusive prefix sum of the ranks' node counts
// (AllPairsShortestPath.h:113-117), global id = prefix[rank] + local id. One rank holding everything is {0, n}.
struct RankSlice {
    std::vector<std::uint64_t> prefix;
    std::uint32_t rank = 0;
    static RankSlice single(std::uint64_t n) { return RankSlice{{0, n}, 0}; }
    std::uint64_t total() const { return prefix.back(); }
    // this rank's own nodes: the sources of compute_apsp_impl's loop (:130) and of BrandesVariant
    std::vector<std::uint32_t> local_sources() const {
        if (rank + 1 >= prefix.size()) throw Exception("rank outside the prefix distribution");
        std::vector<std::uint32_t> s(prefix[rank + 1] - prefix[rank]);
        std::iota(s.begin(), s.end(), static_cast<std::uint32_t>(prefix[rank]));
        return s;
    }
    // all global sources: the loop of compute_apsp_impl_delta_stepping (:220), identical on every rank
    std::vector<std::uint32_t> all_sources() const {
        std::vector<std::uint32_t> s(total());
        std::iota(s.begin(), s.end(), 0u);
        return s;
    }
};

// One rank's running sums of compute_apsp_impl (AllPairsShortestPath.h:121-143), before the reductions.
template <typename d_type>
struct ApspPartial {
    d_type sum{};
    double efficiency = 0.0;
    d_type maximum{};
    std::uint64_t unreachable = 0;
    std::unordered_map<std::uint64_t, std::uint64_t> cluster_sizes{};

    void add(const ng_sssp_summary& s, std::uint64_t total_nodes) {
        if constexpr (std::is_same_v<d_type, std::uint64_t>) {
            sum += s.sum_d_u64;
            maximum = std::max(maximum, s.max_d_u64);
        } else {
            sum += s.sum_d;
            maximum = std::max(maximum, s.max_d);
        }
        efficiency += s.sum_eff;
        unreachable += s.n_unreached;
        ++cluster_sizes[total_nodes - s.n_unreached];
    }
    // reduce_sum / reduce_max / reduce_map over the ranks, then the averages (AllPairsShortestPath.h:145-162)
    static ApspGlobalResult<d_type> reduce(const std::vector<ApspPartial>& ranks, std::uint64_t total_nodes) {
        ApspPartial all;
        for (const auto& p : ranks) {
            all.sum += p.sum;
            all.efficiency += p.efficiency;
            all.maximum = std::max(all.maximum, p.maximum);
            all.unreachable += p.unreachable;
            for (const auto& [k, v] : p.cluster_sizes) all.cluster_sizes[k] += v;
        }
        return all.finish(total_nodes);
    }
    // the averages over reached ordered pairs, without any reduction (compute_apsp_impl_delta_stepping, :238-246)
    ApspGlobalResult<d_type> finish(std::uint64_t total_nodes) const {
        ApspGlobalResult<d_type> r;
        const std::uint64_t reached = total_nodes * total_nodes - total_nodes - unreachable;
        r.average_shortest_path_length = static_cast<double>(sum) / static_cast<double>(reached);
        r.average_efficiency = efficiency / static_cast<double>(reached);
        r.diameter = maximum;
        r.number_disconnected_pairs = unreachable;
        r.cluster_sizes = cluster_sizes;
        return r;
    }
};

class AllPairsShortestPath {
public:
    // AllPairsShortestPath.h:54-57 (distance_type = uint64 for integer weights; real-valued weights give
    // double-valued results through compute_apsp_real). One rank holding every node.
    static ApspGlobalResult<std::uint64_t> compute_apsp(const Graph& graph, DijkstraQueueType queue = DijkstraQueueType::Gpu) {
        return ApspPartial<std::uint64_t>::reduce({compute_apsp_partial(graph, RankSlice::single(graph.number_nodes()), queue)},
                                                  graph.number_nodes());
    }
    // This rank's partial of compute_apsp_impl: its own sources, in one batched call (the loop :130-143).
    static ApspPartial<std::uint64_t> compute_apsp_partial(const Graph& graph, const RankSlice& ranks,
                                                           DijkstraQueueType queue = DijkstraQueueType::Gpu) {
        (void)queue;
        return accumulate<std::uint64_t>(run(graph, ranks.local_sources(), NG_LEN_ABS, 0.0), ranks.total());
    }
    // AllPairsShortestPath.h:76-79: lengths 1/|w|; BucketQueue is refused (Dijkstra.h:89-91).
    static ApspGlobalResult<double> compute_apsp_inverse(const Graph& graph, DijkstraQueueType queue = DijkstraQueueType::Gpu) {
        if (queue == DijkstraQueueType::BucketQueue) {
            throw Exception("Dijkstra::compute_sssp: the bucket queue requires integer distances and cannot be used with the inverse metric");
        }
        const auto one = RankSlice::single(graph.number_nodes());
        return ApspPartial<double>::reduce({accumulate<double>(run(graph, one.local_sources(), NG_LEN_INVERSE, 0.0), one.total())},
                                           graph.number_nodes());
    }
    // AllPairsShortestPath.h:95-98 and :197-247: raw weights (negative weights throw), bucket width delta. Every
    // rank runs all N global sources (DeltaStepping::compute_shortest_distances is collective per source and
    // returns globally reduced results), so the result is complete on every rank and is not reduced again.
    static ApspGlobalResult<std::uint64_t> compute_apsp_delta_stepping(const Graph& graph, std::uint64_t delta,
                                                                      DeltaSteppingEpochType epoch = DeltaSteppingEpochType::Gpu,
                                                                      const RankSlice* ranks = nullptr) {
        (void)epoch;
        if (delta == 0) throw Exception("delta-stepping needs a positive delta");
        const RankSlice one = RankSlice::single(graph.number_nodes());
        const RankSlice& rs = ranks ? *ranks : one;
        return accumulate<std::uint64_t>(run(graph, rs.all_sources(), NG_LEN_RAW, static_cast<double>(delta)), rs.total())
            .finish(rs.total());
    }
    // The same aggregation for real-valued weights (an extension of this library).
    static ApspGlobalResult<double> compute_apsp_real(const Graph& graph, double delta = 0.0) {
        const auto one = RankSlice::single(graph.number_nodes());
        return ApspPartial<double>::reduce({accumulate<double>(run(graph, one.local_sources(), NG_LEN_ABS, delta), one.total())},
                                           graph.number_nodes());
    }

    static std::vector<ng_sssp_summary> run(const Graph& graph, const std::vector<std::uint32_t>& sources, ng_length_kind lkind,
                                            double delta) {
        std::vector<ng_sssp_summary> out(sources.size());
        check(ng_sssp_summaries(graph.get(), sources.data(), static_cast<std::uint32_t>(sources.size()), lkind, delta, out.data(),
                                nullptr),
              "ng_sssp_summaries");
        return out;
    }

    template <typename d_type>
    static ApspPartial<d_type> accumulate(const std::vector<ng_sssp_summary>& per, std::uint64_t total_nodes) {
        ApspPartial<d_type> p;
        for (const auto& s : per) p.add(s, total_nodes);
        return p;
    }
    // compute_apsp_impl aggregation on one rank (AllPairsShortestPath.h:130-162).
    static ApspGlobalResult<std::uint64_t> aggregate_integer(const std::vector<ng_sssp_summary>& per, std::uint64_t n) {
        return ApspPartial<std::uint64_t>::reduce({accumulate<std::uint64_t>(per, n)}, n);
    }
    static ApspGlobalResult<double> aggregate_real(const std::vector<ng_sssp_summary>& per, std::uint64_t n) {
        return ApspPartial<double>::reduce({accumulate<double>(per, n)}, n);
    }
};

// BrandesVariant::accumulate_betweenness (BrandesVariant.h:50-74): this rank's own sources in one batched call,
// accumulated into local_betweenness (N entries, like the reference's), which reduce_componentwise_sum combines
// across ranks (BetweennessCentrality.h:74).
class GpuBrandesVariant {
public:
    static void accumulate_betweenness(const Graph& graph, std::vector<double>& local_betweenness, double delta = 0.0,
                                       const RankSlice* ranks = nullptr) {
        if (local_betweenness.size() != graph.number_nodes()) throw Exception("betweenness vector has the wrong size");
        const auto sources = ranks ? ranks->local_sources() : graph.all_sources();
        check(ng_brandes_accumulate(graph.get(), sources.data(), static_cast<std::uint32_t>(sources.size()), delta,
                                    local_betweenness.data(), nullptr),
              "ng_brandes_accumulate");
    }
};

class BetweennessCentrality {
public:
    // BetweennessCentrality.h:46-80: every algorithm yields the same values; all are served by the GPU variant.
    static double compute_average_betweenness_centrality(const Graph& graph,
                                                         BetweennessCentralityAlgorithm algorithm = BetweennessCentralityAlgorithm::Gpu,
                                                         std::vector<double>* per_node = nullptr) {
        (void)algorithm;
        std::vector<double> bc(graph.number_nodes(), 0.0);
        GpuBrandesVariant::accumulate_betweenness(graph, bc);
        double total = 0.0;
        for (double x : bc) total += x;
        if (per_node) *per_node = bc;
        return total / static_cast<double>(graph.number_nodes());
    }
};

// RandomNumberGenerator (utility/RandomNumberGenerator.h): numbers in [0, 1) for the centers' draws.
class RandomNumberGenerator {
public:
    virtual ~RandomNumberGenerator() = default;
    virtual double draw() = 0;
};

class DiameterApproximation {
public:
    // DiameterApproximation.h:54-58: the default generator seeded with the own rank (0 on one rank).
    static std::uint64_t compute_approximation(const Context& ctx, const ng_host_csr& graph, std::uint64_t r, std::uint32_t rank = 0) {
        std::uint64_t d = 0;
        check(ng_diameter_approximation(ctx.get(), &graph, r, nullptr, nullptr, rank, &d, nullptr), "ng_diameter_approximation");
        return d;
    }
    // DiameterApproximation.h:71-126 with the caller's generator; the contracted graph's APSP (:118) runs on the GPU.
    static std::uint64_t compute_approximation(const Context& ctx, const ng_host_csr& graph, std::uint64_t r,
                                               RandomNumberGenerator& generator) {
        std::uint64_t d = 0;
        check(ng_diameter_approximation(ctx.get(), &graph, r, &draw_from, &generator, 0, &d, nullptr), "ng_diameter_approximation");
        return d;
    }

private:
    static double draw_from(void* generator) { return static_cast<RandomNumberGenerator*>(generator)->draw(); }
};

}  // namespace ng_host
