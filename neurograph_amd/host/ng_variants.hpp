// C++20 host mirror of neurograph's APSP / betweenness drivers over the MI355X path (libng_hip.so).
//
// Same names, argument meaning and error behaviour as the reference's metric classes, with the GPU
// variant added to each selector enum:
//   AllPairsShortestPath::compute_apsp / compute_apsp_inverse / compute_apsp_delta_stepping
//       (source/metrics/paths/AllPairsShortestPath.h:54-98)
//   BetweennessCentrality::compute_average_betweenness_centrality (BetweennessCentrality.h:46-80)
//   DijkstraQueueType / DeltaSteppingEpochType / BetweennessCentralityAlgorithm (source/Types.h:36-60)
// Errors throw ng_host::Exception, as the reference throws utility::Exception (Exception.hpp:76).
// The graph is a device-resident ng_graph built once from host CSR (the loader of neurograph_io.h or
// caller arrays), replacing DistributedGraph's RMA windows on one GPU.
#pragma once

#include <algorithm>
#include <cstdint>
#include <memory>
#include <numeric>
#include <stdexcept>
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

// source/Types.h:36-42 plus GpuMultiSource. All epoch strategies compute the same distances
// (tests/apsp/test_apsp_delta_stepping.cpp), so each is served by the batched delta-stepping kernel.
enum class DeltaSteppingEpochType : std::uint8_t { PushBruteForce, PushCheck, PushShortLong, PushInnerShort, PullModel, GpuMultiSource };

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

class AllPairsShortestPath {
public:
    // AllPairsShortestPath.h:54-57 (distance_type = uint64 for integer weights; real-valued weights give
    // double-valued results through compute_apsp_real).
    static ApspGlobalResult<std::uint64_t> compute_apsp(const Graph& graph, DijkstraQueueType queue = DijkstraQueueType::Gpu) {
        (void)queue;
        return aggregate_integer(run(graph, NG_LEN_ABS, 0.0), graph.number_nodes());
    }
    // AllPairsShortestPath.h:76-79: lengths 1/|w|; BucketQueue is refused (Dijkstra.h:89-91).
    static ApspGlobalResult<double> compute_apsp_inverse(const Graph& graph, DijkstraQueueType queue = DijkstraQueueType::Gpu) {
        if (queue == DijkstraQueueType::BucketQueue) {
            throw Exception("Dijkstra::compute_sssp: the bucket queue requires integer distances and cannot be used with the inverse metric");
        }
        return aggregate_real(run(graph, NG_LEN_INVERSE, 0.0), graph.number_nodes());
    }
    // AllPairsShortestPath.h:95-98: raw weights (negative weights throw), bucket width delta.
    static ApspGlobalResult<std::uint64_t> compute_apsp_delta_stepping(const Graph& graph, std::uint64_t delta,
                                                                      DeltaSteppingEpochType epoch = DeltaSteppingEpochType::GpuMultiSource) {
        (void)epoch;
        if (delta == 0) throw Exception("delta-stepping needs a positive delta");
        return aggregate_integer(run(graph, NG_LEN_RAW, static_cast<double>(delta)), graph.number_nodes());
    }
    // The same aggregation for real-valued weights (an extension of this library).
    static ApspGlobalResult<double> compute_apsp_real(const Graph& graph, double delta = 0.0) {
        return aggregate_real(run(graph, NG_LEN_ABS, delta), graph.number_nodes());
    }

    static std::vector<ng_sssp_summary> run(const Graph& graph, ng_length_kind lkind, double delta) {
        const auto sources = graph.all_sources();
        std::vector<ng_sssp_summary> out(sources.size());
        check(ng_sssp_summaries(graph.get(), sources.data(), static_cast<std::uint32_t>(sources.size()), lkind, delta, out.data(),
                                nullptr),
              "ng_sssp_summaries");
        return out;
    }

    // compute_apsp_impl aggregation, AllPairsShortestPath.h:130-162.
    static ApspGlobalResult<std::uint64_t> aggregate_integer(const std::vector<ng_sssp_summary>& per, std::uint64_t n) {
        ApspGlobalResult<std::uint64_t> r;
        std::uint64_t sum = 0, unreach = 0;
        double eff = 0.0;
        for (const auto& s : per) {
            sum += s.sum_d_u64;
            eff += s.sum_eff;
            unreach += s.n_unreached;
            r.diameter = std::max(r.diameter, s.max_d_u64);
            ++r.cluster_sizes[n - s.n_unreached];
        }
        const std::uint64_t reached = n * n - n - unreach;
        r.average_shortest_path_length = static_cast<double>(sum) / static_cast<double>(reached);
        r.average_efficiency = eff / static_cast<double>(reached);
        r.number_disconnected_pairs = unreach;
        return r;
    }
    static ApspGlobalResult<double> aggregate_real(const std::vector<ng_sssp_summary>& per, std::uint64_t n) {
        ApspGlobalResult<double> r;
        double sum = 0.0, eff = 0.0;
        std::uint64_t unreach = 0;
        for (const auto& s : per) {
            sum += s.sum_d;
            eff += s.sum_eff;
            unreach += s.n_unreached;
            r.diameter = std::max(r.diameter, s.max_d);
            ++r.cluster_sizes[n - s.n_unreached];
        }
        const std::uint64_t reached = n * n - n - unreach;
        r.average_shortest_path_length = sum / static_cast<double>(reached);
        r.average_efficiency = eff / static_cast<double>(reached);
        r.number_disconnected_pairs = unreach;
        return r;
    }
};

// BrandesVariant::accumulate_betweenness (BrandesVariant.h:50-74) over all sources in one batched call.
class GpuBrandesVariant {
public:
    static void accumulate_betweenness(const Graph& graph, std::vector<double>& local_betweenness, double delta = 0.0) {
        if (local_betweenness.size() != graph.number_nodes()) throw Exception("betweenness vector has the wrong size");
        const auto sources = graph.all_sources();
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

}  // namespace ng_host
