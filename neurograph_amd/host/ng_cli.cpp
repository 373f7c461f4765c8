// Command-line front end of the MI355X path, with the reference's flags for the path metrics
// (source/neurograph.cpp:449-778) and MetricStore-compatible output files
// (source/utility/MetricStore.cpp:284-373), written to <input>/metrics[/<output-directory>].
//
//   ng_cli --input DIR --func avg-apsp,avg-apsp-inv,avg-centrality,dia-approx [--apsp-algorithm dijkstra|delta-stepping]
//          [--apsp-delta D] [--apsp-delta-epoch push-check|...|gpu] [--dijkstra-queue priority-queue|...|gpu]
//          [--centrality-algorithm brandes|path-enumeration|gpu] [--undirected] [--all-weights-one]
//          [--remove-autapses] [--prefix P] [--output-directory D] [--threads N]
//          [--weights int|f32|f64] [--device N] [--node-betweenness] [--distance-matrix]
//          [--distance-histogram-width W] [--distance-histogram-bins B] [--csr-cache FILE]
//
// Every selector value runs the GPU variant (they compute identical values); the reference's refusal of
// the bucket queue with avg-apsp-inv is kept (neurograph.cpp:690-693). --weights f32|f64 reads real-valued
// weights (extension); --node-betweenness also writes per-node values (extension). With avg-apsp,
// --distance-matrix writes all_pairs_distances.bin (row s = distances from node s, little-endian u64 for integer
// lengths, f32 for f32 weights, f64 otherwise; unreachable = the type's max) and --distance-histogram-width W
// writes all_pairs_distance_histogram.txt in the reference's histogram text format (MetricStore.cpp:32-40).
// dia-approx runs DiameterApproximation::compute_approximation(graph, 1) as neurograph.cpp:386-391 does (radius 1,
// the default generator seeded with the rank, 0 here), the contracted graph's APSP on the GPU, and writes
// approximate_diameter.txt (MetricStore.cpp:529-543). --csr-cache FILE reads the graph from a binary CSR cache
// written for the same input and load options, or parses the text files and writes the cache (extension).

#include <algorithm>
#include <charconv>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <limits>
#include <map>
#include <numeric>
#include <sstream>
#include <string>
#include <vector>

#include "ng_variants.hpp"

namespace fs = std::filesystem;

namespace {

// Shortest round-trip text of a double, like {fmt}'s "{}".
std::string num(double v) {
    char buf[64];
    auto r = std::to_chars(buf, buf + sizeof buf, v);
    return std::string(buf, r.ptr);
}
std::string num(std::uint64_t v) { return std::to_string(v); }

// FNV-1a over the text that identifies an input and its load options: the key of its CSR cache.
std::uint64_t cache_key(const std::string& text) {
    std::uint64_t h = 0xcbf29ce484222325ull;
    for (unsigned char c : text) {
        h ^= c;
        h *= 0x100000001b3ull;
    }
    return h;
}

std::vector<std::string> split(const std::string& s) {
    std::vector<std::string> out;
    std::stringstream ss(s);
    for (std::string item; std::getline(ss, item, ',');)
        if (!item.empty()) out.push_back(item);
    return out;
}

template <typename R>
std::string apsp_text(const R& r) {
    std::map<std::uint64_t, std::uint64_t> sorted(r.cluster_sizes.begin(), r.cluster_sizes.end());
    std::ostringstream ss;
    for (const auto& [size, occ] : sorted) ss << "Size: " << size << "\tOccurences: " << occ << "\n";
    ss << "The average shortest path is: " << num(r.average_shortest_path_length) << "\n";
    ss << "The global efficiency is: " << num(r.average_efficiency) << "\n";
    ss << "The diameter of the graph is: " << num(r.diameter) << "\n";
    ss << "However, the number of pairs without path is: " << r.number_disconnected_pairs << "\n";
    ss << "The clusters were distributed as follows:\n";
    for (const auto& [size, occ] : sorted) ss << "Size: " << size << "\tOccurences: " << occ << "\n";
    return ss.str();
}

void emit(const fs::path& dir, const std::string& file, const std::string& text) {
    std::cout << text;
    std::ofstream(dir / file) << text;
}

}  // namespace

int main(int argc, char** argv) {
    std::map<std::string, std::string> opt;
    std::vector<std::string> flags;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "--undirected" || a == "--all-weights-one" || a == "--remove-autapses" || a == "--node-betweenness" ||
            a == "--distance-matrix" ||
            a == "--disable-status-report" || a == "--disable-timer-report") {
            flags.push_back(a);
        } else if (a.rfind("--", 0) == 0 && i + 1 < argc) {
            opt[a] = argv[++i];
        } else {
            std::cerr << "unknown argument " << a << "\n";
            return EXIT_FAILURE;
        }
    }
    auto has = [&](const char* f) { return std::find(flags.begin(), flags.end(), f) != flags.end(); };
    auto get = [&](const char* k, const char* d) { return opt.count(k) ? opt[k] : std::string(d); };
    if (!opt.count("--input")) {
        std::cerr << "--input is required\n";
        return EXIT_FAILURE;
    }
    const fs::path input = opt["--input"];
    bool apsp = false, apsp_inv = false, centrality = false, dia_approx = false;
    for (const auto& f : split(get("--func", ""))) {
        if (f == "avg-apsp") apsp = true;
        else if (f == "avg-apsp-inv") apsp_inv = true;
        else if (f == "avg-centrality") centrality = true;
        else if (f == "dia-approx") dia_approx = true;
        else {
            std::cerr << "The argument \"" << f
                      << "\" is not a valid algorithm of this implementation (avg-apsp, avg-apsp-inv, avg-centrality, dia-approx)\n";
            return EXIT_FAILURE;
        }
    }
    const std::string algo = get("--apsp-algorithm", "dijkstra");
    const std::string queue = get("--dijkstra-queue", "gpu");
    const std::string epoch = get("--apsp-delta-epoch", "gpu");
    const std::string calgo = get("--centrality-algorithm", "gpu");
    if (algo != "dijkstra" && algo != "delta-stepping") {
        std::cerr << "The argument \"" << algo << "\" is not a valid --apsp-algorithm\n";
        return EXIT_FAILURE;
    }
    for (const auto& [val, valid] : std::vector<std::pair<std::string, std::vector<std::string>>>{
             {queue, {"priority-queue", "fibonacci-heap", "bucket-queue", "gpu"}},
             {epoch, {"push-brute-force", "push-check", "push-short-long", "push-inner-short", "pull-model", "gpu"}},
             {calgo, {"brandes", "path-enumeration", "gpu"}}}) {
        if (std::find(valid.begin(), valid.end(), val) == valid.end()) {
            std::cerr << "The argument \"" << val << "\" is not a valid selector\n";
            return EXIT_FAILURE;
        }
    }
    if (apsp_inv && queue == "bucket-queue") {
        std::cerr << "--dijkstra-queue bucket-queue cannot be used with avg-apsp-inv: the inverse distances are not integers\n";
        return EXIT_FAILURE;
    }
    const std::string wk = get("--weights", "int");
    const ng_weight_kind wkind = wk == "f64" ? NG_W_F64 : (wk == "f32" ? NG_W_F32 : NG_W_I32);
    fs::path out_dir = input / "metrics";
    if (opt.count("--output-directory")) out_dir /= opt["--output-directory"];

    if (dia_approx && wkind != NG_W_I32) {
        std::cerr << "dia-approx takes integer weights (--weights int)\n";
        return EXIT_FAILURE;
    }

    ng_host_csr csr{};
    try {
        const std::string prefix = get("--prefix", "");
        bool from_cache = false;
        std::uint64_t key = 0;
        if (opt.count("--csr-cache")) {
            std::error_code ec;
            const fs::path canon = fs::weakly_canonical(input, ec);
            key = cache_key((ec ? input : canon).string() + '\n' + prefix + '\n' + std::to_string(has("--remove-autapses")) +
                            std::to_string(has("--undirected")) + std::to_string(has("--all-weights-one")) + '\n' + wk);
            if (fs::exists(opt["--csr-cache"])) {
                const ng_status st = ng_load_csr(opt["--csr-cache"].c_str(), key, &csr);
                if (st == NG_OK) from_cache = true;
                else std::cerr << "ignoring the CSR cache: " << ng_last_error() << "\n";
            }
        }
        if (!from_cache) {
            ng_host::check(ng_load_graph_dir(input.c_str(), prefix.c_str(), 0, has("--remove-autapses"), has("--undirected"),
                                             has("--all-weights-one"), wkind, &csr),
                           "loading the graph");
            if (opt.count("--csr-cache") && ng_save_csr(opt["--csr-cache"].c_str(), &csr, key) != NG_OK) {
                std::cerr << "could not write the CSR cache: " << ng_last_error() << "\n";
            }
        }
        ng_host::Context ctx(std::stoi(get("--device", "0")));
        ng_host::Graph graph(ctx, csr);
        std::cout << "Loaded " << graph.number_nodes() << " nodes and " << graph.number_arcs() << " arcs"
                  << (from_cache ? " from the CSR cache" : "") << "\n";
        fs::create_directories(out_dir);
        if (apsp) {
            if (wkind == NG_W_I32) {
                const auto r = algo == "delta-stepping"
                                   ? ng_host::AllPairsShortestPath::compute_apsp_delta_stepping(graph, std::stoull(get("--apsp-delta", "1")))
                                   : ng_host::AllPairsShortestPath::compute_apsp(graph);
                emit(out_dir, "all_pairs_shortest_paths.txt", apsp_text(r));
            } else {
                emit(out_dir, "all_pairs_shortest_paths.txt", apsp_text(ng_host::AllPairsShortestPath::compute_apsp_real(graph)));
            }
        }
        if (apsp && has("--distance-matrix")) {
            // rows in chunks of at most 256 MiB
            const ng_length_kind lk = algo == "delta-stepping" ? NG_LEN_RAW : NG_LEN_ABS;
            const double dd = algo == "delta-stepping" ? std::stod(get("--apsp-delta", "1")) : 0.0;
            ng_dist_kind dkind{};
            ng_host::check(ng_graph_dist_kind(graph.get(), lk, &dkind), "ng_graph_dist_kind");
            const size_t eb = dkind == NG_DIST_F32 ? 4 : 8;
            const uint32_t n = graph.number_nodes();
            const uint32_t chunk = static_cast<uint32_t>(std::max<size_t>(1, (256ull << 20) / (eb * std::max<uint32_t>(n, 1))));
            std::ofstream f(out_dir / "all_pairs_distances.bin", std::ios::binary);
            std::vector<char> rows(static_cast<size_t>(chunk) * n * eb);
            std::vector<ng_sssp_summary> summ(chunk);
            for (uint32_t b = 0; b < n; b += chunk) {
                const uint32_t k = std::min(chunk, n - b);
                std::vector<uint32_t> src(k);
                std::iota(src.begin(), src.end(), b);
                ng_host::check(ng_sssp_summaries(graph.get(), src.data(), k, lk, dd, summ.data(), rows.data()), "ng_sssp_summaries");
                f.write(rows.data(), static_cast<std::streamsize>(static_cast<size_t>(k) * n * eb));
            }
            std::cout << "Wrote the " << n << " x " << n << " distance matrix ("
                      << (dkind == NG_DIST_U64 ? "u64" : dkind == NG_DIST_F32 ? "f32" : "f64") << ")\n";
        }
        if (apsp && opt.count("--distance-histogram-width")) {
            const ng_length_kind lk = algo == "delta-stepping" ? NG_LEN_RAW : NG_LEN_ABS;
            const double dd = algo == "delta-stepping" ? std::stod(get("--apsp-delta", "1")) : 0.0;
            const double width = std::stod(opt["--distance-histogram-width"]);
            const uint32_t bins = static_cast<uint32_t>(std::stoul(get("--distance-histogram-bins", "256")));
            std::vector<uint64_t> counts(bins);
            const auto sources = graph.all_sources();
            ng_host::check(ng_distance_histogram(graph.get(), sources.data(), static_cast<uint32_t>(sources.size()), lk, dd, width, bins,
                                                 counts.data(), nullptr),
                           "ng_distance_histogram");
            std::ostringstream ss;
            for (uint32_t i = 0; i < bins; ++i) {
                ss << i << ". bin: " << num(width * i) << '-';
                if (i + 1 < bins) ss << num(width * (i + 1));
                else ss << num(std::numeric_limits<double>::max());
                ss << ": " << counts[i] << '\n';
            }
            emit(out_dir, "all_pairs_distance_histogram.txt", ss.str());
        }
        if (apsp_inv) {
            emit(out_dir, "all_pairs_shortest_paths_inverse.txt", apsp_text(ng_host::AllPairsShortestPath::compute_apsp_inverse(graph)));
        }
        if (centrality) {
            std::vector<double> per_node;
            const double avg = ng_host::BetweennessCentrality::compute_average_betweenness_centrality(
                graph, ng_host::BetweennessCentralityAlgorithm::Gpu, &per_node);
            emit(out_dir, "average_betweenness_centrality.txt", "The average betweenness centrality is: " + num(avg) + "\n");
            if (has("--node-betweenness")) {
                std::ofstream f(out_dir / "node_betweenness_centrality.txt");
                for (std::size_t v = 0; v < per_node.size(); ++v) f << (v + 1) << "\t" << num(per_node[v]) << "\n";
            }
        }
        if (dia_approx) {
            const auto d = ng_host::DiameterApproximation::compute_approximation(ctx, csr, 1);
            emit(out_dir, "approximate_diameter.txt", "The approximate diameter is: " + num(d) + "\n");
        }
    } catch (const std::exception& e) {
        ng_host_csr_free(&csr);
        std::cerr << e.what() << "\n";
        return EXIT_FAILURE;
    }
    ng_host_csr_free(&csr);
    return EXIT_SUCCESS;
}
