// Host-side graph I/O: neurograph's on-disk format with its load-time transforms, and deterministic
// generators for the benchmark graphs. See include/neurograph_io.h.

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <sstream>
#include <string>
#include <thread>
#include <unordered_set>
#include <utility>
#include <vector>

#include "../../include/neurograph_io.h"

namespace fs = std::filesystem;

// Errors are reported through ng_last_error(); the API layer (ng_api.cpp) owns that thread-local buffer.
extern "C" void ng_set_error_message(const char* msg);

namespace {

ng_status io_fail(ng_status st, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    ng_set_error_message(buf);
    return st;
}

// ---------------------------------------------------------------------------------------------------
// deterministic random streams: splitmix64 seeding, xoshiro256** generator

uint64_t splitmix64(uint64_t& x) {
    uint64_t z = (x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

struct Xoshiro {
    uint64_t s[4];
    Xoshiro(uint64_t seed, uint64_t stream) {
        uint64_t x = seed ^ (stream * 0xd1b54a32d192ed03ull + 0x8cb92ba72f3d8dd7ull);
        for (auto& v : s) v = splitmix64(x);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }  // [0, 1)
    uint64_t below(uint64_t bound) {  // unbiased integer in [0, bound)
        const uint64_t lim = (~0ull) - ((~0ull) % bound);
        uint64_t r;
        do r = next(); while (r >= lim);
        return r % bound;
    }
    double normal() {  // Box-Muller
        double u1;
        do u1 = uniform(); while (u1 <= 0.0);
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
    uint32_t poisson(double mean) {  // inversion by sequential search
        if (mean <= 0.0) return 0;
        if (mean > 500.0) {
            const double x = std::round(mean + std::sqrt(mean) * normal());
            return x < 0 ? 0u : static_cast<uint32_t>(x);
        }
        const double L = std::exp(-mean);
        double p = L, cdf = L;
        const double u = uniform();
        uint32_t k = 0;
        while (u > cdf && k < 100000) {
            ++k;
            p *= mean / k;
            cdf += p;
        }
        return k;
    }
};

template <class Fn>
void parallel_for(uint64_t count, Fn&& fn) {
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (count < 1024) nt = 1;
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nt; ++t) {
        const uint64_t b = count * t / nt, e = count * (t + 1) / nt;
        if (nt == 1) fn(b, e);
        else pool.emplace_back([b, e, &fn] { fn(b, e); });
    }
    for (auto& th : pool) th.join();
}

size_t wsize(ng_weight_kind k) { return k == NG_W_F64 ? 8 : 4; }

struct Arc {
    uint32_t s, t;
    double w;
};

// Fills `out` from arcs sorted by (s, t) with unique keys.
ng_status to_csr(uint32_t n, const std::vector<Arc>& arcs, ng_weight_kind wk, ng_host_csr* out) {
    std::memset(out, 0, sizeof *out);
    out->n = n;
    out->m = arcs.size();
    out->wkind = wk;
    out->offsets = static_cast<uint64_t*>(std::calloc(static_cast<size_t>(n) + 1, sizeof(uint64_t)));
    out->targets = static_cast<uint32_t*>(std::malloc(std::max<size_t>(arcs.size(), 1) * sizeof(uint32_t)));
    out->weights = std::malloc(std::max<size_t>(arcs.size(), 1) * wsize(wk));
    if (!out->offsets || !out->targets || !out->weights) {
        ng_host_csr_free(out);
        return io_fail(NG_ENOMEM, "host allocation failed");
    }
    for (const Arc& a : arcs) ++out->offsets[a.s + 1];
    for (uint32_t u = 0; u < n; ++u) out->offsets[u + 1] += out->offsets[u];
    for (size_t i = 0; i < arcs.size(); ++i) {
        out->targets[i] = arcs[i].t;
        switch (wk) {
        case NG_W_I32: static_cast<int32_t*>(out->weights)[i] = static_cast<int32_t>(arcs[i].w); break;
        case NG_W_F32: static_cast<float*>(out->weights)[i] = static_cast<float>(arcs[i].w); break;
        default: static_cast<double*>(out->weights)[i] = arcs[i].w; break;
        }
    }
    return NG_OK;
}

std::string rank_prefix(int rank, int n_ranks) {  // FileLoader.cpp:41-49
    const std::string digits_all = std::to_string(n_ranks - 1);
    const std::string mine = std::to_string(rank);
    return "rank_" + std::string(digits_all.size() > mine.size() ? digits_all.size() - mine.size() : 0, '0') + mine;
}

// FileLoader::load_nodes (FileLoader.cpp:96-177): node count and positions of one rank file.
ng_status load_nodes(const fs::path& file, std::vector<double>& pos) {
    std::ifstream in(file);
    if (!in.good()) return io_fail(NG_EINVAL, "file is not good: %s", file.c_str());
    for (std::string line; std::getline(in, line);) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream iss(line);
        uint32_t id = 0;
        double x = 0, y = 0, z = 0;
        std::string signal, area;
        if (!((iss >> id) && (iss >> x) && (iss >> y) && (iss >> z) && (iss >> signal))) continue;
        if (id != pos.size() / 3 + 1) continue;  // expected consecutive 1-based ids (:148-151)
        pos.push_back(x);
        pos.push_back(y);
        pos.push_back(z);
    }
    return NG_OK;
}

// One network file: appends (key node, other node, weight) with the reference's filters.
// in_file: lines are "target_rank target_id source_rank source_id weight [plastic]" for both files.
ng_status load_arcs(const fs::path& file, int my_rank, bool in_file, const std::vector<uint32_t>& counts,
                    const std::vector<uint64_t>& prefix, ng_weight_kind wk, std::vector<Arc>& arcs) {
    std::ifstream in(file);
    if (!in.good()) return io_fail(NG_EINVAL, "file is not good: %s", file.c_str());
    const int n_ranks = static_cast<int>(counts.size());
    for (std::string line; std::getline(in, line);) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream iss(line);
        int t_rank = 0, s_rank = 0;
        uint32_t t_id = 0, s_id = 0;
        double w = 0.0;
        bool ok;
        if (wk == NG_W_I32) {
            int32_t wi = 0;  // FileLoader.cpp:204-206: istream >> int32
            ok = (iss >> t_rank) && (iss >> t_id) && (iss >> s_rank) && (iss >> s_id) && (iss >> wi);
            w = wi;
        } else {
            ok = (iss >> t_rank) && (iss >> t_id) && (iss >> s_rank) && (iss >> s_id) && (iss >> w);
        }
        if (!ok) continue;
        int plastic = 0;
        if ((iss >> plastic) && plastic == 0) continue;  // :212-215
        if (w == 0.0) continue;                           // :217-220
        const int own_rank = in_file ? t_rank : s_rank;
        if (own_rank != my_rank) continue;                // :222-225 / :297-300
        --t_id;                                           // 1-based ids (:227-229)
        --s_id;
        const int other_rank = in_file ? s_rank : t_rank;
        const uint32_t own_id = in_file ? t_id : s_id;
        const uint32_t other_id = in_file ? s_id : t_id;
        if (own_id >= counts[my_rank]) continue;
        if (other_rank < 0 || other_rank >= n_ranks) continue;
        if (other_id >= counts[other_rank]) continue;
        const uint32_t key = static_cast<uint32_t>(prefix[my_rank] + own_id);
        const uint32_t other = static_cast<uint32_t>(prefix[other_rank] + other_id);
        arcs.push_back({key, other, w});
    }
    return NG_OK;
}

// Sorts by (key, other) and merges duplicates by summing (LoadedArcs `+= weight`, FileLoader.cpp:248, 323).
void merge(std::vector<Arc>& a, bool integer) {
    std::stable_sort(a.begin(), a.end(), [](const Arc& x, const Arc& y) { return x.s != y.s ? x.s < y.s : x.t < y.t; });
    size_t k = 0;
    for (size_t i = 0; i < a.size(); ++i) {
        if (k > 0 && a[k - 1].s == a[i].s && a[k - 1].t == a[i].t) {
            if (integer) {
                a[k - 1].w = static_cast<double>(static_cast<int32_t>(static_cast<uint32_t>(static_cast<int64_t>(a[k - 1].w)) +
                                                                      static_cast<uint32_t>(static_cast<int64_t>(a[i].w))));
            } else {
                a[k - 1].w += a[i].w;
            }
        } else {
            a[k++] = a[i];
        }
    }
    a.resize(k);
}

}  // namespace

extern "C" {

void ng_host_csr_free(ng_host_csr* g) {
    if (!g) return;
    std::free(g->offsets);
    std::free(g->targets);
    std::free(g->weights);
    std::free(g->positions);
    g->offsets = nullptr;
    g->targets = nullptr;
    g->weights = nullptr;
    g->positions = nullptr;
    g->n = 0;
    g->m = 0;
}

ng_status ng_load_graph_dir(const char* dir, const char* prefix, int n_ranks, int remove_self_arcs, int undirected,
                            int all_weights_one, ng_weight_kind wkind, ng_host_csr* out) {
    if (!dir || !out) return io_fail(NG_EINVAL, "null argument");
    std::memset(out, 0, sizeof *out);
    if (wkind != NG_W_I32 && wkind != NG_W_F32 && wkind != NG_W_F64) return io_fail(NG_EINVAL, "unknown weight kind");
    const fs::path root(dir);
    const std::string pre = prefix ? prefix : "";
    if (n_ranks <= 0) {
        n_ranks = 0;
        std::error_code ec;
        for (const auto& e : fs::directory_iterator(root / "positions", ec)) {
            const std::string name = e.path().filename().string();
            if (name.rfind("rank_", 0) == 0 && name.size() > 14 && name.substr(name.size() - 14) == "_positions.txt") ++n_ranks;
        }
        if (ec || n_ranks == 0) return io_fail(NG_EINVAL, "no positions files under %s", (root / "positions").c_str());
    }
    std::vector<std::vector<double>> pos(n_ranks);
    std::vector<uint32_t> counts(n_ranks);
    for (int r = 0; r < n_ranks; ++r) {
        const fs::path f = root / "positions" / (rank_prefix(r, n_ranks) + "_positions.txt");
        if (!fs::exists(f)) return io_fail(NG_EINVAL, "the positions for rank %d do not exist: %s", r, f.c_str());
        ng_status st = load_nodes(f, pos[r]);
        if (st != NG_OK) return st;
        counts[r] = static_cast<uint32_t>(pos[r].size() / 3);
    }
    std::vector<uint64_t> pfx(n_ranks + 1, 0);
    for (int r = 0; r < n_ranks; ++r) pfx[r + 1] = pfx[r] + counts[r];
    if (pfx[n_ranks] >= 0xFFFFFFFFull) return io_fail(NG_EUNSUPPORTED, "too many nodes");
    const uint32_t n = static_cast<uint32_t>(pfx[n_ranks]);
    std::vector<Arc> in_arcs, out_arcs;
    for (int r = 0; r < n_ranks; ++r) {
        const fs::path fin = root / "network" / (rank_prefix(r, n_ranks) + pre + "_in_network.txt");
        const fs::path fout = root / "network" / (rank_prefix(r, n_ranks) + pre + "_out_network.txt");
        if (!fs::exists(fin)) return io_fail(NG_EINVAL, "the in arcs for rank %d do not exist: %s", r, fin.c_str());
        if (!fs::exists(fout)) return io_fail(NG_EINVAL, "the out arcs for rank %d do not exist: %s", r, fout.c_str());
        ng_status st = load_arcs(fin, r, true, counts, pfx, wkind, in_arcs);
        if (st != NG_OK) return st;
        st = load_arcs(fout, r, false, counts, pfx, wkind, out_arcs);
        if (st != NG_OK) return st;
    }
    const bool integer = wkind == NG_W_I32;
    merge(in_arcs, integer);   // in_arcs keyed by target: s = target, t = source
    merge(out_arcs, integer);  // out_arcs keyed by source
    // DistributedGraph.cpp:85-101: remove self arcs, then synchronize (+ all weights one) or all weights one
    if (remove_self_arcs) {
        auto self = [](const Arc& a) { return a.s == a.t; };
        in_arcs.erase(std::remove_if(in_arcs.begin(), in_arcs.end(), self), in_arcs.end());
        out_arcs.erase(std::remove_if(out_arcs.begin(), out_arcs.end(), self), out_arcs.end());
    }
    std::vector<Arc> arcs;
    if (undirected) {
        // ArcTransformer::synchronize_arcs (ArcTransformer.cpp:53-67): out[t][s] += w for every in arc t <- s,
        // a resulting 0 becomes 1; the merged map then serves as both in and out arcs.
        arcs = out_arcs;
        std::vector<Arc> add;
        for (const Arc& a : in_arcs) add.push_back({a.s, a.t, a.w});
        std::vector<Arc> all;
        all.reserve(arcs.size() + add.size());
        // merge-by-key with the zero->1 rule applied to keys touched by an in arc
        std::stable_sort(add.begin(), add.end(), [](const Arc& x, const Arc& y) { return x.s != y.s ? x.s < y.s : x.t < y.t; });
        size_t i = 0, j = 0;
        while (i < arcs.size() || j < add.size()) {
            const bool take_out = j >= add.size() || (i < arcs.size() && (arcs[i].s < add[j].s || (arcs[i].s == add[j].s && arcs[i].t <= add[j].t)));
            if (take_out) {
                Arc a = arcs[i++];
                if (j < add.size() && add[j].s == a.s && add[j].t == a.t) {
                    a.w = integer ? static_cast<double>(static_cast<int32_t>(static_cast<uint32_t>(static_cast<int64_t>(a.w)) +
                                                                             static_cast<uint32_t>(static_cast<int64_t>(add[j].w))))
                                  : a.w + add[j].w;
                    if (a.w == 0.0) a.w = 1.0;
                    ++j;
                }
                all.push_back(a);
            } else {
                Arc a = add[j++];
                if (a.w == 0.0) a.w = 1.0;
                all.push_back(a);
            }
        }
        arcs.swap(all);
        if (all_weights_one)
            for (Arc& a : arcs) a.w = 1.0;
    } else {
        arcs = out_arcs;
        if (all_weights_one)
            for (Arc& a : arcs) a.w = 1.0;
    }
    // upload skips weight-0 arcs (DistributedGraph.cpp:222-225)
    arcs.erase(std::remove_if(arcs.begin(), arcs.end(), [](const Arc& a) { return a.w == 0.0; }), arcs.end());
    ng_status st = to_csr(n, arcs, wkind, out);
    if (st != NG_OK) return st;
    out->positions = static_cast<double*>(std::malloc(std::max<size_t>(3ull * n, 1) * sizeof(double)));
    if (out->positions) {
        size_t k = 0;
        for (int r = 0; r < n_ranks; ++r)
            for (double x : pos[r]) out->positions[k++] = x;
    }
    return NG_OK;
}

ng_status ng_write_graph_dir(const char* dir, const ng_host_csr* g) {
    if (!dir || !g) return io_fail(NG_EINVAL, "null argument");
    const fs::path root(dir);
    std::error_code ec;
    fs::create_directories(root / "positions", ec);
    fs::create_directories(root / "network", ec);
    if (ec) return io_fail(NG_EINVAL, "cannot create %s", dir);
    {
        std::FILE* f = std::fopen((root / "positions" / "rank_0_positions.txt").c_str(), "w");
        if (!f) return io_fail(NG_EINVAL, "cannot write positions");
        std::fprintf(f, "# %u of %u\n", g->n, g->n);
        for (uint32_t u = 0; u < g->n; ++u) {
            const double x = g->positions ? g->positions[3 * u] : 0.0, y = g->positions ? g->positions[3 * u + 1] : 0.0,
                         z = g->positions ? g->positions[3 * u + 2] : 0.0;
            std::fprintf(f, "%u %.17g %.17g %.17g area_0 excitatory\n", u + 1, x, y, z);
        }
        std::fclose(f);
    }
    auto wstr = [&](uint64_t a, char* buf) {
        switch (g->wkind) {
        case NG_W_I32: std::snprintf(buf, 64, "%d", static_cast<const int32_t*>(g->weights)[a]); break;
        case NG_W_F32: std::snprintf(buf, 64, "%.9g", static_cast<double>(static_cast<const float*>(g->weights)[a])); break;
        default: std::snprintf(buf, 64, "%.17g", static_cast<const double*>(g->weights)[a]); break;
        }
    };
    const char* header =
        "# Total number neurons: %u\n# Local number neurons: %u\n# Number MPI ranks: 1\n"
        "# <target rank> <target neuron id>\t<source rank> <source neuron id>\t<weight>\t<plastic>\n";
    char buf[64];
    {
        std::FILE* f = std::fopen((root / "network" / "rank_0_out_network.txt").c_str(), "w");
        if (!f) return io_fail(NG_EINVAL, "cannot write out network");
        std::fprintf(f, header, g->n, g->n);
        for (uint32_t u = 0; u < g->n; ++u)
            for (uint64_t a = g->offsets[u]; a < g->offsets[u + 1]; ++a) {
                wstr(a, buf);
                std::fprintf(f, "0 %u\t0 %u\t%s\t1\n", g->targets[a] + 1, u + 1, buf);
            }
        std::fclose(f);
    }
    {
        std::vector<std::pair<uint32_t, uint64_t>> by_target;
        by_target.reserve(g->m);
        for (uint32_t u = 0; u < g->n; ++u)
            for (uint64_t a = g->offsets[u]; a < g->offsets[u + 1]; ++a) by_target.emplace_back(g->targets[a], a);
        std::stable_sort(by_target.begin(), by_target.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
        std::vector<uint32_t> src(g->m);
        for (uint32_t u = 0; u < g->n; ++u)
            for (uint64_t a = g->offsets[u]; a < g->offsets[u + 1]; ++a) src[a] = u;
        std::FILE* f = std::fopen((root / "network" / "rank_0_in_network.txt").c_str(), "w");
        if (!f) return io_fail(NG_EINVAL, "cannot write in network");
        std::fprintf(f, header, g->n, g->n);
        for (const auto& [t, a] : by_target) {
            wstr(a, buf);
            std::fprintf(f, "0 %u\t0 %u\t%s\t1\n", t + 1, src[a] + 1, buf);
        }
        std::fclose(f);
    }
    return NG_OK;
}

ng_status ng_generate_fixed_degree(uint32_t n, uint32_t k, int32_t wlo, int32_t whi, uint64_t seed, ng_host_csr* out) {
    if (!out) return io_fail(NG_EINVAL, "null argument");
    std::memset(out, 0, sizeof *out);
    if (n < 2 && k > 0) return io_fail(NG_EINVAL, "need at least two nodes");
    if (k >= n) return io_fail(NG_EINVAL, "k must be below n");
    if (whi < wlo) return io_fail(NG_EINVAL, "empty weight range");
    std::vector<Arc> arcs(static_cast<size_t>(n) * k);
    parallel_for(n, [&](uint64_t b, uint64_t e) {
        std::vector<uint32_t> picked;
        for (uint64_t u = b; u < e; ++u) {
            Xoshiro rng(seed, u);
            picked.clear();
            while (picked.size() < k) {
                uint32_t t = static_cast<uint32_t>(rng.below(n - 1));
                if (t >= u) ++t;  // uniform over the other nodes, no self-loops
                if (std::find(picked.begin(), picked.end(), t) != picked.end()) continue;  // duplicates rejected
                picked.push_back(t);
            }
            std::sort(picked.begin(), picked.end());
            for (uint32_t i = 0; i < k; ++i) {
                const int32_t w = wlo + static_cast<int32_t>(rng.below(static_cast<uint64_t>(whi - wlo) + 1));
                arcs[u * k + i] = {static_cast<uint32_t>(u), picked[i], static_cast<double>(w)};
            }
        }
    });
    return to_csr(n, arcs, NG_W_I32, out);
}

ng_status ng_generate_spatial(uint32_t n, double box, double lambda, double mean_deg, int wdist, double a, double b,
                              ng_weight_kind wkind, uint64_t seed, ng_host_csr* out) {
    if (!out) return io_fail(NG_EINVAL, "null argument");
    std::memset(out, 0, sizeof *out);
    if (n < 2 || !(box > 0) || !(lambda > 0) || !(mean_deg >= 0)) return io_fail(NG_EINVAL, "bad spatial parameters");
    if (wkind == NG_W_I32) return io_fail(NG_EINVAL, "spatial graphs carry real-valued weights");
    std::vector<double> pos(3ull * n);
    for (uint32_t u = 0; u < n; ++u) {
        Xoshiro rng(seed ^ 0x5a5a5a5a5a5a5a5aull, u);
        for (int d = 0; d < 3; ++d) pos[3ull * u + d] = box * rng.uniform();
    }
    const int G = std::max(1, std::min(24, static_cast<int>(std::ceil(box / (0.66 * lambda)))));
    const double cs = box / G;
    const int ncell = G * G * G;
    std::vector<std::vector<uint32_t>> cells(ncell);
    auto cell_of = [&](double x) { return std::min(G - 1, std::max(0, static_cast<int>(x / cs))); };
    for (uint32_t u = 0; u < n; ++u) {
        const int cx = cell_of(pos[3ull * u]), cy = cell_of(pos[3ull * u + 1]), cz = cell_of(pos[3ull * u + 2]);
        cells[(cz * G + cy) * G + cx].push_back(u);
    }
    // expected uncapped degree (up to the constant p0) from cell centres
    std::vector<double> E(n);
    parallel_for(n, [&](uint64_t b0, uint64_t e0) {
        for (uint64_t u = b0; u < e0; ++u) {
            double acc = 0.0;
            for (int c = 0; c < ncell; ++c) {
                if (cells[c].empty()) continue;
                const int cx = c % G, cy = (c / G) % G, cz = c / (G * G);
                const double dx = (cx + 0.5) * cs - pos[3 * u], dy = (cy + 0.5) * cs - pos[3 * u + 1], dz = (cz + 0.5) * cs - pos[3 * u + 2];
                acc += cells[c].size() * std::exp(-std::sqrt(dx * dx + dy * dy + dz * dz) / lambda);
            }
            E[u] = acc;
        }
    });
    double meanE = 0.0;
    for (double x : E) meanE += x;
    meanE /= n;
    std::vector<std::vector<Arc>> rows(n);
    parallel_for(n, [&](uint64_t b0, uint64_t e0) {
        std::vector<double> cum(ncell), dmin(ncell);
        std::vector<uint32_t> picked;
        for (uint64_t u = b0; u < e0; ++u) {
            Xoshiro rng(seed, u);
            const double px = pos[3 * u], py = pos[3 * u + 1], pz = pos[3 * u + 2];
            double tot = 0.0;
            for (int c = 0; c < ncell; ++c) {
                const int cx = c % G, cy = (c / G) % G, cz = c / (G * G);
                auto gap = [&](double p, int ci) {
                    const double lo = ci * cs, hi = lo + cs;
                    return p < lo ? lo - p : (p > hi ? p - hi : 0.0);
                };
                const double gx = gap(px, cx), gy = gap(py, cy), gz = gap(pz, cz);
                dmin[c] = std::sqrt(gx * gx + gy * gy + gz * gz);
                tot += cells[c].size() * std::exp(-dmin[c] / lambda);
                cum[c] = tot;
            }
            uint32_t K = rng.poisson(mean_deg * E[u] / meanE);
            K = std::min<uint32_t>(K, n - 1);
            picked.clear();
            uint64_t guard = 0;
            while (picked.size() < K && guard < 1000000ull * (K + 1)) {
                ++guard;
                const double r = rng.uniform() * tot;
                const int c = static_cast<int>(std::upper_bound(cum.begin(), cum.end(), r) - cum.begin());
                if (c >= ncell || cells[c].empty()) continue;
                const uint32_t v = cells[c][rng.below(cells[c].size())];
                if (v == u) continue;
                const double dx = pos[3ull * v] - px, dy = pos[3ull * v + 1] - py, dz = pos[3ull * v + 2] - pz;
                const double d = std::sqrt(dx * dx + dy * dy + dz * dz);
                if (rng.uniform() >= std::exp(-(d - dmin[c]) / lambda)) continue;  // exact: accept ~ exp(-d/lambda)
                if (std::find(picked.begin(), picked.end(), v) != picked.end()) continue;
                picked.push_back(v);
            }
            std::sort(picked.begin(), picked.end());
            rows[u].reserve(picked.size());
            for (uint32_t v : picked) {
                double w;
                if (wdist == 0) {
                    w = std::exp(a + b * rng.normal());
                } else {
                    w = a + (b - a) * rng.uniform();
                }
                if (wkind == NG_W_F32) w = static_cast<double>(static_cast<float>(w));
                rows[u].push_back({static_cast<uint32_t>(u), v, w});
            }
        }
    });
    std::vector<Arc> arcs;
    size_t total = 0;
    for (const auto& r : rows) total += r.size();
    arcs.reserve(total);
    for (auto& r : rows) {
        arcs.insert(arcs.end(), r.begin(), r.end());
        std::vector<Arc>().swap(r);
    }
    ng_status st = to_csr(n, arcs, wkind, out);
    if (st != NG_OK) return st;
    out->positions = static_cast<double*>(std::malloc(3ull * n * sizeof(double)));
    if (out->positions) std::memcpy(out->positions, pos.data(), 3ull * n * sizeof(double));
    return NG_OK;
}

// ---------------------------------------------------------------------------------------------------
// binary CSR cache

}  // extern "C"

namespace {

constexpr char kCsrMagic[8] = {'N', 'G', 'C', 'S', 'R', 0, 0, 0};
constexpr uint32_t kCsrVersion = 1;

struct CsrHeader {
    char magic[8];
    uint32_t version;
    uint32_t wkind;
    uint64_t n;
    uint64_t m;
    uint32_t flags;  // bit 0: positions present
    uint32_t zero;
    uint64_t key;
    uint64_t checksum;
};
static_assert(sizeof(CsrHeader) == 56, "header fields");
constexpr size_t kCsrHeaderBytes = 64;  // the header padded with zeros

// 64-bit multiply-xor hash in 8-byte words (a zero-padded tail word), continued across buffers.
struct Hash64 {
    uint64_t h = 0x6a09e667f3bcc909ull;
    void add(const void* p, size_t bytes) {
        const unsigned char* c = static_cast<const unsigned char*>(p);
        size_t i = 0;
        for (; i + 8 <= bytes; i += 8) {
            uint64_t w;
            std::memcpy(&w, c + i, 8);
            mix(w);
        }
        if (i < bytes) {
            uint64_t w = 0;
            std::memcpy(&w, c + i, bytes - i);
            mix(w ^ (static_cast<uint64_t>(bytes - i) << 56));
        }
    }
    void mix(uint64_t w) {
        h ^= w + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
        h *= 0xff51afd7ed558ccdull;
        h ^= h >> 32;
    }
};

uint64_t csr_checksum(const CsrHeader& hd, const ng_host_csr& g) {
    Hash64 hs;
    hs.add(&hd, offsetof(CsrHeader, key) + sizeof(uint64_t));  // the header up to and including the key
    hs.add(g.offsets, (static_cast<size_t>(g.n) + 1) * sizeof(uint64_t));
    hs.add(g.targets, g.m * sizeof(uint32_t));
    hs.add(g.weights, g.m * wsize(g.wkind));
    if (g.positions) hs.add(g.positions, 3ull * g.n * sizeof(double));
    return hs.h;
}

}  // namespace

extern "C" {

ng_status ng_save_csr(const char* path, const ng_host_csr* g, uint64_t key) {
    if (!path || !g) return io_fail(NG_EINVAL, "null argument");
    if (!g->offsets || (g->m && (!g->targets || !g->weights))) return io_fail(NG_EINVAL, "incomplete CSR");
    if (g->wkind != NG_W_I32 && g->wkind != NG_W_F32 && g->wkind != NG_W_F64) return io_fail(NG_EINVAL, "unknown weight kind");
    CsrHeader hd{};
    std::memcpy(hd.magic, kCsrMagic, 8);
    hd.version = kCsrVersion;
    hd.wkind = static_cast<uint32_t>(g->wkind);
    hd.n = g->n;
    hd.m = g->m;
    hd.flags = g->positions ? 1u : 0u;
    hd.key = key;
    hd.checksum = csr_checksum(hd, *g);
    const std::string tmp = std::string(path) + ".tmp";
    {
        std::ofstream f(tmp, std::ios::binary | std::ios::trunc);
        if (!f) return io_fail(NG_EINVAL, "cannot write %s", tmp.c_str());
        char head[kCsrHeaderBytes] = {};
        std::memcpy(head, &hd, sizeof hd);
        f.write(head, kCsrHeaderBytes);
        f.write(reinterpret_cast<const char*>(g->offsets), static_cast<std::streamsize>((static_cast<size_t>(g->n) + 1) * sizeof(uint64_t)));
        if (g->m) {
            f.write(reinterpret_cast<const char*>(g->targets), static_cast<std::streamsize>(g->m * sizeof(uint32_t)));
            f.write(static_cast<const char*>(g->weights), static_cast<std::streamsize>(g->m * wsize(g->wkind)));
        }
        if (g->positions) f.write(reinterpret_cast<const char*>(g->positions), static_cast<std::streamsize>(3ull * g->n * sizeof(double)));
        f.flush();
        if (!f) return io_fail(NG_EINVAL, "writing %s failed", tmp.c_str());
    }
    std::error_code ec;
    fs::rename(tmp, path, ec);
    if (ec) return io_fail(NG_EINVAL, "cannot rename %s to %s: %s", tmp.c_str(), path, ec.message().c_str());
    return NG_OK;
}

ng_status ng_load_csr(const char* path, uint64_t key, ng_host_csr* out) {
    if (!path || !out) return io_fail(NG_EINVAL, "null argument");
    std::memset(out, 0, sizeof *out);
    std::ifstream f(path, std::ios::binary);
    if (!f) return io_fail(NG_EINVAL, "cannot open %s", path);
    char head[kCsrHeaderBytes] = {};
    f.read(head, kCsrHeaderBytes);
    if (f.gcount() != static_cast<std::streamsize>(kCsrHeaderBytes)) return io_fail(NG_EINVAL, "%s: truncated header", path);
    CsrHeader hd{};
    std::memcpy(&hd, head, sizeof hd);
    if (std::memcmp(hd.magic, kCsrMagic, 8) != 0) return io_fail(NG_EINVAL, "%s: not a CSR cache file", path);
    if (hd.version != kCsrVersion) return io_fail(NG_EINVAL, "%s: cache version %u, this library reads %u", path, hd.version, kCsrVersion);
    if (hd.wkind > NG_W_F64 || hd.n >= 0xFFFFFFFFull || hd.m >= 0xFFFFFFFFull || (hd.flags & ~1u) != 0 || hd.zero != 0)
        return io_fail(NG_EINVAL, "%s: malformed header", path);
    if (hd.key != key) return io_fail(NG_EINVAL, "%s: written for other input or load options (key %llx, expected %llx)", path,
                                      (unsigned long long)hd.key, (unsigned long long)key);
    const ng_weight_kind wk = static_cast<ng_weight_kind>(hd.wkind);
    const size_t payload = (hd.n + 1) * 8 + hd.m * 4 + hd.m * wsize(wk) + ((hd.flags & 1u) ? 3 * hd.n * 8 : 0);
    std::error_code ec;
    const uintmax_t fsize = fs::file_size(path, ec);
    if (ec || fsize != kCsrHeaderBytes + payload) return io_fail(NG_EINVAL, "%s: size does not match its header", path);
    out->n = static_cast<uint32_t>(hd.n);
    out->m = hd.m;
    out->wkind = wk;
    out->offsets = static_cast<uint64_t*>(std::malloc((hd.n + 1) * sizeof(uint64_t)));
    out->targets = static_cast<uint32_t*>(std::malloc(std::max<uint64_t>(hd.m, 1) * sizeof(uint32_t)));
    out->weights = std::malloc(std::max<uint64_t>(hd.m, 1) * wsize(wk));
    if ((hd.flags & 1u) != 0) out->positions = static_cast<double*>(std::malloc(std::max<uint64_t>(3 * hd.n, 1) * sizeof(double)));
    if (!out->offsets || !out->targets || !out->weights || ((hd.flags & 1u) && !out->positions)) {
        ng_host_csr_free(out);
        return io_fail(NG_ENOMEM, "host allocation failed");
    }
    f.read(reinterpret_cast<char*>(out->offsets), static_cast<std::streamsize>((hd.n + 1) * sizeof(uint64_t)));
    if (hd.m) {
        f.read(reinterpret_cast<char*>(out->targets), static_cast<std::streamsize>(hd.m * sizeof(uint32_t)));
        f.read(static_cast<char*>(out->weights), static_cast<std::streamsize>(hd.m * wsize(wk)));
    }
    if (out->positions) f.read(reinterpret_cast<char*>(out->positions), static_cast<std::streamsize>(3 * hd.n * sizeof(double)));
    const char* why = nullptr;
    if (!f) why = "truncated payload";
    else if (csr_checksum(hd, *out) != hd.checksum) why = "checksum mismatch (corrupted file)";
    else if (out->offsets[0] != 0 || out->offsets[hd.n] != hd.m) why = "offsets do not start at 0 and end at m";
    if (!why) {
        for (uint64_t u = 0; u < hd.n && !why; ++u)
            if (out->offsets[u] > out->offsets[u + 1]) why = "offsets decrease";
        for (uint64_t a = 0; a < hd.m && !why; ++a)
            if (out->targets[a] >= hd.n) why = "a target is out of range";
    }
    if (why) {
        ng_host_csr_free(out);
        return io_fail(NG_EINVAL, "%s: %s", path, why);
    }
    return NG_OK;
}

}  // extern "C"
