/*
 * neurograph_io.h -- host-side graph I/O of libng_hip.so: neurograph's on-disk format, its load-time
 * transforms, and deterministic generators for the benchmark graphs. Plain C ABI; no device work.
 * Citations are file:line in the reference checkout (tuda-hpclab/neurograph).
 */
#ifndef NEUROGRAPH_IO_H
#define NEUROGRAPH_IO_H

#include <stdint.h>

#include "neurograph_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Host CSR over out-arcs, rows sorted by target id. Arrays are allocated by the library and released
 * with ng_host_csr_free. weights has element type wkind. positions (3*n doubles) may be NULL. */
typedef struct {
    uint32_t n;
    uint64_t m;
    ng_weight_kind wkind;
    uint64_t* offsets;
    uint32_t* targets;
    void* weights;
    double* positions;
} ng_host_csr;

void ng_host_csr_free(ng_host_csr* g);

/* Loads an input directory of the reference format with all its rank files flattened into one graph
 * (global id = prefix[rank] + local id, AllPairsShortestPath.h:115-117):
 *   positions/rank_<r>_positions.txt         FileLoader::load_nodes      (FileLoader.cpp:96-177)
 *   network/rank_<r><prefix>_in_network.txt   FileLoader::load_in_arcs    (FileLoader.cpp:179-252)
 *   network/rank_<r><prefix>_out_network.txt  FileLoader::load_out_arcs   (FileLoader.cpp:254-327)
 * then the transforms of DistributedGraph::construct_graph (DistributedGraph.cpp:85-101) and the upload
 * rule that skips weight-0 arcs (DistributedGraph.cpp:222-225). n_ranks = 0 detects the rank count from
 * the positions files. wkind NG_W_I32 parses weights exactly like the reference (istream >> int32, so
 * "1.5" reads as 1); NG_W_F32 / NG_W_F64 parse real-valued weights (an extension of this library). */
ng_status ng_load_graph_dir(const char* dir, const char* prefix, int n_ranks, int remove_self_arcs, int undirected,
                            int all_weights_one, ng_weight_kind wkind, ng_host_csr* out);

/* Writes g as a one-rank input directory of the reference format (positions, in and out network files,
 * 1-based ids, plastic column 1). Integer weights are written as integers, real weights with %.17g. */
ng_status ng_write_graph_dir(const char* dir, const ng_host_csr* g);

/* Exactly k distinct out-arcs per node, targets uniform over the other nodes, integer weights uniform in
 * [wlo, whi]; xoshiro256** streams seeded per node from `seed` (deterministic, thread-count independent). */
ng_status ng_generate_fixed_degree(uint32_t n, uint32_t k, int32_t wlo, int32_t whi, uint64_t seed, ng_host_csr* out);

/* Spatial connectome-like graph: n nodes uniform in [0, box]^3; arcs u->v drawn with probability
 * p0 * exp(-d(u,v) / lambda) and thinned uniformly to mean out-degree mean_deg. Realised per node as
 * K_u ~ Poisson(mean_deg * E_u / mean E) distinct targets drawn with probability proportional to
 * exp(-d/lambda) (exact rejection sampling over a cell grid), E_u being the node's expected uncapped
 * degree. Weights: wdist 0 = lognormal(a, b) (a = mu, b = sigma), 1 = uniform [a, b]; stored as wkind. */
ng_status ng_generate_spatial(uint32_t n, double box, double lambda, double mean_deg, int wdist, double a, double b,
                              ng_weight_kind wkind, uint64_t seed, ng_host_csr* out);

/* Binary CSR cache: the loaded (and transformed) graph in one file, so that later runs skip the text parse of
 * ng_load_graph_dir (FileLoader.cpp:179-327 reads every arc line of every rank file). No reference counterpart.
 * Layout (little-endian): a 64-byte header -- magic "NGCSR\0\0\0", u32 version (1), u32 weight kind, u64 n,
 * u64 m, u32 flags (bit 0: positions present), u32 zero, u64 key, u64 checksum -- then offsets u64[n+1],
 * targets u32[m], weights[m] of the weight kind, and positions f64[3n] if present. The checksum is a 64-bit
 * multiply-xor hash of the header's first 48 bytes and then of each array in turn, in 8-byte words (each array's
 * tail zero-padded to a word).
 * `key` is the caller's identifier of the source and load options (the CLI hashes the input path, the prefix and
 * the transform flags); ng_load_csr refuses a file whose key differs, whose header, size or checksum is wrong,
 * or whose CSR is malformed (offsets not starting at 0, decreasing, or not ending at m; a target >= n), with
 * NG_EINVAL and a message naming the reason. */
ng_status ng_save_csr(const char* path, const ng_host_csr* g, uint64_t key);
ng_status ng_load_csr(const char* path, uint64_t key, ng_host_csr* out);

/* A source of uniform random numbers in [0, 1) for ng_contract_balls: the reference's RandomNumberGenerator
 * interface (utility/RandomNumberGenerator.h), so callers (tests) can inject deterministic draws. */
typedef double (*ng_draw_fn)(void* user);

/* DiameterApproximation's clustering and contraction (metrics/paths/DiameterApproximation.h:71-113, 154-189 mark
 * centers, 204-284 growing step, 286-295 stabilize, 297-561 contract) for one rank holding the whole graph: in iteration i = 0 .. ceil(log2 n) - 1 every node without a
 * center draws a number and becomes a center when it is below 2^i / n; balls then grow from their centers over
 * arcs of weight <= 2r in synchronous rounds (a claim needs d(u) + w <= (i - generation(u) + 1) 2r and must beat
 * the target's distance; claims of one round apply in order of claiming node, then arc, and a claim at the same
 * distance as an earlier one of the round loses), until no claim succeeds; nodes with a center become stable; the
 * loop ends early when every node is stable. The contracted graph has one node per center (in increasing id) and,
 * for every arc (u, v, w), the arc center(u) -> center(v) of weight w + d(u) + d(v), a later arc between the same
 * centers replacing an earlier one (in node, then target order). Integer weights only (NG_W_I32); a negative
 * weight, a node left without a center, or a contracted weight beyond int32 is NG_EINVAL, as the reference
 * throws. draw == NULL draws from std::mt19937(seed) through std::uniform_real_distribution<double>(0, 1), the
 * reference's DefaultRandomNumberGenerator seeded with the MPI rank (0 on one rank). centers_out (optional)
 * receives each node's center. The contracted graph is allocated like every ng_host_csr (ng_host_csr_free). */
ng_status ng_contract_balls(const ng_host_csr* g, uint64_t r, ng_draw_fn draw, void* user, uint32_t seed, ng_host_csr* out,
                            uint32_t* centers_out);

/* DiameterApproximation::compute_approximation (DiameterApproximation.h:71-126): ng_contract_balls, then the
 * diameter of the contracted graph by AllPairsShortestPath::compute_apsp on the device (ng_sssp_summaries over all
 * its nodes, lengths |w|), plus 2r. The reference's call at DiameterApproximation.h:118 is this inner APSP. */
ng_status ng_diameter_approximation(ng_context* ctx, const ng_host_csr* g, uint64_t r, ng_draw_fn draw, void* user,
                                    uint32_t seed, uint64_t* diameter_out, uint32_t* contracted_nodes_out);

#ifdef __cplusplus
}
#endif

#endif /* NEUROGRAPH_IO_H */
