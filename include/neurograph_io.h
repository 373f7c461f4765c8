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

#ifdef __cplusplus
}
#endif

#endif /* NEUROGRAPH_IO_H */
