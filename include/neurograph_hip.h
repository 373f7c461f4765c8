/*
 * neurograph_hip.h -- C ABI of libng_hip.so, the MI355X (gfx950) path for neurograph's all-pairs
 * shortest paths and Brandes betweenness centrality.
 *
 * Plain C: pointers, sizes and enums only. Every call is blocking: it enqueues on the context's HIP
 * stream and synchronizes before returning. Host buffers are borrowed for the duration of a call;
 * device copies are owned by the returned handles. Errors are status codes; ng_last_error() returns
 * a thread-local message for the last failing call on the calling thread. One host thread per
 * context; calls on one graph handle are not re-entrant.
 *
 * Citations are file:line in the reference checkout (tuda-hpclab/neurograph).
 */
#ifndef NEUROGRAPH_HIP_H
#define NEUROGRAPH_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    NG_OK = 0,
    NG_EINVAL = 1,       /* bad argument: node id out of range, malformed CSR, negative raw weight, ... */
    NG_ENOMEM = 2,       /* device or host allocation failed */
    NG_EDEVICE = 3,      /* HIP runtime error or a device-side consistency check failed */
    NG_EUNSUPPORTED = 4  /* valid request this path does not implement (e.g. zero-length arcs in Brandes) */
} ng_status;

/* Arc weight storage. NG_W_I32 is the reference's weight_type (source/Types.h:31);
 * NG_W_F32 / NG_W_F64 are real-valued weights, an extension of the reference's integer-only loader. */
typedef enum { NG_W_I32 = 0, NG_W_F32 = 1, NG_W_F64 = 2 } ng_weight_kind;

/* Arc length used by the shortest-path metric.
 *   NG_LEN_ABS      |w|, DijkstraCommon::transform_weight<distance_type> (DijkstraCommon.h:43-46)
 *   NG_LEN_INVERSE  1/|w| in double, 0 -> infinitely long (DijkstraCommon.h:48-52)
 *   NG_LEN_RAW      w itself, which must be >= 0: delta-stepping's safe_cast<distance_type>(weight)
 *                   (PushCheckVariant.h:82) throws on negative weights, here NG_EINVAL. */
typedef enum { NG_LEN_ABS = 0, NG_LEN_INVERSE = 1, NG_LEN_RAW = 2 } ng_length_kind;

/* Distance element type of dist_out buffers (ng_graph_dist_kind). Unreachable = the type's max(),
 * as in the reference (Dijkstra.h:63, 71). */
typedef enum { NG_DIST_U64 = 0, NG_DIST_F32 = 1, NG_DIST_F64 = 2 } ng_dist_kind;

/* Per-source result: SsspResult<d_type> (metrics/paths/types/SsspResult.h:19-25) with exact integer
 * channels. sum_eff is the sum of 1/d over reached nodes with d != 0 (SsspResult.h:56-58).
 * Real-valued distance sums are accumulated in double. */
typedef struct {
    double sum_d;
    uint64_t sum_d_u64;   /* integer lengths only, else 0 */
    double max_d;
    uint64_t max_d_u64;   /* integer lengths only, else 0 */
    double sum_eff;
    uint64_t n_unreached;
} ng_sssp_summary;

/* Timing of the most recent call on a graph (HIP events on the context stream, around the batched
 * path kernel only) plus work counters of that call. */
typedef struct {
    double kernel_ms;        /* sum of the hot kernel's launch durations in the last call */
    uint32_t launches;       /* number of hot-kernel launches in the last call */
    uint32_t slots;          /* concurrent source slots (workgroups) used */
    double delta;            /* bucket width used */
    uint64_t arcs_reached;   /* sum over sources of out-arcs of reached nodes (TEPS numerator) */
    uint32_t flags;          /* informational device flags of the last call (2: tight-arc log overflowed
                                and the full-row scan was used; 4: in-bucket work after a settle pass) */
    uint32_t block;          /* lanes per workgroup (one source slot) used */
    /* work counters of the last call, summed over sources */
    uint64_t phases;         /* frontier phases (light-arc rounds) */
    uint64_t buckets;        /* bucket advances */
    uint64_t light_arcs;     /* arcs scanned in frontier phases */
    uint64_t settle_arcs;    /* arcs scanned in settle passes */
    uint64_t logged_arcs;    /* candidate tight arcs logged (betweenness modes) */
    uint64_t layers;         /* Kahn layers of the tight DAG (betweenness modes) */
    uint64_t probes;         /* arc-target distance updates issued (arcs past the LDS filter) */
    uint64_t improvements;   /* successful distance improvements */
    uint64_t section_ticks[8]; /* wall-clock ticks per kernel section (frontier phases, settle passes, bucket
                                  advances, summaries, tight-arc verification, sigma layers, dependency
                                  layers, slot reset), summed over workgroups; zero unless the library was
                                  built with -DNG_SECTION_TIMERS */
    uint64_t tree_sources;   /* betweenness modes: sources whose tight arcs formed a shortest-path tree (every sigma
                                1: the tree layers ran) */
    uint64_t near_tree_sources; /* Brandes mode: sources whose tight arcs formed a tree but for a few nodes with
                                   several tight in-arcs (the near-tree layers ran) */
} ng_call_stats;

typedef struct ng_context ng_context;
typedef struct ng_graph ng_graph;
typedef struct ng_cohort ng_cohort;

/* Opens HIP device `device` and creates one stream (default priority). Replaces the per-rank MPI
 * set-up: MPIWrapper::init (external/mpi-wrapper/MPIWrapper.h:88-109). */
ng_status ng_context_create(int device, ng_context** out);
/* Releases the caller's handle. Graphs and cohorts created on the context keep it alive: its stream and
 * workspace are freed when the context and all of them have been destroyed, in any order. */
void ng_context_destroy(ng_context* ctx);

/* Waits until all work of the context's stream is done (the barrier around timed regions). */
ng_status ng_context_synchronize(ng_context* ctx);

/* Measurement aid, no reference counterpart: one plain device copy of `bytes` bytes (16-byte vector loads
 * and stores) on the context stream, a kernel of known traffic against which profiler byte counters are
 * scaled. ms_out (optional) receives its duration from HIP events. */
ng_status ng_copy_calibration(ng_context* ctx, uint64_t bytes, double* ms_out);

/* Uploads a graph given as host CSR over out-arcs: out_offsets[n+1] (non-decreasing, out_offsets[0]=0,
 * out_offsets[n]=m), out_targets[m] < n, out_weights[m] of `wkind`. Replaces the per-rank RMA windows
 * and remote adjacency cache: DistributedGraph::construct_graph (DistributedGraph.cpp:46-127),
 * get_out_arcs (DistributedGraph.h:275-311). Parallel arcs and self-loops are kept (they are arcs of
 * the reference graph too); m must be < 2^32 (arc_id_type, Types.h:27); real-valued weights must be finite. */
ng_status ng_graph_create(ng_context* ctx, uint32_t n, uint64_t m, const uint64_t* out_offsets,
                          const uint32_t* out_targets, const void* out_weights, ng_weight_kind wkind,
                          ng_graph** out);
void ng_graph_destroy(ng_graph* g);

/* Distance element type of dist_out for this graph and length kind. */
ng_status ng_graph_dist_kind(const ng_graph* g, ng_length_kind lkind, ng_dist_kind* out);

/* Replaces every arc weight in place: weights[m] of the graph's weight kind, indexed like out_targets at creation
 * (the arc structure is unchanged). The device arrays are reused: rows whose (length, target) order the new
 * weights keep are not re-sorted, and when no row changes order only the lengths are uploaded; integer weights
 * that change the distance type (32 <-> 64 bits) rebuild the arrays. Tuned bucket widths and launch shapes are
 * kept (they only affect speed); cohorts holding the graph re-capture on their next replay. A non-finite weight
 * is NG_EINVAL and leaves the graph unchanged. Lets a training loop (the adjoint's torch binding) reuse one
 * handle across steps instead of ng_graph_create per step. No reference counterpart. */
ng_status ng_graph_set_weights(ng_graph* g, const void* weights);
/* Whether the last ng_graph_create / ng_graph_set_weights changed the row order of the |w| view (1) or only
 * the lengths (0). */
ng_status ng_graph_weights_reordered(const ng_graph* g, int* reordered);

/* Launch options of one graph: explicit overrides of the library's automatic launch shape and capacities, for
 * tuning sweeps and for tests that must reach a particular code path. No reference counterpart (the reference
 * has no launch shape). Every option defaults to automatic; results never depend on an option, only speed and
 * which code path runs. The library reads no environment variables.
 *   NG_OPT_BLOCK          lanes per source workgroup for every call on this graph: 64, 256, 512 or 1024;
 *                         0 = the tuned shape (ng_autotune) or the default (256). The autotuner's own trial
 *                         runs are not affected.
 *   NG_OPT_MAX_SLOTS      at most this many concurrent source slots (workgroups); 0 = as many as are resident.
 *   NG_OPT_LDS_MAPS       -1 = automatic; 0..31 = the extra LDS maps for every call (bit flags: 1 settle passes
 *                         in increasing node id, 2 frontier bitmap, 4 16-bit distance bounds, 8 4-bit distance
 *                         codes, 16 the distances themselves in LDS; maps that do not fit are dropped).
 *   NG_OPT_PILE_CAPACITY  parked-node entries per pile list; 0 = 3n + 1024, which is also the largest value
 *                         used; smaller values are raised to n + 2048 (a full pile makes bucket advances reload
 *                         distances: slower, same results).
 *   NG_OPT_LOG_CAPACITY   candidate tight-arc log entries per slot (betweenness modes); 0 = min(m, max(8n, 4096));
 *                         a source whose candidates overflow it rescans all rows of its reached nodes instead
 *                         (ng_call_stats.flags bit 2). */
typedef enum {
    NG_OPT_BLOCK = 0,
    NG_OPT_MAX_SLOTS = 1,
    NG_OPT_LDS_MAPS = 2,
    NG_OPT_PILE_CAPACITY = 3,
    NG_OPT_LOG_CAPACITY = 4
} ng_graph_option;
ng_status ng_graph_set_option(ng_graph* g, ng_graph_option opt, int64_t value);
ng_status ng_graph_get_option(const ng_graph* g, ng_graph_option opt, int64_t* value);

/* Multi-source SSSP with per-source summaries, one batched launch for all sources. Replaces the
 * per-source loop of AllPairsShortestPath::compute_apsp_impl (AllPairsShortestPath.h:130-143) over
 * Dijkstra::compute_sssp (Dijkstra.h:56-98) + summarize_distances (SsspResult.h:43-64), and the
 * delta-stepping dispatcher DeltaStepping::compute_shortest_distances (DeltaStepping.h:36-82).
 * delta <= 0 selects the graph's tuned (or heuristic) bucket width. out[n_sources] is required;
 * dist_out (n_sources * n elements of ng_graph_dist_kind) is optional. */
ng_status ng_sssp_summaries(ng_graph* g, const uint32_t* sources, uint32_t n_sources, ng_length_kind lkind,
                            double delta, ng_sssp_summary* out, void* dist_out);

/* Histogram of the finite shortest-path distances d(s, t), t != s, over `sources` (lengths lkind, as in
 * ng_sssp_summaries): counts[i] = #{d : i <= d / bin_width < i + 1} for i < n_bins - 1; the last bin takes
 * everything beyond. Built on the device from the distance rows of the batched pass, in source chunks of at
 * most 1 GiB of rows; out_opt[n_sources] (optional) receives the per-source summaries of the same pass.
 * The reach-size histogram of the reference's output is ApspGlobalResult::cluster_sizes, rebuilt from the
 * summaries (AllPairsShortestPath.h:140-141); this one is new (no reference counterpart). */
ng_status ng_distance_histogram(ng_graph* g, const uint32_t* sources, uint32_t n_sources, ng_length_kind lkind, double delta,
                                double bin_width, uint32_t n_bins, uint64_t* counts, ng_sssp_summary* out_opt);

/* Brandes betweenness over `sources` with lengths |w|, accumulated into bc_inout[n] like
 * local_betweenness in BrandesVariant::accumulate_betweenness (BrandesVariant.h:50-74, :157) and the
 * dispatcher BetweennessCentrality::compute_average_betweenness_centrality (BetweennessCentrality.h:46-80).
 * out_opt[n_sources] (optional) receives the per-source SSSP summaries of the same forward pass, so
 * one call serves avg-apsp and avg-centrality. Zero-length arcs are NG_EUNSUPPORTED (BrandesVariant.h:30-31). */
ng_status ng_brandes_accumulate(ng_graph* g, const uint32_t* sources, uint32_t n_sources, double delta,
                                double* bc_inout, ng_sssp_summary* out_opt);

/* ng_brandes_accumulate without any device allocation. No counterpart in the reference: it is for callers that
 * cannot allocate device memory (tight serving loops, code that must not call hipMalloc). Same arguments, results
 * and errors as ng_brandes_accumulate; it calls no hipMalloc / hipFree and allocates no host memory. Its working
 * memory (source slots, per-source summaries, the betweenness reduction) is one fixed device array of 256 MiB per
 * GPU, allocated when the library's code is loaded on that GPU and held until the process ends; its address is
 * resolved by ng_context_create. Graphs whose one source slot does not fit it are rejected with NG_EINVAL (a slot
 * takes about 165 bytes per node plus 16 bytes per arc up to 8 arcs per node, 20 with 8-byte distances: roughly
 * 900 000 nodes at mean out-degree 20 and 1.15 million at mean out-degree 4 with 4-byte distances, 800 000 and
 * 1.06 million with 8-byte ones; the 200 000-node spatial graph gets 4 slots, the 50 000-node one 16); fewer slots
 * fit than the GPU could run at once for large graphs, so it is slower there than ng_brandes_accumulate. Every call
 * re-initialises the array, which keeps the last call's contents afterwards; calls that overlap on one GPU (from
 * several threads or contexts) race on it and must be serialised by the caller. Sources are run in launches of
 * at most 16 384. Bucket width and launch shape as in ng_brandes_accumulate, except that 4-bit distance codes are
 * used only when an earlier call on the graph estimated their scale. */
ng_status ng_brandes_accumulate_noalloc(ng_graph* g, const uint32_t* sources, uint32_t n_sources, double delta,
                                        double* bc_inout, ng_sssp_summary* out_opt);

/* Edge-level path counts: edge_inout[m] += sum over sources of sigma(v) * (1 + delta(w)) / sigma(w) for every
 * shortest-path arc (v,w), the per-arc term of BrandesVariant.h:151-154, indexed like out_targets. */
ng_status ng_edge_dependency(ng_graph* g, const uint32_t* sources, uint32_t n_sources, double delta,
                             double* edge_inout);

/* Adjoint of the APSP sums with respect to the arc weights (lengths |w|), tie-split over shortest paths:
 *   grad_sum_d[e]   += d(sum_{s in sources, t reached} d(s,t)) / d w_e
 *   grad_sum_eff[e] += d(sum_{s, t reached, d != 0} 1/d(s,t)) / d w_e
 * indexed like out_targets. The averages of compute_apsp_impl (AllPairsShortestPath.h:154-158) divide
 * these sums by the (piecewise constant) number of reached pairs. sigma is recomputed per source batch;
 * no predecessor DAG is stored. out_opt[n_sources] as in ng_brandes_accumulate. No reference counterpart. */
ng_status ng_apsp_adjoint(ng_graph* g, const uint32_t* sources, uint32_t n_sources, double delta,
                          double* grad_sum_d, double* grad_sum_eff, ng_sssp_summary* out_opt);

/* Picks the bucket width for (lkind, mode) by timing candidate widths on `sample` sources and stores it
 * in the graph (used whenever delta <= 0). mode: 0 = SSSP, 1 = Brandes. The reference takes a fixed
 * integer --apsp-delta (neurograph.cpp:515); tuning is an extension of this library. */
ng_status ng_autotune_delta(ng_graph* g, ng_length_kind lkind, int mode, const uint32_t* sample, uint32_t n_sample,
                            double* delta_out);

/* Picks the bucket width and the lanes per workgroup (64 -- one wavefront per source, tried on graphs of at most
 * 32768 nodes --, 256, 512 or 1024) together for (lkind, mode) by timing
 * the batched kernel on `sample` sources (widths: the 5 % length quantile times 1/4 ... 8 in powers of two,
 * then the best one times 1/sqrt(2) and sqrt(2)), and stores both in the graph (used whenever delta <= 0). Fewer,
 * wider workgroups keep fewer sources in flight, so their per-source state stays cache-resident on large
 * graphs. The sample should hold at least twice as many sources as the GPU runs at once with 256-lane
 * workgroups (5 per CU), so that every candidate is timed at full occupancy. Extension, like
 * ng_autotune_delta. */
ng_status ng_autotune(ng_graph* g, ng_length_kind lkind, int mode, const uint32_t* sample, uint32_t n_sample, double* delta_out,
                      uint32_t* block_out);

/* Stats of the last successful call on `g`. */
ng_status ng_graph_last_stats(const ng_graph* g, ng_call_stats* out);

/* Cohort: many graphs of uneven size, all sources of all graphs, in one persistent launch captured in
 * a hipGraph and replayed. Workgroup b runs the (graph, source) items b, b + S, b + 2S, ... in graph
 * order, so all workgroups sweep the member graphs together; betweenness partials are flushed per
 * (workgroup, graph) and summed in workgroup order, so the results are deterministic. Member graphs are borrowed (they must outlive the cohort), must share one
 * distance type and have no zero-length arcs. No reference counterpart (the reference runs one graph per
 * process); each member's results equal ng_brandes_accumulate over all its sources. */
ng_status ng_cohort_create(ng_context* ctx, ng_graph* const* graphs, uint32_t n_graphs, ng_cohort** out);
void ng_cohort_destroy(ng_cohort* c);
/* Replays the captured launch (built on first use). bc_out: concatenation of the graphs' betweenness
 * vectors (sum of n_i doubles), overwritten; summaries_out (optional): one summary per (graph, source) in
 * graph order; kernel_ms (optional): duration of the whole replay (HIP events). */
ng_status ng_cohort_brandes(ng_cohort* c, double* bc_out, ng_sssp_summary* summaries_out, double* kernel_ms);
/* Autotuner of the cohort launch: per member graph the bucket width (ng_autotune_delta in Brandes mode on
 * n_sample evenly spaced sources), then the launch shape by timing whole replays: 256-lane workgroups at the
 * occupancy maximum and one below it, 512- and 1024-lane workgroups at theirs. The next ng_cohort_brandes
 * recaptures with the chosen shape. */
ng_status ng_cohort_autotune(ng_cohort* c, uint32_t n_sample);
/* Stats of the last cohort replay (kernel_ms covers the whole replay). */
ng_status ng_cohort_last_stats(const ng_cohort* c, ng_call_stats* out);

const char* ng_last_error(void);

/* Library version string, e.g. "ng_hip 0.1 gfx950". */
const char* ng_version(void);

#ifdef __cplusplus
}
#endif

#endif /* NEUROGRAPH_HIP_H */
