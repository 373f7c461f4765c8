// Internal declarations shared by the C-ABI layer (ng_api.cpp) and the kernels (ng_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/neurograph_hip.h"

namespace ng {

// Lanes per workgroup (one source slot, the frontier tile size): the default and the largest supported; the
// batched kernel is built for 64 (one wavefront per source), 256, 512 and 1024.
constexpr int kBlock = 256;
constexpr int kMaxBlock = 1024;
// Largest graph for which the autotuner also times one wavefront per source: smaller graphs' frontiers rarely
// fill wider workgroups, while larger ones would hold thousands of slots of per-node state.
constexpr uint32_t kWaveSourceMaxNodes = 32768;

// Iteration guard of one source in the batched kernel: at most 64 n + 100 000 frontier phases and bucket
// advances, computed in 64 bits and saturated to 2^32 - 1 (the guard counter is 32-bit), so that the bound
// grows with n instead of wrapping for graphs of 2^26 nodes and more.
__host__ __device__ constexpr uint32_t guard_limit(uint32_t n) {
    const uint64_t g = 64ull * n + 100000ull;
    return g > 0xFFFFFFFFull ? 0xFFFFFFFFu : static_cast<uint32_t>(g);
}
static_assert(guard_limit(0) == 100000u, "guard of an empty graph");
static_assert(guard_limit(5000) == 420000u, "guard of a 5000-node graph");
static_assert(guard_limit(67107301u) == 4294967264u, "largest n whose bound fits 32 bits");
static_assert(guard_limit(67107302u) == 0xFFFFFFFFu, "saturates where 32-bit arithmetic would wrap");
static_assert(guard_limit(1u << 26) == 0xFFFFFFFFu, "2^26 nodes: saturated, not wrapped to 100000");
static_assert(guard_limit(0xFFFFFFFEu) == 0xFFFFFFFFu, "saturates at 2^32 - 1");

// Scratch of the allocation-free Brandes entry point (ng_brandes_accumulate_noalloc): a file-scope device array
// of this many bytes, one copy per GPU, held until the process ends; devices 0 .. kScratchMaxDevices - 1.
constexpr size_t kScratchBytes = 256ull << 20;
constexpr int kScratchMaxDevices = 64;
// Sources per launch of the allocation-free entry point (bounds the per-source summaries kept in the scratch).
constexpr uint32_t kScratchChunk = 16384;

// Device distance element type: lengths and distances are kept in one of these.
enum class DKind : int { U32 = 0, U64 = 1, F32 = 2, F64 = 3 };

// Kernel modes (bit flags). SSSP = distances + summaries only.
enum : int { MODE_SSSP = 0, MODE_BC = 1, MODE_ARC1 = 2, MODE_ARC2 = 4 };

// Bounds of row v in one 16-byte record: its first arc, the end of its light prefix (arcs shorter than the
// bucket width) and its end. A relaxation pass touches a node's record with one load instead of two or three
// loads from separate offset arrays, each of which would fetch a line of its own.
struct alignas(16) RowRange {
    uint32_t beg;
    uint32_t light_end;
    uint32_t end;
    uint32_t pad;
};

// Device CSR of one length kind. Rows are sorted by (length, target) so that the arcs lighter than the
// bucket width form a prefix of each row; perm maps a sorted position back to the caller's arc index.
struct View {
    bool built = false;
    DKind dk = DKind::U32;
    uint32_t* off = nullptr;   // [n+1]
    uint32_t* tgt = nullptr;   // [m]
    void* len = nullptr;       // [m], element type dk
    uint32_t* arcs = nullptr;  // [2m] (target, length bits) of every sorted arc, for 4-byte lengths: one 8-byte load
                               // per arc and one range of lines per row instead of two
    uint32_t* perm = nullptr;  // [m]
    double min_len = 0, max_len = 0, mean_len = 0, q05_len = 0;  // over finite lengths
    bool has_zero = false;
    double tuned_delta[2] = {0.0, 0.0};  // per mode (0 SSSP, 1 Brandes); 0 = not tuned
    int tuned_block[2] = {0, 0};         // lanes per workgroup per mode; 0 = not tuned
    double lend_delta = -1.0;
    double bound_inv_q = 0.0;  // scale of the 4-bit distance codes: 15 / (1.1 x largest sampled distance); 0 = not estimated
    bool strong = false;       // the arcs of finite length connect every node to every other (strongly connected)
    bool all_finite = false;   // every length is finite
    RowRange* rows = nullptr;  // [n] row bounds of each node with the end of its light prefix for lend_delta
    std::vector<uint32_t> perm_host;  // host copy of perm: weight updates re-sort only the rows whose order changes
    bool reordered_last = false;      // the last build or weight update changed the row order
};

// Betweenness state of one node in a slot, co-located so that a random touch of a node in the sigma layers
// (sigma + in-degree) and the dependency layers (sigma + dependency + betweenness partial) costs one cache
// line. Distances and frontier stamps stay in compact arrays of their own: the relaxation passes probe them
// far more often, and compact arrays keep more of them cache-resident.
struct alignas(32) BcRec {
    double sigma;   // number of shortest paths
    double delta1;  // dependency (seed 1)
    double bc;      // betweenness partial sum of the slot (kept across sources)
    uint32_t indeg; // tight in-degree still to be released (sigma layers)
    uint32_t spare;
};
static_assert(sizeof(BcRec) == 32, "BcRec must be a quarter of a 128-byte line");

// A candidate tight arc logged by the relaxation passes: arc e from u into t with candidate distance nd (bits).
// The entry holds nd's low 32 bits; 8-byte distances keep their high 32 bits in a parallel array (log_hi).
struct LogEntry {
    uint32_t t;
    uint32_t e;
    uint32_t u;
    uint32_t nd_lo;
};
static_assert(sizeof(LogEntry) == 16, "log entries are written and read as one 16-byte vector");

// Per-slot scratch of the batched kernel, one per context, shared by its graphs and cohorts (calls on a
// context are sequential). Sized for (slots, n, m) = the largest request so far; grown, never shrunk. A
// slot's arrays have stride n; a smaller graph uses a prefix of each.
struct Workspace {
    uint32_t slots = 0, n = 0;
    uint64_t m = 0;
    int mode_bits = 0;      // which optional arrays exist (MODE_* flags)
    uint32_t logcap = 0;
    void* dist = nullptr;       // [slots][n] distance bits (4 or 8 bytes)
    uint32_t* fstamp = nullptr; // [slots][n] phase that last pushed the node to the frontier
    uint32_t* sstamp = nullptr; // [slots][n] settled stamps (used when the LDS bitmap does not fit)
    BcRec* rec = nullptr;       // [slots][n] betweenness state (Brandes modes)
    double* delta2 = nullptr;   // [slots][n] dependency seeded with 1/d^2 (adjoint mode)
    uint32_t* lists = nullptr;  // [slots][3][n]: frontier A/B, settled
    void* pile = nullptr;       // [slots][2][pcap] pile entries (16 bytes reserved per entry)
    uint32_t pcap = 0;
    uint32_t* counters = nullptr;  // [slots][4] persistent stamp generations
    uint32_t* order = nullptr;  // [slots][n]
    uint32_t* layers = nullptr; // [slots][n+1]
    uint32_t* bits = nullptr;   // [slots][words]
    uint64_t words = 0;
    LogEntry* log = nullptr;    // [slots][logcap]
    uint32_t* log_hi = nullptr; // [slots][logcap] high halves of 8-byte candidate distances
    void* summaries = nullptr;  // [max_sources] ng_sssp_summary
    uint32_t summaries_cap = 0;
    uint64_t* arcs_reached = nullptr;  // [16]: arcs reached, phases, buckets, light arcs, settle arcs, logged, layers,
                                       // probes, improvements
    uint32_t* err = nullptr;    // [1]
    uint32_t* sources = nullptr;
    uint32_t sources_cap = 0;
    double* bc_out = nullptr;   // [n]
    double* grad1 = nullptr;    // [m]
    double* grad2 = nullptr;    // [m]
    size_t dist_bytes = 0;         // element size the dist array was allocated for
    DKind dist_kind = DKind::U32;  // distance kind the dist array currently holds +infinity of
    bool dist_valid = false;
    bool bc_zero = false;          // every slot's betweenness partials (rec[].bc) are zero: no zeroing pass before a call
    uint64_t generation = 0;       // bumped whenever a buffer is reallocated (captured launches must re-capture)
};

// Device error bits reported by the batched kernel.
enum : uint32_t {
    ERR_KAHN = 1u,
    ERR_LOG_OVERFLOW = 2u,  // informational: the full-row tight scan was used
    ERR_BUCKET = 4u,        // informational: a settle pass produced in-bucket work
    ERR_GUARD = 8u,         // iteration guard hit: results of that source are incomplete
    ERR_OVERFLOW = 16u      // a queue outgrew its capacity: results of that source are incomplete
};

// One member graph of a cohort launch (device pointers into the graph's |w| view).
struct CohortGraph {
    const uint32_t* off;
    const uint32_t* tgt;
    const void* len;
    const uint32_t* arcs;
    const RowRange* rows;
    const uint32_t* perm;
    uint32_t n;
    uint32_t strong;  // 1: strongly connected (a source ends once every node is settled)
    double delta;     // bucket width (integral for integer lengths)
};

struct LaunchArgs {
    DKind dk;
    int mode;
    uint32_t n;
    const View* view;
    const uint32_t* sources;  // device
    uint32_t n_sources;
    double delta;
    uint32_t grid;            // slots used
    Workspace* ws;
    void* dist_out;           // device, optional
    int dist_out_kind;        // ng_dist_kind
    int maps;                 // extra LDS maps: MAP_SORTED (settle passes in increasing id), MAP_FRONTIER, MAP_BOUND(4)
    double bd_inv_q = 0.0;    // scale of the 4-bit distance codes (MAP_BOUND4)
    int block = kBlock;       // lanes per workgroup: 64, 256, 512 or 1024
    // cohort launch (gtab != nullptr; view, sources and n_sources unused; n = the largest member's n)
    const CohortGraph* gtab = nullptr;
    uint32_t n_graphs = 0;
    const uint32_t* node_prefix = nullptr;  // [n_graphs+1] item prefix (= node prefix) per graph
    double* flush = nullptr;
    const uint64_t* flush_base = nullptr;   // [grid] start of each slot's flush region
};

// Launch wrappers (ng_kernels.hip). All enqueue on `stream`.
hipError_t launch_batch(const LaunchArgs& a, hipStream_t stream);
hipError_t launch_light_end(DKind dk, uint32_t n, const uint32_t* off, const void* len, double delta, RowRange* rows,
                            hipStream_t stream);
// arcs[2k] = tgt[k], arcs[2k+1] = len[k] (4-byte lengths as bits), k < m
hipError_t launch_interleave_arcs(const uint32_t* tgt, const uint32_t* len, uint64_t m, uint32_t* arcs, hipStream_t stream);
// out[v] = sum over the slots of rec[slot][v].bc; with `zero`, each partial is set to 0 once read
hipError_t launch_reduce_slots(BcRec* rec, uint32_t slots, uint32_t n, uint32_t stride, double* out, bool zero, hipStream_t stream);
// Sets every record to the clean state (all zero).
hipError_t launch_rec_init(BcRec* rec, uint64_t count, hipStream_t stream);
// Zeroes the betweenness partial of every record.
hipError_t launch_rec_zero_bc(BcRec* rec, uint64_t count, hipStream_t stream);
hipError_t launch_fill_dist(DKind dk, void* p, uint64_t count, hipStream_t stream);  // fills with +infinity / max
// out[node_prefix[g] + v] = sum over k in [seg_ptr[g], seg_ptr[g+1]) of flush[seg_off[k] + v] (slot order).
hipError_t launch_cohort_reduce(const double* flush, const uint32_t* node_prefix, uint32_t n_graphs, const uint32_t* seg_ptr,
                                const uint64_t* seg_off, uint32_t total_nodes, double* out, hipStream_t stream);
hipError_t launch_fill_u64(uint64_t* p, uint64_t count, uint64_t value, hipStream_t stream);
// counts[min(floor(d / bin_width), n_bins - 1)] += 1 for every finite d = dist[i * n + t], t != sources[i] (dist_kind:
// ng_dist_kind of the rows; unreachable = the kind's max value).
hipError_t launch_distance_histogram(int dist_kind, const void* dist, const uint32_t* sources, uint32_t n_rows, uint32_t n,
                                     double bin_width, uint32_t n_bins, unsigned long long* counts, hipStream_t stream);
hipError_t launch_fill_u32(uint32_t* p, uint64_t count, uint32_t value, hipStream_t stream);
hipError_t launch_fill_f64(double* p, uint64_t count, double value, hipStream_t stream);
hipError_t launch_copy(const void* src, void* dst, uint64_t bytes, hipStream_t stream);  // bytes % 16 == 0
// Resolves the device address of the scratch array on `device`, which must be current (hipGetSymbolAddress, once per
// device; later calls return at once). Called by ng_context_create.
hipError_t scratch_setup(int device);
// The scratch array's device address on `device`, or nullptr before scratch_setup succeeded there.
void* scratch_device_ptr(int device);
// Extra LDS maps beyond the finalized-node bitmap (bit flags of LaunchArgs::maps): bitmaps of each bucket's
// settled nodes and of the next frontier, 16-bit distance bounds per node, and 4-bit distance codes where the
// 16-bit bounds do not fit (MAP_BOUND4 needs LaunchArgs::bd_inv_q).
// MAP_LDIST keeps the current source's distances themselves in LDS (256-lane workgroups; replaces the bounds).
enum : int { MAP_SORTED = 1, MAP_FRONTIER = 2, MAP_BOUND = 4, MAP_BOUND4 = 8, MAP_LDIST = 16 };
// The maps of `maps` that fit next to the static LDS of the kernel with `block` lanes for n nodes of distance kind
// dk, and the bound kind they give (0: none, 1: 16-bit, 2: 4-bit).
int lds_plan_maps(DKind dk, uint32_t n, int maps, int block);
int lds_plan_bound_kind(DKind dk, uint32_t n, int maps, int block);
// Resident workgroups per CU of the batched kernel for (dk, mode, extra LDS bitmaps, lanes per workgroup).
int batch_blocks_per_cu(DKind dk, int mode, uint32_t n, int maps, int block);


}  // namespace ng
