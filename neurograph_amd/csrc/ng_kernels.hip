// HIP kernels for gfx950: multi-source shortest paths, Brandes sigma/delta sweeps and the path adjoint.
//
// One persistent launch serves a whole batch of sources. Workgroup b owns scratch slot b and runs the
// sources b, b + grid, b + 2*grid, ... one after another (a static assignment, so the per-slot
// betweenness partial sums, and therefore the results, do not depend on scheduling).
//
// Per source a workgroup runs delta-stepping (bucket width delta) with the light/heavy arc split:
//   * rows are sorted by length, so the light arcs (length < delta) of a node are a row prefix;
//   * inner phases relax only light arcs of the current frontier until the bucket [lo, hi) is settled;
//   * one settle pass then scans each settled node's row exactly once with its final distance,
//     relaxing heavy arcs (which can only land at or beyond hi) and, for betweenness, logging every arc
//     whose candidate does not exceed the target's current distance: a superset of the tight arcs;
//   * the next bucket starts at the smallest distance parked in the pile (nodes beyond hi).
// Frontier tiles of 256 nodes are staged in LDS with an exclusive scan over their arc counts, so the
// 256 lanes walk consecutive arcs of the concatenated rows (coalesced CSR reads) whatever the degree
// skew; queue appends are wave-aggregated (ballot + popcount, one LDS atomic per wave).
//
// Betweenness: the logged candidates are verified against the final distances (tight arcs set a bit in a
// per-slot arc bitmap and count the target's tight in-degree); Kahn layers over the tight DAG give sigma
// (exact integer sums in double, so order-independent) and a layer order; the reverse layers then pull
// delta(u) = sum over tight arcs (u,t) of sigma(u) * (1 + delta(t)) / sigma(t), the per-arc term of
// BrandesVariant.h:149-154, clearing the bitmap on the way. Nothing is allocated per source: every
// per-slot array is restored to its clean state by touching only the reached nodes.
//
// Reads of data that other lanes modify with atomics go through ldc() (a workgroup-scope relaxed atomic
// load); phase boundaries are bsync() (a workgroup barrier).

#include "ng_internal.h"

#include <algorithm>
#include <cfloat>
#include <cstdint>
#include <mutex>
#include <type_traits>

namespace ng {
namespace {


// ---------------------------------------------------------------------------------------------------
// distance types: distances are stored as their bit patterns so that atomic min on the unsigned bits is
// the min on the value (non-negative IEEE values order like their bits; +inf is the largest).
template <class D>
struct DT;
template <>
struct DT<uint32_t> {
    using B = uint32_t;
    using L = uint32_t;
    static constexpr bool kInt = true;
    static constexpr B kInf = 0xFFFFFFFFu;
    __device__ static B add(B a, L l) { return a + l; }
    __device__ static B sat_add(B a, L l) {
        const uint64_t s = static_cast<uint64_t>(a) + l;
        return s >= kInf ? kInf : static_cast<B>(s);
    }
    __device__ static double val(B b) { return static_cast<double>(b); }
    __device__ static bool light(L l, L delta) { return l < delta; }
};
template <>
struct DT<uint64_t> {
    using B = unsigned long long;
    using L = unsigned long long;
    static constexpr bool kInt = true;
    static constexpr B kInf = ~0ull;
    __device__ static B add(B a, L l) { return a + l; }
    __device__ static B sat_add(B a, L l) { return a > kInf - l ? kInf : a + l; }
    __device__ static double val(B b) { return static_cast<double>(b); }
    __device__ static bool light(L l, L delta) { return l < delta; }
};
template <>
struct DT<float> {
    using B = uint32_t;
    using L = float;
    static constexpr bool kInt = false;
    static constexpr B kInf = 0x7f800000u;
    __device__ static B add(B a, L l) { return __float_as_uint(__uint_as_float(a) + l); }
    __device__ static B sat_add(B a, L l) { return add(a, l); }
    __device__ static double val(B b) { return static_cast<double>(__uint_as_float(b)); }
    __device__ static bool light(L l, L delta) { return l < delta; }
};
template <>
struct DT<double> {
    using B = unsigned long long;
    using L = double;
    static constexpr bool kInt = false;
    static constexpr B kInf = 0x7ff0000000000000ull;
    __device__ static B add(B a, L l) {
        return static_cast<B>(__double_as_longlong(__longlong_as_double(static_cast<long long>(a)) + l));
    }
    __device__ static B sat_add(B a, L l) { return add(a, l); }
    __device__ static double val(B b) { return __longlong_as_double(static_cast<long long>(b)); }
    __device__ static bool light(L l, L delta) { return l < delta; }
};

// ---------------------------------------------------------------------------------------------------
// memory helpers

// Every per-slot array is touched by exactly one workgroup, so workgroup scope is sufficient for all of
// them: the atomics are performed in the XCD's L2 and the coherent loads read there, and no phase
// barrier ever needs an agent-scope fence (which on a multi-XCD device writes back and invalidates L2).
// Only the outputs shared by all workgroups (gradients, counters, error flags) use agent scope.
template <class T>
__device__ __forceinline__ T ldc(const T* p) {
    return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <class T>
__device__ __forceinline__ void stc(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <class T>
__device__ __forceinline__ T wg_min(T* p, T v) {
    return __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <class T>
__device__ __forceinline__ T wg_exch(T* p, T v) {
    return __hip_atomic_exchange(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <class T>
__device__ __forceinline__ T wg_add(T* p, T v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <class T>
__device__ __forceinline__ T wg_sub(T* p, T v) {
    return __hip_atomic_fetch_sub(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <class T>
__device__ __forceinline__ T wg_or(T* p, T v) {
    return __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <class T>
__device__ __forceinline__ T wg_and(T* p, T v) {
    return __hip_atomic_fetch_and(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// Graph arrays live in global memory and are read through pointers of the global address space. A pointer
// loaded from the cohort's graph table would otherwise be generic, and generic (flat) loads also count
// against the LDS wait counter, so every LDS wait would also wait for the loads in flight.
template <class T>
using gptr = const __attribute__((address_space(1))) T*;
template <class T>
__device__ __forceinline__ gptr<T> as_global(const T* p) {
    return (gptr<T>)p;
}
// A node's row record (RowRange: first arc, light-prefix end, end) as one 16-byte load.
using u32x4 = uint32_t __attribute__((ext_vector_type(4)));
using u32x2 = uint32_t __attribute__((ext_vector_type(2)));
// Target and length of sorted arc e: one 8-byte load of the interleaved pair for 4-byte lengths, else two loads.
#ifndef NG_PAIRS_MAX_BS
#define NG_PAIRS_MAX_BS 256
#endif
// Whether a kernel with BS lanes per workgroup and lengths of type L reads the interleaved pairs.
template <class L, int BS>
constexpr bool kPairs = sizeof(L) == 4 && BS <= NG_PAIRS_MAX_BS;
template <class L, int BS>
__device__ __forceinline__ void load_arc(gptr<uint32_t> tgt, gptr<L> len, gptr<u32x2> arcs, uint32_t e, uint32_t& t, L& l) {
    if constexpr (kPairs<L, BS>) {
        const u32x2 p = arcs[e];
        t = p.x;
        // the element is copied to a scalar first: hipcc's __builtin_bit_cast of a vector element (p.y) reads
        // element 0 (checked in the generated code)
        const uint32_t ly = p.y;
        l = __builtin_bit_cast(L, ly);
    } else {
        t = tgt[e];
        l = len[e];
    }
}
__device__ __forceinline__ u32x4 load_row(gptr<RowRange> rows, uint32_t v) { return ((gptr<u32x4>)rows)[v]; }
// Phase boundary: a workgroup barrier with workgroup-scope release/acquire.
__device__ __forceinline__ void bsync() { __syncthreads(); }

// ---------------------------------------------------------------------------------------------------
// workgroup primitives (64-wide waves, BS lanes per workgroup)

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= static_cast<uint32_t>(o)) x += y;
    }
    return x;
}

template <class T, class Op>
__device__ __forceinline__ T wave_reduce(T x, Op op) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x = op(x, __shfl_xor(x, o, 64));
    return x;
}

// Base index of a wave-aggregated append of `n` entries: one LDS atomic by the first active lane, broadcast
// as a scalar (no cross-lane permute).
__device__ __forceinline__ uint32_t wave_reserve(uint32_t* lds_cnt, uint32_t n) {
    uint32_t base = 0;
    if (lane_id() == static_cast<uint32_t>(__ffsll(static_cast<long long>(__activemask())) - 1)) base = atomicAdd(lds_cnt, n);
    return static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(base)));
}

// Wave-aggregated append of `val` for lanes with `pred` into list[cnt++]; all active lanes must call it.
// Appends at or beyond `cap` are dropped (the count still grows, so the caller can detect it).
__device__ __forceinline__ void push(bool pred, uint32_t val, uint32_t* list, uint32_t* lds_cnt, uint32_t cap) {
    const unsigned long long mask = __ballot(pred);
    if (mask == 0ull) return;
    const uint32_t lane = lane_id();
    const uint32_t rank = static_cast<uint32_t>(__popcll(mask & ((1ull << lane) - 1ull)));
    const uint32_t base = wave_reserve(lds_cnt, static_cast<uint32_t>(__popcll(mask)));
    if (pred && base + rank < cap) list[base + rank] = val;
}

// A pile entry: a node parked beyond the current bucket with the distance it had when it was parked.
// Every later improvement that stays beyond the bucket parks it again, so the newest entry of a node
// always carries its current distance and a bucket advance needs no distance loads.
template <class B>
struct PileEntry {
    B val;
    uint32_t node;
};

// Wave-aggregated append of a pile entry (see push()).
template <class B>
__device__ __forceinline__ void push_entry(bool pred, B val, uint32_t node, PileEntry<B>* list, uint32_t* lds_cnt,
                                           uint32_t cap) {
    const unsigned long long mask = __ballot(pred);
    if (mask == 0ull) return;
    const uint32_t lane = lane_id();
    const uint32_t rank = static_cast<uint32_t>(__popcll(mask & ((1ull << lane) - 1ull)));
    const uint32_t base = wave_reserve(lds_cnt, static_cast<uint32_t>(__popcll(mask)));
    if (pred && base + rank < cap) {
        PileEntry<B> e;
        e.val = val;
        e.node = node;
        list[base + rank] = e;
    }
}

template <class B, int BS>
struct Lds {
    static constexpr uint32_t kW = BS / 64;
    uint32_t t_v[BS];
    uint32_t t_beg[BS];
    uint32_t t_off[BS];
    B t_dv[BS];
    // owner (tile index) of the first arc of each chunk of 2^cs tile arcs, plus a sentinel
    uint16_t t_own[2 * BS + 1];
    uint32_t wave[kW];
    uint32_t cnt[8];
    B runmin;
    unsigned long long red_u[3 * kW];
    double red_d[3 * kW];
    // wide workgroups: one queue of 64 filter survivors (target, arc, candidate, tile slot of the arc's tail) per wave
    static constexpr uint32_t kQ = BS >= 512 ? BS : 1;
    uint32_t q_t[kQ];
    uint32_t q_e[kQ];
    B q_n[kQ];
    uint16_t q_o[kQ];
};

template <int BS>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t x, uint32_t* total, uint32_t* lds_wave) {
    const uint32_t inc = wave_incl_scan(x);
    const uint32_t w = threadIdx.x >> 6;
    if (lane_id() == 63) lds_wave[w] = inc;
    __syncthreads();
    uint32_t base = 0, tot = 0;
#pragma unroll
    for (uint32_t i = 0; i < static_cast<uint32_t>(BS / 64); ++i) {
        const uint32_t v = lds_wave[i];
        if (i < w) base += v;
        tot += v;
    }
    __syncthreads();
    *total = tot;
    return base + inc - x;
}

// Expands a list of nodes over per-node arc ranges. node_fn(v, dv&, beg&, cnt&) is called for each listed
// node (by one lane); arc_fn(v, dv, e) is then called once per arc e in [beg, beg+cnt), with the arcs of
// a 256-node tile spread over all lanes. Must be called by the whole workgroup with a uniform count.
template <int BS, class B, class NodeFn, class ArcFn>
__device__ __forceinline__ void expand(const uint32_t* list, uint32_t count, Lds<B, BS>& S, NodeFn node_fn, ArcFn arc_fn) {
    const uint32_t tid = threadIdx.x;
    for (uint32_t base = 0; base < count; base += BS) {
        const uint32_t i = base + tid;
        uint32_t v = 0, beg = 0, cnt = 0;
        B dv = 0;
        if (i < count) {
            v = list[i];
            node_fn(v, dv, beg, cnt);
        }
        uint32_t total;
        const uint32_t ex = block_excl_scan<BS>(cnt, &total, S.wave);
        S.t_v[tid] = v;
        S.t_dv[tid] = dv;
        S.t_beg[tid] = beg;
        S.t_off[tid] = ex;
        __syncthreads();
        for (uint32_t k = tid; k < total; k += BS) {
            uint32_t lo = 0, hi = BS;
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (S.t_off[mid] <= k) lo = mid; else hi = mid;
            }
            arc_fn(S.t_v[lo], S.t_dv[lo], S.t_beg[lo] + (k - S.t_off[lo]));
        }
        __syncthreads();
    }
}

// Arc loop of the relaxation passes with U arcs in flight per lane: owners are located for all U arcs,
// then all U CSR loads are issued, then all U arcs are filtered and all U distance updates (returning
// atomics) issued, and only then are the results applied, so each lane keeps U independent accesses
// outstanding at every step instead of one. The owner of arc k lies between the
// owners of the first arcs of its chunk and of the next chunk (a per-tile table of at most 2 BS chunks), so
// the search is short instead of log2(BS) dependent LDS reads. apply() runs on every active lane (it may
// append to queues with wave ballots), with ok = false for padding slots. tile_fn(valid, dv) runs on every lane
// once the tile's nodes are loaded (valid: the lane holds a node, at distance dv).
#ifndef NG_UNROLL
#define NG_UNROLL 2
#endif
template <int U, int BS, class B, class L, class NodeFn, class TileFn, class FilterFn, class UpdateFn, class ApplyFn>
__device__ __forceinline__ void expand_relax(const uint32_t* list, uint32_t count, Lds<B, BS>& S, gptr<uint32_t> tgt,
                                             gptr<L> len, gptr<u32x2> arcs, NodeFn node_fn, TileFn tile_fn, FilterFn filter,
                                             UpdateFn update, ApplyFn apply) {
    const uint32_t tid = threadIdx.x;
    for (uint32_t base = 0; base < count; base += BS) {
        const uint32_t i = base + tid;
        uint32_t v = 0, beg = 0, cnt = 0;
        B dv = 0;
        if (i < count) {
            v = list[i];
            node_fn(v, dv, beg, cnt);
        }
        tile_fn(i < count, dv);
        uint32_t total;
        const uint32_t ex = block_excl_scan<BS>(cnt, &total, S.wave);
        S.t_v[tid] = v;
        S.t_dv[tid] = dv;
        S.t_beg[tid] = beg;
        S.t_off[tid] = ex;
        // chunks of 2^cs arcs, cs >= 6 and at most 2 BS chunks (uniform: total is the block-wide count)
        uint32_t cs = 6;
        while (((static_cast<uint64_t>(total) + (1ull << cs) - 1ull) >> cs) > 2u * BS) ++cs;
        const uint32_t nq = static_cast<uint32_t>((static_cast<uint64_t>(total) + (1ull << cs) - 1ull) >> cs);
        if (cnt != 0u) {
            // this row owns the chunks whose first arc falls in [ex, ex + cnt)
            const uint64_t end = static_cast<uint64_t>(ex) + cnt;
            for (uint64_t q = (static_cast<uint64_t>(ex) + (1ull << cs) - 1ull) >> cs; (q << cs) < end; ++q)
                S.t_own[q] = static_cast<uint16_t>(tid);
        }
        if (tid == 0) S.t_own[nq] = static_cast<uint16_t>(BS - 1);
        __syncthreads();
        for (uint32_t k0 = tid; k0 < total; k0 += U * BS) {
            uint32_t pv[U], pe[U], pt[U];
            B pdv[U];
            L pl[U];
            bool pok[U], pfin[U];
#pragma unroll
            for (int j = 0; j < U; ++j) {
                const uint32_t k = k0 + static_cast<uint32_t>(j) * BS;
                pok[j] = k < total;
                pv[j] = 0;
                pe[j] = 0;
                pdv[j] = 0;
                pt[j] = 0;
                pl[j] = L(0);
                pfin[j] = true;
                if (pok[j]) {
                    const uint32_t q = k >> cs;
                    uint32_t lo = S.t_own[q], hi = static_cast<uint32_t>(S.t_own[q + 1]) + 1u;
                    while (hi - lo > 1) {
                        const uint32_t mid = (lo + hi) >> 1;
                        if (S.t_off[mid] <= k) lo = mid; else hi = mid;
                    }
                    pv[j] = S.t_v[lo];
                    pdv[j] = S.t_dv[lo];
                    pe[j] = S.t_beg[lo] + (k - S.t_off[lo]);
                }
            }
#pragma unroll
            for (int j = 0; j < U; ++j) {
                if (pok[j]) load_arc<L, BS>(tgt, len, arcs, pe[j], pt[j], pl[j]);
            }
#pragma unroll
            for (int j = 0; j < U; ++j) {
                if (pok[j]) pfin[j] = filter(pt[j], pl[j], pdv[j]);
            }
            B pold[U];
#pragma unroll
            for (int j = 0; j < U; ++j) pold[j] = update(pok[j], pt[j], pdv[j], pl[j], pfin[j]);
#pragma unroll
            for (int j = 0; j < U; ++j) apply(pok[j], pv[j], pdv[j], pe[j], pt[j], pl[j], pfin[j], pold[j]);
        }
        __syncthreads();
    }
}

// Per-node distance bounds in LDS: the distance rounded up to a bfloat16 (the upper 16 bits of an f32), 2 bytes
// per node and scale-free. A stored bound is never below the node's current distance: distances only fall, and
// each improvement stores the bound of its own value, so a late store of an earlier (larger) value is still an
// upper bound. A candidate above a node's bound can neither improve its distance nor be tight, so its global
// distance update is skipped. kBoundInf (+inf) is the bound of an unreached node.
constexpr uint16_t kBoundInf = 0x7F80u;

template <class T>
__device__ __forceinline__ uint16_t bound16(typename T::B b) {
    if constexpr (T::kInt && sizeof(typename T::B) == 8) {
        if (b >= (1ull << 53)) return kBoundInf;  // beyond exact doubles: no bound
    }
    const double d = T::val(b);
    const float f = static_cast<float>(d);
    uint32_t bits = __float_as_uint(f);
    if (static_cast<double>(f) < d) ++bits;  // round up to the next float (d >= 0)
    bits = (bits + 0xFFFFu) & 0xFFFF0000u;   // round up to a bfloat16
    return static_cast<uint16_t>(bits >> 16);
}

// Candidate c is above bound bd. Comparing in double is exact for every bound; a candidate that rounds onto
// the bound is not above it, which only costs an update.
template <class T>
__device__ __forceinline__ bool above_bound(typename T::B c, uint16_t bd) {
    return T::val(c) > static_cast<double>(__uint_as_float(static_cast<uint32_t>(bd) << 16));
}

// Where 2 bytes per node do not fit, 4-bit codes: code(d) = floor(q(d)) capped at 15 (15: no bound), with
// q(d) = (val(d) * inv_q)^2 / 15 and the scale inv_q = 15 / (a little more than the graph's typical largest
// distance). The square spends the codes where the nodes are: most of them lie at the larger distances (a ball
// grows with its radius), where the codes are finest (spatial200k: 11 % fewer probes than with codes linear in
// the distance, 1.5 % less time). A candidate c with q(c) >= code + 1 is above the distance: the rounded products
// are monotone in their argument, so q(c) > q(d) implies c > d, and a code stored for an earlier, larger distance
// only weakens the test. Two nodes share a byte; a byte store that races with its neighbour's may restore the
// neighbour's earlier code, which is still a valid (weaker) bound.
__device__ __forceinline__ uint32_t code4(double x) { return x < 15.0 ? static_cast<uint32_t>(x) : 15u; }

// Arc loop of the relaxation passes (settle passes and frontier phases) in wide workgroups, where few arcs
// survive the LDS filter (bitmap and bounds). Stage A: UA arcs per lane get their owner, CSR entry and
// candidate, and are filtered. The survivors are appended to the wave's LDS queue (ranked slot by slot from
// ballots alone), and stage B (update, apply) runs only when the next survivors would not fit and once at the
// end of the tile, one queued survivor per lane: the settle passes keep a few survivors per iteration, so their
// updates wait for memory once every few iterations instead of every time. More survivors than a queue holds
// run stage B slot by slot. The loop runs wave-uniform iterations, so every lane takes part in each stage B
// (apply uses ballots). Stage B issues the atomic minimum of every survivor at once, without a distance probe
// first: its old value is the probe (the candidates that survive the filter mostly improve their target, so a
// probe would save few atomics and cost a memory round trip). tile_fn as in expand_relax; apply(live, e, t, nd, old, u)
// also receives the arc's tail u.
template <int UA, int BS, class B, class L, class NodeFn, class TileFn, class AddFn, class SkipFn, class UpdateFn, class ApplyFn>
__device__ __forceinline__ void expand_relax_compact(const uint32_t* list, uint32_t count, Lds<B, BS>& S, gptr<uint32_t> tgt,
                                                     gptr<L> len, gptr<u32x2> arcs, NodeFn node_fn, TileFn tile_fn, AddFn add,
                                                     SkipFn skip, UpdateFn update, ApplyFn apply) {
    static_assert(BS >= 512, "the survivor queues are laid out for wide workgroups");
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = lane_id();
    const unsigned long long below = (1ull << lane) - 1ull;
    uint32_t* const qt = S.q_t + (tid & ~63u);
    uint32_t* const qe = S.q_e + (tid & ~63u);
    B* const qn = S.q_n + (tid & ~63u);
    uint16_t* const qo = S.q_o + (tid & ~63u);
    uint32_t qc = 0;  // queued survivors of this wave (wave-uniform)
    // stage B for the queued survivors, one per lane
    auto drain = [&]() {
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");  // the wave's queue stores before its loads
        const bool live = lane < qc;
        uint32_t ct = 0, ce = 0, cu = 0;
        B cnd = 0;
        if (live) {
            ct = qt[lane];
            ce = qe[lane];
            cnd = qn[lane];
            cu = S.t_v[qo[lane]];
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");  // every lane has read before the queue is refilled
        const B old = update(live, ct, cnd);
        apply(live, ce, ct, cnd, old, cu);
        qc = 0;
    };
    for (uint32_t base = 0; base < count; base += BS) {
        const uint32_t i = base + tid;
        uint32_t v = 0, beg = 0, cnt = 0;
        B dv = 0;
        if (i < count) {
            v = list[i];
            node_fn(v, dv, beg, cnt);
        }
        tile_fn(i < count, dv);
        uint32_t total;
        const uint32_t ex = block_excl_scan<BS>(cnt, &total, S.wave);
        S.t_v[tid] = v;
        S.t_dv[tid] = dv;
        S.t_beg[tid] = beg;
        S.t_off[tid] = ex;
        uint32_t cs = 6;
        while (((static_cast<uint64_t>(total) + (1ull << cs) - 1ull) >> cs) > 2u * BS) ++cs;
        const uint32_t nq = static_cast<uint32_t>((static_cast<uint64_t>(total) + (1ull << cs) - 1ull) >> cs);
        if (cnt != 0u) {
            const uint64_t end = static_cast<uint64_t>(ex) + cnt;
            for (uint64_t q = (static_cast<uint64_t>(ex) + (1ull << cs) - 1ull) >> cs; (q << cs) < end; ++q)
                S.t_own[q] = static_cast<uint16_t>(tid);
        }
        if (tid == 0) S.t_own[nq] = static_cast<uint16_t>(BS - 1);
        __syncthreads();
        for (uint32_t kb = tid & ~63u; kb < total; kb += UA * BS) {  // wave-uniform trip count
            const uint32_t k0 = kb + lane;
            uint32_t st[UA], se[UA], so[UA];
            B snd[UA];
            L sl[UA];
            bool sv[UA];
#pragma unroll
            for (int j = 0; j < UA; ++j) {
                const uint32_t k = k0 + static_cast<uint32_t>(j) * BS;
                sv[j] = k < total;
                se[j] = 0;
                so[j] = 0;
                if (sv[j]) {
                    const uint32_t q = k >> cs;
                    uint32_t lo = S.t_own[q], hi = static_cast<uint32_t>(S.t_own[q + 1]) + 1u;
                    while (hi - lo > 1) {
                        const uint32_t mid = (lo + hi) >> 1;
                        if (S.t_off[mid] <= k) lo = mid; else hi = mid;
                    }
                    so[j] = lo;
                    se[j] = S.t_beg[lo] + (k - S.t_off[lo]);
                }
            }
#pragma unroll
            for (int j = 0; j < UA; ++j) {
                st[j] = 0;
                sl[j] = L(0);
                if (sv[j]) load_arc<L, BS>(tgt, len, arcs, se[j], st[j], sl[j]);
            }
#pragma unroll
            for (int j = 0; j < UA; ++j) {
                snd[j] = 0;
                if (sv[j]) {
                    snd[j] = add(S.t_dv[so[j]], sl[j]);
                    sv[j] = !skip(st[j], sl[j], snd[j]);
                }
            }
            unsigned long long m[UA];
            uint32_t o[UA + 1];
            o[0] = 0;
#pragma unroll
            for (int j = 0; j < UA; ++j) {
                m[j] = __ballot(sv[j]);
                o[j + 1] = o[j] + static_cast<uint32_t>(__popcll(m[j]));
            }
            const uint32_t nsv = o[UA];
            if (nsv == 0) continue;
            if (nsv <= 64) {
                if (qc + nsv > 64) drain();
                // slot j's survivors to queue entries [qc + o[j], qc + o[j+1])
#pragma unroll
                for (int j = 0; j < UA; ++j) {
                    if (sv[j]) {
                        const uint32_t q = qc + o[j] + static_cast<uint32_t>(__popcll(m[j] & below));
                        qt[q] = st[j];
                        qe[q] = se[j];
                        qn[q] = snd[j];
                        qo[q] = static_cast<uint16_t>(so[j]);
                    }
                }
                qc += nsv;
            } else {
                B old[UA];
#pragma unroll
                for (int j = 0; j < UA; ++j) old[j] = update(sv[j], st[j], snd[j]);
#pragma unroll
                for (int j = 0; j < UA; ++j) apply(sv[j], se[j], st[j], snd[j], old[j], S.t_v[so[j]]);
            }
        }
        if (qc != 0u) drain();
        __syncthreads();
    }
}

template <class D>
struct KArgs {
    using B = typename DT<D>::B;
    using L = typename DT<D>::L;
    uint32_t n;
    const uint32_t* off;
    const uint32_t* tgt;
    const L* len;
    const uint32_t* arcs;  // interleaved (target, length) pairs, 4-byte lengths only
    const RowRange* rows;
    const uint32_t* perm;
    const uint32_t* sources;
    uint32_t n_sources;
    L delta;
    B* dist;
    uint32_t* fstamp;
    uint32_t* sstamp;
    BcRec* rec;
    double* delta2;
    uint32_t* lists;
    void* pile;
    uint32_t pcap;
    uint32_t strong;  // 1: the graph is strongly connected (a source ends once every node is settled)
    uint32_t* counters;
    uint32_t* order;
    uint32_t* layers;
    uint32_t* bits;
    uint64_t words;
    LogEntry* log;
    uint32_t* log_hi;   // high halves of 8-byte candidate distances, parallel to log
    uint32_t logcap;
    uint32_t stride;    // nodes per slot in the workspace layout (>= n)
    // cohort launch (gtab != nullptr): graph table, item prefix per graph, item range per slot, flush layout
    const CohortGraph* gtab;
    uint32_t n_graphs;
    const uint32_t* node_prefix;
    double* flush;
    const uint64_t* flush_base;
    uint32_t bm_words;  // words of the LDS bitmap of finalized nodes (0: no bitmap)
    uint32_t bd_words;  // words of the LDS distance bounds after the bitmaps (0: no bounds)
    uint32_t bd_kind;   // 1: 16-bit bounds, 2: 4-bit codes (scale bd_inv_q)
    uint32_t ld_off;    // word offset of the distances in dynamic LDS (kernels with LDS distances)
    double bd_inv_q;
    uint32_t maps;      // extra LDS bitmaps: bit 0 orders each bucket's settled nodes by id, bit 1 marks the next frontier
    double* grad1;
    double* grad2;
    ng_sssp_summary* summaries;
    unsigned long long* arcs_reached;
    uint32_t* err;
    void* dist_out;
    int dist_out_kind;
};

template <class D>
__device__ __forceinline__ void write_dist_out(const KArgs<D>& a, size_t idx, typename DT<D>::B b) {
    using T = DT<D>;
    const bool inf = (b == T::kInf);
    switch (a.dist_out_kind) {
    case NG_DIST_U64:
        if constexpr (T::kInt) {
            static_cast<unsigned long long*>(a.dist_out)[idx] = inf ? ~0ull : static_cast<unsigned long long>(b);
        }
        break;
    case NG_DIST_F32:
        static_cast<float*>(a.dist_out)[idx] = inf ? FLT_MAX : static_cast<float>(T::val(b));
        break;
    default:
        static_cast<double*>(a.dist_out)[idx] = inf ? DBL_MAX : T::val(b);
        break;
    }
}

__device__ __forceinline__ uint32_t range_mask(uint32_t wi, uint32_t b, uint32_t e) {
    // bits of word wi that belong to arcs [b, e)
    const uint32_t w0 = wi << 5;
    uint32_t m = 0xFFFFFFFFu;
    if (b > w0) m &= 0xFFFFFFFFu << (b - w0);
    if (e < w0 + 32) m &= (1u << (e - w0)) - 1u;
    return m;
}

// Visits the nonzero words of the tight-arc bitmap covering arcs [b, e_end), in increasing order, with f(wi, w)
// (w masked to the row). Up to four words are loaded before any is visited: they are independent loads, and
// a row of ~100 arcs spans four words.
template <class F>
__device__ __forceinline__ void for_row_words(const uint32_t* bits, uint32_t b, uint32_t e_end, F f) {
    if (b >= e_end) return;
    const uint32_t w1 = (e_end - 1) >> 5;
    for (uint32_t wb = b >> 5; wb <= w1; wb += 4) {
        uint32_t ws[4];
#pragma unroll
        for (uint32_t q = 0; q < 4; ++q) ws[q] = wb + q <= w1 ? ldc(&bits[wb + q]) & range_mask(wb + q, b, e_end) : 0u;
#pragma unroll
        for (uint32_t q = 0; q < 4; ++q)
            if (ws[q] != 0u) f(wb + q, ws[q]);
    }
}

// Optional per-section wall-clock accounting (compile with -DNG_SECTION_TIMERS): workgroup-uniform
// timestamps, summed by lane 0 into arcs_reached[8 + section]. Off in the normal build.
#ifdef NG_SECTION_TIMERS
#define NG_TICK(var) const unsigned long long var = wall_clock64()
#define NG_TOCK(var, sec) \
    do { \
        if (threadIdx.x == 0) atomicAdd(&a.arcs_reached[9 + (sec)], wall_clock64() - (var)); \
    } while (0)
#else
#define NG_TICK(var)
#define NG_TOCK(var, sec)
#endif

// Resident waves per SIMD the hot kernel is compiled for (the register budget follows: 512 / waves VGPRs).
// The path is bound by the latency of random accesses, so occupancy matters more than the few registers
// spilled to fit. Workgroups of 512 or 1024 lanes fill a SIMD with 4 waves from one or two workgroups; 64-lane
// workgroups (one wavefront per source, for small graphs) take the 256-lane kernel's register budget.
#ifndef NG_WAVES_PER_EU
#define NG_WAVES_PER_EU 5
#endif
// Arcs per lane in the filter stage of the wide workgroups' arc loop: frontier phases, and settle passes with 4-
// and 8-byte distances.
#ifndef NG_PHASE_ARCS
#define NG_PHASE_ARCS 3
#endif
#ifndef NG_SETTLE_ARCS_4B
#define NG_SETTLE_ARCS_4B 6
#endif
#ifndef NG_SETTLE_ARCS_8B
#define NG_SETTLE_ARCS_8B 3
#endif
// LD: the current source's distances live in LDS (graphs small enough; 256-lane workgroups), so reads and
// updates are LDS operations; otherwise they live in the slot's global array.
template <class D, int MODE, int BS, bool LD>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(BS <= 256 ? NG_WAVES_PER_EU : 4))) void batch_kernel(
    const KArgs<D> a) {
    using T = DT<D>;
    using B = typename T::B;
    using L = typename T::L;
    __shared__ Lds<B, BS> S;
    constexpr uint32_t kW = BS / 64;
    // Bitmap (dynamic LDS) of the nodes registered in the settled list of the current source (this bucket or
    // an earlier one). It replaces the global settled stamps, and in a settle pass a heavy arc (length >=
    // delta) into a registered node skips its global distance update: the target's distance is final and
    // below hi while the candidate is at least hi, so the arc can neither improve nor be tight.
    // (16-byte aligned: the LDS distances start at an even word offset and are read as 8-byte values)
    extern __shared__ __attribute__((aligned(16))) uint32_t bm[];
    const bool use_bm = a.bm_words != 0;
    // With `sorted`, nodes registered in the current bucket are also marked in bn; before the settle pass
    // the bucket's part of the settled list is written by enumerating bn, i.e. in increasing node id, so
    // that consecutive tiles of the settle pass expand nodes with nearby ids (and, for a locality-ordered
    // numbering, nearby targets).
    const bool sorted = use_bm && (a.maps & 1u) != 0;
    uint32_t* const bn = bm + a.bm_words;
    // With `use_fb`, a third bitmap marks the nodes already pushed to the next frontier (replacing the global
    // frontier stamps): a frontier's bits are cleared when its phase starts, so a node improved again in that
    // phase can be pushed once more.
    const bool use_fb = use_bm && (a.maps & 2u) != 0;
    uint32_t* const fb = bm + a.bm_words * (sorted ? 2u : 1u);
    // With `use_bd`, distance bounds of every node follow the bitmaps: 16-bit (bound16) or 4-bit codes (code4).
    const bool use_bd = use_bm && a.bd_words != 0;
    const bool bd4 = use_bd && a.bd_kind == 2;
    const double inv_q = a.bd_inv_q;
    uint16_t* const fl = reinterpret_cast<uint16_t*>(bm + a.bm_words * (1u + (sorted ? 1u : 0u) + (use_fb ? 1u : 0u)));
    uint8_t* const fl8 = reinterpret_cast<uint8_t*>(fl);
    // candidate c into t is above t's distance
    auto bd_skip = [&](uint32_t t, B c) -> bool {
        if (!use_bd) return false;
        if (bd4) {
            const uint32_t code = (static_cast<uint32_t>(fl8[t >> 1]) >> ((t & 1u) * 4u)) & 15u;
            const double y = T::val(c) * inv_q;
            return code < 15u && y * y * (1.0 / 15.0) >= static_cast<double>(code + 1u);
        }
        return above_bound<T>(c, fl[t]);
    };
    // t's distance was lowered to d
    auto bd_set = [&](uint32_t t, B d) {
        if (!use_bd) return;
        if (bd4) {
            const double y = T::val(d) * inv_q;
            const uint32_t code = code4(y * y * (1.0 / 15.0));
            const uint32_t sh = (t & 1u) * 4u;
            const uint32_t old = fl8[t >> 1];
            if (code < ((old >> sh) & 15u)) fl8[t >> 1] = static_cast<uint8_t>((old & ~(15u << sh)) | (code << sh));
        } else {
            fl[t] = bound16<T>(d);
        }
    };
    auto fb_first = [fb](uint32_t t) -> bool {
        const uint32_t bit = 1u << (t & 31);
        return (atomicOr(&fb[t >> 5], bit) & bit) == 0u;
    };

    const uint32_t tid = threadIdx.x;
    const uint32_t slot = blockIdx.x;
    // slot layout uses the workspace stride; the graph's own size (n) and arrays can change per source in a
    // cohort launch
    const uint32_t stride = a.stride;
    const size_t so = static_cast<size_t>(slot) * stride;
    uint32_t n = a.n;
    gptr<uint32_t> off = as_global(a.off);
    gptr<uint32_t> tgt = as_global(a.tgt);
    gptr<L> len = as_global(a.len);
    gptr<u32x2> arcs = (gptr<u32x2>)as_global(a.arcs);
    // the target of sorted arc e, from the interleaved pairs where they exist: the passes then read one copy of the
    // graph's arcs, which the last-level cache can hold (spatial200k: 160 MB of pairs; pairs plus the separate
    // targets would be 240 MB of the 256 MB MALL)
    auto target_of = [&](uint32_t e) -> uint32_t {
        if constexpr (kPairs<L, BS>) return arcs[e].x;
        else return tgt[e];
    };
    gptr<RowRange> rows = as_global(a.rows);
    gptr<uint32_t> perm = as_global(a.perm);
    uint32_t bmw = a.bm_words;  // bitmap words of the current graph
    B* const dist = LD ? reinterpret_cast<B*>(bm + a.ld_off) : a.dist + so;
    auto DP = [dist](uint32_t v) -> B* { return dist + v; };
    uint32_t* const fstamp = a.fstamp + so;
    uint32_t* sstamp = a.sstamp + so;
    BcRec* const R = (MODE != MODE_SSSP) ? a.rec + so : nullptr;
    double* const delta2 = ((MODE & MODE_ARC2) != 0) ? a.delta2 + so : nullptr;
    // Near trees: Brandes sources whose tight arcs form a tree except at a few nodes with several tight in-arcs
    // (f32 rounding makes two path lengths equal now and then: 13 % of the 200k-node graph's sources). They keep
    // the tree layers' costs: a scan of the internal nodes' tight arcs marks each head in an LDS bitmap and lists
    // the heads it finds marked already; only those nodes take sigma sums and in-degrees. Wide workgroups (their
    // survivor queues hold the lists), Brandes mode, f32 distances (f64 ties are rare, and the kernels of the other
    // types measured slower with the extra code), LDS bounds (their region holds the bitmaps once verified); sigma
    // by order position goes to the second-dependency array, idle in this mode.
    constexpr bool kNear = MODE == MODE_BC && BS >= 512 && !LD && std::is_same_v<D, float>;
    const bool near_ok = kNear && use_bd && a.delta2 != nullptr;
    double* const sigp = near_ok ? a.delta2 + so : nullptr;
    uint32_t* lists = a.lists + so * 3;
    uint32_t* const Fa = lists;
    uint32_t* const Fb = lists + stride;
    uint32_t* const Sl = lists + 2 * static_cast<size_t>(stride);
    const uint32_t pcap = a.pcap;
    PileEntry<B>* const Pa = static_cast<PileEntry<B>*>(a.pile) + static_cast<size_t>(slot) * 2 * pcap;
    PileEntry<B>* const Pb = Pa + pcap;
    uint32_t* order = nullptr;
    uint32_t* layers = nullptr;
    uint32_t* bits = nullptr;
    LogEntry* log = nullptr;
    uint32_t* log_hi = nullptr;
    if constexpr (MODE != MODE_SSSP) {
        order = a.order + so;
        layers = a.layers + static_cast<size_t>(slot) * (stride + 1);
        bits = a.bits + static_cast<size_t>(slot) * a.words;
        log = a.log + static_cast<size_t>(slot) * a.logcap;
        if constexpr (sizeof(B) == 8) log_hi = a.log_hi + static_cast<size_t>(slot) * a.logcap;
    }
    uint32_t* ctr = a.counters + static_cast<size_t>(slot) * 4;
    uint32_t phase = ctr[0], gen = ctr[1], pgen = ctr[2];
    L delta = a.delta;
    // On a strongly connected graph every node is reached: once all are settled, and the current bucket is
    // [lo, lo + delta) with lo + delta > lo, the settle pass needs no heavy arcs (each candidate is at least hi,
    // above every distance, so it can neither improve a distance nor be tight) and the pile holds only stale
    // entries: the source ends without reading them.
    uint32_t g_strong = a.strong;
    // work counters (uniform per workgroup, except the arc counts which lane 0 accumulates from the tiles)
    unsigned long long c_phases = 0, c_buckets = 0, c_logged = 0, c_layers = 0;
    unsigned long long c_trees = 0;  // sources whose tight arcs formed a shortest-path tree
    unsigned long long c_near = 0;   // sources whose tight arcs formed a near tree
    // per-lane counts of one launch (32 bits: a lane handles far fewer than 2^32 arcs per launch)
    uint32_t c_light = 0, c_settle = 0, c_probe = 0, c_improve = 0;

    // Sources blockIdx.x, +gridDim.x, ... of a.sources (plain launch), or of the cohort's (graph, source) items
    // in graph order (cohort launch, a.gtab != nullptr): all slots sweep the member graphs together, so the
    // CSR in use stays cache-resident and the tail of one graph overlaps the start of the next. When a slot's
    // graph changes, its betweenness partials of the previous graph are flushed to the slot's flush region
    // (and cleared); a reduction sums each graph's regions in slot order.
    const bool cohort = a.gtab != nullptr;
    uint32_t cur_g = 0xFFFFFFFFu;
    size_t fpos = cohort ? a.flush_base[slot] : 0;
    auto flush_bc = [&]() {
        if constexpr ((MODE & MODE_BC) != 0) {
            for (uint32_t v = tid; v < n; v += BS) {
                a.flush[fpos + v] = R[v].bc;
                R[v].bc = 0.0;
            }
            fpos += n;
        }
    };
    for (uint32_t si = blockIdx.x; si < a.n_sources; si += gridDim.x) {
        uint32_t s;
        if (cohort) {
            uint32_t lo = 0, hi = a.n_graphs;  // graph g holds items [node_prefix[g], node_prefix[g+1])
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (a.node_prefix[mid] <= si) lo = mid; else hi = mid;
            }
            if (lo != cur_g) {
                if (cur_g != 0xFFFFFFFFu) flush_bc();
                const CohortGraph& cg = a.gtab[lo];
                n = cg.n;
                off = as_global(cg.off);
                tgt = as_global(cg.tgt);
                len = as_global(static_cast<const L*>(cg.len));
                arcs = (gptr<u32x2>)as_global(cg.arcs);
                rows = as_global(cg.rows);
                perm = as_global(cg.perm);
                delta = static_cast<L>(cg.delta);
                g_strong = cg.strong;
                bmw = use_bm ? (n + 31) / 32 : 0u;
                cur_g = lo;
                bsync();
            }
            s = si - a.node_prefix[lo];
        } else {
            s = a.sources[si];
        }
        if (phase > 0xE0000000u || gen > 0xE0000000u || pgen > 0xE0000000u) {
            for (uint32_t v = tid; v < stride; v += BS) {
                stc(&fstamp[v], 0u);
                stc(&sstamp[v], 0u);
            }
            phase = gen = pgen = 1;
            bsync();
        }
        ++gen;
        ++phase;
        if (use_bm)
            for (uint32_t w = tid; w < bmw; w += BS) bm[w] = 0u;
        if (sorted)
            for (uint32_t w = tid; w < bmw; w += BS) bn[w] = 0u;
        if (use_fb)
            for (uint32_t w = tid; w < bmw; w += BS) fb[w] = 0u;
        if constexpr (LD) {
            for (uint32_t v = tid; v < n; v += BS) dist[v] = T::kInf;
        }
        if (use_bd) {
            uint32_t* const flw = reinterpret_cast<uint32_t*>(fl);
            const uint32_t bdw = bd4 ? (n + 7) / 8 : (n + 1) / 2;
            const uint32_t fillv = bd4 ? 0xFFFFFFFFu : (static_cast<uint32_t>(kBoundInf) << 16) | kBoundInf;
            for (uint32_t w = tid; w < bdw; w += BS) flw[w] = fillv;
        }
        if (tid < 3 * kW) {  // the waves' summary sums (the previous source's were read before its reset)
            S.red_u[tid] = 0;
            S.red_d[tid] = 0.0;
        }
        bsync();
        if (tid == 0) {
#pragma unroll
            for (int k = 0; k < 8; ++k) S.cnt[k] = 0;
            S.runmin = T::kInf;
            stc(DP(s), B(0));
            bd_set(s, B(0));
            Fa[0] = s;
        }
        uint32_t* Fc = Fa;
        uint32_t* Fn = Fb;
        PileEntry<B>* Pc = Pa;
        PileEntry<B>* Pn = Pb;
        uint32_t nF = 1;
        uint32_t settle_from = 0;
        B hi = T::sat_add(B(0), delta);
        // heavy arcs land at or beyond hi only while lo + delta does not round back to lo
        bool skip_ok = true;
        // Uniform iteration guard: phases and bucket advances are bounded, so a runaway loop (a bug, or a
        // pathological input) ends with ERR_GUARD instead of occupying the device.
        uint32_t guard = 0;
        const uint32_t guard_max = guard_limit(n);
        bool aborted = false;
        // set when the Brandes tree layers left every record but the source's sigma restored (uniform)
        bool rec_clean = false;
        auto clampn = [&](uint32_t c) -> uint32_t {
            if (c > n) {
                if (tid == 0) atomicOr(a.err, ERR_OVERFLOW);
                aborted = true;
                return 0u;
            }
            return c;
        };
        bsync();

        // filter of arc (v, t) with length l from v at distance dv (fin: skipped): a heavy arc into a registered
        // node, or a candidate above the target's bound. The arcs that pass take part in the distance update.
        auto filter = [&](uint32_t t, L l, B dv) -> bool {
            bool fin = use_bm && skip_ok && !(l < delta) && ((bm[t >> 5] >> (t & 31)) & 1u);
            if (!fin) fin = bd_skip(t, T::add(dv, l));
            c_probe += fin ? 0u : 1u;
            return fin;
        };
        // distance update of arc (v, t) with length l from v at distance dv that passed the filter: the atomic
        // minimum, whose old value (t's previous distance) also serves as the probe; the candidate itself for a
        // padding slot or a filtered arc
        auto update = [&](bool ok, uint32_t t, B dv, L l, bool final_t) -> B {
            const B nd = T::add(dv, l);
            return (ok && !final_t) ? wg_min(DP(t), nd) : nd;
        };
        // relaxation of arc e into t with candidate distance nd, given the update (live: updated; old: t's
        // distance before it); new frontier entries are stamped `nph`. In the betweenness modes every candidate
        // that could be tight (nd <= old) is logged: light arcs by the frontier phases (a node is expanded again
        // after each improvement, so also with its final distance), heavy arcs by the settle passes.
        auto apply_core = [&](bool live, uint32_t e, uint32_t t, B nd, B old, uint32_t nph, uint32_t u, bool heavy) {
            bool pf = false, pp = false, parked = false;
            if (nd < old) {
                ++c_improve;
                bd_set(t, nd);
                if (nd < hi) {
                    pf = use_fb ? fb_first(t) : wg_exch(&fstamp[t], nph) != nph;
                } else {
                    parked = true;
                    // first parking of t in this source (its first reach: every later reach finds a finite
                    // distance) always fits, the pile keeps n entries in reserve; a re-parking with the
                    // improved value only while the reserve is untouched, else the next bucket advance
                    // reloads the distances (exact mode)
                    if (old == T::kInf) {
                        pp = true;
                    } else if (S.cnt[1] + BS < pcap - n) {
                        pp = true;
                    } else {
                        S.cnt[5] = 1u;
                    }
                }
            }
            push(pf, t, Fn, &S.cnt[0], n);
            push_entry(pp, nd, t, Pc, &S.cnt[1], pcap);
            if (__ballot(parked) != 0ull) {  // the smallest parked value, one LDS atomic per wave
                auto minb = [](B x, B y) { return x < y ? x : y; };
                const B m = wave_reduce(parked ? nd : T::kInf, minb);
                if (lane_id() == 0) atomicMin(&S.runmin, m);
            }
            if constexpr (MODE != MODE_SSSP) {
                const bool lg = live && nd <= old;
                const unsigned long long mask = __ballot(lg);
                if (mask != 0ull) {
                    const uint32_t lane = lane_id();
                    const uint32_t rank = static_cast<uint32_t>(__popcll(mask & ((1ull << lane) - 1ull)));
                    // light arcs (frontier phases, possibly logged more than once) fill the log from the front,
                    // heavy arcs (settle passes, each examined once per source) from the back
                    const uint32_t base = wave_reserve(heavy ? &S.cnt[4] : &S.cnt[3], static_cast<uint32_t>(__popcll(mask)));
                    const uint32_t pos = heavy ? a.logcap - 1u - (base + rank) : base + rank;
                    // the log is written once here and read once by the verification: non-temporal stores (and
                    // loads there), so its lines leave the caches first and the distances and rows stay longer
                    if (lg && base + rank < a.logcap) {
                        const unsigned long long ndu = static_cast<unsigned long long>(nd);
                        u32x4 q;  // LogEntry{t, e, u, low half of nd} as one 16-byte store
                        q.x = t;
                        q.y = e;
                        q.z = u;
                        q.w = static_cast<uint32_t>(ndu);
                        __builtin_nontemporal_store(q, reinterpret_cast<u32x4*>(log + pos));
                        if constexpr (sizeof(B) == 8) __builtin_nontemporal_store(static_cast<uint32_t>(ndu >> 32), log_hi + pos);
                    }
                }
            }
        };
        auto apply = [&](bool ok, uint32_t v, B dv, uint32_t e, uint32_t t, L l, bool final_t, B old, uint32_t nph, bool heavy) {
            apply_core(ok && !final_t, e, t, T::add(dv, l), old, nph, v, heavy);
        };
        // the same steps for the compacted arc loop, given the candidate: the filter (heavy arc into a registered
        // node, or candidate above the target's bound) and the update
        auto skip_arc = [&](uint32_t t, L l, B nd) -> bool {
            if (use_bm && skip_ok && !(l < delta) && ((bm[t >> 5] >> (t & 31)) & 1u)) return true;
            return bd_skip(t, nd);
        };
        // the update of a survivor: t's previous distance (the candidate itself for a lane with no survivor)
        auto update_nd = [&](bool live, uint32_t t, B nd) -> B {
            if (!live) return nd;
            ++c_probe;
            return wg_min(DP(t), nd);
        };

        while (true) {
            // ---- inner phases: light arcs of the frontier until the bucket [lo, hi) is settled
            while (nF > 0) {
                if (++guard > guard_max) {
                    if (tid == 0) atomicOr(a.err, ERR_GUARD);
                    aborted = true;
                    break;
                }
                const uint32_t nph = phase + 1;
                ++c_phases;
                NG_TICK(t_inner);
                if (use_fb) {
                    for (uint32_t i = tid; i < nF; i += BS) {
                        const uint32_t v = Fc[i];
                        atomicAnd(&fb[v >> 5], ~(1u << (v & 31)));
                    }
                    bsync();
                }
                auto light_row = [&](uint32_t v, B& dv, uint32_t& beg, uint32_t& cnt) {
                    dv = ldc(DP(v));
                    const u32x4 r = load_row(rows, v);  // x: first arc, y: light-prefix end
                    beg = r.x;
                    cnt = r.y - r.x;
                    c_light += cnt;
                };
                if constexpr (BS >= 512) {
                    expand_relax_compact<NG_PHASE_ARCS, BS>(
                        Fc, nF, S, tgt, len, arcs, light_row, [](bool, B) {}, [](B dv, L l) -> B { return T::add(dv, l); }, skip_arc, update_nd,
                        [&](bool live, uint32_t e, uint32_t t, B nd, B old, uint32_t u) { apply_core(live, e, t, nd, old, nph, u, false); });
                } else {
                    expand_relax<NG_UNROLL, BS>(
                        Fc, nF, S, tgt, len, arcs, light_row, [](bool, B) {}, filter, update,
                        [&](bool ok, uint32_t v, B dv, uint32_t e, uint32_t t, L l, bool fin, B old) {
                            apply(ok, v, dv, e, t, l, fin, old, nph, false);
                        });
                }
                // first visit in this bucket -> settled list
                for (uint32_t base = 0; base < nF; base += BS) {
                    const uint32_t i = base + tid;
                    bool fresh = false;
                    uint32_t v = 0;
                    if (i < nF) {
                        v = Fc[i];
                        if (use_bm) {
                            const uint32_t bit = 1u << (v & 31);
                            fresh = (atomicOr(&bm[v >> 5], bit) & bit) == 0u;
                            if (fresh && sorted) atomicOr(&bn[v >> 5], bit);
                        } else {
                            fresh = wg_exch(&sstamp[v], gen) != gen;
                        }
                    }
                    // sorted: only counted here; the list is written from bn before the settle pass
                    push(fresh, v, Sl, &S.cnt[2], sorted ? 0u : n);
                }
                bsync();
                nF = clampn(S.cnt[0]);
                uint32_t* tmp = Fc;
                Fc = Fn;
                Fn = tmp;
                phase = nph;
                bsync();
                if (tid == 0) S.cnt[0] = 0;
                bsync();
                NG_TOCK(t_inner, 0);
            }
            if (aborted) break;
            // ---- settle pass: the heavy arcs of every settled node once, with its final distance
            const uint32_t nS = clampn(S.cnt[2]);
            if (aborted) break;
            const bool finishing = g_strong != 0u && nS == n && skip_ok;  // every node settled: summary only
            if (sorted && nS > settle_from) {
                uint32_t pos = settle_from;
                for (uint32_t wb = 0; wb < bmw; wb += BS) {
                    const uint32_t wi = wb + tid;
                    uint32_t w = wi < bmw ? bn[wi] : 0u;
                    uint32_t tot;
                    const uint32_t ex = block_excl_scan<BS>(static_cast<uint32_t>(__popc(w)), &tot, S.wave);
                    if (w != 0u) {
                        bn[wi] = 0u;
                        uint32_t k = pos + ex;
                        while (w) {
                            const uint32_t b = static_cast<uint32_t>(__ffs(w) - 1);
                            w &= w - 1;
                            if (k < n) Sl[k] = (wi << 5) + b;
                            ++k;
                        }
                    }
                    pos += tot;
                    if (pos >= nS) break;  // uniform: tot is the block-wide count
                }
                bsync();
            }
            if (nS > settle_from) {
                NG_TICK(t_settle);
                const uint32_t nph = phase + 1;
                // heavy arcs only: the light ones were relaxed, and their candidates logged, by the phases
                uint32_t deg = 0;  // out-degree of the lane's node
                auto settle_row = [&](uint32_t v, B& dv, uint32_t& beg, uint32_t& cnt) {
                    dv = ldc(DP(v));
                    const u32x4 r = load_row(rows, v);  // x: first arc, y: light-prefix end, z: end
                    beg = r.y;
                    cnt = finishing ? 0u : r.z - beg;
                    c_settle += cnt;
                    deg = r.z - r.x;
                };
                // Every reached node is settled exactly once, with its final distance: the per-source summary is
                // accumulated per tile (wave sums added to the wave's summary slots) instead of in a pass of its own
                // over the distances in global memory. With LDS distances that pass is cheap and many small tiles
                // would each pay the reductions, so the summary keeps its own pass.
                auto settle_tile = [&](bool valid, B dv) {
                    if constexpr (LD) return;
                    unsigned long long su = 0, mx = 0, arcs = valid ? deg : 0u;
                    double sd = 0.0, se = 0.0;
                    if (valid) {
                        if constexpr (T::kInt) su = static_cast<unsigned long long>(dv);
                        else sd = T::val(dv);
                        if (dv != B(0)) se = 1.0 / T::val(dv);
                        mx = static_cast<unsigned long long>(dv);
                    }
                    auto addu = [](unsigned long long x, unsigned long long y) { return x + y; };
                    auto maxu = [](unsigned long long x, unsigned long long y) { return x > y ? x : y; };
                    auto addd = [](double x, double y) { return x + y; };
                    if constexpr (T::kInt) su = wave_reduce(su, addu);
                    else sd = wave_reduce(sd, addd);
                    se = wave_reduce(se, addd);
                    mx = wave_reduce(mx, maxu);
                    arcs = wave_reduce(arcs, addu);
                    const uint32_t w = tid >> 6;
                    if (lane_id() == 0) {
                        S.red_u[w] += su;
                        S.red_u[kW + w] = mx > S.red_u[kW + w] ? mx : S.red_u[kW + w];
                        S.red_u[2 * kW + w] += arcs;
                        S.red_d[w] += sd;
                        S.red_d[kW + w] += se;
                    }
                };
                if constexpr (BS >= 512) {
                    // wide workgroups: compacted survivors (the 256-lane kernel's register budget is too small)
                    // arcs per lane in the filter stage: fewer for 8-byte distances (registers; measured 3 vs 4:
                    // +1 % on spatial50k), more for 4-byte ones (6 vs 4: +0.6 % on spatial200k)
                    constexpr int kFilterArcs = sizeof(B) == 8 ? NG_SETTLE_ARCS_8B : NG_SETTLE_ARCS_4B;
                    expand_relax_compact<kFilterArcs, BS>(
                        Sl + settle_from, nS - settle_from, S, tgt, len, arcs, settle_row, settle_tile,
                        [](B dv, L l) -> B { return T::add(dv, l); }, skip_arc, update_nd,
                        [&](bool live, uint32_t e, uint32_t t, B nd, B old, uint32_t u) { apply_core(live, e, t, nd, old, nph, u, true); });
                } else {
                    expand_relax<NG_UNROLL, BS>(
                        Sl + settle_from, nS - settle_from, S, tgt, len, arcs, settle_row, settle_tile, filter, update,
                        [&](bool ok, uint32_t v, B dv, uint32_t e, uint32_t t, L l, bool fin, B old) {
                            apply(ok, v, dv, e, t, l, fin, old, nph, true);
                        });
                }
                bsync();
                NG_TOCK(t_settle, 1);
                settle_from = nS;
                nF = clampn(S.cnt[0]);
                if (aborted) break;
                if (nF > 0) {  // cannot happen with monotone rounding; handled for safety
                    if (tid == 0) atomicOr(a.err, ERR_BUCKET);
                    uint32_t* tmp = Fc;
                    Fc = Fn;
                    Fn = tmp;
                    phase = nph;
                    bsync();
                    if (tid == 0) S.cnt[0] = 0;
                    bsync();
                    continue;
                }
            }
            if (finishing) break;  // the pile holds only entries of settled nodes
            const uint32_t nP = S.cnt[1] > pcap ? (clampn(n + 1), 0u) : S.cnt[1];
            if (aborted || nP == 0) break;
            const bool exact = S.cnt[5] != 0u;
            if (++guard > guard_max) {
                if (tid == 0) atomicOr(a.err, ERR_GUARD);
                aborted = true;
                break;
            }
            // ---- next bucket: [max(pile minimum, hi), +delta)
            NG_TICK(t_split);
            ++c_buckets;
            const B rm = S.runmin;
            const B lo = rm > hi ? rm : hi;
            hi = T::sat_add(lo, delta);
            skip_ok = hi > lo;
            // lo + delta may round back to lo for real-valued distances: the window must not be empty,
            // or the pile would never drain. The next representable value is bits + 1.
            if (hi <= lo && lo != T::kInf) hi = lo + 1;
            const uint32_t nph = phase + 1;
            bsync();
            if (tid == 0) {
                S.runmin = T::kInf;
                S.cnt[0] = 0;
                S.cnt[1] = 0;
                S.cnt[5] = 0;
            }
            bsync();
            // split the pile: entries of settled nodes are dropped, values below hi join the frontier
            // (duplicates removed by the frontier stamp), the rest is carried over with its value. Each lane
            // has kSplit entries in flight; the minimum carried value is reduced per wave (one LDS atomic).
            constexpr uint32_t kSplit = 2;
            B lmin = T::kInf;
            for (uint32_t base = 0; base < nP; base += kSplit * BS) {
                PileEntry<B> pe[kSplit];
#pragma unroll
                for (uint32_t j = 0; j < kSplit; ++j) {
                    const uint32_t i = base + j * BS + tid;
                    if (i < nP) pe[j] = Pc[i];
                }
#pragma unroll
                for (uint32_t j = 0; j < kSplit; ++j) {
                    const uint32_t i = base + j * BS + tid;
                    bool tf = false, tp = false;
                    uint32_t t = 0;
                    B d = 0;
                    if (i < nP) {
                        t = pe[j].node;
                        const bool settled = use_bm ? ((bm[t >> 5] >> (t & 31)) & 1u) != 0u : ldc(&sstamp[t]) == gen;
                        if (!settled) {
                            d = exact ? ldc(DP(t)) : pe[j].val;
                            if (d < hi) {
                                tf = use_fb ? fb_first(t) : wg_exch(&fstamp[t], nph) != nph;
                            } else {
                                lmin = d < lmin ? d : lmin;
                                tp = true;
                            }
                        }
                    }
                    push(tf, t, Fn, &S.cnt[0], n);
                    push_entry(tp, d, t, Pn, &S.cnt[1], pcap);
                }
            }
            {
                auto minb = [](B x, B y) { return x < y ? x : y; };
                lmin = wave_reduce(lmin, minb);
                if (lane_id() == 0 && lmin != T::kInf) atomicMin(&S.runmin, lmin);
            }
            bsync();
            nF = clampn(S.cnt[0]);
            if (aborted) break;
            {
                uint32_t* tmp = Fc;
                Fc = Fn;
                Fn = tmp;
                PileEntry<B>* ptmp = Pc;
                Pc = Pn;
                Pn = ptmp;
            }
            phase = nph;
            bsync();
            if (tid == 0) S.cnt[0] = 0;
            bsync();
            NG_TOCK(t_split, 2);
        }

        NG_TICK(t_summary);
        const uint32_t nS = S.cnt[2] > n ? n : S.cnt[2];
        // Where at least half of the nodes were reached, the tight-arc bitmap is cleared by one coalesced pass at
        // the slot restore (every line of it holds tight arcs of some internal node) instead of an atomic and
        // per word of each internal node's row where the Brandes layers read it.
        const bool bits_linear = 2ull * nS >= n;
        // ---- per-source summary (SsspResult.h:43-64) over the reached nodes: the waves' sums of the settle passes,
        // or with LDS distances a pass over the reached nodes
        {
            if constexpr (LD) {
                unsigned long long su = 0, mx = 0, arcs = 0;
                double sd = 0.0, se = 0.0;
                for (uint32_t i = tid; i < nS; i += BS) {
                    const uint32_t v = Sl[i];
                    const B d = ldc(DP(v));
                    if constexpr (T::kInt) su += static_cast<unsigned long long>(d);
                    else sd += T::val(d);
                    if (d != B(0)) se += 1.0 / T::val(d);
                    mx = (static_cast<unsigned long long>(d) > mx) ? static_cast<unsigned long long>(d) : mx;
                    const u32x4 r = load_row(rows, v);
                    arcs += r.z - r.x;
                }
                auto addu = [](unsigned long long x, unsigned long long y) { return x + y; };
                auto maxu = [](unsigned long long x, unsigned long long y) { return x > y ? x : y; };
                auto addd = [](double x, double y) { return x + y; };
                su = wave_reduce(su, addu);
                mx = wave_reduce(mx, maxu);
                arcs = wave_reduce(arcs, addu);
                sd = wave_reduce(sd, addd);
                se = wave_reduce(se, addd);
                const uint32_t w = tid >> 6;
                if (lane_id() == 0) {
                    S.red_u[w] = su;
                    S.red_u[kW + w] = mx;
                    S.red_u[2 * kW + w] = arcs;
                    S.red_d[w] = sd;
                    S.red_d[kW + w] = se;
                }
                __syncthreads();
            }
            if (tid == 0) {
                unsigned long long tsu = 0, tmx = 0, tar = 0;
                double tsd = 0.0, tse = 0.0;
                for (uint32_t k = 0; k < kW; ++k) {
                    tsu += S.red_u[k];
                    tmx = S.red_u[kW + k] > tmx ? S.red_u[kW + k] : tmx;
                    tar += S.red_u[2 * kW + k];
                    tsd += S.red_d[k];
                    tse += S.red_d[kW + k];
                }
                ng_sssp_summary r;
                if constexpr (T::kInt) {
                    r.sum_d = static_cast<double>(tsu);
                    r.sum_d_u64 = tsu;
                    r.max_d = static_cast<double>(tmx);
                    r.max_d_u64 = tmx;
                } else {
                    r.sum_d = tsd;
                    r.sum_d_u64 = 0;
                    r.max_d = T::val(static_cast<B>(tmx));
                    r.max_d_u64 = 0;
                }
                r.sum_eff = tse;
                r.n_unreached = static_cast<uint64_t>(n) - nS;
                a.summaries[si] = r;
                atomicAdd(a.arcs_reached, tar);
            }
        }

        NG_TOCK(t_summary, 3);
        if constexpr (MODE != MODE_SSSP) if (!aborted) {
            NG_TICK(t_verify);
            // ---- tight arcs: verify the logged candidates against the final distances
            const uint32_t nlog_l = S.cnt[3];  // light-arc entries [0, nlog_l)
            const uint32_t nlog_h = S.cnt[4];  // heavy-arc entries [logcap - nlog_h, logcap)
            const bool log_ok = static_cast<uint64_t>(nlog_l) + nlog_h <= a.logcap;
            c_logged += static_cast<unsigned long long>(nlog_l) + nlog_h;
            // arc e into t is tight when its candidate c equals the final distance of t
            uint32_t n_tight = 0;
            auto tight = [&](uint32_t t, B c) -> bool { return !bd_skip(t, c) && c == ldc(DP(t)); };
            // Once the distances are final the settled-node bitmap is free; it becomes the map of the nodes with
            // a tight out-arc (`hc`). About two thirds to three quarters of a shortest-path tree's nodes are
            // leaves, and the layers below skip a leaf's row bounds, bitmap words and dependency record.
            uint32_t* const hc = bm;
            if (use_bm) {
                for (uint32_t w = tid; w < bmw; w += BS) hc[w] = 0u;
                bsync();
            }
            auto internal = [&](uint32_t u) -> bool { return !use_bm || ((hc[u >> 5] >> (u & 31)) & 1u) != 0u; };
            // the frontier phases may log a light arc more than once: an arc counts when its bit is new
            auto mark = [&](uint32_t t, uint32_t e, B c, uint32_t u) {
                if (tight(t, c)) {
                    const uint32_t bit = 1u << (e & 31);
                    if ((wg_or(&bits[e >> 5], bit) & bit) == 0u) {
                        ++n_tight;
                        if (use_bm) atomicOr(&hc[u >> 5], 1u << (u & 31));
                    }
                }
            };
            // a heavy arc is logged at most once: its bit is set without waiting for the old word
            auto mark_once = [&](uint32_t t, uint32_t e, B c, uint32_t u) {
                if (tight(t, c)) {
                    __hip_atomic_fetch_or(&bits[e >> 5], 1u << (e & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    ++n_tight;
                    if (use_bm) atomicOr(&hc[u >> 5], 1u << (u & 31));
                }
            };
            // every candidate (from the log: light entries through fn, heavy ones through fn_once; or from the full
            // rows through fn when the log overflowed)
            auto each_candidate = [&](auto fn, auto fn_once) {
                if (log_ok) {
                    auto entry = [&](uint32_t i, auto f) {
                        const u32x4 q = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(log + i));  // t, e, u, nd
                        unsigned long long nd = q.w;
                        if constexpr (sizeof(B) == 8)
                            nd |= static_cast<unsigned long long>(__builtin_nontemporal_load(log_hi + i)) << 32;
                        f(q.x, q.y, static_cast<B>(nd), q.z);
                    };
                    for (uint32_t i = tid; i < nlog_l; i += BS) entry(i, fn);
                    for (uint32_t i = a.logcap - nlog_h + tid; i < a.logcap; i += BS) entry(i, fn_once);
                } else {
                    expand<BS>(
                        Sl, nS, S,
                        [&](uint32_t v, B& dv, uint32_t& beg, uint32_t& cnt) {
                            dv = 0;
                            beg = off[v];
                            cnt = off[v + 1] - beg;
                        },
                        [&](uint32_t v, B, uint32_t e) {
                            uint32_t t;
                            L l;
                            load_arc<L, BS>(tgt, len, arcs, e, t, l);
                            fn(t, e, T::add(ldc(DP(v)), l), v);
                        });
                }
            };
            if (!log_ok && tid == 0) atomicOr(a.err, ERR_LOG_OVERFLOW);
            each_candidate(mark, mark_once);
            {
                auto add32 = [](uint32_t x, uint32_t y) { return x + y; };
                n_tight = wave_reduce(n_tight, add32);
                if (lane_id() == 0) atomicAdd(&S.cnt[6], n_tight);
            }
            bsync();
            // Every reached node but the source has at least one tight in-arc. With exactly one each (the rule
            // for continuous weights) the tight arcs form a tree: every sigma is 1, the sigma layers need no
            // in-degrees or sums, and the dependency term is 1 + delta(t) (the general formula with sigma = 1,
            // bit for bit). Otherwise a second pass counts the tight in-degrees for the general layers.
            const bool tree = S.cnt[6] + 1u == nS;
            c_trees += tree ? 1u : 0u;
            // Brandes on a tree pushes dependencies to parents: every delta is a subtree size, an integer, so the
            // atomic sums are exact in any order. The sigma layers record each node's parent beside its order
            // position (in the second frontier list, idle after the distances are final) and clear the bitmap.
            const bool push_dep = MODE == MODE_BC && tree;
            uint32_t* const par = Fb;
            // ... and each internal node's number of children beside its own order position (in the first frontier
            // list): a leaf's dependency is 0, so its term 1 reaches the parent as part of that count, not as a push
            uint32_t* const nch = Fa;
            // A near tree: its extra tight arcs (every reached node but the source has one tight in-arc at least) fit
            // the list, and the nodes with several tight in-arcs with the arcs into them fit the second list (at most
            // twice as many entries). The bounds' LDS, free once verified, holds the bitmap of the heads seen by the
            // scan below and then the map of the nodes with several tight in-arcs (`multi`), whose in-degrees go to
            // their records.
            const uint32_t n_extra = S.cnt[6] + 1u - nS;
            const bool near = kNear && near_ok && !tree && 2u * n_extra <= Lds<B, BS>::kQ;
            uint32_t* const multi = reinterpret_cast<uint32_t*>(fl);
            uint32_t* const seen = multi + bmw;
            auto is_multi = [&](uint32_t v) -> bool { return ((multi[v >> 5] >> (v & 31)) & 1u) != 0u; };
            if constexpr (kNear) if (near) {
                for (uint32_t w = tid; w < bmw; w += BS) {
                    multi[w] = 0u;
                    seen[w] = 0u;
                }
                if (tid == 0) {
                    S.cnt[1] = 0;  // entries of the list of arcs into multi nodes (Kahn)
                    S.cnt[7] = 0;  // listed heads
                }
                bsync();
                // every tight arc once (the bitmap rows of the internal nodes): a head seen before is listed
                for (uint32_t i = tid; i < nS; i += BS) {
                    const uint32_t u = Sl[i];
                    if (!internal(u)) continue;
                    for_row_words(bits, off[u], off[u + 1], [&](uint32_t wi, uint32_t w) {
                        while (w) {
                            const uint32_t e = (wi << 5) + static_cast<uint32_t>(__ffs(w) - 1);
                            w &= w - 1;
                            const uint32_t t = target_of(e);
                            const uint32_t bit = 1u << (t & 31);
                            if ((atomicOr(&seen[t >> 5], bit) & bit) != 0u) {
                                const uint32_t k = atomicAdd(&S.cnt[7], 1u);
                                if (k < Lds<B, BS>::kQ) S.q_t[k] = t;
                            }
                        }
                    });
                }
                bsync();
                // the scan lists exactly n_extra heads (each tight arc beyond a node's first); bounded all the same
                const uint32_t n_list = S.cnt[7] < Lds<B, BS>::kQ ? S.cnt[7] : Lds<B, BS>::kQ;
                for (uint32_t k = tid; k < n_list; k += BS) {
                    const uint32_t t = S.q_t[k];
                    const uint32_t bit = 1u << (t & 31);
                    // in-degree = listed entries + the first tight in-arc
                    wg_add(&R[t].indeg, (atomicOr(&multi[t >> 5], bit) & bit) == 0u ? 2u : 1u);
                }
                bsync();
            }
            c_near += near ? 1u : 0u;
            if (!tree && !near) {  // tight in-degrees from the bitmap rows (each tight arc once)
                for (uint32_t i = tid; i < nS; i += BS) {
                    const uint32_t u = Sl[i];
                    if (!internal(u)) continue;
                    for_row_words(bits, off[u], off[u + 1], [&](uint32_t wi, uint32_t w) {
                        while (w) {
                            const uint32_t e = (wi << 5) + static_cast<uint32_t>(__ffs(w) - 1);
                            w &= w - 1;
                            wg_add(&R[target_of(e)].indeg, 1u);
                        }
                    });
                }
                bsync();
            }
            NG_TOCK(t_verify, 4);
            NG_TICK(t_kahn);
            // fn(i, v) for the order positions i in [lb, le) whose node v = order[i] is internal (has a tight out-arc). In
            // wide workgroups each wave packs the internal positions of its share into its LDS queue and runs them 64
            // at a time, so that its lanes all work on internal nodes (a quarter to a third of a tree's nodes) instead
            // of waiting, one memory chain at a time, for the few lanes that hold one.
            auto for_internal = [&](uint32_t lb, uint32_t le, auto fn) {
                if constexpr (BS >= 512) {
                    const uint32_t lane = lane_id();
                    const unsigned long long below = (1ull << lane) - 1ull;
                    uint32_t* const qi = S.q_t + (tid & ~63u);
                    uint32_t* const qv = S.q_e + (tid & ~63u);
                    uint32_t qc = 0;  // queued positions of this wave (wave-uniform)
                    auto run = [&]() {
                        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");  // the queue stores before its loads
                        uint32_t i = 0, v = 0;
                        const bool live = lane < qc;
                        if (live) {
                            i = qi[lane];
                            v = qv[lane];
                        }
                        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");  // every lane has read before a refill
                        if (live) fn(i, v);
                        qc = 0;
                    };
                    for (uint32_t pb = lb + (tid & ~63u); pb < le; pb += BS) {  // wave-uniform trip count
                        const uint32_t i = pb + lane;
                        uint32_t v = 0;
                        bool keep = false;
                        if (i < le) {
                            v = order[i];
                            keep = internal(v);
                        }
                        const unsigned long long mk = __ballot(keep);
                        const uint32_t c = static_cast<uint32_t>(__popcll(mk));
                        if (c == 0u) continue;
                        if (qc + c > 64u) run();
                        if (keep) {
                            const uint32_t q = qc + static_cast<uint32_t>(__popcll(mk & below));
                            qi[q] = i;
                            qv[q] = v;
                        }
                        qc += c;
                    }
                    if (qc != 0u) run();
                } else {
                    for (uint32_t i = lb + tid; i < le; i += BS) {
                        const uint32_t v = order[i];
                        if (internal(v)) fn(i, v);
                    }
                }
            };
            // ---- Kahn layers over the tight DAG: sigma and a topological order
            if (tid == 0) {
                order[0] = s;
                stc(&R[s].sigma, 1.0);
                S.cnt[4] = 1;
            }
            bsync();
            uint32_t lb = 0, le = 1, nl = 0;
            while (lb < le && nl < n) {
                if (tid == 0) layers[nl] = lb;
                ++nl;
                // one node u at order position i: a leaf (not internal) releases no children
                auto kahn_node = [&](uint32_t i, uint32_t u) {
                    if (tree) {
                        // a tree node releases each tight child: up to four children's targets are loaded
                        // together and appended with one LDS atomic
                        uint32_t e0 = 0, e1 = 0, e2 = 0, e3 = 0, k = 0, kids = 0;
                        auto flush = [&]() {
                            kids += k;
                            const uint32_t t0 = target_of(e0);
                            const uint32_t t1 = k > 1 ? target_of(e1) : 0u;
                            const uint32_t t2 = k > 2 ? target_of(e2) : 0u;
                            const uint32_t t3 = k > 3 ? target_of(e3) : 0u;
                            const uint32_t pos = atomicAdd(&S.cnt[4], k);
                            auto put = [&](uint32_t j, uint32_t t) {
                                if (pos + j < n) {
                                    // order and parents are read once, by the dependency layers (non-temporal)
                                    __builtin_nontemporal_store(t, order + pos + j);
                                    if (push_dep) __builtin_nontemporal_store(u, par + pos + j);
                                }
                            };
                            put(0, t0);
                            if (k > 1) put(1, t1);
                            if (k > 2) put(2, t2);
                            if (k > 3) put(3, t3);
                            k = 0;
                        };
                        for_row_words(bits, off[u], off[u + 1], [&](uint32_t wi, uint32_t w) {
                            if (push_dep && !bits_linear) wg_and(&bits[wi], ~w);
                            while (w) {
                                const uint32_t e = (wi << 5) + static_cast<uint32_t>(__ffs(w) - 1);
                                w &= w - 1;
                                e0 = k == 0 ? e : e0;
                                e1 = k == 1 ? e : e1;
                                e2 = k == 2 ? e : e2;
                                e3 = k == 3 ? e : e3;
                                if (++k == 4u) flush();
                            }
                        });
                        if (k != 0u) flush();
                        if (push_dep) __builtin_nontemporal_store(kids, nch + i);
                        return;
                    }
                    if constexpr (kNear) if (near) {
                        // A near-tree node: sigma by order position (the source's is 1; a multi node's, marked -1, is
                        // the sum in its record). A child with one tight in-arc is released at once with the
                        // parent's sigma; a multi child adds the parent's sigma to its record, lists the arc and is
                        // released by its last parent. The bitmap row is cleared here. nch: single-parent leaves.
                        double su = i == 0 ? 1.0 : sigp[i];
                        if (su < 0.0) su = ldc(&R[u].sigma);
                        uint32_t nsl = 0;
                        auto release = [&](uint32_t t, double st) {
                            const uint32_t pos = atomicAdd(&S.cnt[4], 1u);
                            if (pos < n) {
                                order[pos] = t;
                                par[pos] = u;
                                sigp[pos] = st;
                            }
                        };
                        for_row_words(bits, off[u], off[u + 1], [&](uint32_t wi, uint32_t w) {
                            if (!bits_linear) wg_and(&bits[wi], ~w);
                            while (w) {
                                const uint32_t e = (wi << 5) + static_cast<uint32_t>(__ffs(w) - 1);
                                w &= w - 1;
                                const uint32_t t = target_of(e);
                                if (is_multi(t)) {
                                    wg_add(&R[t].sigma, su);
                                    const uint32_t k = atomicAdd(&S.cnt[1], 1u);
                                    if (k < Lds<B, BS>::kQ) {
                                        S.q_e[k] = t;
                                        S.q_n[k] = static_cast<B>(u);
                                    }
                                    if (wg_sub(&R[t].indeg, 1u) == 1u) release(t, -1.0);
                                } else {
                                    nsl += internal(t) ? 0u : 1u;
                                    release(t, su);
                                }
                            }
                        });
                        nch[i] = nsl;
                        return;
                    }
                    const double su = ldc(&R[u].sigma);
                    for_row_words(bits, off[u], off[u + 1], [&](uint32_t wi, uint32_t w) {
                        while (w) {
                            const uint32_t e = (wi << 5) + static_cast<uint32_t>(__ffs(w) - 1);
                            w &= w - 1;
                            const uint32_t t = target_of(e);
                            wg_add(&R[t].sigma, su);
                            const bool released = wg_sub(&R[t].indeg, 1u) == 1u;
                            if (released) {
                                const uint32_t pos = atomicAdd(&S.cnt[4], 1u);
                                if (pos < n) order[pos] = t;
                            }
                        }
                    });
                };
                if (kNear && near) {
                    for (uint32_t i = lb + tid; i < le; i += BS) {
                        const uint32_t u = order[i];
                        if (internal(u)) kahn_node(i, u);
                    }
                } else {
                    for_internal(lb, le, kahn_node);
                }
                bsync();
                lb = le;
                le = S.cnt[4] > n ? n : S.cnt[4];
                bsync();
            }
            if (tid == 0) layers[nl] = le;
            c_layers += nl;
            if (le != nS) {  // a zero-length tight cycle kept some nodes from being released
                if (tid == 0) atomicOr(a.err, ERR_KAHN);
                for (uint32_t i = tid; i < nS; i += BS) {
                    const uint32_t u = Sl[i];
                    stc(&R[u].indeg, 0u);
                    const uint32_t b = off[u], e_end = off[u + 1];
                    if (b < e_end)
                        for (uint32_t wi = b >> 5; wi <= (e_end - 1) >> 5; ++wi) wg_and(&bits[wi], ~range_mask(wi, b, e_end));
                }
            }
            bsync();
            NG_TOCK(t_kahn, 5);
            NG_TICK(t_back);
            // ---- reverse layers: dependencies, per-node and per-arc accumulation, bitmap cleared
            if (push_dep) {
                // Only internal nodes take part: delta(t) = (number of children) + the dependencies its internal
                // children pushed, all integers, so the atomic sums are exact in any order. A node's record is
                // written only then (the betweenness sum, and the pushed part cleared for the next source, in one
                // line): a leaf's record is neither read nor written, and the reset below leaves every record alone.
                rec_clean = le == nS;
                for (int ly = static_cast<int>(nl) - 1; ly >= 1; --ly) {
                    const uint32_t l0 = layers[ly], l1 = layers[ly + 1];
                    for_internal(l0, l1, [&](uint32_t i, uint32_t t) {
                        const uint32_t p = __builtin_nontemporal_load(par + i);
                        const double pushed = ldc(&R[t].delta1);
                        const double d = pushed + static_cast<double>(__builtin_nontemporal_load(nch + i));
                        R[t].bc += d;
                        if (rec_clean && pushed != 0.0) stc(&R[t].delta1, 0.0);
                        if (ly > 1) wg_add(&R[p].delta1, d);
                    });
                    bsync();
                }
            } else if (kNear && near) {
                // Near tree, pushed the other way: a node t other than a single-parent leaf takes
                // delta(t) = sigma(t) (pushed + nch / sigma(t)), where `pushed` holds the terms
                // y(c) = (1 + delta(c)) / sigma(c) of its internal single-parent children and its multi children, and
                // each single-parent leaf child contributes 1 / sigma(t); it then sends y(t) to its parent, or to each
                // of its parents from the list when it is a multi node: the per-arc term sigma(p) (1 + delta(t)) /
                // sigma(t) of BrandesVariant.h:151-154, with sigma(p) applied at p. Records are cleaned as on a tree.
                rec_clean = le == nS;
                const uint32_t n_arcs = S.cnt[1];
                for (int ly = static_cast<int>(nl) - 1; ly >= 1; --ly) {
                    const uint32_t l0 = layers[ly], l1 = layers[ly + 1];
                    for (uint32_t i = l0 + tid; i < l1; i += BS) {
                        const uint32_t t = order[i];
                        const bool mt = is_multi(t), it = internal(t);
                        if (!it && !mt) continue;
                        double st = sigp[i];
                        if (st < 0.0) st = ldc(&R[t].sigma);
                        double dlt = 0.0;
                        if (it) {
                            const double pushed = ldc(&R[t].delta1);
                            dlt = st * (pushed + static_cast<double>(nch[i]) / st);
                            R[t].bc += dlt;
                            if (rec_clean && pushed != 0.0) stc(&R[t].delta1, 0.0);
                        }
                        const double y = (1.0 + dlt) / st;
                        if (mt) {
                            for (uint32_t k = 0; k < n_arcs; ++k) {
                                if (S.q_e[k] != t) continue;
                                const uint32_t p = static_cast<uint32_t>(S.q_n[k]);
                                if (p != s) wg_add(&R[p].delta1, y);
                            }
                            if (rec_clean) stc(&R[t].sigma, 0.0);
                        } else if (ly > 1) {
                            wg_add(&R[par[i]].delta1, y);
                        }
                    }
                    bsync();
                }
            } else {
                // On a tree each node has one parent: the parent clears a child's dependencies once it has read
                // them, a node's own dependencies are written only when nonzero, and the reset below leaves the
                // records alone, so a leaf's lines are read and never dirtied.
                rec_clean = tree && le == nS;
                for (int ly = static_cast<int>(nl) - 1; ly >= 0; --ly) {
                    const uint32_t l0 = layers[ly], l1 = layers[ly + 1];
                    // a leaf has no tight children: its dependencies stay 0 and its bitmap row is clear
                    for_internal(l0, l1, [&](uint32_t i, uint32_t u) {
                        const double su = tree ? 1.0 : ldc(&R[u].sigma);
                        double acc1 = 0.0, acc2 = 0.0;
                        // the tight children in increasing arc order, two at a time: both children's loads are
                        // in flight together, and the sums still add the terms in arc order
                        auto visit = [&](uint32_t ea, uint32_t eb, bool two) {
                            const uint32_t ta = target_of(ea);
                            const uint32_t tb = two ? target_of(eb) : ta;
                            const double sta = tree ? 1.0 : ldc(&R[ta].sigma);
                            const double stb = tree ? 1.0 : ldc(&R[tb].sigma);
                            // a leaf child's dependencies are 0 (on a tree its record is not read at all)
                            const bool ia = internal(ta), ib = internal(tb);
                            const double da = ia ? ldc(&R[ta].delta1) : 0.0;
                            const double db = ib ? ldc(&R[tb].delta1) : 0.0;
                            uint32_t pa = 0, pb = 0;
                            if constexpr ((MODE & MODE_ARC1) != 0) {
                                pa = perm[ea];
                                pb = two ? perm[eb] : pa;
                            }
                            double dta = 0.0, dtb = 0.0, xa = 0.0, xb = 0.0;
                            if constexpr ((MODE & MODE_ARC2) != 0) {
                                dta = T::val(ldc(DP(ta)));
                                dtb = T::val(ldc(DP(tb)));
                                xa = ia ? ldc(&delta2[ta]) : 0.0;
                                xb = ib ? ldc(&delta2[tb]) : 0.0;
                            }
                            if (rec_clean) {
                                if (da != 0.0) stc(&R[ta].delta1, 0.0);
                                if (two && db != 0.0) stc(&R[tb].delta1, 0.0);
                                if constexpr ((MODE & MODE_ARC2) != 0) {
                                    if (xa != 0.0) stc(&delta2[ta], 0.0);
                                    if (two && xb != 0.0) stc(&delta2[tb], 0.0);
                                }
                            }
                            auto term = [&](uint32_t pe, double st, double dl, double dt, double x) {
                                const double term1 = su * ((1.0 + dl) / st);
                                acc1 += term1;
                                if constexpr ((MODE & MODE_ARC1) != 0) unsafeAtomicAdd(&a.grad1[pe], term1);
                                if constexpr ((MODE & MODE_ARC2) != 0) {
                                    const double term2 = su * ((1.0 / (dt * dt) + x) / st);
                                    acc2 += term2;
                                    unsafeAtomicAdd(&a.grad2[pe], term2);
                                }
                            };
                            term(pa, sta, da, dta, xa);
                            if (two) term(pb, stb, db, dtb, xb);
                        };
                        uint32_t e_held = 0;
                        bool held = false;
                        for_row_words(bits, off[u], off[u + 1], [&](uint32_t wi, uint32_t w) {
                            if (!bits_linear) wg_and(&bits[wi], ~w);
                            while (w) {
                                const uint32_t e = (wi << 5) + static_cast<uint32_t>(__ffs(w) - 1);
                                w &= w - 1;
                                if (held) visit(e_held, e, true);
                                else e_held = e;
                                held = !held;
                            }
                        });
                        if (held) visit(e_held, e_held, false);
                        if (ly > 0) {
                            if (!rec_clean || acc1 != 0.0) {
                                stc(&R[u].delta1, acc1);
                                if constexpr ((MODE & MODE_BC) != 0) R[u].bc += acc1;
                            }
                            if constexpr ((MODE & MODE_ARC2) != 0) if (!rec_clean || acc2 != 0.0) stc(&delta2[u], acc2);
                        }
                    });
                    bsync();
                }
            }
            NG_TOCK(t_back, 6);
        }

        NG_TICK(t_reset);
        if (a.dist_out != nullptr) {
            for (uint32_t v = tid; v < n; v += BS)
                write_dist_out(a, static_cast<size_t>(si) * n + v, ldc(DP(v)));
            bsync();  // every lane has read its distances before any lane resets them below
        }
        // ---- restore the slot: only reached nodes were touched (after an abort, frontier and pile nodes
        // may hold distances too, so the whole array is reset). Where at least half of the nodes were reached,
        // one coalesced pass over the whole distance array costs less than a scattered store per reached node.
        const bool linear = aborted || 2ull * nS >= n;
        if (linear) {
            for (uint32_t v = tid; v < n; v += BS) stc(DP(v), T::kInf);
        }
        if constexpr (MODE != MODE_SSSP) {
            if (bits_linear) {
                const uint32_t words = (off[n] + 31u) >> 5;
                for (uint32_t w = tid; w < words; w += BS) stc(&bits[w], 0u);
            }
        }
        bool per_node = !linear;
        if constexpr (MODE != MODE_SSSP) {
            if (rec_clean && tid == 0) stc(&R[s].sigma, 0.0);  // on a tree only the source has a sigma
            per_node = per_node || !rec_clean;
        }
        if (per_node) {
            for (uint32_t i = tid; i < nS; i += BS) {
                const uint32_t v = Sl[i];
                if (!linear) stc(DP(v), T::kInf);
                if constexpr (MODE != MODE_SSSP) if (!rec_clean) {
                    stc(&R[v].sigma, 0.0);
                    stc(&R[v].delta1, 0.0);
                    if constexpr ((MODE & MODE_ARC2) != 0) stc(&delta2[v], 0.0);
                }
            }
        }
        bsync();
        NG_TOCK(t_reset, 7);
    }
    if (cohort && cur_g != 0xFFFFFFFFu) flush_bc();
    if (tid == 0) {
        ctr[0] = phase;
        ctr[1] = gen;
        ctr[2] = pgen;
        atomicAdd(&a.arcs_reached[1], c_phases);
        atomicAdd(&a.arcs_reached[2], c_buckets);
        atomicAdd(&a.arcs_reached[5], c_logged);
        atomicAdd(&a.arcs_reached[6], c_layers);
        atomicAdd(&a.arcs_reached[17], c_trees);
        atomicAdd(&a.arcs_reached[18], c_near);
    }
    // per-lane arc counts (each lane counted the rows it loaded)
    auto add64 = [](unsigned long long x, unsigned long long y) { return x + y; };
    const unsigned long long w_light = wave_reduce(static_cast<unsigned long long>(c_light), add64);
    const unsigned long long w_settle = wave_reduce(static_cast<unsigned long long>(c_settle), add64);
    const unsigned long long w_probe = wave_reduce(static_cast<unsigned long long>(c_probe), add64);
    const unsigned long long w_improve = wave_reduce(static_cast<unsigned long long>(c_improve), add64);
    if (lane_id() == 0) {
        atomicAdd(&a.arcs_reached[3], w_light);
        atomicAdd(&a.arcs_reached[4], w_settle);
        atomicAdd(&a.arcs_reached[7], w_probe);
        atomicAdd(&a.arcs_reached[8], w_improve);
    }
}

// end of the light prefix of each (length-sorted) row
template <class L>
__global__ void light_end_kernel(uint32_t n, const uint32_t* off, const L* len, L delta, RowRange* rows) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const uint32_t beg = off[v], end = off[v + 1];
    uint32_t lo = beg, hi = end;
    while (lo < hi) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        if (len[mid] < delta) lo = mid + 1; else hi = mid;
    }
    rows[v] = RowRange{beg, lo, end, 0u};
}

__global__ void interleave_arcs_kernel(const uint32_t* tgt, const uint32_t* len, uint64_t m, u32x2* arcs) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < m; i += stride) {
        u32x2 p;
        p.x = tgt[i];
        p.y = len[i];
        arcs[i] = p;
    }
}

// out[v] = sum over slots of the betweenness partial of node v, in slot order (deterministic)
// (with `zero`, each partial is cleared once read, so that the next call needs no zeroing pass of its own)
__global__ void reduce_slots_kernel(BcRec* rec, uint32_t slots, uint32_t n, uint32_t stride, double* out, bool zero) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    double acc = 0.0;
    for (uint32_t k = 0; k < slots; ++k) {
        double* const p = &rec[static_cast<size_t>(k) * stride + v].bc;
        acc += *p;
        if (zero) *p = 0.0;
    }
    out[v] = acc;
}

__global__ void rec_init_kernel(BcRec* rec, uint64_t count) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride) rec[i] = BcRec{};
}

__global__ void rec_zero_bc_kernel(BcRec* rec, uint64_t count) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride) rec[i].bc = 0.0;
}

template <class T>
__global__ void fill_kernel(T* p, uint64_t count, T value) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride) p[i] = value;
}

template <class T>
hipError_t fill(T* p, uint64_t count, T value, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    const uint64_t blocks64 = (count + 255) / 256;
    const uint32_t blocks = static_cast<uint32_t>(blocks64 < 65536 ? blocks64 : 65536);
    hipLaunchKernelGGL(fill_kernel<T>, dim3(blocks), dim3(256), 0, stream, p, count, value);
    return hipGetLastError();
}

// LDS bitmaps of the batched kernel: the finalized-node bitmap plus one per bit of `maps` (MAP_SORTED,
// MAP_FRONTIER).
uint32_t lds_maps(int maps) { return 1u + (maps & MAP_SORTED ? 1u : 0u) + (maps & MAP_FRONTIER ? 1u : 0u); }

// Static LDS of the batched kernel: the expansion tile, reductions and, in wide workgroups, the survivor queues.
size_t lds_static_bytes(uint32_t dsize, int block) {
    const bool w = dsize == 8;
    switch (block) {
    case 64: return w ? sizeof(Lds<unsigned long long, 64>) : sizeof(Lds<uint32_t, 64>);
    case 512: return w ? sizeof(Lds<unsigned long long, 512>) : sizeof(Lds<uint32_t, 512>);
    case 1024: return w ? sizeof(Lds<unsigned long long, 1024>) : sizeof(Lds<uint32_t, 1024>);
    default: return w ? sizeof(Lds<unsigned long long, 256>) : sizeof(Lds<uint32_t, 256>);
    }
}

// Dynamic LDS of the batched kernel for n nodes: the bitmaps (none when they exceed 96 KB), then with
// MAP_BOUND the 16-bit distance bounds, or with MAP_BOUND4 the 4-bit codes where those do not fit, as long
// as everything stays within 136 KB and, with the kernel's static LDS, within the 160 KB of a CU. The plan's
// maps are the requested ones that fit.
struct LdsPlan {
    uint32_t bm_words = 0;
    uint32_t bd_words = 0;
    uint32_t bd_kind = 0;
    uint32_t ld_off = 0;
    int maps = 0;
    size_t bytes = 0;
};
// dsize: bytes per distance. With MAP_LDIST, the distances (n x dsize, 8-byte aligned) follow the bitmaps
// instead of the bounds.
LdsPlan lds_plan(uint32_t n, int maps, uint32_t dsize, int block) {
    LdsPlan p;
    const uint32_t w = (n + 31) / 32;
    const size_t bm_bytes = static_cast<size_t>(w) * 4 * lds_maps(maps);
    if (bm_bytes > 96 * 1024) return p;
    p.bm_words = w;
    p.maps = maps & (MAP_SORTED | MAP_FRONTIER);
    p.bytes = bm_bytes;
    const size_t kBudget = std::min<size_t>(136 * 1024, 160 * 1024 - lds_static_bytes(dsize, block) - 256);
    if ((maps & MAP_LDIST) != 0) {
        const uint32_t off = (w * lds_maps(maps) + 1u) & ~1u;
        const size_t bytes = static_cast<size_t>(off) * 4 + static_cast<size_t>(n) * dsize;
        if (bytes <= kBudget) {
            p.ld_off = off;
            p.maps |= MAP_LDIST;
            p.bytes = bytes;
        }
        return p;
    }
    const uint32_t b16 = (n + 1) / 2, b4 = (n + 7) / 8;
    if ((maps & MAP_BOUND) != 0 && bm_bytes + static_cast<size_t>(b16) * 4 <= kBudget) {
        p.bd_words = b16;
        p.bd_kind = 1;
        p.maps |= maps & (MAP_BOUND | MAP_BOUND4);
    } else if ((maps & MAP_BOUND4) != 0 && bm_bytes + static_cast<size_t>(b4) * 4 <= kBudget) {
        p.bd_words = b4;
        p.bd_kind = 2;
        p.maps |= maps & (MAP_BOUND | MAP_BOUND4);
    }
    p.bytes += static_cast<size_t>(p.bd_words) * 4;
    return p;
}

template <class D, int MODE, int BS, bool LD>
hipError_t launch_typed(const LaunchArgs& la, hipStream_t stream) {
    using T = DT<D>;
    using B = typename T::B;
    using L = typename T::L;
    KArgs<D> k{};
    Workspace& ws = *la.ws;
    k.n = la.n;
    k.stride = ws.n;
    if (la.gtab == nullptr) {
        const View& vw = *la.view;
        k.off = vw.off;
        k.tgt = vw.tgt;
        k.len = static_cast<const L*>(vw.len);
        k.arcs = vw.arcs;
        k.rows = vw.rows;
        k.perm = vw.perm;
        k.strong = vw.strong ? 1u : 0u;
    } else {
        k.gtab = la.gtab;
        k.n_graphs = la.n_graphs;
        k.node_prefix = la.node_prefix;
        k.flush = la.flush;
        k.flush_base = la.flush_base;
    }
    k.sources = la.sources;
    k.n_sources = la.n_sources;
    if constexpr (T::kInt) {
        double d = la.delta < 1.0 ? 1.0 : la.delta;
        k.delta = static_cast<L>(d + 0.5);
    } else {
        k.delta = static_cast<L>(la.delta);
    }
    k.dist = static_cast<B*>(ws.dist);
    k.fstamp = ws.fstamp;
    k.sstamp = ws.sstamp;
    k.rec = ws.rec;
    k.delta2 = ws.delta2;
    k.lists = ws.lists;
    k.pile = ws.pile;
    k.pcap = ws.pcap;
    k.counters = ws.counters;
    k.order = ws.order;
    k.layers = ws.layers;
    k.bits = ws.bits;
    k.words = ws.words;
    k.log = ws.log;
    k.log_hi = ws.log_hi;
    k.logcap = ws.logcap;
    k.grad1 = ws.grad1;
    k.grad2 = ws.grad2;
    k.summaries = static_cast<ng_sssp_summary*>(ws.summaries);
    k.arcs_reached = reinterpret_cast<unsigned long long*>(ws.arcs_reached);
    k.err = ws.err;
    k.dist_out = la.dist_out;
    k.dist_out_kind = la.dist_out_kind;
    const LdsPlan lp = lds_plan(la.n, LD ? la.maps : (la.maps & ~MAP_LDIST), static_cast<uint32_t>(sizeof(B)), BS);
    if (LD && (lp.maps & MAP_LDIST) == 0) return hipErrorInvalidValue;
    k.bm_words = lp.bm_words;
    k.bd_words = lp.bd_words;
    k.bd_kind = lp.bd_kind;
    k.ld_off = lp.ld_off;
    k.bd_inv_q = la.bd_inv_q;
    k.maps = static_cast<uint32_t>(lp.maps);
    hipLaunchKernelGGL((batch_kernel<D, MODE, BS, LD>), dim3(la.grid), dim3(BS), lp.bytes, stream, k);
    return hipGetLastError();
}

template <class D, int BS, bool LD>
hipError_t launch_mode(const LaunchArgs& a, hipStream_t stream) {
    switch (a.mode) {
    case MODE_SSSP: return launch_typed<D, MODE_SSSP, BS, LD>(a, stream);
    case MODE_BC: return launch_typed<D, MODE_BC, BS, LD>(a, stream);
    case MODE_ARC1: return launch_typed<D, MODE_ARC1, BS, LD>(a, stream);
    case MODE_ARC1 | MODE_ARC2: return launch_typed<D, MODE_ARC1 | MODE_ARC2, BS, LD>(a, stream);
    default: return hipErrorInvalidValue;
    }
}

// LDS distances are built for 256-lane workgroups (small graphs) and 1024-lane ones (one workgroup per CU
// holds up to ~110 KB of distances); 64-lane workgroups (one wavefront per source) keep them in the slot's
// global array, since a few KB of distances per wave would already cut the resident waves
template <class D>
hipError_t launch_block(const LaunchArgs& a, hipStream_t stream) {
    switch (a.block) {
    case 64: return launch_mode<D, 64, false>(a, stream);
    case 256: return (a.maps & MAP_LDIST) ? launch_mode<D, 256, true>(a, stream) : launch_mode<D, 256, false>(a, stream);
    case 512: return launch_mode<D, 512, false>(a, stream);
    case 1024: return (a.maps & MAP_LDIST) ? launch_mode<D, 1024, true>(a, stream) : launch_mode<D, 1024, false>(a, stream);
    default: return hipErrorInvalidValue;
    }
}

template <class D, int BS, bool LD>
int blocks_per_cu_bs(int mode, size_t dyn) {
    int nb = 0;
    hipError_t e = hipErrorInvalidValue;
    switch (mode) {
    case MODE_SSSP: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_kernel<D, MODE_SSSP, BS, LD>, BS, dyn); break;
    case MODE_BC: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_kernel<D, MODE_BC, BS, LD>, BS, dyn); break;
    case MODE_ARC1: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_kernel<D, MODE_ARC1, BS, LD>, BS, dyn); break;
    default:
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_kernel<D, MODE_ARC1 | MODE_ARC2, BS, LD>, BS, dyn);
        break;
    }
    return (e == hipSuccess && nb > 0) ? nb : 1;
}

template <class D>
int blocks_per_cu_mode(int mode, uint32_t n, int maps, int block) {
    using B = typename DT<D>::B;
    const bool ld = (block == 256 || block == 1024) && (maps & MAP_LDIST) != 0;
    const size_t dyn = lds_plan(n, ld ? maps : (maps & ~MAP_LDIST), static_cast<uint32_t>(sizeof(B)), block).bytes;
    switch (block) {
    case 64: return blocks_per_cu_bs<D, 64, false>(mode, dyn);
    case 512: return blocks_per_cu_bs<D, 512, false>(mode, dyn);
    case 1024: return ld ? blocks_per_cu_bs<D, 1024, true>(mode, dyn) : blocks_per_cu_bs<D, 1024, false>(mode, dyn);
    default: return ld ? blocks_per_cu_bs<D, 256, true>(mode, dyn) : blocks_per_cu_bs<D, 256, false>(mode, dyn);
    }
}

}  // namespace

hipError_t launch_batch(const LaunchArgs& a, hipStream_t stream) {
    switch (a.dk) {
    case DKind::U32: return launch_block<uint32_t>(a, stream);
    case DKind::U64: return launch_block<uint64_t>(a, stream);
    case DKind::F32: return launch_block<float>(a, stream);
    case DKind::F64: return launch_block<double>(a, stream);
    }
    return hipErrorInvalidValue;
}

int lds_plan_maps(DKind dk, uint32_t n, int maps, int block) {
    return lds_plan(n, maps, (dk == DKind::U32 || dk == DKind::F32) ? 4u : 8u, block).maps;
}
int lds_plan_bound_kind(DKind dk, uint32_t n, int maps, int block) {
    return static_cast<int>(lds_plan(n, maps, (dk == DKind::U32 || dk == DKind::F32) ? 4u : 8u, block).bd_kind);
}

int batch_blocks_per_cu(DKind dk, int mode, uint32_t n, int maps, int block) {
    switch (dk) {
    case DKind::U32: return blocks_per_cu_mode<uint32_t>(mode, n, maps, block);
    case DKind::U64: return blocks_per_cu_mode<uint64_t>(mode, n, maps, block);
    case DKind::F32: return blocks_per_cu_mode<float>(mode, n, maps, block);
    case DKind::F64: return blocks_per_cu_mode<double>(mode, n, maps, block);
    }
    return 1;
}

hipError_t launch_light_end(DKind dk, uint32_t n, const uint32_t* off, const void* len, double delta, RowRange* rows,
                            hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const dim3 grid((n + 255) / 256), block(256);
    switch (dk) {
    case DKind::U32: {
        const uint32_t d = static_cast<uint32_t>((delta < 1.0 ? 1.0 : delta) + 0.5);
        hipLaunchKernelGGL(light_end_kernel<uint32_t>, grid, block, 0, stream, n, off, static_cast<const uint32_t*>(len), d, rows);
        break;
    }
    case DKind::U64: {
        const unsigned long long d = static_cast<unsigned long long>((delta < 1.0 ? 1.0 : delta) + 0.5);
        hipLaunchKernelGGL(light_end_kernel<unsigned long long>, grid, block, 0, stream, n, off,
                           static_cast<const unsigned long long*>(len), d, rows);
        break;
    }
    case DKind::F32:
        hipLaunchKernelGGL(light_end_kernel<float>, grid, block, 0, stream, n, off, static_cast<const float*>(len),
                           static_cast<float>(delta), rows);
        break;
    case DKind::F64:
        hipLaunchKernelGGL(light_end_kernel<double>, grid, block, 0, stream, n, off, static_cast<const double*>(len), delta,
                           rows);
        break;
    }
    return hipGetLastError();
}

namespace {
__global__ void cohort_reduce_kernel(const double* flush, const uint32_t* node_prefix, uint32_t n_graphs, const uint32_t* seg_ptr,
                                     const uint64_t* seg_off, uint32_t total, double* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    uint32_t lo = 0, hi = n_graphs;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (node_prefix[mid] <= i) lo = mid; else hi = mid;
    }
    const uint32_t v = i - node_prefix[lo];
    double acc = 0.0;
    for (uint32_t k = seg_ptr[lo]; k < seg_ptr[lo + 1]; ++k) acc += flush[seg_off[k] + v];
    out[i] = acc;
}
}  // namespace

namespace {
// Block-privatised histogram of one tile of distance rows: LDS bins for up to kHistLds bins, global atomics
// beyond. Row i (source sources[i]) holds n distances; the source itself and unreachable nodes are skipped.
constexpr uint32_t kHistLds = 4096;

template <class T>
__device__ __forceinline__ bool hist_bin(T d, double bw, uint32_t nb, uint32_t* bin) {
    double x;
    if constexpr (sizeof(T) == 8 && !std::is_floating_point_v<T>) {
        if (d == ~0ull) return false;
        x = static_cast<double>(d);
    } else if constexpr (std::is_same_v<T, float>) {
        if (d == FLT_MAX) return false;
        x = static_cast<double>(d);
    } else {
        if (d == DBL_MAX) return false;
        x = d;
    }
    const double q = floor(x / bw);
    *bin = q >= static_cast<double>(nb - 1) ? nb - 1 : static_cast<uint32_t>(q);
    return true;
}

template <class T>
__global__ void distance_histogram_kernel(const T* dist, const uint32_t* sources, uint32_t n_rows, uint32_t n, double bw,
                                          uint32_t nb, unsigned long long* counts) {
    __shared__ uint32_t lds[kHistLds];
    const bool local = nb <= kHistLds;
    if (local)
        for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x) lds[i] = 0u;
    __syncthreads();
    const uint64_t total = static_cast<uint64_t>(n_rows) * n;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += stride) {
        const uint32_t r = static_cast<uint32_t>(i / n), t = static_cast<uint32_t>(i - static_cast<uint64_t>(r) * n);
        if (t == sources[r]) continue;
        uint32_t bin;
        if (!hist_bin<T>(dist[i], bw, nb, &bin)) continue;
        if (local) atomicAdd(&lds[bin], 1u);
        else atomicAdd(&counts[bin], 1ull);
    }
    __syncthreads();
    if (local)
        for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x)
            if (lds[i]) atomicAdd(&counts[i], static_cast<unsigned long long>(lds[i]));
}
}  // namespace

hipError_t launch_distance_histogram(int dist_kind, const void* dist, const uint32_t* sources, uint32_t n_rows, uint32_t n,
                                     double bin_width, uint32_t n_bins, unsigned long long* counts, hipStream_t stream) {
    const uint64_t total = static_cast<uint64_t>(n_rows) * n;
    if (total == 0) return hipSuccess;
    const uint32_t blocks = static_cast<uint32_t>(std::min<uint64_t>((total + 255) / 256, 4096));
    switch (dist_kind) {
    case NG_DIST_U64:
        hipLaunchKernelGGL(distance_histogram_kernel<unsigned long long>, dim3(blocks), dim3(256), 0, stream,
                           static_cast<const unsigned long long*>(dist), sources, n_rows, n, bin_width, n_bins, counts);
        break;
    case NG_DIST_F32:
        hipLaunchKernelGGL(distance_histogram_kernel<float>, dim3(blocks), dim3(256), 0, stream, static_cast<const float*>(dist),
                           sources, n_rows, n, bin_width, n_bins, counts);
        break;
    default:
        hipLaunchKernelGGL(distance_histogram_kernel<double>, dim3(blocks), dim3(256), 0, stream, static_cast<const double*>(dist),
                           sources, n_rows, n, bin_width, n_bins, counts);
        break;
    }
    return hipGetLastError();
}

hipError_t launch_cohort_reduce(const double* flush, const uint32_t* node_prefix, uint32_t n_graphs, const uint32_t* seg_ptr,
                                const uint64_t* seg_off, uint32_t total_nodes, double* out, hipStream_t stream) {
    if (total_nodes == 0) return hipSuccess;
    hipLaunchKernelGGL(cohort_reduce_kernel, dim3((total_nodes + 255) / 256), dim3(256), 0, stream, flush, node_prefix, n_graphs,
                       seg_ptr, seg_off, total_nodes, out);
    return hipGetLastError();
}

hipError_t launch_interleave_arcs(const uint32_t* tgt, const uint32_t* len, uint64_t m, uint32_t* arcs, hipStream_t stream) {
    if (m == 0) return hipSuccess;
    const uint64_t blocks64 = (m + 255) / 256;
    const uint32_t blocks = static_cast<uint32_t>(blocks64 < 65536 ? blocks64 : 65536);
    hipLaunchKernelGGL(interleave_arcs_kernel, dim3(blocks), dim3(256), 0, stream, tgt, len, m, reinterpret_cast<u32x2*>(arcs));
    return hipGetLastError();
}

hipError_t launch_reduce_slots(BcRec* rec, uint32_t slots, uint32_t n, uint32_t stride, double* out, bool zero, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(reduce_slots_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, rec, slots, n, stride, out, zero);
    return hipGetLastError();
}

__global__ void copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, uint64_t count) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride) dst[i] = src[i];
}

hipError_t launch_copy(const void* src, void* dst, uint64_t bytes, hipStream_t stream) {
    const uint64_t count = bytes / 16;
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(copy_kernel, dim3(8192), dim3(256), 0, stream, static_cast<const uint4*>(src), static_cast<uint4*>(dst),
                       count);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint64_t count, uint32_t value, hipStream_t stream) {
    return fill<uint32_t>(p, count, value, stream);
}
hipError_t launch_fill_u64(uint64_t* p, uint64_t count, uint64_t value, hipStream_t stream) {
    return fill<unsigned long long>(reinterpret_cast<unsigned long long*>(p), count, static_cast<unsigned long long>(value), stream);
}
hipError_t launch_fill_f64(double* p, uint64_t count, double value, hipStream_t stream) {
    return fill<double>(p, count, value, stream);
}
namespace {
uint64_t grid_for(uint64_t count) {
    const uint64_t b = (count + 255) / 256;
    return b < 65536 ? (b ? b : 1) : 65536;
}
}  // namespace

hipError_t launch_rec_init(BcRec* rec, uint64_t count, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(rec_init_kernel, dim3(static_cast<uint32_t>(grid_for(count))), dim3(256), 0, stream, rec, count);
    return hipGetLastError();
}

hipError_t launch_fill_dist(DKind dk, void* p, uint64_t count, hipStream_t stream) {
    switch (dk) {
    case DKind::U32: return fill<uint32_t>(static_cast<uint32_t*>(p), count, 0xFFFFFFFFu, stream);
    case DKind::F32: return fill<uint32_t>(static_cast<uint32_t*>(p), count, 0x7f800000u, stream);
    case DKind::U64: return fill<unsigned long long>(static_cast<unsigned long long*>(p), count, ~0ull, stream);
    case DKind::F64: return fill<unsigned long long>(static_cast<unsigned long long*>(p), count, 0x7ff0000000000000ull, stream);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_rec_zero_bc(BcRec* rec, uint64_t count, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(rec_zero_bc_kernel, dim3(static_cast<uint32_t>(grid_for(count))), dim3(256), 0, stream, rec, count);
    return hipGetLastError();
}

// Scratch of the allocation-free entry point: one file-scope device array, of which every GPU that loads this
// library holds its own copy (at its own address) until the process ends. In host code the name `scratch` is a
// host-side stand-in, not a device address: it is never passed to a kernel, a fill or a copy. scratch_setup()
// resolves the device address once per device (with that device current) and every host-side use goes through
// scratch_device_ptr().
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[kScratchMaxDevices] = {};
}  // namespace

hipError_t scratch_setup(int device) {
    if (device < 0 || device >= kScratchMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device] != nullptr) return hipSuccess;
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device) return hipErrorInvalidDevice;  // the caller makes `device` current first
    void* ptr = nullptr;
    e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (ptr == nullptr) return hipErrorInvalidSymbol;
    g_scratch_ptr[device] = ptr;
    return hipSuccess;
}

void* scratch_device_ptr(int device) {
    if (device < 0 || device >= kScratchMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_ptr[device];
}

}  // namespace ng
