"""ORACLE -- test infrastructure only.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import this
module, and only as the reference check (or as the timed CPU baseline). The product (``neurograph_amd``)
never imports, loads or calls anything under ``oracle/``.

ctypes front end of ``oracle/oracle.cpp``, a CPU restatement of neurograph's APSP / Brandes hot path
(paths relative to the reference checkout):

==========================  =============================================  ==========================================
function                    restates                                       pinned by
==========================  =============================================  ==========================================
``sssp``                    Dijkstra.h:56-98, PriorityQueueVariant.h:47-93  reference per-source tables
                            DijkstraCommon.h:36-54, SsspResult.h:43-64      (test_push_check.cpp:32-36, 104-117,
                                                                            seven_rank_expectations.h:91-175);
                                                                            closed forms (cycle, complete graph);
                                                                            scipy.sparse.csgraph.dijkstra (FP);
                                                                            Floyd-Warshall brute force (tiny)
``apsp``                    AllPairsShortestPath.h:111-163                  reference APSP results
                                                                            (test_apsp.cpp:67-84 ... 739), incl.
                                                                            the inverse metric (:274-291)
``brandes``                 BrandesVariant.h:50-159,                        reference averages (test_brandes.cpp);
                            BetweennessCentrality.h:46-80                   networkx betweenness / edge betweenness
                                                                            (FP weights); path-enumeration brute
                                                                            force on tiny graphs; cycle closed form;
                                                                            invariants sum BC = sum (d-1) (unit w)
``brandes(seed=1)``         the 1/d^2-seeded dependency (adjoint of the     central finite differences of the
                            average efficiency, a library extension)        oracle's own APSP (FP64, h=1e-6);
                                                                            homogeneity sum_e w_e dE/dw_e = -E
``delta_stepping``          DeltaStepping.h:36-82, EpochCommon.h:46-49,     reference per-source tables for every
                            PushCheckVariant.h:44-126                       delta in {1,4,10,10000}; equality with
                                                                            ``sssp``; negative weight -> error
``contract_balls``,         DiameterApproximation.h:71-126, 154-561         the reference's expectations
``diameter_approximation``  (one rank, Python loops: small graphs)          (test_diameter_approximation.cpp:57-322);
                                                                            invariants on random graphs (ball
                                                                            distances >= shortest distances from the
                                                                            center, within the grown radius; every
                                                                            contracted arc comes from an original
                                                                            arc); the all-centers and one-center
                                                                            closed forms
``Mt19937Canonical``        RandomNumberGenerator.h (std::mt19937 +          libstdc++'s own draws, compiled and run
                            uniform_real_distribution<double>)              beside the test; the standard's 10000th
                                                                            mt19937 output (4123659995)
==========================  =============================================  ==========================================

Ties that the reference leaves to unordered_map iteration order (DistributedGraph.cpp:222-230; two balls claiming
one node at the same distance in one round, several arcs joining the same two centers) are resolved here in node,
then target order; outside the reference's own expectations such tie cases are "parity unpinned" against the
reference itself, and pinned between this restatement and the library's host contraction.

Every function is pinned ("parity pinned"); none is left "parity unpinned".
"""

from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liborc.so")

W_I32, W_F32, W_F64 = 0, 1, 2
L_ABS, L_INVERSE, L_RAW = 0, 1, 2
SEED_ONE, SEED_INV_D2 = 0, 1


class Summary(ctypes.Structure):
    _fields_ = [
        ("sum_d", ctypes.c_double),
        ("sum_d_u64", ctypes.c_uint64),
        ("max_d", ctypes.c_double),
        ("max_d_u64", ctypes.c_uint64),
        ("sum_eff", ctypes.c_double),
        ("n_unreached", ctypes.c_uint64),
    ]


class ApspResult(ctypes.Structure):
    _fields_ = [
        ("avg_spl", ctypes.c_double),
        ("avg_eff", ctypes.c_double),
        ("diameter", ctypes.c_double),
        ("diameter_u64", ctypes.c_uint64),
        ("n_disconnected", ctypes.c_uint64),
        ("sum_d_u64", ctypes.c_uint64),
        ("sum_d", ctypes.c_double),
        ("sum_eff", ctypes.c_double),
    ]


SUMMARY_DTYPE = np.dtype(
    [("sum_d", "<f8"), ("sum_d_u64", "<u8"), ("max_d", "<f8"), ("max_d_u64", "<u8"), ("sum_eff", "<f8"), ("n_unreached", "<u8")]
)

_lib = None


def build() -> str:
    """Compile the oracle with its Makefile (g++). Returns the library path."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp, u32, u64, i32, dp = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p
        L.orc_sssp.argtypes = [u32, vp, vp, vp, i32, i32, u32, vp, ctypes.POINTER(Summary)]
        L.orc_apsp.argtypes = [u32, vp, vp, vp, i32, i32, vp, u32, i32, vp, ctypes.POINTER(ApspResult), vp]
        L.orc_brandes.argtypes = [u32, vp, vp, vp, i32, vp, u32, i32, i32, dp, dp, vp, dp, vp]
        L.orc_delta_stepping.argtypes = [u32, vp, vp, vp, u64, u32, vp, ctypes.POINTER(Summary)]
        L.orc_dist_kind.argtypes = [i32, i32]
        for f in ("orc_sssp", "orc_apsp", "orc_brandes", "orc_delta_stepping", "orc_dist_kind", "orc_hardware_threads"):
            getattr(L, f).restype = ctypes.c_int
        _lib = L
    return _lib


class OracleError(RuntimeError):
    pass


@dataclass
class Csr:
    """Host CSR: n nodes, off (uint64[n+1]), tgt (uint32[m]), w (int32/float32/float64[m])."""

    n: int
    off: np.ndarray
    tgt: np.ndarray
    w: np.ndarray

    @property
    def m(self) -> int:
        return int(self.off[-1])

    @property
    def wkind(self) -> int:
        return {np.dtype(np.int32): W_I32, np.dtype(np.float32): W_F32, np.dtype(np.float64): W_F64}[self.w.dtype]

    @staticmethod
    def from_edges(n: int, src, dst, w, dtype=np.int32) -> "Csr":
        src = np.asarray(src, dtype=np.int64)
        dst = np.asarray(dst, dtype=np.int64)
        w = np.asarray(w, dtype=dtype)
        order = np.lexsort((dst, src))
        src, dst, w = src[order], dst[order], w[order]
        off = np.zeros(n + 1, dtype=np.uint64)
        np.add.at(off, src + 1, 1)
        off = np.cumsum(off).astype(np.uint64)
        return Csr(n, off, dst.astype(np.uint32), np.ascontiguousarray(w))


def _ptr(a):
    return a.ctypes.data if a is not None else None


def _dist_dtype(wkind: int, lkind: int):
    return {0: np.uint64, 1: np.float32, 2: np.float64}[lib().orc_dist_kind(wkind, lkind)]


def _check(rc: int, what: str):
    if rc != 0:
        raise OracleError(f"{what} failed with status {rc}")


def sssp(g: Csr, root: int, lkind: int = L_ABS):
    dist = np.empty(g.n, dtype=_dist_dtype(g.wkind, lkind))
    s = Summary()
    _check(lib().orc_sssp(g.n, _ptr(g.off), _ptr(g.tgt), _ptr(g.w), g.wkind, lkind, root, _ptr(dist), ctypes.byref(s)), "orc_sssp")
    return dist, _summary_dict(s)


def _summary_dict(s: Summary) -> dict:
    return {f: getattr(s, f) for f, _ in Summary._fields_}


def apsp(g: Csr, lkind: int = L_ABS, sources=None, threads: int = 1, per_source: bool = False) -> dict:
    src = np.arange(g.n, dtype=np.uint32) if sources is None else np.ascontiguousarray(sources, dtype=np.uint32)
    res = ApspResult()
    clusters = np.zeros(g.n + 1, dtype=np.uint64)
    per = np.zeros(len(src), dtype=SUMMARY_DTYPE) if per_source else None
    rc = lib().orc_apsp(g.n, _ptr(g.off), _ptr(g.tgt), _ptr(g.w), g.wkind, lkind, _ptr(src), len(src), threads, _ptr(per),
                        ctypes.byref(res), _ptr(clusters))
    _check(rc, "orc_apsp")
    out = {f: getattr(res, f) for f, _ in ApspResult._fields_}
    out["cluster_sizes"] = {int(k): int(v) for k, v in enumerate(clusters) if v}
    if per is not None:
        out["per_source"] = per
    return out


def brandes(g: Csr, sources=None, threads: int = 1, seed: int = SEED_ONE, edge: bool = False, per_source: bool = False,
            want_sigma: bool = False, want_dist: bool = False) -> dict:
    src = np.arange(g.n, dtype=np.uint32) if sources is None else np.ascontiguousarray(sources, dtype=np.uint32)
    bc = np.zeros(g.n, dtype=np.float64)
    ed = np.zeros(g.m, dtype=np.float64) if edge else None
    per = np.zeros(len(src), dtype=SUMMARY_DTYPE) if per_source else None
    sig = np.zeros((len(src), g.n), dtype=np.float64) if want_sigma else None
    dist = np.zeros((len(src), g.n), dtype=_dist_dtype(g.wkind, L_ABS)) if want_dist else None
    rc = lib().orc_brandes(g.n, _ptr(g.off), _ptr(g.tgt), _ptr(g.w), g.wkind, _ptr(src), len(src), threads, seed, _ptr(bc),
                           _ptr(ed), _ptr(per), _ptr(sig), _ptr(dist))
    _check(rc, "orc_brandes")
    out = {"bc": bc, "avg_bc": float(np.sum(bc)) / g.n if g.n else float("nan")}
    if ed is not None:
        out["edge"] = ed
    if per is not None:
        out["per_source"] = per
    if sig is not None:
        out["sigma"] = sig
    if dist is not None:
        out["dist"] = dist
    return out


def delta_stepping(g: Csr, root: int, delta: int):
    if g.wkind != W_I32:
        raise OracleError("delta-stepping is an integer-weight path in the reference")
    dist = np.empty(g.n, dtype=np.uint64)
    s = Summary()
    _check(lib().orc_delta_stepping(g.n, _ptr(g.off), _ptr(g.tgt), _ptr(g.w), delta, root, _ptr(dist), ctypes.byref(s)),
           "orc_delta_stepping")
    return dist, _summary_dict(s)


def hardware_threads() -> int:
    return int(lib().orc_hardware_threads())


# ---------------------------------------------------------------------------------------------------------------
# DiameterApproximation (one rank), restated in plain Python loops for small graphs


class Mt19937Canonical:
    """Draws of std::mt19937(seed) through std::uniform_real_distribution<double>(0, 1) (the reference's
    DefaultRandomNumberGenerator, utility/RandomNumberGenerator.h). The C++ standard's generate_canonical for 53
    bits from a 32-bit engine takes two outputs x1, x2 and returns (x1 + x2 * 2^32) / 2^64, or the largest double
    below 1 if that rounds to 1 ([rand.util.canonical]); numpy's legacy integer seeding of MT19937 is the
    standard's mersenne_twister_engine seed sequence."""

    def __init__(self, seed: int):
        self._bits = np.random.RandomState(seed)._bit_generator

    def __call__(self) -> float:
        x1 = int(self._bits.random_raw())
        x2 = int(self._bits.random_raw())
        v = (x1 + x2 * 4294967296.0) / 18446744073709551616.0
        return v if v < 1.0 else float(np.nextafter(1.0, 0.0))


def contract_balls(g: Csr, r: int, draws=None, seed: int = 0):
    """DiameterApproximation::compute_approximation up to the contraction (DiameterApproximation.h:71-113):
    mark_as_centers (:154-179), perform_growing_step (:204-284) in synchronous rounds, stabilize_nodes (:286-295),
    check_all_stable (:191-202), contract (:297-561). Ties of one round (two claims at the same distance) go to the
    claim of the lower claiming node / earlier arc; several arcs between the same two centers keep the last one in
    (node, arc) order. Returns (contracted Csr with int32 weights, center of every node, ball distance of every
    node). Raises OracleError where the reference throws."""
    if g.w.dtype != np.int32:
        raise OracleError("integer weights only")
    n = g.n
    draws = Mt19937Canonical(seed) if draws is None else draws
    off = [int(x) for x in g.off]
    tgt = [int(x) for x in g.tgt]
    w = [int(x) for x in g.w]
    if any(x < 0 for x in w):
        raise OracleError("negative weight")
    none = -1
    center = [none] * n
    dist = [None] * n  # None = unreached (the reference's max distance)
    gen = [None] * n
    stable = [False] * n
    last_iteration = int(math.ceil(math.log2(n))) if n > 0 else 0
    for it in range(last_iteration):
        thr = 2.0 ** it / n
        any_center = False
        for v in range(n):
            if center[v] != none:
                any_center = True
                continue
            if draws() >= thr:
                continue
            center[v], dist[v], gen[v], stable[v] = v, 0, it, True
            any_center = True
        if not any_center:
            continue
        while True:  # growing rounds: questions from the state at the round's start, answers in order
            questions = []
            for u in range(n):
                if center[u] == none:
                    continue
                limit = (it - gen[u] + 1) * 2 * r
                for a in range(off[u], off[u + 1]):
                    t = tgt[a]
                    if w[a] == 0 or w[a] > 2 * r or stable[t]:
                        continue
                    nd = dist[u] + w[a]
                    if limit < nd or (dist[t] is not None and nd >= dist[t]):
                        continue
                    questions.append((t, center[u], nd, gen[u]))
            changed = False
            for t, c, nd, gq in questions:
                if dist[t] is not None and dist[t] <= nd:
                    continue
                center[t], dist[t], gen[t], stable[t] = c, nd, gq, False
                changed = True
            if not changed:
                break
        for v in range(n):
            if center[v] != none:
                stable[v] = True
        if all(stable):
            break
    for v in range(n):
        if center[v] == none:
            raise OracleError(f"node {v} was reached by no ball")
    new_id = {}
    for v in range(n):
        if center[v] == v:
            new_id[v] = len(new_id)
    arcs = {}
    for u in range(n):
        for a in range(off[u], off[u + 1]):
            if w[a] == 0:
                continue
            t = tgt[a]
            nw = w[a] + dist[u] + dist[t]
            if nw > 2**31 - 1:
                raise OracleError("contracted weight beyond int32")
            arcs[(new_id[center[u]], new_id[center[t]])] = nw
    keys = sorted(arcs)
    k = len(new_id)
    small = Csr.from_edges(k, [s for s, _ in keys], [t for _, t in keys], [arcs[x] for x in keys], dtype=np.int32)
    return small, np.array(center, dtype=np.uint32), dist


def diameter_approximation(g: Csr, r: int, draws=None, seed: int = 0) -> int:
    """The approximate diameter (DiameterApproximation.h:71-126): the contracted graph's APSP diameter plus 2r."""
    small, _, _ = contract_balls(g, r, draws, seed)
    return int(apsp(small)["diameter_u64"]) + 2 * r
