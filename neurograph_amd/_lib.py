"""ctypes binding of libng_hip.so (include/neurograph_hip.h, include/neurograph_io.h).

The library is the product: this module only marshals numpy arrays across the C ABI. There is no
fallback -- if the library is missing or no GPU is present, the calls fail loudly.
"""

from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libng_hip.so")

NG_OK, NG_EINVAL, NG_ENOMEM, NG_EDEVICE, NG_EUNSUPPORTED = 0, 1, 2, 3, 4
NG_W_I32, NG_W_F32, NG_W_F64 = 0, 1, 2
NG_LEN_ABS, NG_LEN_INVERSE, NG_LEN_RAW = 0, 1, 2
NG_DIST_U64, NG_DIST_F32, NG_DIST_F64 = 0, 1, 2
NG_OPT_BLOCK, NG_OPT_MAX_SLOTS, NG_OPT_LDS_MAPS, NG_OPT_PILE_CAPACITY, NG_OPT_LOG_CAPACITY = 0, 1, 2, 3, 4
_OPTIONS = {"block": NG_OPT_BLOCK, "max_slots": NG_OPT_MAX_SLOTS, "lds_maps": NG_OPT_LDS_MAPS,
            "pile_capacity": NG_OPT_PILE_CAPACITY, "log_capacity": NG_OPT_LOG_CAPACITY}

_WKIND = {np.dtype(np.int32): NG_W_I32, np.dtype(np.float32): NG_W_F32, np.dtype(np.float64): NG_W_F64}
_WDTYPE = {NG_W_I32: np.int32, NG_W_F32: np.float32, NG_W_F64: np.float64}
_DDTYPE = {NG_DIST_U64: np.uint64, NG_DIST_F32: np.float32, NG_DIST_F64: np.float64}


class NgError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"[status {status}] {message}")
        self.status = status


class Summary(ctypes.Structure):
    _fields_ = [("sum_d", ctypes.c_double), ("sum_d_u64", ctypes.c_uint64), ("max_d", ctypes.c_double),
                ("max_d_u64", ctypes.c_uint64), ("sum_eff", ctypes.c_double), ("n_unreached", ctypes.c_uint64)]


SUMMARY_DTYPE = np.dtype([("sum_d", "<f8"), ("sum_d_u64", "<u8"), ("max_d", "<f8"), ("max_d_u64", "<u8"),
                          ("sum_eff", "<f8"), ("n_unreached", "<u8")])
assert SUMMARY_DTYPE.itemsize == ctypes.sizeof(Summary)


class CallStats(ctypes.Structure):
    _fields_ = [("kernel_ms", ctypes.c_double), ("launches", ctypes.c_uint32), ("slots", ctypes.c_uint32),
                ("delta", ctypes.c_double), ("arcs_reached", ctypes.c_uint64), ("flags", ctypes.c_uint32),
                ("block", ctypes.c_uint32), ("phases", ctypes.c_uint64), ("buckets", ctypes.c_uint64),
                ("light_arcs", ctypes.c_uint64), ("settle_arcs", ctypes.c_uint64), ("logged_arcs", ctypes.c_uint64),
                ("layers", ctypes.c_uint64), ("probes", ctypes.c_uint64), ("improvements", ctypes.c_uint64),
                ("section_ticks", ctypes.c_uint64 * 8), ("tree_sources", ctypes.c_uint64),
                ("near_tree_sources", ctypes.c_uint64)]


class HostCsrC(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint32), ("m", ctypes.c_uint64), ("wkind", ctypes.c_int),
                ("offsets", ctypes.POINTER(ctypes.c_uint64)), ("targets", ctypes.POINTER(ctypes.c_uint32)),
                ("weights", ctypes.c_void_p), ("positions", ctypes.POINTER(ctypes.c_double))]


DRAW_FN = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_void_p)

_lib = None


def lib() -> ctypes.CDLL:
    """Loads libng_hip.so (built in-tree by neurograph_amd.build). Raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run `python -c 'import __graft_entry__ as g; g.build()'` first")
    L = ctypes.CDLL(LIB_PATH)
    vp, u32, u64, i32, dbl = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, ctypes.c_double
    P = ctypes.POINTER
    sig = {
        "ng_context_create": (i32, [i32, P(vp)]),
        "ng_context_destroy": (None, [vp]),
        "ng_context_synchronize": (i32, [vp]),
        "ng_copy_calibration": (i32, [vp, u64, P(dbl)]),
        "ng_graph_create": (i32, [vp, u32, u64, vp, vp, vp, i32, P(vp)]),
        "ng_graph_destroy": (None, [vp]),
        "ng_graph_dist_kind": (i32, [vp, i32, P(i32)]),
        "ng_graph_set_option": (i32, [vp, i32, ctypes.c_int64]),
        "ng_graph_set_weights": (i32, [vp, vp]),
        "ng_graph_weights_reordered": (i32, [vp, P(i32)]),
        "ng_graph_get_option": (i32, [vp, i32, P(ctypes.c_int64)]),
        "ng_sssp_summaries": (i32, [vp, vp, u32, i32, dbl, vp, vp]),
        "ng_brandes_accumulate": (i32, [vp, vp, u32, dbl, vp, vp]),
        "ng_brandes_accumulate_noalloc": (i32, [vp, vp, u32, dbl, vp, vp]),
        "ng_distance_histogram": (i32, [vp, vp, u32, i32, dbl, dbl, u32, vp, vp]),
        "ng_edge_dependency": (i32, [vp, vp, u32, dbl, vp]),
        "ng_apsp_adjoint": (i32, [vp, vp, u32, dbl, vp, vp, vp]),
        "ng_autotune_delta": (i32, [vp, i32, i32, vp, u32, P(dbl)]),
        "ng_autotune": (i32, [vp, i32, i32, vp, u32, P(dbl), P(u32)]),
        "ng_graph_last_stats": (i32, [vp, P(CallStats)]),
        "ng_cohort_create": (i32, [vp, vp, u32, P(vp)]),
        "ng_cohort_destroy": (None, [vp]),
        "ng_cohort_brandes": (i32, [vp, vp, vp, P(dbl)]),
        "ng_cohort_autotune": (i32, [vp, u32]),
        "ng_cohort_last_stats": (i32, [vp, P(CallStats)]),
        "ng_last_error": (ctypes.c_char_p, []),
        "ng_version": (ctypes.c_char_p, []),
        "ng_host_csr_free": (None, [P(HostCsrC)]),
        "ng_load_graph_dir": (i32, [ctypes.c_char_p, ctypes.c_char_p, i32, i32, i32, i32, i32, P(HostCsrC)]),
        "ng_write_graph_dir": (i32, [ctypes.c_char_p, P(HostCsrC)]),
        "ng_generate_fixed_degree": (i32, [u32, u32, ctypes.c_int32, ctypes.c_int32, u64, P(HostCsrC)]),
        "ng_generate_spatial": (i32, [u32, dbl, dbl, dbl, i32, dbl, dbl, i32, u64, P(HostCsrC)]),
        "ng_save_csr": (i32, [ctypes.c_char_p, P(HostCsrC), u64]),
        "ng_load_csr": (i32, [ctypes.c_char_p, u64, P(HostCsrC)]),
        "ng_contract_balls": (i32, [P(HostCsrC), u64, DRAW_FN, vp, u32, P(HostCsrC), vp]),
        "ng_diameter_approximation": (i32, [vp, P(HostCsrC), u64, DRAW_FN, vp, u32, P(u64), P(u32)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(status: int):
    if status != NG_OK:
        raise NgError(status, lib().ng_last_error().decode(errors="replace"))


def _ptr(a):
    return None if a is None else a.ctypes.data


# ------------------------------------------------------------------------------------------------
# host CSR


@dataclass
class HostGraph:
    """Host CSR over out-arcs: offsets uint64[n+1], targets uint32[m], weights int32/float32/float64[m]."""

    n: int
    offsets: np.ndarray
    targets: np.ndarray
    weights: np.ndarray
    positions: np.ndarray | None = None

    @property
    def m(self) -> int:
        return int(self.offsets[-1])

    @property
    def wkind(self) -> int:
        return _WKIND[self.weights.dtype]

    @staticmethod
    def from_edges(n, src, dst, w, dtype=None) -> "HostGraph":
        src = np.asarray(src, dtype=np.int64)
        dst = np.asarray(dst, dtype=np.int64)
        w = np.asarray(w) if dtype is None else np.asarray(w, dtype=dtype)
        order = np.lexsort((dst, src))
        off = np.zeros(n + 1, dtype=np.uint64)
        np.add.at(off, src + 1, 1)
        return HostGraph(n, np.cumsum(off).astype(np.uint64), dst[order].astype(np.uint32), np.ascontiguousarray(w[order]))

    def _as_c(self) -> HostCsrC:
        c = HostCsrC()
        c.n, c.m, c.wkind = self.n, self.m, self.wkind
        c.offsets = self.offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))
        c.targets = self.targets.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
        c.weights = self.weights.ctypes.data
        c.positions = None if self.positions is None else self.positions.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        return c


def _take_csr(c: HostCsrC) -> HostGraph:
    try:
        n, m = int(c.n), int(c.m)
        off = np.ctypeslib.as_array(c.offsets, shape=(n + 1,)).copy()
        tgt = np.ctypeslib.as_array(c.targets, shape=(m,)).copy() if m else np.zeros(0, np.uint32)
        wdt = _WDTYPE[c.wkind]
        if m:
            w = np.frombuffer((ctypes.c_char * (m * np.dtype(wdt).itemsize)).from_address(c.weights), dtype=wdt).copy()
        else:
            w = np.zeros(0, wdt)
        pos = np.ctypeslib.as_array(c.positions, shape=(3 * n,)).copy().reshape(n, 3) if (c.positions and n) else None
        return HostGraph(n, off, tgt, w, pos)
    finally:
        lib().ng_host_csr_free(ctypes.byref(c))


def load_graph_dir(path: str, prefix: str = "", n_ranks: int = 0, remove_self_arcs=False, undirected=False,
                   all_weights_one=False, weights=np.int32) -> HostGraph:
    c = HostCsrC()
    _check(lib().ng_load_graph_dir(path.encode(), prefix.encode(), n_ranks, int(remove_self_arcs), int(undirected),
                                   int(all_weights_one), _WKIND[np.dtype(weights)], ctypes.byref(c)))
    return _take_csr(c)


def write_graph_dir(path: str, g: HostGraph) -> None:
    c = g._as_c()
    _check(lib().ng_write_graph_dir(path.encode(), ctypes.byref(c)))


def generate_fixed_degree(n: int, k: int, wlo: int = 1, whi: int = 10, seed: int = 0) -> HostGraph:
    c = HostCsrC()
    _check(lib().ng_generate_fixed_degree(n, k, wlo, whi, seed, ctypes.byref(c)))
    return _take_csr(c)


def generate_spatial(n: int, mean_deg: float, box: float = 1000.0, lam: float = 150.0, wdist: str = "lognormal",
                     a: float = 0.0, b: float = 0.5, weights=np.float64, seed: int = 0) -> HostGraph:
    c = HostCsrC()
    wd = {"lognormal": 0, "uniform": 1}[wdist]
    _check(lib().ng_generate_spatial(n, box, lam, mean_deg, wd, a, b, _WKIND[np.dtype(weights)], seed, ctypes.byref(c)))
    return _take_csr(c)


def save_csr(path: str, g: HostGraph, key: int = 0) -> None:
    """Binary CSR cache of a host graph (ng_save_csr); `key` identifies its source and load options."""
    c = g._as_c()
    _check(lib().ng_save_csr(path.encode(), ctypes.byref(c), key))


def load_csr(path: str, key: int = 0) -> HostGraph:
    c = HostCsrC()
    _check(lib().ng_load_csr(path.encode(), key, ctypes.byref(c)))
    return _take_csr(c)


def _draw_fn(draws):
    """A C draw callback from a Python callable returning floats in [0, 1), or None (the default generator)."""
    if draws is None:
        return DRAW_FN(), None
    return DRAW_FN(lambda _user: float(draws())), None


def contract_balls(g: HostGraph, r: int, draws=None, seed: int = 0):
    """DiameterApproximation's clustering and contraction (ng_contract_balls): (contracted graph, center of each
    node). draws: a callable returning the next number in [0, 1), or None for std::mt19937(seed)."""
    fn, _ = _draw_fn(draws)
    out = HostCsrC()
    centers = np.zeros(g.n, dtype=np.uint32)
    c = g._as_c()
    _check(lib().ng_contract_balls(ctypes.byref(c), int(r), fn, None, int(seed), ctypes.byref(out), _ptr(centers)))
    return _take_csr(out), centers


# ------------------------------------------------------------------------------------------------
# device handles


class Context:
    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        _check(lib().ng_context_create(device, ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().ng_context_destroy(self.h)
            self.h = None

    def synchronize(self):
        _check(lib().ng_context_synchronize(self.h))

    def diameter_approximation(self, g: HostGraph, r: int, draws=None, seed: int = 0):
        """DiameterApproximation::compute_approximation with the contracted graph's APSP on the device
        (ng_diameter_approximation): returns (approximate diameter, contracted node count)."""
        fn, _ = _draw_fn(draws)
        d = ctypes.c_uint64()
        k = ctypes.c_uint32()
        c = g._as_c()
        _check(lib().ng_diameter_approximation(self.h, ctypes.byref(c), int(r), fn, None, int(seed), ctypes.byref(d),
                                               ctypes.byref(k)))
        return d.value, k.value

    def copy_calibration(self, nbytes: int) -> float:
        """Runs one plain device copy of nbytes; returns its duration in ms."""
        ms = ctypes.c_double()
        _check(lib().ng_copy_calibration(self.h, nbytes, ctypes.byref(ms)))
        return ms.value

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """A device-resident graph (ng_graph). Keeps its context alive."""

    def __init__(self, ctx: Context, g: HostGraph):
        self.ctx = ctx
        self.n, self.m = g.n, g.m
        self.host = g
        h = ctypes.c_void_p()
        off = np.ascontiguousarray(g.offsets, dtype=np.uint64)
        tgt = np.ascontiguousarray(g.targets, dtype=np.uint32)
        w = np.ascontiguousarray(g.weights)
        _check(lib().ng_graph_create(ctx.h, g.n, g.m, _ptr(off), _ptr(tgt), _ptr(w), _WKIND[w.dtype], ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().ng_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _sources(self, sources):
        if sources is None:
            return np.arange(self.n, dtype=np.uint32)
        return np.ascontiguousarray(sources, dtype=np.uint32)

    def set_options(self, **opts) -> "Graph":
        """Launch options (ng_graph_set_option): block, max_slots, lds_maps, pile_capacity, log_capacity."""
        for k, v in opts.items():
            _check(lib().ng_graph_set_option(self.h, _OPTIONS[k], int(v)))
        return self

    def set_weights(self, weights) -> bool:
        """Replaces the arc weights in place (ng_graph_set_weights), same dtype and order as at creation; returns
        whether the rows had to be re-sorted."""
        w = np.ascontiguousarray(weights, dtype=self.host.weights.dtype)
        if w.shape != (self.m,):
            raise ValueError(f"expected {self.m} weights, got shape {w.shape}")
        _check(lib().ng_graph_set_weights(self.h, _ptr(w)))
        self.host = HostGraph(self.n, self.host.offsets, self.host.targets, w, self.host.positions)
        r = ctypes.c_int()
        _check(lib().ng_graph_weights_reordered(self.h, ctypes.byref(r)))
        return bool(r.value)

    def get_option(self, name: str) -> int:
        v = ctypes.c_int64()
        _check(lib().ng_graph_get_option(self.h, _OPTIONS[name], ctypes.byref(v)))
        return v.value

    def dist_dtype(self, lkind: int = NG_LEN_ABS):
        k = ctypes.c_int()
        _check(lib().ng_graph_dist_kind(self.h, lkind, ctypes.byref(k)))
        return _DDTYPE[k.value]

    def sssp(self, sources=None, lkind: int = NG_LEN_ABS, delta: float = 0.0, want_dist: bool = False):
        src = self._sources(sources)
        out = np.zeros(len(src), dtype=SUMMARY_DTYPE)
        dist = np.zeros((len(src), self.n), dtype=self.dist_dtype(lkind)) if want_dist else None
        _check(lib().ng_sssp_summaries(self.h, _ptr(src), len(src), lkind, float(delta), _ptr(out), _ptr(dist)))
        return (out, dist) if want_dist else out

    def distance_histogram(self, bin_width: float, n_bins: int, sources=None, lkind: int = NG_LEN_ABS, delta: float = 0.0,
                           summaries: bool = False):
        """Counts of finite d(s, t), t != s, in bins [i*w, (i+1)*w); the last bin takes everything beyond."""
        src = self._sources(sources)
        counts = np.zeros(n_bins, dtype=np.uint64)
        out = np.zeros(len(src), dtype=SUMMARY_DTYPE) if summaries else None
        _check(lib().ng_distance_histogram(self.h, _ptr(src), len(src), lkind, float(delta), float(bin_width), n_bins,
                                           _ptr(counts), _ptr(out)))
        return (counts, out) if summaries else counts

    def brandes(self, sources=None, delta: float = 0.0, bc=None, summaries: bool = False):
        src = self._sources(sources)
        bc = np.zeros(self.n, dtype=np.float64) if bc is None else bc
        out = np.zeros(len(src), dtype=SUMMARY_DTYPE) if summaries else None
        _check(lib().ng_brandes_accumulate(self.h, _ptr(src), len(src), float(delta), _ptr(bc), _ptr(out)))
        return (bc, out) if summaries else bc

    def brandes_noalloc(self, sources=None, delta: float = 0.0, bc=None, summaries: bool = False):
        """As brandes(), through the allocation-free entry point (fixed device scratch array per GPU)."""
        src = self._sources(sources)
        bc = np.zeros(self.n, dtype=np.float64) if bc is None else bc
        out = np.zeros(len(src), dtype=SUMMARY_DTYPE) if summaries else None
        _check(lib().ng_brandes_accumulate_noalloc(self.h, _ptr(src), len(src), float(delta), _ptr(bc), _ptr(out)))
        return (bc, out) if summaries else bc

    def edge_dependency(self, sources=None, delta: float = 0.0):
        src = self._sources(sources)
        edge = np.zeros(self.m, dtype=np.float64)
        _check(lib().ng_edge_dependency(self.h, _ptr(src), len(src), float(delta), _ptr(edge)))
        return edge

    def adjoint(self, sources=None, delta: float = 0.0, summaries: bool = False):
        src = self._sources(sources)
        g1 = np.zeros(self.m, dtype=np.float64)
        g2 = np.zeros(self.m, dtype=np.float64)
        out = np.zeros(len(src), dtype=SUMMARY_DTYPE) if summaries else None
        _check(lib().ng_apsp_adjoint(self.h, _ptr(src), len(src), float(delta), _ptr(g1), _ptr(g2), _ptr(out)))
        return (g1, g2, out) if summaries else (g1, g2)

    def autotune(self, sample, lkind: int = NG_LEN_ABS, mode: int = 0) -> float:
        src = np.ascontiguousarray(sample, dtype=np.uint32)
        d = ctypes.c_double()
        _check(lib().ng_autotune_delta(self.h, lkind, mode, _ptr(src), len(src), ctypes.byref(d)))
        return d.value

    def autotune_shape(self, sample, lkind: int = NG_LEN_ABS, mode: int = 0):
        """Tunes bucket width and lanes per workgroup together; returns (delta, block)."""
        src = np.ascontiguousarray(sample, dtype=np.uint32)
        d = ctypes.c_double()
        b = ctypes.c_uint32()
        _check(lib().ng_autotune(self.h, lkind, mode, _ptr(src), len(src), ctypes.byref(d), ctypes.byref(b)))
        return d.value, b.value

    def last_stats(self) -> dict:
        s = CallStats()
        _check(lib().ng_graph_last_stats(self.h, ctypes.byref(s)))
        out = {f: getattr(s, f) for f, _ in CallStats._fields_}
        out["section_ticks"] = list(s.section_ticks)
        return out


class Cohort:
    """All sources of many graphs in one persistent, hipGraph-captured launch (ng_cohort_*). The member
    Graph objects must stay alive while the cohort is used."""

    def __init__(self, ctx: Context, graphs):
        self.ctx = ctx
        self.graphs = list(graphs)
        arr = (ctypes.c_void_p * len(self.graphs))(*[g.h.value for g in self.graphs])
        h = ctypes.c_void_p()
        _check(lib().ng_cohort_create(ctx.h, arr, len(self.graphs), ctypes.byref(h)))
        self.h = h
        self.sizes = [g.n for g in self.graphs]
        self.prefix = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)

    def close(self):
        if self.h:
            lib().ng_cohort_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def autotune(self, n_sample: int = 256):
        _check(lib().ng_cohort_autotune(self.h, n_sample))

    def brandes(self, summaries: bool = False):
        """Returns (list of per-graph betweenness vectors, summaries or None, replay ms)."""
        total = int(self.prefix[-1])
        bc = np.zeros(total, dtype=np.float64)
        out = np.zeros(total, dtype=SUMMARY_DTYPE) if summaries else None
        ms = ctypes.c_double()
        _check(lib().ng_cohort_brandes(self.h, _ptr(bc), _ptr(out), ctypes.byref(ms)))
        parts = [bc[self.prefix[i]:self.prefix[i + 1]] for i in range(len(self.graphs))]
        return parts, out, ms.value

    def last_stats(self) -> dict:
        s = CallStats()
        _check(lib().ng_cohort_last_stats(self.h, ctypes.byref(s)))
        out = {f: getattr(s, f) for f, _ in CallStats._fields_}
        out["section_ticks"] = list(s.section_ticks)
        return out


def version() -> str:
    return lib().ng_version().decode()
