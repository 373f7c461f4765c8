"""CPU-only checks of libng_hip.so: every declared entry point is exported, host-side errors are clean
without a GPU, the loader follows the reference's file-format rules, and the generators are deterministic.
"""

import os
import re

import numpy as np
import pytest

import neurograph_amd as ng
from neurograph_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_INPUT = "/root/reference/input"


def declared_functions():
    names = set()
    for h in ("neurograph_hip.h", "neurograph_io.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(ng_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    L = ng.lib()
    names = declared_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(L, name), name
    assert ng.version().startswith("ng_hip")


def test_host_side_argument_errors_need_no_gpu():
    L = ng.lib()
    import ctypes

    h = ctypes.c_void_p()
    off = np.array([0, 1], dtype=np.uint64)
    tgt = np.array([0], dtype=np.uint32)
    w = np.array([1], dtype=np.int32)
    assert L.ng_graph_create(None, 1, 1, off.ctypes.data, tgt.ctypes.data, w.ctypes.data, 0, ctypes.byref(h)) == _lib.NG_EINVAL
    assert b"context" in L.ng_last_error()
    assert L.ng_sssp_summaries(None, None, 0, 0, 0.0, None, None) == _lib.NG_EINVAL
    st = L.ng_context_create(0, ctypes.byref(h))
    if st == _lib.NG_OK:  # a GPU is present: clean up
        L.ng_context_destroy(h)
    else:
        assert st in (_lib.NG_EDEVICE, _lib.NG_EINVAL)
        assert L.ng_last_error()


def _write(path, text):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        f.write(text)


def _positions(path, n):
    _write(os.path.join(path, "positions", "rank_0_positions.txt"),
           "".join(f"{i + 1} 0.{i} 0.5 0.5 area_1 excitatory\n" for i in range(n)))


def test_loader_node_file_rules(tmp_path):
    # tests/file_loader/test_file_loader.cpp:26-61: comments, blanks, malformed lines and wrong ids are skipped
    d = str(tmp_path)
    _write(os.path.join(d, "positions", "rank_0_positions.txt"),
           "# a comment line\n\nnot a valid node line\n1 0.1 0.2 0.3 excitatory\n5 0.4 0.5 0.6 area_1 inhibitory\n"
           "2 0.7 0.8 0.9 area_1 excitatory\n")
    _write(os.path.join(d, "network", "rank_0_in_network.txt"), "")
    _write(os.path.join(d, "network", "rank_0_out_network.txt"), "")
    g = ng.load_graph_dir(d)
    assert g.n == 2 and g.m == 0
    np.testing.assert_allclose(g.positions, [[0.1, 0.2, 0.3], [0.7, 0.8, 0.9]])


def test_loader_arc_file_rules(tmp_path):
    # tests/file_loader/test_file_loader.cpp:63-131: the same invalid lines are skipped in both files
    d = str(tmp_path)
    _positions(d, 3)
    _write(os.path.join(d, "network", "rank_0_out_network.txt"),
           "# comment\n\ngarbage line\n0 1 0 1 1 0\n0 1 0 1 0 1\n0 1 5 1 1 1\n0 1 0 99 1 1\n5 1 0 1 1 1\n0 99 0 1 1 1\n"
           "0 2 0 1 7 1\n")
    _write(os.path.join(d, "network", "rank_0_in_network.txt"),
           "# comment\n\ngarbage line\n0 1 0 2 1 0\n0 1 0 2 0 1\n1 1 0 2 1 1\n0 99 0 2 1 1\n0 1 5 1 1 1\n0 1 0 99 1 1\n"
           "0 1 0 2 4 1\n")
    g = ng.load_graph_dir(d)
    assert g.n == 3 and g.m == 1
    assert (int(g.targets[0]), int(g.weights[0])) == (1, 7) and int(g.offsets[1]) == 1  # node 0 -> node 1, weight 7
    # undirected: the in arc 0 <- 1 (weight 4) is added to the out arc 0 -> 1 (ArcTransformer.cpp:53-67)
    gu = ng.load_graph_dir(d, undirected=True)
    assert gu.m == 1 and int(gu.weights[0]) == 11
    g1 = ng.load_graph_dir(d, all_weights_one=True)
    assert int(g1.weights[0]) == 1


def test_loader_missing_file(tmp_path):
    d = str(tmp_path)
    _positions(d, 3)
    with pytest.raises(ng.NgError):
        ng.load_graph_dir(d)


def test_loader_merge_transforms_and_integer_parse(tmp_path):
    d = str(tmp_path)
    _positions(d, 4)
    # duplicates merge by summing, a self-loop, an integer parse of "2.7" (reads 2), a -3/+3 pair summing to 0
    _write(os.path.join(d, "network", "rank_0_out_network.txt"),
           "0 2\t0 1\t3\t1\n0 2\t0 1\t4\t1\n0 1\t0 1\t5\t1\n0 3\t0 2\t2.7\n0 4\t0 3\t-3\t1\n0 4\t0 3\t3\t1\n")
    _write(os.path.join(d, "network", "rank_0_in_network.txt"), "0 1\t0 2\t6\t1\n")
    g = ng.load_graph_dir(d)
    arcs = {(u, int(g.targets[e])): int(g.weights[e]) for u in range(g.n) for e in range(int(g.offsets[u]), int(g.offsets[u + 1]))}
    assert arcs == {(0, 1): 7, (0, 0): 5, (1, 2): 2}  # the merged 0-weight arc 2->3 is not uploaded
    g2 = ng.load_graph_dir(d, remove_self_arcs=True)
    assert int(g2.m) == 2
    gf = ng.load_graph_dir(d, weights=np.float64)
    arcs_f = {(u, int(gf.targets[e])): float(gf.weights[e]) for u in range(gf.n) for e in range(int(gf.offsets[u]), int(gf.offsets[u + 1]))}
    assert arcs_f[(1, 2)] == 2.7
    gu = ng.load_graph_dir(d, undirected=True)
    arcs_u = {(u, int(gu.targets[e])): int(gu.weights[e]) for u in range(gu.n) for e in range(int(gu.offsets[u]), int(gu.offsets[u + 1]))}
    # the in arc of node 0 from node 1 (file line "0 1 0 2 6") is added to the out arc 0 -> 1: 7 + 6
    assert arcs_u[(0, 1)] == 13


@pytest.mark.skipif(not os.path.isdir(REF_INPUT), reason="reference input directory not present")
def test_bundled_inputs_load_and_flatten():
    base = ng.load_graph_dir(os.path.join(REF_INPUT, "rank_1_nodes_5000_arcs_6"))
    assert base.n == 5000 and base.m == 29987
    w = base.weights
    assert int(np.sum(w == 1)) == 29974 and int(np.sum(w == 2)) == 13  # duplicates merged by summing
    # independent parse of the well-formed out file: merge duplicate (source, target) pairs by summing
    merged = {}
    with open(os.path.join(REF_INPUT, "rank_1_nodes_5000_arcs_6", "network", "rank_0_out_network.txt")) as f:
        for line in f:
            if not line.strip() or line.startswith("#"):
                continue
            tr, ti, sr, si, wt = line.split()[:5]
            key = (int(si) - 1, int(ti) - 1)
            merged[key] = merged.get(key, 0) + int(wt)
    src = np.repeat(np.arange(base.n), np.diff(base.offsets).astype(np.int64))
    got = {(int(u), int(t)): int(x) for u, t, x in zip(src, base.targets, base.weights)}
    assert got == {k: v for k, v in merged.items() if v != 0}
    # "nodes_5000" counts nodes per rank: the r-rank inputs flatten to 5000 * r nodes (global id = prefix + local)
    for r in (2, 4, 8):
        other = ng.load_graph_dir(os.path.join(REF_INPUT, f"rank_{r}_nodes_5000_arcs_6"))
        assert other.n == 5000 * r
        assert abs(other.m - 6 * other.n) < 0.01 * other.n
        assert other.targets.max() < other.n
        # arcs cross rank boundaries after flattening
        osrc = np.repeat(np.arange(other.n), np.diff(other.offsets).astype(np.int64))
        assert np.any((osrc // 5000) != (other.targets // 5000))


def test_writer_loader_round_trip(tmp_path):
    for g in (ng.generate_fixed_degree(300, 6, 1, 10, 3), ng.generate_spatial(500, 8, seed=4, weights=np.float64),
              ng.generate_spatial(500, 8, seed=4, weights=np.float32)):
        d = str(tmp_path / f"g{g.weights.dtype}")
        ng.write_graph_dir(d, g)
        back = ng.load_graph_dir(d, weights=g.weights.dtype)
        np.testing.assert_array_equal(back.offsets, g.offsets)
        np.testing.assert_array_equal(back.targets, g.targets)
        np.testing.assert_array_equal(back.weights, g.weights)


def test_fixed_degree_generator():
    g = ng.generate_fixed_degree(5000, 6, 1, 10, 0)
    assert g.n == 5000 and g.m == 30000
    deg = np.diff(g.offsets)
    assert np.all(deg == 6)
    src = np.repeat(np.arange(g.n), 6)
    assert not np.any(src == g.targets)  # no self-loops
    pairs = src.astype(np.int64) * g.n + g.targets
    assert len(np.unique(pairs)) == g.m  # duplicates rejected
    assert g.weights.min() == 1 and g.weights.max() == 10
    h = ng.generate_fixed_degree(5000, 6, 1, 10, 0)
    np.testing.assert_array_equal(h.targets, g.targets)
    np.testing.assert_array_equal(h.weights, g.weights)


def test_spatial_generator():
    g = ng.generate_spatial(20000, 30, seed=5, weights=np.float64, wdist="uniform", a=0.5, b=1.5)
    assert g.n == 20000
    assert abs(g.m / g.n - 30) < 0.5
    assert g.weights.min() >= 0.5 and g.weights.max() < 1.5
    assert np.all((g.positions >= 0) & (g.positions <= 1000))
    src = np.repeat(np.arange(g.n), np.diff(g.offsets).astype(np.int64))
    assert not np.any(src == g.targets)
    d = np.linalg.norm(g.positions[src] - g.positions[g.targets], axis=1)
    # arc lengths follow the exp(-d/150) profile: far shorter than random pairs (~660 um in a 1000^3 box)
    assert np.median(d) < 400
    h = ng.generate_spatial(20000, 30, seed=5, weights=np.float64, wdist="uniform", a=0.5, b=1.5)
    np.testing.assert_array_equal(h.targets, g.targets)
    ln = ng.generate_spatial(5000, 20, seed=1, weights=np.float32)
    lw = np.log(ln.weights.astype(np.float64))
    assert abs(lw.mean()) < 0.02 and abs(lw.std() - 0.5) < 0.02
