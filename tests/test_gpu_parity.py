"""GPU parity: the HIP path (through the C ABI of libng_hip.so) against the reference's own known-answer
values and against the CPU oracle on the same seeded inputs.

Bars: integer distances, sums, maxima, unreachable counts, reach histograms and sigma-derived exact
quantities are bit-exact; real-valued distances are bit-exact too (same single IEEE add per relaxation,
the fixpoint min over the same candidates); order-dependent double sums (efficiency sums, betweenness,
dependencies, gradients) are compared with a relative tolerance stated per test (the reference's own
tests use 1e-6 absolute on these).
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import oracle as orc
from tests.conftest import csr_from_fixture

DELTAS = [0.0, 1, 3, 4, 10, 10000]


@pytest.fixture(scope="module")
def ctx():
    import neurograph_amd as ng

    return ng.Context(0)


def dev_graph(ctx, c, **opts):
    """Device graph of an oracle CSR; opts are launch options (block, max_slots, lds_maps, pile_capacity,
    log_capacity) set through ng_graph_set_option."""
    import neurograph_amd as ng

    return ng.Graph(ctx, ng.HostGraph(c.n, c.off, c.tgt, c.w)).set_options(**opts)


def rand_graph(n, m, seed, wdtype=np.int32, wlo=1, whi=10, parallel=True, self_loops=True):
    rng = np.random.default_rng(seed)
    src = rng.integers(0, n, size=m)
    # skewed out-degrees: a few hubs
    hubs = rng.integers(0, n, size=max(1, n // 50))
    src[: m // 4] = rng.choice(hubs, size=m // 4)
    dst = rng.integers(0, n, size=m)
    if not self_loops:
        keep = src != dst
        src, dst = src[keep], dst[keep]
    if not parallel:
        _, idx = np.unique(src * n + dst, return_index=True)
        src, dst = src[idx], dst[idx]
    if np.issubdtype(wdtype, np.integer):
        w = rng.integers(wlo, whi + 1, size=len(src)).astype(wdtype)
    else:
        w = np.exp(rng.normal(0.0, 0.5, size=len(src))).astype(wdtype)
    return orc.Csr.from_edges(n, src, dst, w, dtype=wdtype)


FIXTURES = ["standard", "standard_uu", "full", "negative", "four", "four_uu", "full36", "seven", "seven_uu", "full63"]


@pytest.mark.parametrize("name", FIXTURES)
def test_fixture_per_source_tables(ctx, reference_fixtures, name):
    fx = reference_fixtures["fixtures"][name]
    g = dev_graph(ctx, csr_from_fixture(fx))
    rows = fx["per_source"]["rows"]
    for delta in DELTAS:
        s = g.sssp(None, delta=delta)
        for i, (sd, mx, eff, un) in enumerate(rows):
            assert int(s["sum_d_u64"][i]) == sd, (name, delta, i)
            assert int(s["max_d_u64"][i]) == mx, (name, delta, i)
            assert int(s["n_unreached"][i]) == un, (name, delta, i)
            assert s["sum_eff"][i] == pytest.approx(eff, abs=fx["per_source"]["eff_tol"]), (name, delta, i)


@pytest.mark.parametrize("name", [f for f in FIXTURES if f != "negative"] + ["disconnected"])
def test_fixture_apsp_results(ctx, reference_fixtures, name):
    from neurograph_amd import NG_LEN_ABS, NG_LEN_INVERSE
    from neurograph_amd.metrics import apsp

    fx = reference_fixtures["fixtures"][name]
    g = dev_graph(ctx, csr_from_fixture(fx))
    for key, lk in (("apsp", NG_LEN_ABS), ("apsp_inverse", NG_LEN_INVERSE)):
        if key not in fx:
            continue
        exp = fx[key]
        got = apsp(g, lk)
        if exp["spl_exact"] and lk == NG_LEN_ABS:
            assert got.average_shortest_path_length == exp["avg_spl"]
        else:
            assert got.average_shortest_path_length == pytest.approx(exp["avg_spl"], abs=1e-6)
        assert got.average_efficiency == pytest.approx(exp["avg_eff"], abs=1e-6)
        assert got.diameter == pytest.approx(exp["diameter"], abs=1e-6)
        assert got.number_disconnected_pairs == exp["n_disconnected"]
        assert got.cluster_sizes == {int(k): v for k, v in exp["cluster_sizes"].items()}


@pytest.mark.parametrize("name", [f for f in FIXTURES if f != "negative"])
def test_fixture_brandes(ctx, reference_fixtures, name):
    fx = reference_fixtures["fixtures"][name]
    c = csr_from_fixture(fx)
    g = dev_graph(ctx, c)
    ref = orc.brandes(c)
    for delta in (0.0, 1, 4, 10000):
        bc = g.brandes(None, delta=delta)
        assert np.sum(bc) / c.n == pytest.approx(fx["brandes_avg"]["value"], abs=1e-6)
        np.testing.assert_allclose(bc, ref["bc"], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("seed,n,m", [(1, 60, 400), (2, 300, 1500), (3, 1000, 12000), (4, 2500, 7000)])
def test_random_integer_graphs_bit_exact(ctx, seed, n, m):
    c = rand_graph(n, m, seed)
    g = dev_graph(ctx, c)
    ref = orc.apsp(c, per_source=True, threads=8)["per_source"]
    ref_dist = np.stack([orc.sssp(c, s)[0] for s in range(min(n, 64))])
    for delta in (0.0, 1, 2, 7, 100):
        s, dist = g.sssp(np.arange(min(n, 64)), delta=delta, want_dist=True)
        np.testing.assert_array_equal(dist, ref_dist)
        full = g.sssp(None, delta=delta)
        np.testing.assert_array_equal(full["sum_d_u64"], ref["sum_d_u64"])
        np.testing.assert_array_equal(full["max_d_u64"], ref["max_d_u64"])
        np.testing.assert_array_equal(full["n_unreached"], ref["n_unreached"])
        np.testing.assert_allclose(full["sum_eff"], ref["sum_eff"], rtol=1e-12)


@pytest.mark.parametrize("wdtype", [np.float32, np.float64])
def test_random_real_graphs_bit_exact(ctx, wdtype):
    c = rand_graph(1500, 15000, 11, wdtype=wdtype, parallel=False, self_loops=False)
    g = dev_graph(ctx, c)
    srcs = np.arange(0, 1500, 7)
    ref_dist = np.stack([orc.sssp(c, int(s))[0] for s in srcs])
    ref_sum = orc.apsp(c, sources=srcs, per_source=True)["per_source"]
    for delta in (0.0, 0.05, 0.3, 1.0, 5.0):
        s, dist = g.sssp(srcs, delta=delta, want_dist=True)
        np.testing.assert_array_equal(dist, ref_dist)
        np.testing.assert_allclose(s["sum_d"], ref_sum["sum_d"], rtol=1e-12)
        np.testing.assert_array_equal(s["max_d"], ref_sum["max_d"])
        np.testing.assert_array_equal(s["n_unreached"], ref_sum["n_unreached"])


def test_inverse_metric_bit_exact(ctx):
    from neurograph_amd import NG_LEN_INVERSE

    c = rand_graph(800, 6000, 5, wlo=-4, whi=9)  # zero weights become infinitely long, negatives use |w|
    g = dev_graph(ctx, c)
    srcs = np.arange(0, 800, 5)
    ref = np.stack([orc.sssp(c, int(s), orc.L_INVERSE)[0] for s in srcs])
    _, dist = g.sssp(srcs, lkind=NG_LEN_INVERSE, want_dist=True)
    np.testing.assert_array_equal(dist, ref)


@pytest.mark.parametrize("wdtype,rtol", [(np.int32, 1e-9), (np.float32, 1e-9), (np.float64, 1e-9)])
def test_brandes_random_graphs(ctx, wdtype, rtol):
    c = rand_graph(700, 5000, 21, wdtype=wdtype, wlo=1, whi=4, parallel=True, self_loops=True)
    g = dev_graph(ctx, c)
    ref = orc.brandes(c, threads=8, per_source=True)
    for delta in (0.0, 1, 3):
        bc, summ = g.brandes(None, delta=delta, summaries=True)
        np.testing.assert_allclose(bc, ref["bc"], rtol=rtol, atol=1e-9)
        np.testing.assert_array_equal(summ["n_unreached"], ref["per_source"]["n_unreached"])
        np.testing.assert_allclose(summ["sum_d"], ref["per_source"]["sum_d"], rtol=1e-12)


def test_brandes_is_deterministic_and_accumulates(ctx):
    c = rand_graph(900, 9000, 8, wdtype=np.float32)
    g = dev_graph(ctx, c)
    a = g.brandes(None)
    b = g.brandes(None)
    np.testing.assert_array_equal(a, b)
    half = g.brandes(np.arange(450))
    both = g.brandes(np.arange(450, 900), bc=half.copy())
    np.testing.assert_allclose(both, a, rtol=1e-12, atol=1e-12)


def test_betweenness_partials_stay_clean_across_calls(ctx):
    """Calls of every kind on one context's workspace, with different slot counts, in a row: each Brandes call starts
    from clean per-slot betweenness partials (cleared by the previous Brandes call's slot reduction, or by a zeroing
    pass after edge-dependency and adjoint calls and cohort replays, which leave partials of their own)."""
    import neurograph_amd as ng

    own = ng.Context(0)
    c = rand_graph(2000, 16000, 61, wdtype=np.float64)
    c2 = rand_graph(900, 7000, 62, wdtype=np.float64)
    g = dev_graph(own, c)
    g3 = dev_graph(own, c, max_slots=3)
    h = dev_graph(own, c2)
    srcs = np.arange(0, 2000, 9)
    s2 = np.arange(0, 900, 5)
    ref = orc.brandes(c, sources=srcs)["bc"]
    ref2 = orc.brandes(c2, sources=s2)["bc"]
    edge = orc.brandes(c, sources=srcs, edge=True)["edge"]
    full2 = orc.brandes(c2, threads=8)["bc"]
    co = ng.Cohort(own, [h])
    for _ in range(2):
        np.testing.assert_allclose(g.brandes(srcs), ref, rtol=1e-9, atol=1e-9)
        np.testing.assert_allclose(g3.brandes(srcs), ref, rtol=1e-9, atol=1e-9)
        np.testing.assert_allclose(g.brandes(srcs), ref, rtol=1e-9, atol=1e-9)
        np.testing.assert_allclose(g.edge_dependency(srcs), edge, rtol=1e-9, atol=1e-12)
        np.testing.assert_allclose(h.brandes(s2), ref2, rtol=1e-9, atol=1e-9)
        g.adjoint(srcs)
        np.testing.assert_allclose(g.brandes(srcs), ref, rtol=1e-9, atol=1e-9)
        parts, _, _ = co.brandes()
        np.testing.assert_allclose(parts[0], full2, rtol=1e-9, atol=1e-9)
        np.testing.assert_allclose(h.brandes(s2), ref2, rtol=1e-9, atol=1e-9)


def test_edge_dependency_and_adjoint_match_oracle(ctx):
    c = rand_graph(400, 3000, 31, wdtype=np.float64, parallel=False, self_loops=False)
    g = dev_graph(ctx, c)
    one = orc.brandes(c, edge=True)
    inv = orc.brandes(c, edge=True, seed=orc.SEED_INV_D2)
    edge = g.edge_dependency(None)
    np.testing.assert_allclose(edge, one["edge"], rtol=1e-9, atol=1e-12)
    g1, g2 = g.adjoint(None)
    np.testing.assert_allclose(g1, one["edge"], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(g2, -inv["edge"], rtol=1e-9, atol=1e-12)


def test_integer_ties_edge_dependency(ctx):
    c = rand_graph(300, 2500, 41, wlo=1, whi=3)  # many ties: sigma > 1
    g = dev_graph(ctx, c)
    ref = orc.brandes(c, edge=True)
    np.testing.assert_allclose(g.edge_dependency(None), ref["edge"], rtol=1e-9, atol=1e-12)


def test_error_paths(ctx):
    import neurograph_amd as ng

    c = rand_graph(50, 200, 2)
    g = dev_graph(ctx, c)
    with pytest.raises(ng.NgError) as e:
        g.sssp([50])
    assert e.value.status == 1
    neg = orc.Csr.from_edges(3, [0, 1], [1, 2], [-1, 2])
    gn = dev_graph(ctx, neg)
    with pytest.raises(ng.NgError) as e:
        gn.sssp([0], lkind=ng.NG_LEN_RAW)
    assert e.value.status == 1
    assert int(gn.sssp([0])["sum_d_u64"][0]) == 1 + 3  # |w| for the Dijkstra metric
    zero = orc.Csr.from_edges(3, [0, 1], [1, 2], [0.0, 2.0], dtype=np.float64)
    gz = dev_graph(ctx, zero)
    with pytest.raises(ng.NgError) as e:
        gz.brandes(None)
    assert e.value.status == 4
    empty = g.sssp(np.zeros(0, np.uint32))
    assert len(empty) == 0


def test_fixed_degree_config_parity(ctx):
    """5000 nodes, exactly 6 distinct out-arcs, integer weights in [1, 10], seed 0: every source."""
    import neurograph_amd as ng
    from neurograph_amd.metrics import aggregate_apsp

    hg = ng.generate_fixed_degree(5000, 6, 1, 10, 0)
    c = orc.Csr(hg.n, hg.offsets, hg.targets, hg.weights)
    g = ng.Graph(ctx, hg)
    s = g.sssp(None)
    ref = orc.apsp(c, threads=8, per_source=True)
    np.testing.assert_array_equal(s["sum_d_u64"], ref["per_source"]["sum_d_u64"])
    np.testing.assert_array_equal(s["max_d_u64"], ref["per_source"]["max_d_u64"])
    np.testing.assert_array_equal(s["n_unreached"], ref["per_source"]["n_unreached"])
    agg = aggregate_apsp(s, hg.n, True)
    assert agg.average_shortest_path_length == ref["avg_spl"]
    assert agg.diameter == ref["diameter_u64"]
    assert agg.number_disconnected_pairs == ref["n_disconnected"]
    assert agg.cluster_sizes == ref["cluster_sizes"]
    assert agg.average_efficiency == pytest.approx(ref["avg_eff"], rel=1e-12)


def test_small_pile_forces_exact_bucket_advance(ctx):
    """With the pile reserve exhausted, bucket advances reload distances instead of trusting parked values."""
    import neurograph_amd as ng

    own = ng.Context(0)  # a fresh context, so that the workspace stride is this graph's n
    for wdtype in (np.int32, np.float32):
        c = rand_graph(3000, 30000, 77, wdtype=wdtype, wlo=1, whi=20)
        # clamped to n + 2048 entries: re-parking is refused almost always
        g = dev_graph(own, c, pile_capacity=1)
        assert g.get_option("pile_capacity") == 1
        srcs = np.arange(0, 3000, 37)
        ref = np.stack([orc.sssp(c, int(s))[0] for s in srcs])
        for delta in (0.0, 0.5, 3.0):
            _, dist = g.sssp(srcs, delta=delta, want_dist=True)
            np.testing.assert_array_equal(dist, ref)
        bc = g.brandes(srcs)
        np.testing.assert_allclose(bc, orc.brandes(c, sources=srcs)["bc"], rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("block", ["64", "256", "1024"])
@pytest.mark.parametrize("maps", [str(k) for k in range(20)])
def test_lds_bitmap_variants(ctx, maps, block):
    """Every combination of the optional LDS maps (settle passes in discovery order or in increasing id;
    frontier de-duplication by global stamps or by an LDS bitmap; probes with no bounds, 16-bit bounds or, for
    maps 8-15, the 4-bit distance codes; for maps 16-19 in 256-lane workgroups, the distances themselves in LDS)
    gives the oracle's distances bit-exactly and its betweenness within
    1e-9, for integer, FP32 and FP64 weights and both length kinds (|w| and 1/|w|), in 64- and 256-lane
    workgroups (plain arc loop) and 1024-lane ones (filter stage, survivors compacted by lane permutes)."""
    opts = {"lds_maps": int(maps), "block": int(block)}
    # int32 weights up to 2^31 - 1 take the 64-bit integer distances, whose bounds round
    for wdtype, seed, wlo, whi in ((np.int32, 5, 1, 9), (np.int32, 9, 1 << 20, 2**31 - 1), (np.float32, 7, 1, 9),
                                   (np.float64, 6, 1, 9)):
        c = rand_graph(4000, 40000, seed, wdtype=wdtype, wlo=wlo, whi=whi)
        g = dev_graph(ctx, c, **opts)
        srcs = np.arange(0, 4000, 29)
        for lkind in (orc.L_ABS, orc.L_INVERSE):
            ref = np.stack([orc.sssp(c, int(s), lkind)[0] for s in srcs])
            scale = (whi if lkind == orc.L_ABS else 1.0 / whi) / 9.0
            for delta in (0.0, 0.7, 5.0):
                _, dist = g.sssp(srcs, lkind=lkind, delta=delta * scale, want_dist=True)
                np.testing.assert_array_equal(dist, ref)
        bc, summ = g.brandes(srcs, summaries=True)
        np.testing.assert_allclose(bc, orc.brandes(c, sources=srcs)["bc"], rtol=1e-9, atol=1e-9)
    # a graph of unit weights ties everywhere (sigma > 1, the general dependency layers)
    c = rand_graph(3000, 30000, 8, wdtype=np.int32, wlo=1, whi=2)
    g = dev_graph(ctx, c, **opts)
    srcs = np.arange(0, 3000, 37)
    np.testing.assert_allclose(g.brandes(srcs), orc.brandes(c, sources=srcs)["bc"], rtol=1e-9, atol=1e-9)
    assert g.last_stats()["block"] == int(block)


def test_shared_workspace_across_graphs(ctx):
    """Graphs of one context share its workspace (grown to the largest graph, never shrunk): alternating a large
    and a small graph, and distance kinds, keeps every result equal to the oracle's."""
    big = rand_graph(6000, 50000, 31, wdtype=np.float64)
    small = rand_graph(700, 5000, 32, wdtype=np.int32, wlo=1, whi=5)
    mid = rand_graph(2500, 20000, 33, wdtype=np.float32)
    gs = {"small": (small, dev_graph(ctx, small)), "big": (big, dev_graph(ctx, big)), "mid": (mid, dev_graph(ctx, mid))}
    for name in ("small", "big", "small", "mid", "big", "small"):
        c, g = gs[name]
        srcs = np.arange(0, c.n, max(1, c.n // 90))
        ref = np.stack([orc.sssp(c, int(s))[0] for s in srcs])
        _, dist = g.sssp(srcs, want_dist=True)
        np.testing.assert_array_equal(dist, ref)
        np.testing.assert_allclose(g.brandes(srcs), orc.brandes(c, sources=srcs)["bc"], rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("wdtype", [np.int32, np.float32, np.float64])
def test_distance_histogram_and_matrix(ctx, wdtype):
    """The device distance histogram equals numpy's binning of the oracle's distance matrix, count for count;
    the matrix itself is bit-exact (fixed-degree C1 generator for integer weights)."""
    import neurograph_amd as ng

    if wdtype == np.int32:
        hg = ng.generate_fixed_degree(1500, 6, 1, 10, 0)
        width, bins = 1.0, 64
    else:
        hg = ng.generate_spatial(1500, 12, weights=wdtype, seed=3)
        width, bins = 0.25, 40
    c = orc.Csr(hg.n, hg.offsets, hg.targets, hg.weights)
    g = ng.Graph(ctx, hg)
    srcs = np.arange(0, hg.n, 3)
    ref = orc.brandes(c, sources=srcs, want_dist=True, threads=8)["dist"]
    _, dist = g.sssp(srcs, want_dist=True)
    np.testing.assert_array_equal(dist, ref)
    counts, summ = g.distance_histogram(width, bins, srcs, summaries=True)
    unreach = np.iinfo(np.uint64).max if ref.dtype == np.uint64 else np.finfo(ref.dtype).max
    mask = ref != unreach
    mask[np.arange(len(srcs)), srcs] = False
    d = ref[mask].astype(np.float64)
    b = np.minimum(np.floor(d / width), bins - 1).astype(np.int64)
    np.testing.assert_array_equal(counts, np.bincount(b, minlength=bins).astype(np.uint64))
    assert int(counts.sum()) == len(srcs) * (hg.n - 1) - int(summ["n_unreached"].sum())


@pytest.mark.parametrize("block", ["64", "256", "512", "1024"])
def test_workgroup_sizes(ctx, block):
    """Every workgroup size the kernel is built for (one source per wavefront, or per workgroup of 256, 512 or
    1024 lanes) gives
    the oracle's distances bit-exactly and its betweenness and edge path counts within 1e-9."""
    import neurograph_amd as ng

    for wdtype, seed in ((np.int32, 15), (np.float32, 16), (np.float64, 17)):
        c = rand_graph(3000, 30000, seed, wdtype=wdtype, wlo=1, whi=9)
        g = dev_graph(ctx, c, block=int(block))
        srcs = np.arange(0, 3000, 23)
        ref = np.stack([orc.sssp(c, int(s))[0] for s in srcs])
        _, dist = g.sssp(srcs, delta=0.8, want_dist=True)
        np.testing.assert_array_equal(dist, ref)
        assert g.last_stats()["block"] == int(block)
        bc = g.brandes(srcs)
        np.testing.assert_allclose(bc, orc.brandes(c, sources=srcs)["bc"], rtol=1e-9, atol=1e-9)
        edge = g.edge_dependency(srcs)
        np.testing.assert_allclose(edge, orc.brandes(c, sources=srcs, edge=True)["edge"], rtol=1e-9, atol=1e-12)


def test_autotune_shape(ctx):
    import neurograph_amd as ng

    c = rand_graph(5000, 60000, 21, wdtype=np.float32)
    g = dev_graph(ctx, c)
    d, b = g.autotune_shape(np.arange(0, 5000, 2), mode=1)
    assert d > 0 and b in (64, 256, 512, 1024)
    srcs = np.arange(0, 5000, 41)
    g.brandes(srcs)
    assert g.last_stats()["block"] == b and g.last_stats()["delta"] == d


def test_random_configurations(ctx):
    """30 seeded random cases over graph size and density (uniform or local arcs), weight kind (small and
    2^20..2^31 integers, FP32, FP64), length kind and workgroup size, with the library's own LDS map choice
    (bitmaps, 16-bit bounds, 4-bit codes or LDS distances): distances bit-exact, betweenness within 1e-9."""
    import neurograph_amd as ng

    rng = np.random.default_rng(2024)
    for case in range(30):
        n = int(rng.choice([60, 300, 2000, 9000]))
        m = max(1, int(n * float(rng.choice([1.5, 4, 12, 40]))))
        kind = str(rng.choice(["i32s", "i32l", "f32", "f64"]))
        block = str(rng.choice([256, 512, 1024]))
        lkind = int(rng.choice([orc.L_ABS, orc.L_INVERSE]))
        src = rng.integers(0, n, m)
        dst = rng.integers(0, n, m) if rng.random() < 0.5 else (src + rng.integers(1, max(2, n // 50), m)) % n
        if kind == "i32s":
            w = rng.integers(1, 10, m).astype(np.int32)
        elif kind == "i32l":
            w = rng.integers(1 << 20, 2**31 - 1, m).astype(np.int32)
        else:
            w = np.exp(rng.normal(0.0, 0.5, m)).astype(np.float32 if kind == "f32" else np.float64)
        c = orc.Csr.from_edges(n, src, dst, w, dtype=w.dtype)
        srcs = np.unique(rng.integers(0, n, 16)).astype(np.uint32)
        g = ng.Graph(ctx, ng.HostGraph(c.n, c.off, c.tgt, c.w)).set_options(block=int(block))
        ref = np.stack([orc.sssp(c, int(s), lkind)[0] for s in srcs])
        _, dist = g.sssp(srcs, lkind=lkind, want_dist=True)
        np.testing.assert_array_equal(dist, ref, err_msg=f"case {case}: n={n} m={m} {kind} block {block} lkind {lkind}")
        if lkind == orc.L_ABS:
            np.testing.assert_allclose(g.brandes(srcs), orc.brandes(c, sources=srcs)["bc"], rtol=1e-9, atol=1e-9,
                                       err_msg=f"case {case}")
        g.close()


@pytest.mark.parametrize("block", ["64", "256", "1024"])
def test_many_sources_per_slot(ctx, block):
    """Few slots, so each runs dozens of sources one after another: every source must find its slot restored
    (distances, records, second dependencies, bitmaps, lists) by the previous one, whether the previous source's
    shortest paths formed a tree (continuous weights: the tree layers restore the nodes, parents clear their
    children's dependencies) or not (small integer weights: ties), or a mix (wide integer weights). Distances
    bit-exact; betweenness, edge path counts and both adjoint dependencies within 1e-9."""
    for wdtype, seed, wlo, whi in ((np.float64, 51, 1, 9), (np.int32, 52, 1 << 20, 2**31 - 1), (np.int32, 53, 1, 4),
                                   (np.float32, 54, 1, 9)):
        c = rand_graph(1500, 12000, seed, wdtype=wdtype, wlo=wlo, whi=whi)
        g = dev_graph(ctx, c, max_slots=3, block=int(block))
        srcs = np.arange(0, 1500, 7)
        _, dist = g.sssp(srcs, want_dist=True)
        np.testing.assert_array_equal(dist, np.stack([orc.sssp(c, int(s))[0] for s in srcs]))
        assert g.last_stats()["slots"] == 3
        ref = orc.brandes(c, sources=srcs, edge=True)
        np.testing.assert_allclose(g.brandes(srcs), ref["bc"], rtol=1e-9, atol=1e-9)
        np.testing.assert_allclose(g.edge_dependency(srcs), ref["edge"], rtol=1e-9, atol=1e-12)
        # adjoint mode (both dependencies): d(sum d)/dw = edge path counts, -d(sum 1/d)/dw = the 1/d^2-seeded ones
        g1, g2 = g.adjoint(srcs)
        assert g.last_stats()["slots"] == 3
        np.testing.assert_allclose(g1, ref["edge"], rtol=1e-9, atol=1e-12)
        ref2 = orc.brandes(c, sources=srcs, edge=True, seed=orc.SEED_INV_D2)["edge"]
        np.testing.assert_allclose(g2, -ref2, rtol=1e-9, atol=1e-12)
        _, dist2 = g.sssp(srcs, want_dist=True)  # the slots are clean after Brandes and the adjoint too
        np.testing.assert_array_equal(dist2, dist)
        np.testing.assert_allclose(g.brandes(srcs), ref["bc"], rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("block", [256, 1024])
def test_log_overflow_falls_back_to_row_scan(ctx, block):
    """A tight-arc log too small for the candidates (NG_OPT_LOG_CAPACITY 64): sources that overflow it rescan every
    row of their reached nodes instead (flags bit 2). Trees (continuous weights), ties (small integer weights) and
    mixes: distances bit-exact, betweenness, edge path counts and both adjoint dependencies within 1e-9."""
    for wdtype, seed, wlo, whi in ((np.float64, 61, 1, 9), (np.int32, 62, 1, 3), (np.int32, 63, 1 << 20, 2**31 - 1)):
        c = rand_graph(3000, 30000, seed, wdtype=wdtype, wlo=wlo, whi=whi)
        g = dev_graph(ctx, c, log_capacity=64, block=block)
        srcs = np.arange(0, 3000, 31)
        ref = orc.brandes(c, sources=srcs, edge=True, want_dist=True)
        bc = g.brandes(srcs)
        assert g.last_stats()["flags"] & 2
        np.testing.assert_allclose(bc, ref["bc"], rtol=1e-9, atol=1e-9)
        np.testing.assert_allclose(g.edge_dependency(srcs), ref["edge"], rtol=1e-9, atol=1e-12)
        g1, g2 = g.adjoint(srcs)
        np.testing.assert_allclose(g1, ref["edge"], rtol=1e-9, atol=1e-12)
        ref2 = orc.brandes(c, sources=srcs, edge=True, seed=orc.SEED_INV_D2)["edge"]
        np.testing.assert_allclose(g2, -ref2, rtol=1e-9, atol=1e-12)
        _, dist = g.sssp(srcs, want_dist=True)
        np.testing.assert_array_equal(dist, ref["dist"])
        # the default capacity does not overflow on these graphs: same values
        g0 = dev_graph(ctx, c, block=block)
        np.testing.assert_allclose(g0.brandes(srcs), bc, rtol=1e-12, atol=1e-12)
        assert not (g0.last_stats()["flags"] & 2)


def test_graph_options_are_validated(ctx):
    import neurograph_amd as ng

    c = rand_graph(500, 3000, 71)
    g = dev_graph(ctx, c)
    for name, bad in (("block", 128), ("lds_maps", 32), ("lds_maps", -2), ("max_slots", -1)):
        with pytest.raises(ng.NgError) as e:
            g.set_options(**{name: bad})
        assert e.value.status == ng.NG_EINVAL
    assert g.get_option("lds_maps") == -1 and g.get_option("block") == 0


def gadget_graph(n, m, n_gadgets, seed, wdtype):
    """A random graph of lognormal lengths (all >= 0.2) plus `n_gadgets` tie gadgets: from a random node x, arcs of
    length 1/8 into fresh nodes a and b (and c for every third gadget) and from each of them into a fresh node t, so
    that every source reaching x finds t at the same distance over two or three paths (sums of exact binary
    fractions): t has several tight in-arcs, while the rest of the shortest-path DAG stays a tree. Every t gets
    out-arcs into the random part, and every fifth gadget hangs below the previous one's t, so that multi nodes are
    internal, leaves, and nested (sigma 4 and more)."""
    rng = np.random.default_rng(seed)
    src = list(rng.integers(0, n, size=m))
    dst = list(rng.integers(0, n, size=m))
    w = list(0.2 + np.exp(rng.normal(0.0, 0.5, size=m)))
    total = n
    prev_t = None
    for k in range(n_gadgets):
        x = prev_t if (k % 5 == 4 and prev_t is not None) else int(rng.integers(0, n))
        mids = [total + j for j in range(3 if k % 3 == 2 else 2)]
        t = total + len(mids)
        total = t + 1
        for a in mids:
            src += [x, a]
            dst += [a, t]
            w += [0.125, 0.125]
        if k % 2 == 0:  # internal multi node
            for y in rng.integers(0, n, size=3):
                src.append(t)
                dst.append(int(y))
                w.append(0.2 + float(np.exp(rng.normal(0.0, 0.5))))
        prev_t = t
    return orc.Csr.from_edges(total, np.array(src), np.array(dst), np.array(w, dtype=wdtype), dtype=wdtype)


@pytest.mark.parametrize("wdtype", [np.float32, np.float64])
def test_near_tree_sources(ctx, wdtype):
    """Sources whose tight arcs form a tree except at a few nodes with several tight in-arcs (tie gadgets) take the
    near-tree layers in 1024-lane workgroups with LDS bounds and f32 distances; with more such nodes than the lists
    hold, or f64 distances, they take the general layers. All against the oracle: betweenness within 1e-9, distances
    bit-exact, and the slots clean for the next calls (a second Brandes call and SSSP give the same values)."""
    for n_gadgets, expect_near in ((7, wdtype == np.float32), (700, False)):
        c = gadget_graph(3000, 24000, n_gadgets, 90 + n_gadgets, wdtype)
        srcs = np.arange(0, 3000, 11, dtype=np.uint32)
        ref = orc.brandes(c, sources=srcs, want_dist=True)
        for maps in (4, 8):  # 16-bit bounds, 4-bit codes
            g = dev_graph(ctx, c, block=1024, lds_maps=maps)
            bc = g.brandes(srcs)
            st = g.last_stats()
            assert st["block"] == 1024
            if expect_near:
                assert st["near_tree_sources"] > 0.5 * len(srcs), st
            else:
                assert st["near_tree_sources"] == 0, st
            np.testing.assert_allclose(bc, ref["bc"], rtol=1e-9, atol=1e-9)
            np.testing.assert_allclose(g.brandes(srcs), bc, rtol=1e-12, atol=1e-12)
            _, dist = g.sssp(srcs, want_dist=True)
            np.testing.assert_array_equal(dist, ref["dist"])
            # edge path counts (general layers, no near-tree path) after the near-tree calls
            np.testing.assert_allclose(g.edge_dependency(srcs), orc.brandes(c, sources=srcs, edge=True)["edge"],
                                       rtol=1e-9, atol=1e-12)
