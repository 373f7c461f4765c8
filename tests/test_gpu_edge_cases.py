"""Edge cases of the GPU path against the oracle and closed forms: empty and one-node graphs, isolated nodes,
self-loops and parallel arcs, repeated sources, a 20 000-node directed path (one bucket per node), extreme
degree skew (a 60 000-arc hub), a graph too large for the LDS maps, integer weights that need 64-bit distances,
and rejected inputs."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import oracle as orc


@pytest.fixture(scope="module")
def ctx():
    import neurograph_amd as ng

    return ng.Context(0)


def dev(ctx, c, **opts):
    import neurograph_amd as ng

    return ng.Graph(ctx, ng.HostGraph(c.n, c.off, c.tgt, c.w)).set_options(**opts)


def check_all(ctx, c, sources, rtol=1e-9, **opts):
    g = dev(ctx, c, **opts)
    ref = np.stack([orc.sssp(c, int(s))[0] for s in sources]) if len(sources) else None
    summ, dist = g.sssp(sources, want_dist=True)
    if len(sources):
        np.testing.assert_array_equal(dist, ref)
    bc = g.brandes(sources)
    np.testing.assert_allclose(bc, orc.brandes(c, sources=np.asarray(sources, np.uint32))["bc"], rtol=rtol, atol=1e-9)
    return g, summ


def test_empty_graph(ctx):
    import neurograph_amd as ng

    g = ng.Graph(ctx, ng.HostGraph(0, np.zeros(1, np.uint64), np.zeros(0, np.uint32), np.zeros(0, np.int32)))
    assert len(g.sssp(np.zeros(0, np.uint32))) == 0
    assert len(g.brandes(np.zeros(0, np.uint32))) == 0
    with pytest.raises(ng.NgError):
        g.sssp([0])


def test_one_and_two_node_graphs(ctx):
    one = orc.Csr.from_edges(1, [], [], np.zeros(0, np.int32))
    _, s = check_all(ctx, one, [0])
    assert int(s["sum_d_u64"][0]) == 0 and int(s["n_unreached"][0]) == 0
    loop = orc.Csr.from_edges(1, [0], [0], [5])
    _, s = check_all(ctx, loop, [0, 0])
    assert int(s["max_d_u64"].max()) == 0
    two = orc.Csr.from_edges(2, [], [], np.zeros(0, np.int32))
    _, s = check_all(ctx, two, [0, 1])
    assert list(s["n_unreached"]) == [1, 1] and s["sum_eff"].sum() == 0.0


def test_parallel_arcs_and_self_loops(ctx):
    # 0->1 twice (3 and 1), 1->2, 0->2 of length 2 (ties with 0->1->2), self-loops everywhere
    c = orc.Csr.from_edges(3, [0, 0, 1, 0, 0, 1, 2], [1, 1, 2, 2, 0, 1, 2], [3, 1, 1, 2, 4, 1, 7])
    g, s = check_all(ctx, c, [0, 1, 2])
    assert int(s["sum_d_u64"][0]) == 1 + 2
    edge = g.edge_dependency([0, 1, 2])
    np.testing.assert_allclose(edge, orc.brandes(c, edge=True)["edge"], rtol=1e-12, atol=1e-12)


def test_repeated_sources_accumulate(ctx):
    rng = np.random.default_rng(4)
    src, dst = rng.integers(0, 500, 4000), rng.integers(0, 500, 4000)
    c = orc.Csr.from_edges(500, src, dst, rng.integers(1, 6, 4000).astype(np.int32))
    g, s = check_all(ctx, c, [7, 7, 7, 3, 499, 7])
    assert s[0] == s[1] and s[1] == s[2] and s[2] == s[5]


@pytest.mark.parametrize("block", ["256", "1024"])
def test_long_directed_path(ctx, block):
    n = 20000
    c = orc.Csr.from_edges(n, np.arange(n - 1), np.arange(1, n), np.ones(n - 1, np.int32))
    g = dev(ctx, c, block=int(block))
    sources = np.array([0, 1, 5000, 19998, 19999], np.uint32)
    s = g.sssp(sources, delta=1)
    k = (n - 1 - sources).astype(np.uint64)
    np.testing.assert_array_equal(s["sum_d_u64"], k * (k + 1) // 2)  # closed form: 1 + 2 + ... + k
    np.testing.assert_array_equal(s["n_unreached"], sources.astype(np.uint64))
    bc = g.brandes(sources, delta=1)
    # node v lies strictly inside the path from source s to every t > v: (n - 1 - v) pairs per source s < v
    v = np.arange(n)
    expect = sum(np.where(v > s, n - 1 - v, 0) for s in sources.astype(np.int64)).astype(np.float64)
    np.testing.assert_array_equal(bc, expect)


@pytest.mark.parametrize("block", ["256", "1024"])
def test_hub_degree_skew(ctx, block):
    """A 60 000-arc hub row: expansion tiles whose arc count outgrows the chunk owner table's 64-arc chunks."""
    n = 60001
    leaves = np.arange(1, n)
    rng = np.random.default_rng(8)
    src = np.concatenate([np.zeros(n - 1, np.int64), leaves, rng.integers(1, n, 30000)])
    dst = np.concatenate([leaves, np.zeros(n - 1, np.int64), rng.integers(1, n, 30000)])
    w = rng.integers(1, 20, len(src)).astype(np.int32)
    c = orc.Csr.from_edges(n, src, dst, w)
    g, _ = check_all(ctx, c, [0, 1, 17, 60000], block=int(block))
    assert g.last_stats()["block"] == int(block)


def test_graph_beyond_lds_maps(ctx):
    """800 000 nodes: no LDS bitmap or bound fits, so the settled and frontier stamps live in global memory and
    the wide workgroups' filter stage passes every arc to the slot-by-slot form."""
    n, m = 800_000, 2_400_000
    rng = np.random.default_rng(12)
    src = rng.integers(0, n, m)
    dst = (src + rng.integers(1, 2000, m)) % n  # local arcs: long shortest paths, many buckets
    w = rng.integers(1, 50, m).astype(np.int32)
    c = orc.Csr.from_edges(n, src, dst, w)
    g, _ = check_all(ctx, c, [0, 123_457, 799_999], block=1024)
    assert g.last_stats()["block"] == 1024


def test_large_integer_weights_use_64_bit_distances(ctx):
    import neurograph_amd as ng

    rng = np.random.default_rng(9)
    n, m = 5000, 40000
    src, dst = rng.integers(0, n, m), rng.integers(0, n, m)
    w = rng.integers(1 << 20, (1 << 31) - 1, m).astype(np.int32)
    c = orc.Csr.from_edges(n, src, dst, w)
    g, s = check_all(ctx, c, np.arange(0, n, 97))
    assert g.dist_dtype() == np.uint64
    assert int(s["max_d_u64"].max()) > (1 << 32)


def test_rejected_inputs(ctx):
    import neurograph_amd as ng

    for bad in (np.nan, np.inf, -np.inf):
        w = np.array([1.0, bad], np.float64)
        with pytest.raises(ng.NgError) as e:
            ng.Graph(ctx, ng.HostGraph(3, np.array([0, 1, 2, 2], np.uint64), np.array([1, 2], np.uint32), w))
        assert e.value.status == 1
    with pytest.raises(ng.NgError):  # a target out of range
        ng.Graph(ctx, ng.HostGraph(2, np.array([0, 1, 1], np.uint64), np.array([2], np.uint32), np.array([1], np.int32)))
    with pytest.raises(ng.NgError):  # offsets that decrease
        ng.Graph(ctx, ng.HostGraph(2, np.array([0, 1, 0], np.uint64), np.zeros(0, np.uint32), np.zeros(0, np.int32)))


def test_context_destroyed_before_its_graphs():
    """ng_context_destroy releases only the caller's handle: graphs and cohorts created on the context keep it
    alive and keep working until they are destroyed themselves, in any order. This is also the order in which the
    garbage collector may finalize a reference cycle that holds a context and its graphs."""
    import gc

    import neurograph_amd as ng

    c = orc.Csr.from_edges(4, [0, 1, 2, 0], [1, 2, 3, 3], [1, 1, 1, 5])
    own = ng.Context(0)
    g = dev(own, c)
    g2 = dev(own, c)
    co = ng.Cohort(own, [g, g2])
    own.close()
    s = g.sssp([0])
    assert int(s["sum_d_u64"][0]) == 1 + 2 + 3
    parts, _, _ = co.brandes()
    np.testing.assert_array_equal(parts[0], [0.0, 2.0, 2.0, 0.0])
    co.close()
    g.close()
    np.testing.assert_array_equal(g2.brandes([0]), [0.0, 2.0, 1.0, 0.0])
    g2.close()

    class Holder:
        pass

    h = Holder()
    h.ctx = ng.Context(0)
    h.g = dev(h.ctx, c)
    h.me = h  # a cycle: context and graph are finalized together, in the collector's order
    del h
    gc.collect()
