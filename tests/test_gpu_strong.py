"""GPU parity of the early end of a source on strongly connected graphs: once every node is settled (and the current
bucket's end lies above its start) the settle pass reads no heavy arcs and the source ends without draining the
pile of stale entries. The results must be the oracle's (distances bit-exact; summaries, betweenness, edge
dependencies and adjoint gradients within 1e-9) in every workgroup width and LDS map variant, and the same graph with
one node that no arc reaches (not strongly connected: no early end) must give the same values with more bucket
advances.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import oracle as orc


@pytest.fixture(scope="module")
def ctx():
    import neurograph_amd as ng

    return ng.Context(0)


def strong_graph(n, m, seed, wdtype, extra_source=False):
    """Random arcs plus a cycle through every node (strongly connected): lognormal lengths, or integers in [1, 9] for
    integer types. With extra_source, one more node n with an arc into node 0 and none into it: no longer strongly
    connected."""
    rng = np.random.default_rng(seed)
    src = np.concatenate([rng.integers(0, n, size=m), np.arange(n)])
    dst = np.concatenate([rng.integers(0, n, size=m), (np.arange(n) + 1) % n])
    if np.issubdtype(wdtype, np.integer):
        w = rng.integers(1, 10, size=len(src)).astype(np.float64)
    else:
        w = np.exp(rng.normal(0.0, 0.5, size=len(src)))
    if extra_source:
        src = np.append(src, n)
        dst = np.append(dst, 0)
        w = np.append(w, 1.0)
    return orc.Csr.from_edges(n + (1 if extra_source else 0), src, dst, w.astype(wdtype), dtype=wdtype)


def dev_graph(ctx, c, **opts):
    import neurograph_amd as ng

    return ng.Graph(ctx, ng.HostGraph(c.n, c.off, c.tgt, c.w)).set_options(**opts)


@pytest.mark.parametrize("block,maps", [(1024, 0), (1024, 4), (1024, 11), (512, 8), (256, 4), (256, 17), (64, 2)])
@pytest.mark.parametrize("wdtype", [np.int32, np.float32, np.float64])
def test_strongly_connected_sources_end_early(ctx, wdtype, block, maps):
    n = 3000
    c = strong_graph(n, 12 * n, 11, wdtype)
    c_open = strong_graph(n, 12 * n, 11, wdtype, extra_source=True)
    srcs = np.arange(0, n, 13, dtype=np.uint32)
    ref_dist = np.stack([orc.sssp(c, int(s))[0] for s in srcs])
    ref_inv = np.stack([orc.sssp(c, int(s), orc.L_INVERSE)[0] for s in srcs])
    ref_sum = orc.apsp(c, sources=srcs, per_source=True)["per_source"]
    ref_bc = orc.brandes(c, sources=srcs)["bc"]
    scale = 5.0 if np.issubdtype(wdtype, np.integer) else 1.0
    g = dev_graph(ctx, c, block=block, lds_maps=maps)
    g_open = dev_graph(ctx, c_open, block=block, lds_maps=maps)
    for delta in (0.0, 0.3 * scale, 0.9 * scale):
        s, dist = g.sssp(srcs, delta=delta, want_dist=True)
        np.testing.assert_array_equal(dist, ref_dist)
        np.testing.assert_allclose(s["sum_d"], ref_sum["sum_d"], rtol=1e-12)
        np.testing.assert_array_equal(s["max_d"], ref_sum["max_d"])
        np.testing.assert_array_equal(s["n_unreached"], np.zeros(len(srcs)))
    _, dist = g.sssp(srcs, lkind=orc.L_INVERSE, want_dist=True)
    np.testing.assert_array_equal(dist, ref_inv)
    bc = g.brandes(srcs, delta=0.3 * scale)
    st = g.last_stats()
    assert st["block"] == block
    np.testing.assert_allclose(bc, ref_bc, rtol=1e-9, atol=1e-9)
    # the open graph: the same sources reach the same nodes (node n is reached by none of them)
    bc_open = g_open.brandes(srcs, delta=0.3 * scale)
    st_open = g_open.last_stats()
    np.testing.assert_allclose(bc_open[:n], ref_bc, rtol=1e-9, atol=1e-9)
    assert bc_open[n] == 0.0
    assert st["buckets"] < st_open["buckets"], (st["buckets"], st_open["buckets"])
    assert st["settle_arcs"] <= st_open["settle_arcs"]
    # the slots are clean for the next calls
    np.testing.assert_allclose(g.brandes(srcs, delta=0.3 * scale), bc, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("wdtype", [np.float32, np.float64])
def test_strongly_connected_edge_dependency_and_adjoint(ctx, wdtype):
    n = 1200
    c = strong_graph(n, 10 * n, 23, wdtype)
    g = dev_graph(ctx, c, block=1024)
    one = orc.brandes(c, edge=True)
    inv = orc.brandes(c, edge=True, seed=orc.SEED_INV_D2)
    np.testing.assert_allclose(g.edge_dependency(None, delta=0.4), one["edge"], rtol=1e-9, atol=1e-12)
    g1, g2 = g.adjoint(None, delta=0.4)
    np.testing.assert_allclose(g1, one["edge"], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(g2, -inv["edge"], rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("block", [256, 1024])
@pytest.mark.parametrize("wdtype", [np.int32, np.float32])
def test_partial_reach_restores_slots_per_node(ctx, wdtype, block):
    """Sources that reach fewer than half of the nodes restore their slot node by node (distances, tight-arc bitmap
    words), the others by one coalesced pass over the arrays. Four disconnected components (each source reaches
    under a quarter of the nodes), run again and again on the same slots and alternated with a connected graph on
    the same workspace, keep every result the oracle's."""
    import neurograph_amd as ng

    own = ng.Context(0)
    n4 = 900
    rng = np.random.default_rng(77)
    src, dst = [], []
    for k in range(4):
        base = k * n4
        src.append(base + rng.integers(0, n4, size=8 * n4))
        dst.append(base + rng.integers(0, n4, size=8 * n4))
    src, dst = np.concatenate(src), np.concatenate(dst)
    if np.issubdtype(wdtype, np.integer):
        w = rng.integers(1, 10, size=len(src)).astype(wdtype)
    else:
        w = np.exp(rng.normal(0.0, 0.5, size=len(src))).astype(wdtype)
    parts = orc.Csr.from_edges(4 * n4, src, dst, w, dtype=wdtype)
    whole = strong_graph(4 * n4, 8 * 4 * n4, 78, wdtype)
    gp = dev_graph(own, parts, block=block)
    gw = dev_graph(own, whole, block=block)
    sp = np.arange(0, 4 * n4, 23, dtype=np.uint32)
    ref_p = orc.brandes(parts, sources=sp, want_dist=True)
    ref_w = orc.brandes(whole, sources=sp, want_dist=True)
    for _ in range(2):
        for g, ref in ((gp, ref_p), (gw, ref_w), (gp, ref_p)):
            _, dist = g.sssp(sp, want_dist=True)
            np.testing.assert_array_equal(dist, ref["dist"])
            np.testing.assert_allclose(g.brandes(sp), ref["bc"], rtol=1e-9, atol=1e-9)
    d = ref_p["dist"]
    top = np.finfo(d.dtype).max if d.dtype.kind == "f" else np.iinfo(d.dtype).max  # the oracle's "unreached"
    assert (d == top).mean() > 0.7  # most nodes unreached from each source
