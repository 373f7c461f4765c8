"""Differentiation (ng_apsp_adjoint and its torch.autograd binding): gradients of characteristic path length and
global efficiency with respect to the arc weights, checked against central finite differences of the forward
path (FP64, h = 1e-6) on a 20k-node spatial graph with continuous weights, against the oracle's edge-level path
counts (the per-arc Brandes term, BrandesVariant.h:151-154), and against the homogeneity identities
sum_e w_e dD/dw_e = D and sum_e w_e dE/dw_e = -E (D = sum of distances, E = sum of efficiencies)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import oracle as orc


@pytest.fixture(scope="module")
def ctx():
    import neurograph_amd as ng

    return ng.Context(0)


@pytest.fixture(scope="module")
def c5(ctx):
    """20k nodes, mean out-degree 30, FP64 weights uniform in [0.5, 1.5]."""
    import neurograph_amd as ng

    return ng.generate_spatial(20000, 30, wdist="uniform", a=0.5, b=1.5, weights=np.float64, seed=5)


def host_sums(dist: np.ndarray):
    """Sum of finite distances and of 1/d over finite non-zero distances, in a fixed order on the host."""
    finite = dist < np.finfo(np.float64).max
    d = dist[finite]
    nz = d[d != 0.0]
    return float(np.sum(d)), float(np.sum(1.0 / nz)), int(np.count_nonzero(~finite))


def test_autograd_matches_oracle_edge_counts(ctx):
    import torch

    import neurograph_amd as ng
    from neurograph_amd.autograd import Topology, apsp_metrics

    hg = ng.generate_spatial(800, 10, wdist="uniform", a=0.5, b=1.5, weights=np.float64, seed=9)
    topo = Topology.of(ctx, hg)
    w = torch.tensor(hg.weights, dtype=torch.float64, requires_grad=True)
    y = apsp_metrics(w, topo)
    c = orc.Csr(hg.n, hg.offsets, hg.targets, hg.weights)
    ap = orc.apsp(c, threads=8)
    reached = hg.n * (hg.n - 1) - ap["n_disconnected"]
    assert float(y.detach()[0]) == pytest.approx(ap["avg_spl"], rel=1e-12)
    assert float(y.detach()[1]) == pytest.approx(ap["avg_eff"], rel=1e-12)
    (gc,) = torch.autograd.grad(y[0], w, retain_graph=True)
    (ge,) = torch.autograd.grad(y[1], w)
    one = orc.brandes(c, edge=True, threads=8)["edge"]
    inv = orc.brandes(c, edge=True, seed=orc.SEED_INV_D2, threads=8)["edge"]
    np.testing.assert_allclose(gc.numpy() * reached, one, rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(ge.numpy() * reached, -inv, rtol=1e-9, atol=1e-9)
    # a weighted combination goes through the same backward
    w2 = torch.tensor(hg.weights, dtype=torch.float64, requires_grad=True)
    (apsp_metrics(w2, topo) * torch.tensor([2.0, -3.0], dtype=torch.float64)).sum().backward()
    np.testing.assert_allclose(w2.grad.numpy(), 2.0 * gc.numpy() - 3.0 * ge.numpy(), rtol=1e-12, atol=1e-15)


def test_homogeneity_full_c5(ctx, c5):
    import neurograph_amd as ng

    g = ng.Graph(ctx, c5)
    g1, g2, summ = g.adjoint(None, summaries=True)
    D = float(np.sum(summ["sum_d"]))
    E = float(np.sum(summ["sum_eff"]))
    w = c5.weights
    assert float(np.dot(w, g1)) == pytest.approx(D, rel=1e-9)
    assert float(np.dot(w, g2)) == pytest.approx(-E, rel=1e-9)
    # the per-arc path usage is the Brandes edge term, so its sum over arcs is the number of (s, t, hop) triples
    assert np.all(g1 >= 0.0) and np.all(g2 <= 0.0)


def test_finite_differences_c5(ctx, c5):
    """Central differences (h = 1e-6) of the forward path's CPL and GE on 256 arcs: 128 drawn among the arcs
    with a non-zero gradient and 128 uniformly (mostly off every shortest path, gradient 0)."""
    import neurograph_amd as ng
    from neurograph_amd.autograd import Topology, apsp_metrics_and_gradients

    topo = Topology.of(ctx, c5)
    rng = np.random.default_rng(11)
    sources = np.sort(rng.choice(c5.n, size=64, replace=False)).astype(np.uint32)
    cpl, ge, dcpl, dge = apsp_metrics_and_gradients(topo, c5.weights, sources)
    used = np.flatnonzero(dcpl != 0.0)
    arcs = np.concatenate([rng.choice(used, size=128, replace=False), rng.choice(c5.m, size=128, replace=False)])
    h = 1e-6
    base = ng.Graph(ctx, c5)
    _, d0 = base.sssp(sources, want_dist=True)
    D0, E0, unreached = host_sums(d0)
    reached = len(sources) * (c5.n - 1) - unreached
    assert D0 / reached == pytest.approx(cpl, rel=1e-12)
    assert E0 / reached == pytest.approx(ge, rel=1e-12)
    fd_c, fd_e = [], []
    for e in arcs:
        vals = []
        for sgn in (1.0, -1.0):
            w = c5.weights.copy()
            w[e] += sgn * h
            g = ng.Graph(ctx, ng.HostGraph(c5.n, c5.offsets, c5.targets, w))
            _, d = g.sssp(sources, want_dist=True)
            vals.append(host_sums(d)[:2])
            g.close()
        fd_c.append((vals[0][0] - vals[1][0]) / (2 * h) / reached)
        fd_e.append((vals[0][1] - vals[1][1]) / (2 * h) / reached)
    fd_c, fd_e = np.array(fd_c), np.array(fd_e)
    # the sums are ~1e6; their rounding over 2h bounds the difference quotient's error to ~1e-9 after / reached
    np.testing.assert_allclose(dcpl[arcs], fd_c, rtol=1e-5, atol=2e-9)
    np.testing.assert_allclose(dge[arcs], fd_e, rtol=1e-5, atol=2e-9)
    assert np.count_nonzero(fd_c[:128]) == 128


def test_set_weights_matches_fresh_graph(ctx):
    """ng_graph_set_weights: results equal a graph created with the new weights and the oracle's; scaling every
    weight by a positive constant keeps every row's order (lengths uploaded only), new random weights re-sort rows;
    integer weights crossing the 32-bit distance bound switch the distance type and back; non-finite weights are
    refused and leave the graph unchanged."""
    import neurograph_amd as ng

    rng = np.random.default_rng(3)
    n, m = 3000, 30000
    src, dst = rng.integers(0, n, m), rng.integers(0, n, m)
    c0 = orc.Csr.from_edges(n, src, dst, np.exp(rng.normal(0, 0.5, m)), dtype=np.float64)
    g = ng.Graph(ctx, ng.HostGraph(c0.n, c0.off, c0.tgt, c0.w))
    srcs = np.arange(0, n, 37, dtype=np.uint32)
    for w, reorder in ((c0.w * 2.5, False), (np.exp(rng.normal(0, 0.5, m)), True), (c0.w.copy(), True)):
        assert g.set_weights(w) == reorder
        c = orc.Csr(n, c0.off, c0.tgt, np.ascontiguousarray(w))
        _, dist = g.sssp(srcs, want_dist=True)
        np.testing.assert_array_equal(dist, np.stack([orc.sssp(c, int(s))[0] for s in srcs]))
        fresh = ng.Graph(ctx, ng.HostGraph(n, c0.off, c0.tgt, np.ascontiguousarray(w)))
        np.testing.assert_array_equal(g.brandes(srcs), fresh.brandes(srcs))  # same lengths, width and slots: bit-identical
        np.testing.assert_allclose(g.brandes(srcs), orc.brandes(c, sources=srcs)["bc"], rtol=1e-9, atol=1e-9)
        g1, g2 = g.adjoint(srcs)
        np.testing.assert_allclose(g1, orc.brandes(c, sources=srcs, edge=True)["edge"], rtol=1e-9, atol=1e-12)
    bad = c0.w.copy()
    bad[5] = np.nan
    with pytest.raises(ng.NgError):
        g.set_weights(bad)
    np.testing.assert_array_equal(g.sssp(srcs)["sum_d"], ng.Graph(ctx, ng.HostGraph(n, c0.off, c0.tgt, c0.w)).sssp(srcs)["sum_d"])
    # integer weights: small (u32 distances) -> up to 2^31 - 1 (u64) -> small again
    ci = orc.Csr.from_edges(n, src, dst, rng.integers(1, 10, m).astype(np.int32))
    gi = ng.Graph(ctx, ng.HostGraph(ci.n, ci.off, ci.tgt, ci.w))
    assert gi.dist_dtype() == np.uint64  # dist_out buffers are u64 for integer lengths either way
    for w in (rng.integers(1 << 20, 2**31 - 1, m).astype(np.int32), ci.w.copy()):
        gi.set_weights(w)
        c = orc.Csr(n, ci.off, ci.tgt, w)
        _, dist = gi.sssp(srcs, want_dist=True)
        np.testing.assert_array_equal(dist, np.stack([orc.sssp(c, int(s))[0] for s in srcs]))
        np.testing.assert_allclose(gi.brandes(srcs), orc.brandes(c, sources=srcs)["bc"], rtol=1e-9, atol=1e-9)


def test_set_weights_recaptures_cohort(ctx):
    import neurograph_amd as ng

    hs = [ng.generate_spatial(n, 10, weights=np.float32, seed=40 + n) for n in (1200, 2500)]
    gs = [ng.Graph(ctx, h) for h in hs]
    co = ng.Cohort(ctx, gs)
    co.brandes()
    w = (hs[1].weights[::-1]).copy()  # same arcs, weights permuted: new lengths and row orders
    gs[1].set_weights(w)
    parts, _, _ = co.brandes()
    c = orc.Csr(hs[1].n, hs[1].offsets, hs[1].targets, w)
    np.testing.assert_allclose(parts[1], orc.brandes(c, threads=8)["bc"], rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(parts[1], gs[1].brandes(), rtol=1e-9, atol=1e-9)


def test_training_step_reuses_one_graph(ctx, c5):
    """One gradient step through the torch binding on the C5 graph: the topology keeps one device graph, updated in
    place; the gradient after the step matches central differences at the new weights and the oracle's edge counts
    on a 64-source sample."""
    import torch

    import neurograph_amd as ng
    from neurograph_amd.autograd import Topology, apsp_metrics

    topo = Topology.of(ctx, c5)
    rng = np.random.default_rng(17)
    sources = np.sort(rng.choice(c5.n, size=64, replace=False)).astype(np.uint32)
    w = torch.tensor(c5.weights, dtype=torch.float64, requires_grad=True)
    y = apsp_metrics(w, topo, sources)
    handle = topo.graph.h.value
    y[0].backward()
    with torch.no_grad():
        w -= 0.05 * w.grad / w.grad.abs().max()
    w.grad = None
    y2 = apsp_metrics(w, topo, sources)
    assert topo.graph.h.value == handle
    (g,) = torch.autograd.grad(y2[0], w)
    wn = w.detach().numpy().copy()
    c = orc.Csr(c5.n, c5.offsets, c5.targets, wn)
    ap = orc.apsp(c, sources=sources, threads=8)
    reached = len(sources) * (c5.n - 1) - ap["n_disconnected"]  # CPL averages over the sources' reached pairs
    assert float(y2.detach()[0]) == pytest.approx(ap["sum_d"] / reached, rel=1e-12)
    np.testing.assert_allclose(g.numpy() * reached, orc.brandes(c, sources=sources, edge=True, threads=8)["edge"],
                               rtol=1e-9, atol=1e-9)
    used = np.flatnonzero(g.numpy() != 0.0)
    arcs = rng.choice(used, size=16, replace=False)
    h = 1e-6
    for e in arcs:
        vals = []
        for sgn in (1.0, -1.0):
            ww = wn.copy()
            ww[e] += sgn * h
            gg = ng.Graph(ctx, ng.HostGraph(c5.n, c5.offsets, c5.targets, ww))
            _, d = gg.sssp(sources, want_dist=True)
            vals.append(host_sums(d)[0])
            gg.close()
        fd = (vals[0] - vals[1]) / (2 * h) / reached
        assert float(g[e]) == pytest.approx(fd, rel=1e-5, abs=2e-9)
