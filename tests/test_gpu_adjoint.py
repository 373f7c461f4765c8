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
