"""Differentiable APSP metrics: characteristic path length and global efficiency as functions of the arc
weights, with gradients from the hand-written adjoint kernel (ng_apsp_adjoint).

For a source set S (all nodes by default, as in AllPairsShortestPath::compute_apsp_impl,
AllPairsShortestPath.h:130-162) with lengths |w|:

    CPL(w) = sum_{s in S, t reached, t != s} d(s, t) / R
    GE(w)  = sum_{s in S, t reached, t != s} 1 / d(s, t) / R,     R = sum_s (n - 1 - unreached(s))

R only depends on reachability, which positive weights do not change, so the derivatives come from the two
sums. d(s, t) is piecewise linear in w; where several shortest paths tie, the adjoint splits the derivative
over them in proportion to their path counts (the Brandes dependency of the arc, BrandesVariant.h:151-154),
which is the average of the one-sided derivatives. The forward pass only computes per-source summaries; the
backward pass recomputes distances and path counts per source batch (no predecessor DAG is stored).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from ._lib import Context, Graph, HostGraph


@dataclass
class Topology:
    """The arc structure the weights live on: CSR over out-arcs (offsets[n+1], targets[m]). It owns one device graph,
    created on first use and updated in place (ng_graph_set_weights) whenever the weights change, so a training loop
    does not rebuild the device CSR per step."""

    ctx: Context
    n: int
    offsets: np.ndarray
    targets: np.ndarray
    graph: Graph | None = field(default=None, repr=False)
    weights: np.ndarray | None = field(default=None, repr=False)  # the weights the device graph holds

    @property
    def m(self) -> int:
        return int(self.offsets[-1])

    @staticmethod
    def of(ctx: Context, g: HostGraph) -> "Topology":
        return Topology(ctx, g.n, np.ascontiguousarray(g.offsets, dtype=np.uint64),
                        np.ascontiguousarray(g.targets, dtype=np.uint32))


def _device_graph(topo: Topology, w: np.ndarray) -> Graph:
    """The topology's device graph holding weights w (created once, then updated in place when w differs)."""
    w = np.ascontiguousarray(w, dtype=np.float64)
    if topo.graph is None:
        topo.graph = Graph(topo.ctx, HostGraph(topo.n, topo.offsets, topo.targets, w.copy()))
        topo.weights = w.copy()
    elif not np.array_equal(topo.weights, w):
        topo.graph.set_weights(w)
        topo.weights = w.copy()
    return topo.graph


def _metrics(summ: np.ndarray, n: int):
    reached = float(len(summ)) * (n - 1) - float(np.sum(summ["n_unreached"], dtype=np.uint64))
    return float(np.sum(summ["sum_d"])) / reached, float(np.sum(summ["sum_eff"])) / reached, reached


def apsp_metrics_and_gradients(topo: Topology, weights: np.ndarray, sources=None):
    """(CPL, GE, dCPL/dw, dGE/dw) in one adjoint pass, without torch (the topology's device graph is reused)."""
    g = _device_graph(topo, weights)
    g1, g2, summ = g.adjoint(sources, summaries=True)
    cpl, ge, reached = _metrics(summ, topo.n)
    return cpl, ge, g1 / reached, g2 / reached


def _function():
    import torch

    class ApspMetrics(torch.autograd.Function):
        """weights (float64 [m]) -> tensor [CPL, GE]; backward through ng_apsp_adjoint."""

        @staticmethod
        def forward(ctx, weights, topo: Topology, sources):
            w = weights.detach().to("cpu", torch.float64).numpy().copy()
            g = _device_graph(topo, w)
            summ = g.sssp(sources)
            cpl, ge, reached = _metrics(summ, topo.n)
            ctx.topo = topo
            ctx.weights = w  # the backward re-runs on these weights (another forward may have moved the graph on)
            ctx.sources = sources
            ctx.reached = reached
            return torch.tensor([cpl, ge], dtype=torch.float64, device=weights.device)

        @staticmethod
        def backward(ctx, grad_out):
            g1, g2 = _device_graph(ctx.topo, ctx.weights).adjoint(ctx.sources)
            go = grad_out.detach().to("cpu", torch.float64).numpy()
            grad = (go[0] * g1 + go[1] * g2) / ctx.reached
            return torch.from_numpy(grad).to(grad_out.device), None, None

    return ApspMetrics


_FN = None


def apsp_metrics(weights, topo: Topology, sources=None):
    """Differentiable [CPL, GE] of the graph `topo` with arc weights `weights` (a float64 torch tensor of m
    elements, any device), over `sources` (default: all nodes)."""
    global _FN
    if _FN is None:
        _FN = _function()
    src = None if sources is None else np.ascontiguousarray(sources, dtype=np.uint32)
    return _FN.apply(weights, topo, src)
