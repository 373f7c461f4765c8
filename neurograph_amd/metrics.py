"""Host-side aggregation of per-source results, mirroring the reference's metric drivers.

* ``aggregate_apsp``  -- AllPairsShortestPath::compute_apsp_impl, AllPairsShortestPath.h:121-162: sums, max,
  unreachable count and ``cluster_sizes[N - unreachable] += 1`` over the per-source summaries, then the
  averages over reached ordered pairs (N*N - N - unreachable).
* ``average_betweenness`` -- BetweennessCentrality.h:74-77: sum of the per-node values over N.
* ``apsp`` / ``betweenness`` -- convenience drivers over a device graph (one batched call each).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from ._lib import NG_LEN_ABS, Graph


@dataclass
class ApspGlobalResult:
    """ApspGlobalResult<d_type>, source/Types.h:237-244."""

    average_shortest_path_length: float
    average_efficiency: float
    diameter: float
    number_disconnected_pairs: int
    cluster_sizes: dict = field(default_factory=dict)
    sum_distances: float = 0.0
    sum_efficiency: float = 0.0


def aggregate_apsp(summaries: np.ndarray, n: int, integer: bool) -> ApspGlobalResult:
    if integer:
        total = int(np.sum(summaries["sum_d_u64"], dtype=np.uint64))
        diameter = int(np.max(summaries["max_d_u64"])) if len(summaries) else 0
    else:
        total = float(np.sum(summaries["sum_d"]))
        diameter = float(np.max(summaries["max_d"])) if len(summaries) else 0.0
    eff = float(np.sum(summaries["sum_eff"]))
    unreach = int(np.sum(summaries["n_unreached"], dtype=np.uint64))
    clusters: dict = {}
    for u in summaries["n_unreached"]:
        k = n - int(u)
        clusters[k] = clusters.get(k, 0) + 1
    reached = n * n - n - unreach
    avg = (float(total) / float(reached)) if reached else float("nan")
    avg_eff = eff / float(reached) if reached else float("nan")
    return ApspGlobalResult(avg, avg_eff, diameter, unreach, clusters, float(total), eff)


def average_betweenness(bc: np.ndarray) -> float:
    return float(np.sum(bc)) / len(bc) if len(bc) else float("nan")


def apsp(g: Graph, lkind: int = NG_LEN_ABS, delta: float = 0.0) -> ApspGlobalResult:
    s = g.sssp(None, lkind=lkind, delta=delta)
    integer = np.issubdtype(g.dist_dtype(lkind), np.integer)
    return aggregate_apsp(s, g.n, integer)


def betweenness(g: Graph, delta: float = 0.0) -> float:
    return average_betweenness(g.brandes(None, delta=delta))
