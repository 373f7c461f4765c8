"""Pins the CPU oracle (oracle/oracle.cpp) before anything is compared against it.

Each check is chosen so that a plausible slip in the restatement fails it: the reference's own
known-answer values (tests/golden/reference_fixtures.json), closed forms, brute force on tiny graphs,
independent library routines (scipy Dijkstra, networkx betweenness) and invariants.
"""

import itertools
import math

import numpy as np
import pytest

from oracle import oracle as orc
from tests.conftest import csr_from_fixture

FIXTURES_WITH_APSP = ["standard", "standard_uu", "full", "disconnected", "four", "four_uu", "full36", "seven", "seven_uu",
                      "full63"]


def _check_apsp(got, exp):
    if exp["spl_exact"]:
        assert got["avg_spl"] == exp["avg_spl"]
    else:
        assert got["avg_spl"] == pytest.approx(exp["avg_spl"], abs=1e-6)
    assert got["avg_eff"] == pytest.approx(exp["avg_eff"], abs=1e-6)
    assert got["diameter"] == pytest.approx(exp["diameter"], abs=1e-6)
    assert got["n_disconnected"] == exp["n_disconnected"]
    assert got["cluster_sizes"] == {int(k): v for k, v in exp["cluster_sizes"].items()}


@pytest.mark.parametrize("name", FIXTURES_WITH_APSP)
def test_apsp_matches_reference_assertions(reference_fixtures, name):
    fx = reference_fixtures["fixtures"][name]
    g = csr_from_fixture(fx)
    _check_apsp(orc.apsp(g, orc.L_ABS), fx["apsp"])
    # the threaded runner reduces the same partial sums
    _check_apsp(orc.apsp(g, orc.L_ABS, threads=3), fx["apsp"])


@pytest.mark.parametrize("name", [n for n in FIXTURES_WITH_APSP if n != "disconnected"])
def test_apsp_inverse_matches_reference_assertions(reference_fixtures, name):
    fx = reference_fixtures["fixtures"][name]
    _check_apsp(orc.apsp(csr_from_fixture(fx), orc.L_INVERSE), fx["apsp_inverse"])


@pytest.mark.parametrize("name", ["standard", "standard_uu", "full", "four", "four_uu", "full36", "seven", "seven_uu", "full63"])
def test_brandes_matches_reference_assertions(reference_fixtures, name):
    fx = reference_fixtures["fixtures"][name]
    g = csr_from_fixture(fx)
    got = orc.brandes(g)
    assert got["avg_bc"] == pytest.approx(fx["brandes_avg"]["value"], abs=1e-6)
    got3 = orc.brandes(g, threads=4)
    assert got3["avg_bc"] == pytest.approx(fx["brandes_avg"]["value"], abs=1e-6)


def _check_rows(rows, summaries, tol):
    for i, (s, m, e, u) in enumerate(rows):
        got = summaries[i]
        assert got["sum_d_u64"] == s, i
        assert got["max_d_u64"] == m, i
        assert got["sum_eff"] == pytest.approx(e, abs=tol), i
        assert got["n_unreached"] == u, i


@pytest.mark.parametrize("name", ["standard", "standard_uu", "full", "negative", "four", "four_uu", "full36", "seven",
                                  "seven_uu", "full63"])
def test_per_source_summaries_match_reference_tables(reference_fixtures, name):
    fx = reference_fixtures["fixtures"][name]
    g = csr_from_fixture(fx)
    rows = fx["per_source"]["rows"]
    sums = [orc.sssp(g, r)[1] for r in range(g.n)]
    _check_rows(rows, sums, fx["per_source"]["eff_tol"])
    per = orc.apsp(g, per_source=True)["per_source"]
    _check_rows(rows, [dict(zip(per.dtype.names, p)) for p in per], fx["per_source"]["eff_tol"])
    if name != "negative":
        for delta in reference_fixtures["deltas"]:
            _check_rows(rows, [orc.delta_stepping(g, r, delta)[1] for r in range(g.n)], fx["per_source"]["eff_tol"])


def test_delta_stepping_rejects_negative_weights(reference_fixtures):
    g = csr_from_fixture(reference_fixtures["fixtures"]["negative"])
    with pytest.raises(orc.OracleError):
        orc.delta_stepping(g, 0, 4)


def random_graph(n, m, rng, wdtype=np.int32, wlo=1, whi=10, parallel=False, self_loops=True):
    src = rng.integers(0, n, size=m)
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
        w = rng.uniform(wlo, whi, size=len(src)).astype(wdtype)
    return orc.Csr.from_edges(n, src, dst, w, dtype=wdtype)


def floyd_warshall(g, inverse=False):
    n = g.n
    INF = math.inf
    d = [[INF] * n for _ in range(n)]
    for u in range(n):
        d[u][u] = 0
        for e in range(int(g.off[u]), int(g.off[u + 1])):
            aw = abs(int(g.w[e])) if g.wkind == orc.W_I32 else abs(float(g.w[e]))
            if inverse:
                if aw == 0:
                    continue
                ln = 1.0 / aw
            else:
                ln = aw
            t = int(g.tgt[e])
            d[u][t] = min(d[u][t], ln)
    for k in range(n):
        for i in range(n):
            for j in range(n):
                if d[i][k] + d[k][j] < d[i][j]:
                    d[i][j] = d[i][k] + d[k][j]
    return d


@pytest.mark.parametrize("seed", range(6))
def test_sssp_matches_floyd_warshall(seed):
    rng = np.random.default_rng(seed)
    g = random_graph(12, 40, rng, parallel=True)
    fw = floyd_warshall(g)
    for s in range(g.n):
        dist, summ = orc.sssp(g, s)
        for t in range(g.n):
            if math.isinf(fw[s][t]):
                assert dist[t] == np.iinfo(np.uint64).max
            else:
                assert int(dist[t]) == fw[s][t]
        finite = [fw[s][t] for t in range(g.n) if not math.isinf(fw[s][t])]
        assert summ["sum_d_u64"] == sum(finite)
        assert summ["max_d_u64"] == max(finite)
        assert summ["n_unreached"] == g.n - len(finite)


def test_sssp_fp_matches_scipy():
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import dijkstra

    rng = np.random.default_rng(7)
    g = random_graph(300, 3000, rng, wdtype=np.float64, wlo=0.1, whi=2.0, self_loops=False)
    rows = np.repeat(np.arange(g.n), np.diff(g.off).astype(np.int64))
    mat = csr_matrix((g.w, (rows, g.tgt.astype(np.int64))), shape=(g.n, g.n))
    ref = dijkstra(mat, directed=True, indices=np.arange(20))
    for s in range(20):
        dist, _ = orc.sssp(g, s)
        finite = np.isfinite(ref[s])
        # same IEEE adds in the same order along each path: exact
        np.testing.assert_array_equal(dist[finite], ref[s][finite])
        assert np.all(dist[~finite] == np.finfo(np.float64).max)


def test_inverse_metric_zero_weight_is_infinitely_long():
    g = orc.Csr.from_edges(3, [0, 0, 1], [1, 2, 2], [0, 4, 2], dtype=np.float64)
    dist, summ = orc.sssp(g, 0, orc.L_INVERSE)
    assert dist[1] == np.finfo(np.float64).max  # arc 0->1 has weight 0: never relaxed (PriorityQueueVariant.h:75-79)
    assert dist[2] == 0.25
    assert summ["n_unreached"] == 1


def test_directed_cycle_closed_forms():
    n = 9
    g = orc.Csr.from_edges(n, list(range(n)), [(i + 1) % n for i in range(n)], [1] * n)
    res = orc.apsp(g)
    assert res["sum_d_u64"] == n * n * (n - 1) // 2  # per source 1 + ... + (n-1)
    assert res["diameter_u64"] == n - 1
    bc = orc.brandes(g)["bc"]
    np.testing.assert_allclose(bc, (n - 1) * (n - 2) / 2)


def _enumerate_bc(g):
    """PathEnumerationVariant semantics (BetweennessCentralityVariants/PathEnumerationVariant.h:80-170) by brute
    force: every shortest path is listed explicitly, parallel arcs giving distinct paths."""
    fw = floyd_warshall(g)
    n = g.n
    bc = np.zeros(n)
    edge = np.zeros(g.m)
    for s in range(n):
        for t in range(n):
            if s == t or math.isinf(fw[s][t]):
                continue
            paths = []

            def dfs(u, length, nodes, arcs):
                if length > fw[s][t] + 1e-12:
                    return
                if u == t:
                    if abs(length - fw[s][t]) <= 1e-12:
                        paths.append((list(nodes), list(arcs)))
                    return
                for e in range(int(g.off[u]), int(g.off[u + 1])):
                    v = int(g.tgt[e])
                    if v in nodes:
                        continue
                    w = abs(float(g.w[e]))
                    nodes.append(v)
                    arcs.append(e)
                    dfs(v, length + w, nodes, arcs)
                    nodes.pop()
                    arcs.pop()

            dfs(s, 0.0, [s], [])
            for nodes, arcs in paths:
                for v in nodes[1:-1]:
                    bc[v] += 1.0 / len(paths)
                for e in arcs:
                    edge[e] += 1.0 / len(paths)
    return bc, edge


@pytest.mark.parametrize("seed", range(5))
def test_brandes_matches_path_enumeration(seed):
    rng = np.random.default_rng(100 + seed)
    g = random_graph(7, 18, rng, wlo=1, whi=3, parallel=(seed % 2 == 1))
    bc_ref, edge_ref = _enumerate_bc(g)
    got = orc.brandes(g, edge=True)
    np.testing.assert_allclose(got["bc"], bc_ref, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(got["edge"], edge_ref, rtol=1e-12, atol=1e-12)


def test_brandes_fp_matches_networkx():
    import networkx as nx

    rng = np.random.default_rng(11)
    # dyadic weights: path sums are exact, so ties (sigma > 1) really occur in floating point
    g = random_graph(60, 500, rng, wdtype=np.float64, wlo=0, whi=1, self_loops=False)
    g.w[:] = np.round(g.w * 4 + 1) / 4
    G = nx.DiGraph()
    G.add_nodes_from(range(g.n))
    for u in range(g.n):
        for e in range(int(g.off[u]), int(g.off[u + 1])):
            G.add_edge(u, int(g.tgt[e]), weight=float(g.w[e]), idx=e)
    ref = nx.betweenness_centrality(G, normalized=False, weight="weight")
    eref = nx.edge_betweenness_centrality(G, normalized=False, weight="weight")
    got = orc.brandes(g, edge=True, want_sigma=True)
    np.testing.assert_allclose(got["bc"], [ref[v] for v in range(g.n)], rtol=1e-9, atol=1e-9)
    for (u, v), val in eref.items():
        assert got["edge"][G[u][v]["idx"]] == pytest.approx(val, rel=1e-9, abs=1e-9)
    assert got["sigma"].max() > 1  # ties were exercised


def test_unit_weight_bc_invariant():
    rng = np.random.default_rng(3)
    g = random_graph(80, 400, rng, wlo=1, whi=1)
    bc = orc.brandes(g)["bc"]
    total = 0
    for s in range(g.n):
        dist, _ = orc.sssp(g, s)
        fin = dist[dist != np.iinfo(np.uint64).max]
        total += int(np.sum(fin[fin > 0] - 1))
    assert np.sum(bc) == pytest.approx(total, rel=1e-12)


def test_edge_dependency_invariant_and_efficiency_adjoint():
    rng = np.random.default_rng(5)
    g = random_graph(30, 200, rng, wdtype=np.float64, wlo=0.5, whi=1.5, self_loops=False)
    one = orc.brandes(g, edge=True)
    inv = orc.brandes(g, edge=True, seed=orc.SEED_INV_D2)
    base = orc.apsp(g)
    # Euler homogeneity: sum d is degree-1 homogeneous in w, sum 1/d degree -1
    assert np.dot(g.w, one["edge"]) == pytest.approx(base["sum_d"], rel=1e-10)
    assert np.dot(g.w, inv["edge"]) == pytest.approx(base["sum_eff"], rel=1e-10)
    # central finite differences of the oracle's own APSP sums
    h = 1e-6
    for e in rng.choice(g.m, size=12, replace=False):
        wp = g.w.copy()
        wm = g.w.copy()
        wp[e] += h
        wm[e] -= h
        fp = orc.apsp(orc.Csr(g.n, g.off, g.tgt, wp))
        fm = orc.apsp(orc.Csr(g.n, g.off, g.tgt, wm))
        assert (fp["sum_d"] - fm["sum_d"]) / (2 * h) == pytest.approx(one["edge"][e], abs=1e-5)
        assert (fp["sum_eff"] - fm["sum_eff"]) / (2 * h) == pytest.approx(-inv["edge"][e], abs=1e-5)


def test_sampled_sources_and_errors():
    rng = np.random.default_rng(9)
    g = random_graph(50, 300, rng)
    full = orc.apsp(g, per_source=True)["per_source"]
    sub = orc.apsp(g, sources=[3, 17, 42], per_source=True)["per_source"]
    for i, s in enumerate([3, 17, 42]):
        assert tuple(sub[i]) == tuple(full[s])
    with pytest.raises(orc.OracleError):
        orc.apsp(g, sources=[50])
    with pytest.raises(orc.OracleError):
        orc.sssp(g, 50)
