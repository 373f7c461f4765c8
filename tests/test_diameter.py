"""DiameterApproximation on one rank (CPU): the oracle's restatement against the reference's own expectations
(tests/diameter_approximation/test_diameter_approximation.cpp) and against invariants of the clustering; its
std::mt19937 + uniform_real_distribution draws against libstdc++ itself; and the library's host contraction
(ng_contract_balls, used before the contracted graph's APSP runs on the GPU) against the oracle, node for node and
arc for arc. The device APSP of the contracted graphs is checked in test_gpu_diameter.py."""

import os
import subprocess
import textwrap

import numpy as np
import pytest

from oracle import oracle as orc
from tests.conftest import csr_from_fixture


def ring_graph(n, k, wlo, whi, seed):
    """A directed ring (every node has an in-arc, so balls can reach it) plus k random out-arcs per node, integer
    weights uniform in [wlo, whi] (parallel arcs merged by the generator's CSR)."""
    import neurograph_amd as ng

    rng = np.random.default_rng(seed)
    src = np.concatenate([np.arange(n), np.repeat(np.arange(n), k)])
    dst = np.concatenate([(np.arange(n) + 1) % n, rng.integers(0, n, n * k)])
    keep = src != dst
    src, dst = src[keep], dst[keep]
    key = np.unique(src * n + dst)
    src, dst = key // n, key % n
    w = rng.integers(wlo, whi + 1, len(src)).astype(np.int32)
    return ng.HostGraph.from_edges(n, src, dst, w)


def draws_of(kind):
    """The reference test generators: ConstantRandomNumberGenerator{0.0} and SingleCenterRandomNumberGenerator
    (test_diameter_approximation.cpp:20-56); "default" = the library's default generator."""
    if kind == "constant0":
        return lambda: 0.0
    if kind == "single_center":
        state = {"first": True}

        def draw():
            if state["first"]:
                state["first"] = False
                return 0.0
            return 1.0

        return draw
    return None


def fixture_cases(reference_fixtures):
    for name, fx in reference_fixtures["fixtures"].items():
        for case in fx.get("diameter_approximation", []):
            yield name, fx, case


def test_oracle_matches_reference_expectations(reference_fixtures):
    seen = 0
    for name, fx, case in fixture_cases(reference_fixtures):
        c = csr_from_fixture(fx)
        got = orc.diameter_approximation(c, case["r"], draws_of(case["draws"]))
        if case["expected"] is None:  # the default generator: two runs agree (test_diameter_approximation.cpp:182-195)
            assert got == orc.diameter_approximation(c, case["r"], None), name
        else:
            assert got == case["expected"], (name, case)
        seen += 1
    assert seen == 19


def test_all_centers_and_one_center_closed_forms(reference_fixtures):
    """constant 0.0: every node its own center at distance 0, the contracted graph is the graph itself (parallel arcs
    merged, the last weight kept); one center that reaches everything: a single node, diameter 0."""
    fx = reference_fixtures["fixtures"]["four"]
    c = csr_from_fixture(fx)
    small, centers, dist = orc.contract_balls(c, 2, lambda: 0.0)
    assert small.n == c.n and np.array_equal(centers, np.arange(c.n)) and all(d == 0 for d in dist)
    np.testing.assert_array_equal(small.off, c.off)
    np.testing.assert_array_equal(small.tgt, c.tgt)
    np.testing.assert_array_equal(small.w, c.w)
    # a directed cycle with one center and r at least half the cycle length: one ball swallows it in iteration 0
    n = 12
    cyc = orc.Csr.from_edges(n, np.arange(n), (np.arange(n) + 1) % n, np.ones(n, np.int32))
    first = draws_of("single_center")
    small, centers, dist = orc.contract_balls(cyc, n, first)
    assert small.n == 1 and np.all(centers == 0) and list(dist) == list(range(n))
    assert orc.diameter_approximation(cyc, n, draws_of("single_center")) == 2 * n


@pytest.mark.parametrize("seed", [0, 1, 5])
def test_clustering_invariants(seed):
    """On random unit-weight graphs (every node reached): a node's ball distance is the length of a path from its
    center, so at least the shortest distance from the center and within the grown radius; centers are their own
    centers at distance 0; every contracted arc is w + d(u) + d(v) of an original arc between the two balls."""
    hg = ring_graph(1200, 4, 1, 1, seed)
    c = orc.Csr(hg.n, hg.offsets, hg.targets, hg.weights)
    r = 1
    small, centers, dist = orc.contract_balls(c, r, None, seed)
    last = int(np.ceil(np.log2(c.n)))
    cs = sorted(set(int(x) for x in centers))
    assert small.n == len(cs)
    for ctr in cs:
        assert centers[ctr] == ctr and dist[ctr] == 0
        dref, _ = orc.sssp(c, ctr)
        members = np.flatnonzero(centers == ctr)
        assert np.all(np.array([dist[v] for v in members], np.uint64) >= dref[members])
    assert max(dist) <= last * 2 * r
    new_id = {v: i for i, v in enumerate(cs)}
    arcs = {}
    for u in range(c.n):
        for a in range(int(c.off[u]), int(c.off[u + 1])):
            t = int(c.tgt[a])
            arcs[(new_id[int(centers[u])], new_id[int(centers[t])])] = int(c.w[a]) + dist[u] + dist[t]
    got = {(s, int(small.tgt[a])): int(small.w[a]) for s in range(small.n) for a in range(int(small.off[s]), int(small.off[s + 1]))}
    assert got == arcs


def test_mt19937_draws_match_libstdcxx(tmp_path):
    """The oracle's default-generator draws equal std::uniform_real_distribution<double>(0, 1) over std::mt19937
    from libstdc++, compiled and run here; and the standard's check value of mt19937 (its 10000th output from the
    default seed is 4123659995, [rand.predef])."""
    src = tmp_path / "draws.cpp"
    src.write_text(textwrap.dedent(r"""
        #include <cstdio>
        #include <random>
        int main() {
            for (unsigned seed : {0u, 1u, 7u, 4000000000u}) {
                std::mt19937 g(seed);
                std::uniform_real_distribution<double> d(0.0, 1.0);
                for (int i = 0; i < 64; ++i) std::printf("%.17g\n", d(g));
            }
            std::mt19937 def;
            def.discard(9999);
            std::printf("%u\n", static_cast<unsigned>(def()));
        }
    """))
    exe = tmp_path / "draws"
    subprocess.run(["g++", "-std=c++20", "-O1", str(src), "-o", str(exe)], check=True, capture_output=True, timeout=120)
    lines = subprocess.run([str(exe)], check=True, capture_output=True, text=True, timeout=60).stdout.split()
    assert lines[-1] == "4123659995"
    k = 0
    for seed in (0, 1, 7, 4000000000):
        d = orc.Mt19937Canonical(seed)
        for _ in range(64):
            assert float(lines[k]) == d(), (seed, k)
            k += 1


def _same_contraction(hg, r, draws_kind, seed=0):
    import neurograph_amd as ng

    c = orc.Csr(hg.n, hg.offsets, hg.targets, hg.weights)
    try:
        so, co, _ = orc.contract_balls(c, r, draws_of(draws_kind), seed)
        err_o = None
    except orc.OracleError as e:
        err_o = str(e)
    try:
        sp, cp = ng.contract_balls(hg, r, draws_of(draws_kind), seed)
        err_p = None
    except ng.NgError as e:
        err_p = str(e)
    if err_o or err_p:
        # both refuse, at the same node (the reference throws there, DiameterApproximation.h:329)
        assert err_o and err_p and err_o.split(" was ")[0] in err_p, (err_o, err_p)
        return None
    np.testing.assert_array_equal(cp, co)
    np.testing.assert_array_equal(sp.offsets, so.off)
    np.testing.assert_array_equal(sp.targets, so.tgt)
    np.testing.assert_array_equal(sp.weights, so.w)
    return sp


def test_library_contraction_matches_oracle_on_fixtures(reference_fixtures):
    import neurograph_amd as ng

    for name, fx, case in fixture_cases(reference_fixtures):
        c = csr_from_fixture(fx)
        hg = ng.HostGraph(c.n, c.off, c.tgt, c.w)
        assert _same_contraction(hg, case["r"], case["draws"]) is not None, name


@pytest.mark.parametrize("n,k,wlo,whi,r,seed,complete", [(2000, 5, 1, 1, 1, 0, True), (3000, 3, 1, 1, 1, 3, True),
                                                         (1500, 9, 1, 3, 2, 2, True), (5000, 5, 1, 1, 1, 0, True),
                                                         (2000, 5, 1, 10, 1, 0, False), (1500, 5, 1, 10, 5, 4, None)])
def test_library_contraction_matches_oracle_random(n, k, wlo, whi, r, seed, complete):
    """Default generator (std::mt19937 seeded like rank `seed`) on ring-plus-random graphs: small weights, where every
    node is reached, and weights above 2r, where some node is not and both refuse at the same node."""
    small = _same_contraction(ring_graph(n, k, wlo, whi, seed), r, "default", seed)
    if complete is not None:
        assert (small is not None) == complete
    if small is not None:
        assert 1 <= small.n < n


def test_contraction_rejects_inputs():
    import neurograph_amd as ng

    neg = ng.HostGraph(3, np.array([0, 1, 2, 2], np.uint64), np.array([1, 2], np.uint32), np.array([1, -2], np.int32))
    with pytest.raises(ng.NgError) as e:
        ng.contract_balls(neg, 1, draws_of("constant0"))
    assert e.value.status == ng.NG_EINVAL
    real = ng.HostGraph(2, np.array([0, 1, 1], np.uint64), np.array([1], np.uint32), np.array([1.5], np.float64))
    with pytest.raises(ng.NgError) as e:
        ng.contract_balls(real, 1)
    assert e.value.status == ng.NG_EUNSUPPORTED
    one = ng.HostGraph(1, np.array([0, 0], np.uint64), np.zeros(0, np.uint32), np.zeros(0, np.int32))
    with pytest.raises(ng.NgError):  # ceil(log2 1) = 0 iterations: the node never becomes a center
        ng.contract_balls(one, 1, draws_of("constant0"))
