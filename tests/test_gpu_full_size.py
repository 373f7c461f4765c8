"""GPU parity at the full workload sizes: the 50k-node spatial graph (FP64 lognormal weights, all 50 000
sources), the 200k-node spatial graph (FP32 lognormal weights, ~20M arcs, all 8192 of its seeded sources in
batches of 1024), the 5000-node fixed-degree graph (all-pairs distance matrix and histogram), the 20k-node adjoint
graph (the gradient of every arc) and the 64-graph cohort (all 683 944 sources).

The oracle cannot run every source of these graphs in a test's time, so parity here rests on:
  * the oracle on a seeded sample of sources: distances bit-exact, per-source summaries (unreached counts and
    maxima exact, double sums to 1e-12 relative) and the sample's betweenness to 1e-9 relative;
  * identities that hold for every source set, whatever its size:
      - edge form: sum over arcs of EB(a) * len(a) = sum over sources of sum_d (every shortest s-t path has
        length d(s, t), and the path counts of one pair sum to 1 over its paths);
      - node form: sum_v BC(v) = sum_a EB(a) - sum_s reached(s), with reached(s) = n - 1 - unreached(s) (a
        shortest path with h arcs has h - 1 interior nodes);
      - linearity and determinism: BC(S1 u S2) = BC(S1) + BC(S2); a repeated call is bit-identical;
  * the cohort launch against 64 plain launches of the same members.

Tolerances: FP64 weights 1e-9 relative; FP32 weights 1e-5 relative for the edge form (the distances are
rounded f32 add chains, the identity is exact real arithmetic over the f32 lengths); betweenness sums 1e-9.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import oracle as orc


@pytest.fixture(scope="module")
def ctx():
    import neurograph_amd as ng

    return ng.Context(0)


def spatial50k():
    import neurograph_amd as ng

    hg = ng.generate_spatial(50000, 20, box=1000.0, lam=150.0, wdist="lognormal", a=0.0, b=0.5, weights=np.float64,
                             seed=0)
    return hg, np.arange(hg.n, dtype=np.uint32)


def spatial200k():
    """The 200k-node workload and all 8192 of its seeded sources (the bench runs them in batches of 1024)."""
    import neurograph_amd as ng

    hg = ng.generate_spatial(200000, 100, box=1000.0, lam=150.0, wdist="lognormal", a=0.0, b=0.5,
                             weights=np.float32, seed=0)
    rng = np.random.default_rng(0)
    src = np.sort(rng.choice(hg.n, size=8192, replace=False)).astype(np.uint32)
    return hg, src


def check_identities(hg, g, sources, rtol_edge):
    """Edge and node forms, linearity and determinism on one source set."""
    bc, summ = g.brandes(sources, summaries=True)
    edge = g.edge_dependency(sources)
    lens = np.abs(hg.weights.astype(np.float64))
    assert np.all(edge >= 0.0)
    np.testing.assert_allclose(np.dot(edge, lens), np.sum(summ["sum_d"]), rtol=rtol_edge)
    reached = (hg.n - 1 - summ["n_unreached"].astype(np.int64)).sum()
    np.testing.assert_allclose(np.sum(bc), np.sum(edge) - float(reached), rtol=1e-9)
    again = g.brandes(sources)
    np.testing.assert_array_equal(again, bc)
    half = len(sources) // 3
    part = g.brandes(sources[:half])
    both = g.brandes(sources[half:], bc=part.copy())
    np.testing.assert_allclose(both, bc, rtol=1e-9, atol=1e-9)
    return bc, summ


def check_oracle_sample(hg, g, sample, threads):
    c = orc.Csr(hg.n, hg.offsets, hg.targets, hg.weights)
    ref = orc.brandes(c, sources=sample, threads=threads, per_source=True, want_dist=True)
    summ, dist = g.sssp(sample, want_dist=True)
    np.testing.assert_array_equal(dist, ref["dist"])
    rs = ref["per_source"]
    np.testing.assert_array_equal(summ["n_unreached"], rs["n_unreached"])
    np.testing.assert_array_equal(summ["max_d"], rs["max_d"])
    np.testing.assert_allclose(summ["sum_d"], rs["sum_d"], rtol=1e-12)
    np.testing.assert_allclose(summ["sum_eff"], rs["sum_eff"], rtol=1e-12)
    bc, bsumm = g.brandes(sample, summaries=True)
    np.testing.assert_array_equal(bsumm["n_unreached"], rs["n_unreached"])
    np.testing.assert_allclose(bc, ref["bc"], rtol=1e-9, atol=1e-9)


def test_spatial50k_all_sources(ctx):
    import neurograph_amd as ng

    hg, src = spatial50k()
    g = ng.Graph(ctx, hg)
    check_identities(hg, g, src, rtol_edge=1e-9)
    rng = np.random.default_rng(11)
    check_oracle_sample(hg, g, np.sort(rng.choice(src, size=48, replace=False)).astype(np.uint32), threads=8)


def test_spatial200k_all_sources(ctx):
    """All 8192 sources in the bench's 8 batches of 1024: the edge and node identities, linearity and determinism
    over the whole set and per batch, and the oracle on 64 sources, 8 from every batch (distances bit-exact,
    summaries, betweenness within 1e-9 on that sample: the f32 lengths are summed in the same order)."""
    import neurograph_amd as ng

    hg, src = spatial200k()
    g = ng.Graph(ctx, hg)
    batches = [src[i:i + 1024] for i in range(0, len(src), 1024)]
    bc_all = np.zeros(hg.n)
    for b in batches:
        bc_b, _ = check_identities(hg, g, b, rtol_edge=1e-5)
        bc_all += bc_b
    assert g.last_stats()["launches"] == 1
    whole = np.zeros(hg.n)
    for b in batches:  # accumulate-into over the batches, like the bench loop
        g.brandes(b, bc=whole)
    np.testing.assert_allclose(whole, bc_all, rtol=1e-9, atol=1e-9)
    rng = np.random.default_rng(7)
    sample = np.concatenate([np.sort(rng.choice(b, size=8, replace=False)) for b in batches]).astype(np.uint32)
    check_oracle_sample(hg, g, sample, threads=16)


def test_cohort64_matches_plain_launches(ctx):
    import neurograph_amd as ng

    rng = np.random.default_rng(0)
    ns = rng.integers(1000, 20001, size=64)
    ds = rng.uniform(10.0, 60.0, size=64)
    hgs = [ng.generate_spatial(int(n), float(d), box=1000.0, lam=150.0, wdist="lognormal", a=0.0, b=0.5,
                               weights=np.float32, seed=1000 + i) for i, (n, d) in enumerate(zip(ns, ds))]
    gs = [ng.Graph(ctx, h) for h in hgs]
    co = ng.Cohort(ctx, gs)
    parts, summ, _ = co.brandes(summaries=True)
    for i, g in enumerate(gs):
        bc_plain, s_plain = g.brandes(summaries=True)
        np.testing.assert_allclose(parts[i], bc_plain, rtol=1e-9, atol=1e-9)
        s_co = summ[co.prefix[i]:co.prefix[i + 1]]
        np.testing.assert_array_equal(s_co["n_unreached"], s_plain["n_unreached"])
        np.testing.assert_array_equal(s_co["max_d"], s_plain["max_d"])
        np.testing.assert_allclose(s_co["sum_d"], s_plain["sum_d"], rtol=1e-12)
    # one member against the oracle on a seeded sample of its sources
    h = hgs[int(np.argmin(ns))]
    g = gs[int(np.argmin(ns))]
    check_oracle_sample(h, g, np.arange(0, h.n, 97, dtype=np.uint32), threads=8)


def test_fixed5k_distance_matrix_and_histogram(ctx, tmp_path):
    """The 5000-node, 6-out-arc integer graph: the full 5000 x 5000 distance matrix through the CLI's
    --distance-matrix and through ng_sssp_summaries, bit-exact against the oracle, and the distance histogram
    (ng_distance_histogram, 1-wide bins) against numpy's binning of the oracle's matrix, count for count."""
    import os
    import subprocess

    import neurograph_amd as ng

    hg = ng.generate_fixed_degree(5000, 6, 1, 10, 0)
    c = orc.Csr(hg.n, hg.offsets, hg.targets, hg.weights)
    ref = orc.brandes(c, want_dist=True, threads=16)["dist"]
    g = ng.Graph(ctx, hg)
    summ, dist = g.sssp(None, want_dist=True)
    np.testing.assert_array_equal(dist, ref)
    counts = g.distance_histogram(1.0, 96)
    off = ref[~np.eye(hg.n, dtype=bool)]
    fin = off[off != np.iinfo(np.uint64).max]
    np.testing.assert_array_equal(counts, np.bincount(np.minimum(fin, 95).astype(np.int64), minlength=96).astype(np.uint64))
    d = str(tmp_path / "c1")
    ng.write_graph_dir(d, hg)
    cli = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "neurograph_amd", "ng_cli")
    r = subprocess.run([cli, "--input", d, "--func", "avg-apsp", "--distance-matrix"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    mat = np.fromfile(os.path.join(d, "metrics", "all_pairs_distances.bin"), dtype=np.uint64).reshape(hg.n, hg.n)
    np.testing.assert_array_equal(mat, ref)


def test_adjoint20k_every_arc_against_oracle(ctx):
    """The 20k-node FP64 graph: the gradient of every one of its ~600k arcs over a 64-source sample against the
    oracle's edge path counts (the per-arc term of BrandesVariant.h:151-154) and its 1/d^2-seeded dependencies."""
    import neurograph_amd as ng

    hg = ng.generate_spatial(20000, 30, wdist="uniform", a=0.5, b=1.5, weights=np.float64, seed=5)
    c = orc.Csr(hg.n, hg.offsets, hg.targets, hg.weights)
    rng = np.random.default_rng(23)
    src = np.sort(rng.choice(hg.n, size=64, replace=False)).astype(np.uint32)
    g = ng.Graph(ctx, hg)
    g1, g2 = g.adjoint(src)
    one = orc.brandes(c, sources=src, edge=True, threads=16)["edge"]
    inv = orc.brandes(c, sources=src, edge=True, seed=orc.SEED_INV_D2, threads=16)["edge"]
    assert np.count_nonzero(one) > 0.2 * hg.m
    np.testing.assert_allclose(g1, one, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(g2, -inv, rtol=1e-9, atol=1e-15)
