"""Cohort launch (ng_cohort_*): all sources of many graphs of uneven size in one hipGraph-captured persistent
launch. Each member's betweenness and per-source summaries must equal the plain path's (and the oracle's);
replays must repeat exactly, also with other eager work on the stream between them."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import oracle as orc


@pytest.fixture(scope="module")
def ctx():
    import neurograph_amd as ng

    return ng.Context(0)


def members(ng, wdtype, sizes, degs, seed):
    return [ng.generate_spatial(n, d, weights=wdtype, seed=seed + i) for i, (n, d) in enumerate(zip(sizes, degs))]


@pytest.mark.parametrize("wdtype", [np.float32, np.float64])
def test_cohort_matches_plain_and_oracle(ctx, wdtype):
    import neurograph_amd as ng

    hgs = members(ng, wdtype, [700, 2500, 300, 1800, 1200, 50], [8, 20, 5, 30, 12, 4], 40)
    gs = [ng.Graph(ctx, h) for h in hgs]
    co = ng.Cohort(ctx, gs)
    parts, summ, ms = co.brandes(summaries=True)
    assert ms > 0.0
    for i, (h, g) in enumerate(zip(hgs, gs)):
        bc_plain, s_plain = g.brandes(summaries=True)
        np.testing.assert_allclose(parts[i], bc_plain, rtol=1e-12, atol=1e-12)
        s_co = summ[co.prefix[i]:co.prefix[i + 1]]
        np.testing.assert_array_equal(s_co["n_unreached"], s_plain["n_unreached"])
        np.testing.assert_array_equal(s_co["max_d"], s_plain["max_d"])
        np.testing.assert_allclose(s_co["sum_d"], s_plain["sum_d"], rtol=1e-12)
        np.testing.assert_allclose(s_co["sum_eff"], s_plain["sum_eff"], rtol=1e-12)
        if h.n <= 1200:
            c = orc.Csr(h.n, h.offsets, h.targets, h.weights)
            ref = orc.brandes(c, threads=8)
            np.testing.assert_allclose(parts[i], ref["bc"], rtol=1e-9, atol=1e-9)


def test_cohort_integer_weights_bit_exact(ctx):
    import neurograph_amd as ng

    hgs = [ng.generate_fixed_degree(n, k, 1, 10, 7 + n) for n, k in ((400, 3), (1500, 6), (90, 2), (800, 9))]
    gs = [ng.Graph(ctx, h) for h in hgs]
    co = ng.Cohort(ctx, gs)
    parts, summ, _ = co.brandes(summaries=True)
    for i, h in enumerate(hgs):
        c = orc.Csr(h.n, h.offsets, h.targets, h.weights)
        ref = orc.brandes(c, threads=8, per_source=True)
        s_co = summ[co.prefix[i]:co.prefix[i + 1]]
        for f in ("sum_d_u64", "max_d_u64", "n_unreached"):
            np.testing.assert_array_equal(s_co[f], ref["per_source"][f])
        np.testing.assert_allclose(s_co["sum_eff"], ref["per_source"]["sum_eff"], rtol=1e-12)
        np.testing.assert_allclose(parts[i], ref["bc"], rtol=1e-9, atol=1e-9)


def test_cohort_replays_repeat(ctx):
    import neurograph_amd as ng

    hgs = members(ng, np.float32, [900, 1600, 400], [10, 25, 6], 90)
    gs = [ng.Graph(ctx, h) for h in hgs]
    co = ng.Cohort(ctx, gs)
    first, s0, _ = co.brandes(summaries=True)
    eager = [g.brandes() for g in gs]
    for a, b in zip(first, eager):
        np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12)
    for rep in range(3):
        gs[rep % 3].brandes(np.arange(0, hgs[rep % 3].n, 3))  # eager work on the same stream between replays
        ctx.synchronize()
        parts, s1, _ = co.brandes(summaries=True)
        for a, b in zip(first, parts):
            np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(s0["n_unreached"], s1["n_unreached"])
        np.testing.assert_allclose(s0["sum_d"], s1["sum_d"], rtol=1e-12)


def test_cohort_autotune_keeps_results(ctx):
    import neurograph_amd as ng

    hgs = members(ng, np.float32, [1000, 300, 2000], [15, 6, 25], 120)
    gs = [ng.Graph(ctx, h) for h in hgs]
    co = ng.Cohort(ctx, gs)
    before, _, _ = co.brandes()
    co.autotune(64)
    after, _, _ = co.brandes()
    for a, b in zip(before, after):
        np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12)


def test_cohort_errors(ctx):
    import neurograph_amd as ng

    g32 = ng.Graph(ctx, ng.generate_spatial(200, 5, weights=np.float32, seed=1))
    g64 = ng.Graph(ctx, ng.generate_spatial(200, 5, weights=np.float64, seed=2))
    with pytest.raises(ng.NgError, match="distance type"):
        ng.Cohort(ctx, [g32, g64]).brandes()
    z = ng.HostGraph.from_edges(3, [0, 1], [1, 2], np.array([0.0, 1.0], dtype=np.float32))
    with pytest.raises(ng.NgError, match="zero-length"):
        ng.Cohort(ctx, [g32, ng.Graph(ctx, z)]).brandes()


def test_cohort_recaptures_after_workspace_growth(ctx):
    """A plain call on a larger graph grows the context's shared workspace between replays; the next replay
    re-captures against the new buffers and gives the same results."""
    import neurograph_amd as ng

    hgs = members(ng, np.float32, [500, 900], [8, 12], 150)
    gs = [ng.Graph(ctx, h) for h in hgs]
    co = ng.Cohort(ctx, gs)
    first, _, _ = co.brandes()
    large = ng.Graph(ctx, ng.generate_spatial(30000, 10, weights=np.float64, seed=151))
    large.brandes(np.arange(0, 30000, 7))
    again, _, _ = co.brandes()
    for a, b in zip(first, again):
        np.testing.assert_array_equal(a, b)
