"""The allocation-free Brandes entry point (ng_brandes_accumulate_noalloc) against ng_brandes_accumulate, through
the C ABI: two calls in a row, the second on a smaller graph with fewer sources, each compared with the allocating
entry point (per-source summaries: integer fields and maxima exact, double sums within 1e-12 relative;
betweenness within 1e-9 relative, since the per-slot partial sums are added in another grouping) and with the
oracle; the accumulate-into semantics of bc_inout; more sources than one launch takes; a forced single slot; and
the rejection of a graph whose one source slot does not fit the scratch array."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import oracle as orc


def rand_csr(n, m, seed, wdtype, wlo=1, whi=9):
    rng = np.random.default_rng(seed)
    src = rng.integers(0, n, m)
    dst = rng.integers(0, n, m)
    if np.issubdtype(wdtype, np.integer):
        w = rng.integers(wlo, whi + 1, m).astype(wdtype)
    else:
        w = np.exp(rng.normal(0.0, 0.5, m)).astype(wdtype)
    return orc.Csr.from_edges(n, src, dst, w, dtype=wdtype)


def dev(ctx, c, **opts):
    import neurograph_amd as ng

    return ng.Graph(ctx, ng.HostGraph(c.n, c.off, c.tgt, c.w)).set_options(**opts)


def check_pair(g, c, srcs, oracle=True):
    bc_ref, s_ref = g.brandes(srcs, summaries=True)
    plain_slots = g.last_stats()["slots"]
    start = np.full(g.n, 0.5)
    bc, s = g.brandes_noalloc(srcs, bc=start.copy(), summaries=True)
    st = g.last_stats()
    assert st["slots"] >= 1 and st["launches"] == (len(srcs) + 16383) // 16384
    for f in ("n_unreached", "max_d", "sum_d_u64", "max_d_u64"):
        np.testing.assert_array_equal(s[f], s_ref[f], err_msg=f)
    np.testing.assert_allclose(s["sum_d"], s_ref["sum_d"], rtol=1e-12)
    np.testing.assert_allclose(s["sum_eff"], s_ref["sum_eff"], rtol=1e-12)
    np.testing.assert_allclose(bc - 0.5, bc_ref, rtol=1e-9, atol=1e-9)
    if oracle:
        np.testing.assert_allclose(bc - 0.5, orc.brandes(c, sources=srcs, threads=8)["bc"], rtol=1e-9, atol=1e-9)
    return st, plain_slots


def test_noalloc_two_calls_second_smaller():
    import neurograph_amd as ng

    ctx = ng.Context(0)
    big = rand_csr(6000, 60000, 81, np.float64)
    st, _ = check_pair(dev(ctx, big), big, np.arange(0, 6000, 7, dtype=np.uint32))
    small = rand_csr(1500, 9000, 82, np.int32, 1, 3)  # small integer weights: ties, the general layers
    st2, _ = check_pair(dev(ctx, small), small, np.arange(0, 1500, 11, dtype=np.uint32))
    assert st2["slots"] <= 137


def test_noalloc_shapes_and_chunks():
    import neurograph_amd as ng

    ctx = ng.Context(0)
    c = rand_csr(2000, 16000, 83, np.float32)
    for opts in ({"block": 1024}, {"block": 256, "max_slots": 1}, {"block": 512}, {"block": 64}):
        g = dev(ctx, c, **opts)
        st, _ = check_pair(g, c, np.arange(0, 2000, 9, dtype=np.uint32))
        assert st["block"] == opts["block"]
        if "max_slots" in opts:
            assert st["slots"] == 1
    # more sources than one launch takes: two launches, sources repeated
    g = dev(ctx, c, block=1024)
    srcs = np.tile(np.arange(2000, dtype=np.uint32), 9)
    st, _ = check_pair(g, c, srcs, oracle=False)
    assert st["launches"] == 2


def test_noalloc_spatial_fills_fewer_slots():
    """The 50k-node spatial graph: its slots take ~15 MB each, so the scratch holds a few dozen, not the hundreds the
    allocating entry point runs; same results."""
    import neurograph_amd as ng

    ctx = ng.Context(0)
    hg = ng.generate_spatial(50000, 20, wdist="lognormal", a=0.0, b=0.5, weights=np.float64, seed=0)
    c = orc.Csr(hg.n, hg.offsets, hg.targets, hg.weights)
    g = ng.Graph(ctx, hg)
    st, plain = check_pair(g, c, np.arange(0, hg.n, 397, dtype=np.uint32))
    assert st["slots"] < plain


def test_noalloc_rejects_graph_beyond_scratch():
    import neurograph_amd as ng

    ctx = ng.Context(0)
    n = 3_000_000
    c = orc.Csr.from_edges(n, np.arange(n), (np.arange(n) + 1) % n, np.ones(n, np.int32))
    g = dev(ctx, c)
    with pytest.raises(ng.NgError) as e:
        g.brandes_noalloc([0, 1])
    assert e.value.status == ng.NG_EINVAL and "scratch" in str(e.value)
