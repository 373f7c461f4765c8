"""DiameterApproximation with its inner APSP on the GPU (ng_diameter_approximation: the host contraction of
ng_contract_balls, then AllPairsShortestPath::compute_apsp of the contracted graph through ng_sssp_summaries, as at
DiameterApproximation.h:118): the reference's own expectations (test_diameter_approximation.cpp, with its constant
and single-center generators), the oracle's approximation on random graphs with the default generator, and the
contracted graphs' device distances against the oracle, bit-exact."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import oracle as orc
from tests.conftest import csr_from_fixture
from tests.test_diameter import draws_of, ring_graph


@pytest.fixture(scope="module")
def ctx():
    import neurograph_amd as ng

    return ng.Context(0)


def test_reference_expectations_on_gpu(ctx, reference_fixtures):
    import neurograph_amd as ng

    seen = 0
    for name, fx in reference_fixtures["fixtures"].items():
        for case in fx.get("diameter_approximation", []):
            c = csr_from_fixture(fx)
            hg = ng.HostGraph(c.n, c.off, c.tgt, c.w)
            got, k = ctx.diameter_approximation(hg, case["r"], draws_of(case["draws"]))
            if case["expected"] is None:
                assert got == ctx.diameter_approximation(hg, case["r"])[0] == orc.diameter_approximation(c, case["r"])
            else:
                assert got == case["expected"], (name, case)
            seen += 1
    assert seen == 19


@pytest.mark.parametrize("n,k,wlo,whi,r,seed", [(5000, 5, 1, 1, 1, 0), (5000, 5, 1, 1, 1, 1), (4000, 8, 1, 3, 2, 2),
                                                (3000, 4, 1, 2, 3, 3)])
def test_default_generator_matches_oracle(ctx, n, k, wlo, whi, r, seed):
    import neurograph_amd as ng

    hg = ring_graph(n, k, wlo, whi, seed)
    c = orc.Csr(hg.n, hg.offsets, hg.targets, hg.weights)
    got, kn = ctx.diameter_approximation(hg, r, None, seed)
    small_o, _, _ = orc.contract_balls(c, r, None, seed)
    assert kn == small_o.n and 1 < kn < n
    assert got == int(orc.apsp(small_o, threads=8)["diameter_u64"]) + 2 * r
    # the contracted graph itself on the device: every distance bit-exact against the oracle
    small, _ = ng.contract_balls(hg, r, None, seed)
    g = ng.Graph(ctx, small)
    _, dist = g.sssp(None, want_dist=True)
    np.testing.assert_array_equal(dist, np.stack([orc.sssp(small_o, s)[0] for s in range(small_o.n)]))


def test_refusal_when_a_node_is_unreached(ctx):
    """Weights above 2r leave nodes without a ball: the reference throws (DiameterApproximation.h:329), and so does
    the device path, before any device work."""
    import neurograph_amd as ng

    hg = ring_graph(2000, 5, 1, 10, 0)
    with pytest.raises(ng.NgError) as e:
        ctx.diameter_approximation(hg, 1)
    assert e.value.status == ng.NG_EINVAL and "reached by no ball" in str(e.value)
