"""Writes tests/golden/reference_fixtures.json: the reference's test graphs as arc lists, each with the
expected values its own test suite asserts for the APSP / Brandes / delta-stepping hot path.

The graphs are rebuilt here from their construction in the reference's test fixtures (paths relative to
the reference checkout) and flattened to one rank:

* ``standard``        tests/GraphTestOneRank.cpp:25-105      (10 nodes; out arcs only matter for paths)
* ``standard_uu``     tests/GraphTestOneRank.cpp:129-212     (synchronize_arcs + all_weights_one)
* ``full``            tests/GraphTestOneRank.cpp:214-275     (all 100 ordered pairs, self-loops included)
* ``negative``        tests/GraphTestOneRank.cpp:277-326     (weights used by absolute value)
* ``disconnected``    tests/apsp/test_apsp.cpp:31-64         (directed 3-cycle + isolated node)
* ``four``            tests/GraphTestOneRank.cpp:328-544     (4 copies of standard + ring of weight 11)
* ``four_uu``         tests/GraphTestOneRank.cpp:546-765
* ``full36``          tests/GraphTestOneRank.cpp:767-929     (36 nodes, all pairs without self)
* ``seven``           tests/GraphTestSevenRanks.cpp:353-377  (70 nodes, two rings + a source/sink block)
* ``seven_uu``        tests/GraphTestSevenRanks.cpp:353-377
* ``full63``          tests/GraphTestSevenRanks.cpp:379-390  (63 nodes, all pairs without self)

Expected values are transcribed from the reference's assertions (file:line given per entry). Only
``negative`` carries hand-derived values (the reference has no path assertion on it).

Run:  python tests/golden/make_fixtures.py
"""

from __future__ import annotations

import json
import os
from collections import OrderedDict

HERE = os.path.dirname(os.path.abspath(__file__))

# Arcs of the 10-node "standard" graph as (source, target, weight), tests/GraphTestOneRank.cpp:89-99.
STANDARD_OUT = [
    (0, 6, 1), (0, 8, 2), (0, 9, 1),
    (1, 0, 1), (1, 6, 3),
    (2, 3, 1), (2, 4, 2),
    (3, 1, 1), (3, 5, 2), (3, 7, 1),
    (4, 1, 1), (4, 2, 2), (4, 6, 2), (4, 8, 1),
    (5, 5, 1), (5, 6, 9), (5, 7, 2),
    (6, 6, 2), (6, 9, 1),
    (7, 2, 1), (7, 3, 2),
    (8, 3, 1), (8, 9, 2),
    (9, 1, 2), (9, 4, 1),
]
# In arcs of the same graph as (target, source, weight), tests/GraphTestOneRank.cpp:77-87.
STANDARD_IN = [
    (0, 1, 1),
    (1, 3, 1), (1, 9, 2), (1, 4, 1),
    (2, 7, 1), (2, 4, 2),
    (3, 2, 1), (3, 7, 2), (3, 8, 1),
    (4, 9, 1), (4, 2, 2),
    (5, 5, 1), (5, 3, 2),
    (6, 0, 1), (6, 4, 2), (6, 1, 3), (6, 5, 9), (6, 6, 2),
    (7, 3, 1), (7, 5, 2),
    (8, 4, 1), (8, 0, 2),
    (9, 6, 1), (9, 8, 2), (9, 0, 1),
]


def _maps(n):
    return [OrderedDict() for _ in range(n)]


def synchronize(in_maps, out_maps):
    """ArcTransformer::synchronize_arcs (source/graph/ArcTransformer.cpp:53-67): every in arc t <- s with
    weight w adds w to the out arc t -> s; a sum of 0 becomes 1."""
    for t, m in enumerate(in_maps):
        for s, w in m.items():
            cur = out_maps[t].get(s, 0) + w
            out_maps[t][s] = 1 if cur == 0 else cur
    return out_maps


def all_weights_one(maps):
    """ArcTransformer::all_weights_one (source/graph/ArcTransformer.cpp:47-50)."""
    for m in maps:
        for k in m:
            m[k] = 1
    return maps


def to_arcs(out_maps):
    """Upload order does not matter for the checked values; weight-0 arcs are skipped at upload
    (source/graph/DistributedGraph.cpp:222-225)."""
    arcs = []
    for s, m in enumerate(out_maps):
        for t, w in sorted(m.items()):
            if w != 0:
                arcs.append((s, t, w))
    return arcs


def standard_block(offset, in_maps, out_maps):
    for s, t, w in STANDARD_OUT:
        out_maps[s + offset][t + offset] = w
    for t, s, w in STANDARD_IN:
        in_maps[t + offset][s + offset] = w


def g_standard(uu=False):
    in_maps, out_maps = _maps(10), _maps(10)
    standard_block(0, in_maps, out_maps)
    if uu:
        out_maps = all_weights_one(synchronize(in_maps, out_maps))
    return 10, to_arcs(out_maps)


def g_full():
    return 10, [(s, t, 1) for s in range(10) for t in range(10)]


def g_full_n(n):
    return n, [(s, t, 1) for s in range(n) for t in range(n) if s != t]


def g_four(uu=False):
    """get_standard_four_rank_graph_on_one_rank: 4 standard blocks; ring arcs i+10r -> i+10((r+1)%4), w=11."""
    in_maps, out_maps = _maps(40), _maps(40)
    for r in range(4):
        standard_block(10 * r, in_maps, out_maps)
        for i in range(10):
            in_maps[i + 10 * r][i + 10 * ((r + 3) % 4)] = 11
            out_maps[i + 10 * r][i + 10 * ((r + 1) % 4)] = 11
    if uu:
        out_maps = all_weights_one(synchronize(in_maps, out_maps))
    return 40, to_arcs(out_maps)


# tests/GraphTestSevenRanks.cpp:36-57: rank -> (predecessor, successor); rank 3 is the source/sink block.
SEVEN_RING = {0: (6, 1), 1: (0, 4), 2: (5, 5), 4: (1, 6), 5: (2, 2), 6: (4, 0)}


def g_seven(uu=False):
    in_maps, out_maps = _maps(70), _maps(70)
    for r in range(7):
        off = 10 * r
        if r == 3:
            # get_source_sink_arcs, tests/GraphTestSevenRanks.cpp:189-206
            for s, t, w in ((6, 3, 1), (7, 4, 2), (8, 5, 3), (9, 3, 4)):
                out_maps[off + s][off + t] = w
                in_maps[off + t][off + s] = w
            continue
        pred, succ = SEVEN_RING[r]
        standard_block(off, in_maps, out_maps)
        for i in range(10):
            in_maps[off + i][10 * pred + i] = 11
            out_maps[off + i][10 * succ + i] = 11
    if uu:
        out_maps = all_weights_one(synchronize(in_maps, out_maps))
    return 70, to_arcs(out_maps)


def g_negative():
    return 4, [(0, 1, -3), (0, 2, 2), (1, 2, -1), (1, 3, 2), (2, 3, -4)]


def g_disconnected():
    return 4, [(0, 1, 1), (1, 2, 1), (2, 0, 1)]


def apsp_expect(spl, eff, diam, disc, clusters, spl_exact=False, src=""):
    return {"avg_spl": spl, "avg_eff": eff, "diameter": diam, "n_disconnected": disc,
            "cluster_sizes": {str(k): v for k, v in clusters.items()}, "spl_exact": spl_exact, "source": src}


LARGE = dict(
    sums=[760, 784, 756, 748, 740, 840, 788, 784, 752, 764],
    maxima=[38, 39, 37, 37, 37, 40, 39, 38, 37, 38],
    eff=[5.8086783847071217, 4.8829981275875758, 5.4463030539884745, 6.1299867303575732, 6.3949926703635134,
         3.4097103354784561, 4.6742945932261470, 4.7133019123031259, 5.5366875043729253, 5.2153791587224756],
)
LARGE_UU = dict(
    sums=[104, 92, 104, 92, 92, 100, 92, 108, 96, 96],
    maxima=[5, 4, 5, 4, 4, 4, 4, 5, 4, 5],
    eff=[18.150000000000002, 19.833333333333336, 17.616666666666667, 19.833333333333332, 19.833333333333332, 18.0,
         19.833333333333332, 17.233333333333334, 18.916666666666668, 19.449999999999999],
)


def per_source_seven():
    """tests/delta_stepping/seven_rank_expectations.h:91-119 (block order ranks 0..6)."""
    small = dict(sums=[160, 172, 158, 154, 150, 200, 174, 172, 156, 162], maxima=[16, 17, 15, 15, 15, 18, 17, 16, 15, 16],
                 eff=[5.1167790542790522, 4.2047692503574865, 4.7526307026306999, 5.4312021312021292, 5.6914585414585390,
                      2.7620665282429986, 3.9986795067677430, 4.0356893106893117, 4.8407259407259392, 4.5263028638028633])
    rows = []
    for r in range(7):
        if r == 3:
            rows += [(0, 0, 0.0, 69)] * 6 + [(1, 1, 1.0, 68), (2, 2, 0.5, 68), (3, 3, 1.0 / 3.0, 68), (4, 4, 0.25, 68)]
        elif r in (2, 5):
            rows += [(small["sums"][i], small["maxima"][i], small["eff"][i], 50) for i in range(10)]
        else:
            rows += [(LARGE["sums"][i], LARGE["maxima"][i], LARGE["eff"][i], 30) for i in range(10)]
    return rows


def per_source_seven_uu():
    """tests/delta_stepping/seven_rank_expectations.h:126-159."""
    small = dict(sums=[42, 36, 42, 36, 36, 40, 36, 44, 38, 38], maxima=[4, 3, 4, 3, 3, 3, 3, 4, 3, 4],
                 eff=[10.666666666666668, 11.833333333333336, 10.250000000000002, 11.833333333333334, 11.833333333333334,
                      10.5, 11.833333333333336, 10.0, 11.166666666666668, 11.583333333333334])
    three = dict(sums=[0, 0, 0, 2, 1, 1, 3, 1, 1, 3], maxima=[0, 0, 0, 1, 1, 1, 2, 1, 1, 2],
                 eff=[0.0, 0.0, 0.0, 2.0, 1.0, 1.0, 1.5, 1.0, 1.0, 1.5], unreach=[69, 69, 69, 67, 68, 68, 67, 68, 68, 67])
    large_eff = list(LARGE_UU["eff"])
    large_eff[6] = 19.833333333333336  # seven_rank_expectations.h:130-131 differs in the last digit from the 4-rank table
    rows = []
    for r in range(7):
        if r == 3:
            rows += [(three["sums"][i], three["maxima"][i], three["eff"][i], three["unreach"][i]) for i in range(10)]
        elif r in (2, 5):
            rows += [(small["sums"][i], small["maxima"][i], small["eff"][i], 50) for i in range(10)]
        else:
            rows += [(LARGE_UU["sums"][i], LARGE_UU["maxima"][i], large_eff[i], 30) for i in range(10)]
    return rows


def rows_from(block, reps, unreach=0):
    return [(block["sums"][i], block["maxima"][i], block["eff"][i], unreach) for _ in range(reps) for i in range(10)]


DIA = "tests/diameter_approximation/test_diameter_approximation.cpp"


def dia(*cases):
    """Diameter approximation expectations: (draws, r, expected, lines). draws "constant0" always returns 0.0 (every
    node becomes its own center), "single_center" returns 0.0 once and 1.0 afterwards (one center), "default" is the
    default generator, for which the reference asserts only that two runs agree (expected None)."""
    return [{"draws": d, "r": r, "expected": e, "source": f"{DIA}:{lines}"} for d, r, e, lines in cases]


def main():
    fixtures = OrderedDict()

    def add(name, graph, **expect):
        n, arcs = graph
        fixtures[name] = {"n": n, "arcs": arcs, **expect}

    add("standard", g_standard(),
        apsp=apsp_expect(3.1, 0.4222486772486772, 7, 0, {10: 10}, True, "tests/apsp/test_apsp.cpp:67-84"),
        apsp_inverse=apsp_expect(1.7043209876543211, 0.89050315072273489, 4.5, 0, {10: 10}, False,
                                 "tests/apsp/test_apsp.cpp:274-291"),
        brandes_avg={"value": 13.583333333333332, "source": "tests/centrality/test_brandes.cpp:23-34"},
        diameter_approximation=dia(("constant0", 2, 11, "60-73"), ("single_center", 2, 4, "150-164")),
        per_source={"rows": [(s, m, e, 0) for s, m, e in zip(
            [25, 31, 24, 22, 20, 45, 32, 31, 23, 26], [5, 6, 4, 4, 4, 7, 6, 5, 4, 5],
            [4.3666666, 3.4833333, 4.0, 4.6666666, 4.9166666, 2.1023809, 3.2833333, 3.3166666, 4.0833333, 3.7833333])],
            "eff_tol": 1e-6, "source": "tests/delta_stepping/test_push_check.cpp:27-50"})
    add("standard_uu", g_standard(uu=True),
        apsp=apsp_expect(1.6, 0.7222222222222222, 3, 0, {10: 10}, False, "tests/apsp/test_apsp.cpp:110-127"),
        apsp_inverse=apsp_expect(1.6, 0.7222222222222222, 3, 0, {10: 10}, False, "tests/apsp/test_apsp.cpp:293-310"),
        brandes_avg={"value": 5.4, "source": "tests/centrality/test_brandes.cpp:36-47"},
        diameter_approximation=dia(("constant0", 1, 5, "75-88"), ("single_center", 1, 2, "166-180"),
                                   ("default", 2, None, "182-195")),
        per_source={"rows": [(s, m, e, 0) for s, m, e in zip(
            [16, 13, 16, 13, 13, 15, 13, 17, 14, 14], [3, 2, 3, 2, 2, 2, 2, 3, 2, 3],
            [6.1666666, 7.0, 5.8333333, 7.0, 7.0, 6.0, 7.0, 5.6666666, 6.5, 6.8333333])],
            "eff_tol": 1e-6, "source": "tests/delta_stepping/test_push_check.cpp:52-75"})
    add("full", g_full(),
        apsp=apsp_expect(1.0, 1.0, 1, 0, {10: 10}, True, "tests/apsp/test_apsp.cpp:129-146"),
        apsp_inverse=apsp_expect(1.0, 1.0, 1, 0, {10: 10}, True, "tests/apsp/test_apsp.cpp:312-329"),
        brandes_avg={"value": 0.0, "source": "tests/centrality/test_brandes.cpp:49-60"},
        diameter_approximation=dia(("constant0", 2, 5, "90-103")),
        per_source={"rows": [(9, 1, 9.0, 0)] * 10, "eff_tol": 1e-6, "source": "tests/delta_stepping/test_push_check.cpp:77-95"})
    add("negative", g_negative(),
        per_source={"rows": [(10, 5, 1 / 3 + 1 / 2 + 1 / 5, 0), (3, 2, 1 + 1 / 2, 1), (4, 4, 0.25, 2), (0, 0, 0.0, 3)],
                    "eff_tol": 1e-12, "source": "hand-derived from |w| (GraphTestOneRank.cpp:277-326)"})
    add("disconnected", g_disconnected(),
        apsp=apsp_expect(1.5, 0.75, 2, 6, {3: 3, 1: 1}, True, "tests/apsp/test_apsp.cpp:86-108"))
    add("four", g_four(),
        apsp=apsp_expect(19.784615384615385, 0.13387777556694200, 40, 0, {40: 40}, True, "tests/apsp/test_apsp.cpp:148-165"),
        apsp_inverse=apsp_expect(1.7130795130795136, 1.2097533633823983, 4.7727272727272725, 0, {40: 40}, False,
                                 "tests/apsp/test_apsp.cpp:331-348"),
        brandes_avg={"value": 112.12194180319423, "source": "tests/centrality/test_brandes.cpp:62-73"},
        diameter_approximation=dia(("constant0", 2, 44, "105-118"), ("constant0", 2, 44, "197-212")),
        per_source={"rows": rows_from(LARGE, 4), "eff_tol": 1e-6, "source": "tests/delta_stepping/test_push_check.cpp:97-129"})
    add("four_uu", g_four(uu=True),
        apsp=apsp_expect(2.5025641025641026, 0.48384615384615381, 5, 0, {40: 40}, False, "tests/apsp/test_apsp.cpp:167-184"),
        apsp_inverse=apsp_expect(2.5025641025641026, 0.48384615384615381, 5, 0, {40: 40}, False,
                                 "tests/apsp/test_apsp.cpp:350-367"),
        brandes_avg={"value": 58.6, "source": "tests/centrality/test_brandes.cpp:75-86"},
        diameter_approximation=dia(("constant0", 2, 9, "120-133"), ("constant0", 2, 9, "214-229"),
                                   ("default", 2, None, "248-263")),
        per_source={"rows": rows_from(LARGE_UU, 4), "eff_tol": 1e-6, "source": "tests/delta_stepping/test_push_check.cpp:131-175"})
    add("full36", g_full_n(36),
        apsp=apsp_expect(1.0, 1.0, 1, 0, {36: 36}, True, "tests/apsp/test_apsp.cpp:186-203"),
        apsp_inverse=apsp_expect(1.0, 1.0, 1, 0, {36: 36}, True, "tests/apsp/test_apsp.cpp:369-386"),
        brandes_avg={"value": 0.0, "source": "tests/centrality/test_brandes.cpp:88-99"},
        diameter_approximation=dia(("constant0", 2, 5, "135-148"), ("constant0", 2, 5, "231-246")),
        per_source={"rows": [(35, 1, 35.0, 0)] * 36, "eff_tol": 1e-6, "source": "tests/delta_stepping/test_push_check.cpp:177-195"})
    add("seven", g_seven(),
        apsp=apsp_expect(17.587448559670783, 0.15517143563687033, 40, 2886, {1: 6, 2: 4, 20: 20, 40: 40}, True,
                         "tests/apsp/test_apsp.cpp:461-483"),
        apsp_inverse=apsp_expect(1.7007328947915260, 1.2430679013912713, 4.7727272727272725, 2886,
                                 {1: 6, 2: 4, 20: 20, 40: 40}, False, "tests/apsp/test_apsp.cpp:603-624"),
        brandes_avg={"value": 74.447050803480930, "source": "tests/centrality/test_brandes.cpp:146-158"},
        diameter_approximation=dia(("constant0", 2, 44, "265-280"), ("constant0", 2, 44, "314-328")),
        per_source={"rows": per_source_seven(), "eff_tol": 1e-6, "source": "tests/delta_stepping/seven_rank_expectations.h:91-119"})
    add("seven_uu", g_seven(uu=True),
        apsp=apsp_expect(2.4061538461538463, 0.50605128205128700, 5, 2880, {1: 3, 2: 4, 3: 3, 20: 20, 40: 40}, False,
                         "tests/apsp/test_apsp.cpp:485-508"),
        apsp_inverse=apsp_expect(2.4061538461538463, 0.50605128205128700, 5.0, 2880, {1: 3, 2: 4, 3: 3, 20: 20, 40: 40},
                                 False, "tests/apsp/test_apsp.cpp:626-647"),
        brandes_avg={"value": 39.171428571428571, "source": "tests/centrality/test_brandes.cpp:160-171"},
        diameter_approximation=dia(("constant0", 2, 9, "282-296"), ("constant0", 2, 9, "330-344")),
        per_source={"rows": per_source_seven_uu(), "eff_tol": 1e-6,
                    "source": "tests/delta_stepping/seven_rank_expectations.h:126-159"})
    add("full63", g_full_n(63),
        apsp=apsp_expect(1.0, 1.0, 1, 0, {63: 63}, True, "tests/apsp/test_apsp.cpp:510-528"),
        apsp_inverse=apsp_expect(1.0, 1.0, 1.0, 0, {63: 63}, False, "tests/apsp/test_apsp.cpp:649-666"),
        brandes_avg={"value": 0.0, "source": "tests/centrality/test_brandes.cpp:173-185"},
        diameter_approximation=dia(("constant0", 2, 5, "298-312"), ("constant0", 2, 5, "346-360")),
        per_source={"rows": [(62, 1, 62.0, 0)] * 63, "eff_tol": 1e-6,
                    "source": "tests/delta_stepping/seven_rank_expectations.h:166-175"})

    path = os.path.join(HERE, "reference_fixtures.json")
    with open(path, "w") as f:
        json.dump({"generator": "tests/golden/make_fixtures.py", "deltas": [1, 3, 4, 10, 10000], "fixtures": fixtures}, f,
                  separators=(",", ":"))
        f.write("\n")
    print(f"wrote {path}: {len(fixtures)} fixtures")


if __name__ == "__main__":
    main()
