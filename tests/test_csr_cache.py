"""Binary CSR cache (ng_save_csr / ng_load_csr, CPU): a graph parsed from the reference's text format round-trips
through the cache bit for bit (offsets, targets, weights of every kind, positions); a cache written for other load
options, a corrupted payload or header, a truncated file and a malformed CSR are refused with NG_EINVAL."""

import os

import numpy as np
import pytest


def equal(a, b):
    assert a.n == b.n and a.m == b.m and a.weights.dtype == b.weights.dtype
    np.testing.assert_array_equal(a.offsets, b.offsets)
    np.testing.assert_array_equal(a.targets, b.targets)
    np.testing.assert_array_equal(a.weights, b.weights)
    if a.positions is None:
        assert b.positions is None
    else:
        np.testing.assert_array_equal(a.positions, b.positions)


@pytest.mark.parametrize("wdtype", [np.int32, np.float32, np.float64])
def test_text_to_cache_round_trip(tmp_path, wdtype):
    import neurograph_amd as ng

    if wdtype == np.int32:
        g = ng.generate_fixed_degree(3000, 7, 1, 10, 4)
    else:
        g = ng.generate_spatial(3000, 15, weights=wdtype, seed=4)
    d = str(tmp_path / "in")
    ng.write_graph_dir(d, g)
    text = ng.load_graph_dir(d, weights=wdtype)
    path = str(tmp_path / "g.ngcsr")
    ng.save_csr(path, text, key=0x1234)
    back = ng.load_csr(path, key=0x1234)
    equal(text, back)
    assert os.path.getsize(path) == 64 + (g.n + 1) * 8 + g.m * (4 + np.dtype(wdtype).itemsize) + (
        3 * g.n * 8 if text.positions is not None else 0)


def test_cache_without_positions_and_empty_graph(tmp_path):
    import neurograph_amd as ng

    g = ng.HostGraph(5, np.array([0, 2, 2, 3, 3, 3], np.uint64), np.array([1, 4, 0], np.uint32), np.array([3, 1, 2], np.int32))
    path = str(tmp_path / "small.ngcsr")
    ng.save_csr(path, g)
    equal(g, ng.load_csr(path))
    e = ng.HostGraph(0, np.zeros(1, np.uint64), np.zeros(0, np.uint32), np.zeros(0, np.float64))
    ng.save_csr(path, e)
    equal(e, ng.load_csr(path))


def test_cache_refusals(tmp_path):
    import neurograph_amd as ng

    g = ng.generate_fixed_degree(500, 4, 1, 9, 1)
    path = str(tmp_path / "g.ngcsr")
    ng.save_csr(path, g, key=7)
    raw = open(path, "rb").read()

    def refused(data, key=7, what=""):
        p = str(tmp_path / "bad.ngcsr")
        with open(p, "wb") as f:
            f.write(data)
        with pytest.raises(ng.NgError) as e:
            ng.load_csr(p, key=key)
        assert e.value.status == ng.NG_EINVAL
        assert what in str(e.value), str(e.value)

    refused(raw, key=8, what="other input or load options")
    flipped = bytearray(raw)
    flipped[64 + 8 * 100] ^= 0x01  # one bit of an offset
    refused(bytes(flipped), what="checksum")
    flipped = bytearray(raw)
    flipped[-1] ^= 0x80  # the last weight
    refused(bytes(flipped), what="checksum")
    flipped = bytearray(raw)
    flipped[16] ^= 0x01  # n in the header: the size no longer matches
    refused(bytes(flipped), what="size")
    refused(raw[:-4], what="size")
    refused(raw[:40], what="truncated header")
    refused(b"NOTACSR!" + raw[8:], what="not a CSR cache")
    bad = bytearray(raw)
    bad[8] = 9  # version
    refused(bytes(bad), what="version")
    # a well-formed file of a malformed CSR (a target out of range) is refused after the checksum passes
    t = g.targets.copy()
    t[3] = g.n
    ng.save_csr(str(tmp_path / "oob.ngcsr"), ng.HostGraph(g.n, g.offsets, t, g.weights), key=7)
    with pytest.raises(ng.NgError) as e:
        ng.load_csr(str(tmp_path / "oob.ngcsr"), key=7)
    assert "out of range" in str(e.value)
    with pytest.raises(ng.NgError):
        ng.load_csr(str(tmp_path / "missing.ngcsr"))
