"""The C++20 host mirror of the reference's drivers (neurograph_amd/host/ng_variants.hpp) compiles with the host
compiler against include/*.h, links libng_hip.so, and reports library errors by throwing ng_host::Exception,
as the reference throws utility::Exception (Exception.hpp:76). CPU only: no device call succeeds here."""

import os
import subprocess
import textwrap

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "neurograph_amd")

PROGRAM = textwrap.dedent(r"""
    #include <cstdio>
    #include <cstring>
    #include <string>
    #include "neurograph_amd/host/ng_variants.hpp"

    int main() {
        // selector enums carry the GPU variant next to the reference's values (Types.h:36-60)
        static_assert(static_cast<int>(ng_host::DijkstraQueueType::Gpu) == 3);
        static_assert(static_cast<int>(ng_host::DeltaSteppingEpochType::GpuMultiSource) == 5);
        static_assert(static_cast<int>(ng_host::BetweennessCentralityAlgorithm::Gpu) == 2);
        if (std::strncmp(ng_version(), "ng_hip", 6) != 0) return 2;
        // aggregation of per-source summaries, compute_apsp_impl (AllPairsShortestPath.h:130-162)
        std::vector<ng_sssp_summary> per(3);
        per[0] = {6.0, 6, 3.0, 3, 1.5, 0};
        per[1] = {4.0, 4, 2.0, 2, 2.0, 1};
        per[2] = {0.0, 0, 0.0, 0, 0.0, 3};
        const auto r = ng_host::AllPairsShortestPath::aggregate_integer(per, 4);
        if (r.number_disconnected_pairs != 4 || r.diameter != 3) return 3;
        if (r.average_shortest_path_length != 10.0 / 8.0) return 4;
        if (r.cluster_sizes.at(4) != 1 || r.cluster_sizes.at(3) != 1 || r.cluster_sizes.at(1) != 1) return 5;
        // without a usable device the context constructor throws the library's message
        try {
            ng_host::Context ctx(12345);
        } catch (const ng_host::Exception& e) {
            std::printf("%s\n", e.what());
            return std::string(e.what()).find("ng_context_create") == 0 ? 0 : 6;
        }
        return 7;
    }
""")


def test_host_mirror_compiles_links_and_throws(tmp_path):
    if not os.path.exists(os.path.join(LIBDIR, "libng_hip.so")):
        pytest.skip("libng_hip.so is not built")
    src = tmp_path / "mirror.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "mirror"
    cmd = ["g++", "-std=c++20", "-O1", "-Wall", f"-I{ROOT}", f"-I{os.path.join(ROOT, 'include')}", str(src), "-o", str(exe),
           f"-L{LIBDIR}", "-lng_hip", f"-Wl,-rpath,{LIBDIR}"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert run.returncode == 0, (run.returncode, run.stdout, run.stderr)
    assert run.stdout.startswith("ng_context_create")
