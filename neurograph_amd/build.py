"""Builds the in-tree native artefacts for gfx950.

* ``neurograph_amd/libng_hip.so`` -- HIP kernels + C ABI (include/neurograph_hip.h, include/neurograph_io.h)
* ``neurograph_amd/ng_cli``         -- command-line front end mirroring the reference's flags

Every HIP translation unit is compiled with ``-ffp-contract=off`` (distances and dependencies are plain
IEEE adds/multiplies, never fused) and with packed FP32 instructions disabled.
"""

from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libng_hip.so")
CLI = os.path.join(HERE, "ng_cli")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = ["-O3", "-std=c++20", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-parameter",
          f"-I{os.path.join(ROOT, 'include')}"]
DEVICE = [f"--offload-arch={ARCH}", "-munsafe-fp-atomics", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

SOURCES = ["ng_kernels.hip", "ng_api.cpp", "ng_io.cpp", "ng_diameter.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[-1])}")
    return r


def _obj(src):
    out = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON + DEVICE + ["-c", os.path.join(CSRC, src), "-o", out]
    _run(cmd)
    return out


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_obj, SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC] + DEVICE + ["-shared", "-fPIC", "-o", tmp] + objs + ["-pthread"])
    os.replace(tmp, LIB)
    cli_src = os.path.join(HOST, "ng_cli.cpp")
    if os.path.exists(cli_src):
        tmp = CLI + ".tmp"
        _run([HIPCC] + COMMON + [cli_src, "-o", tmp, f"-L{HERE}", "-lng_hip", "-Wl,-rpath,$ORIGIN", "-pthread"])
        os.replace(tmp, CLI)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
