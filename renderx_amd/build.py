"""Build librxhip.so for gfx950 in-tree (renderx_amd/librxhip.so).

HIP sources are compiled with hipcc --offload-arch=gfx950; geometry code needs
-ffp-contract=off and correctly rounded f32 division / square root so the
traversal reproduces the reference's float results bit for bit.  Packed FP32
instructions are disabled (-packed-fp32-ops).  Host-only C++ is compiled with g++.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "librxhip.so")
BUILD = os.path.join(HERE, "_build")

HIP_SOURCES = ["rx_capi.hip", "rx_traverse.hip", "rx_render.hip", "rx_bpt.hip", "rx_batch.hip", "rx_guide.hip"]
CXX_SOURCES = ["rx_tree.cpp", "rx_film.cpp", "rx_scene_io.cpp"]
CMP_OUT = os.path.join(HERE, "rx_cmp")  # the cmp tool (host only)

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
             "-fhip-fp32-correctly-rounded-divide-sqrt", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-Wall", "-Wno-unused-function", "-Wno-unused-result"]
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-pthread", "-Wall"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build(verbose: bool = False, out: str = OUT, defines: tuple = ()) -> str:
    """Compile every source and link `out`; `defines` (e.g. "RX_LDS_DEPTH=12")
    go to the HIP compiles, with their own object folder."""
    build_dir = BUILD if not defines else BUILD + "_" + "_".join(d.replace("=", "") for d in defines)
    os.makedirs(build_dir, exist_ok=True)
    dflags = ["-D" + d for d in defines]
    jobs = []
    for s in HIP_SOURCES:
        jobs.append([HIPCC, *HIP_FLAGS, *dflags, "-c", os.path.join(CSRC, s), "-o", os.path.join(build_dir, s + ".o")])
    for s in CXX_SOURCES:
        jobs.append(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, s), "-o", os.path.join(build_dir, s + ".o")])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for r in ex.map(_run, jobs):
            if verbose and (r.stdout or r.stderr):
                print(r.stdout, r.stderr, file=sys.stderr)
    objs = [os.path.join(build_dir, s + ".o") for s in HIP_SOURCES + CXX_SOURCES]
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", "-o", tmp, *objs])
    os.replace(tmp, out)
    if not defines:
        _run(["g++", *CXX_FLAGS, "-o", CMP_OUT + ".tmp", os.path.join(CSRC, "rx_cmp.cpp"),
              os.path.join(build_dir, "rx_film.cpp.o")])
        os.replace(CMP_OUT + ".tmp", CMP_OUT)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
