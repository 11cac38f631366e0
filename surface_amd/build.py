"""Compile libsrf.so (the HIP path) in-tree with hipcc for gfx950.

    python -m surface_amd.build            # or __graft_entry__.build()

Floating-point flags are part of the contract: -ffp-contract=off (the
reference's x86-64 build has no FMA), correctly rounded float division and
square root, denormals kept. Packed-FP32 instruction selection is disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libsrf.so")
SOURCES = ["srf_api.hip", "srf_build.hip", "srf_trace.hip", "srf_scene.hip", "srf_volume.hip",
           "srf_isect.hip", "srf_host.hip"]

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm required to build libsrf.so)")


def build(verbose=False, extra=()):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *HIP_FLAGS, *extra, "-o", tmp, *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libsrf.so")
    if verbose and r.stderr.strip():
        sys.stderr.write(r.stderr)
    os.replace(tmp, LIB)
    return LIB


def build_cpp_test(verbose=False):
    """The C++ host tests of include/srf_accel.hpp and include/srf_frame.hpp
    (tests/cpp), linked against libsrf.so and the oracle; run by
    tests/test_gpu_cpp.py on a GPU."""
    root = os.path.dirname(HERE)
    oracle = os.path.join(root, "oracle")
    outs = []
    for name, extra in (("test_accel_cpp", ["-pthread"]),
                        ("test_frame_cpp", ["-fopenmp", "-pthread"]),
                        ("test_gl_cpp", ["-fopenmp", "-pthread"])):
        src = os.path.join(root, "tests", "cpp", name + ".cpp")
        out = os.path.join(root, "tests", "cpp", name)
        cmd = [hipcc(), "--offload-arch=gfx950", "-O2", "-std=c++17", *extra, "-o", out, src,
               "-L" + HERE, "-lsrf", "-L" + oracle, "-loracle",
               "-Wl,-rpath," + HERE, "-Wl,-rpath," + oracle,
               "-Wl,-rpath,$ORIGIN/../../surface_amd", "-Wl,-rpath,$ORIGIN/../../oracle"]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"hipcc failed building tests/cpp/{name}")
        outs.append(out)
    return outs


if __name__ == "__main__":
    print(build(verbose=True))
