"""Builds the HIP library libbm3d_amd.so in-tree for gfx950.

    python vapoursynth-bm3d_amd/build.py

Every translation unit goes through hipcc.  -ffp-contract=off keeps each
multiply and add separately rounded, as the reference's baseline-x86-64 build
computes them; packed FP32 instructions are disabled for the device code.
"""
from __future__ import annotations

import os
import shlex
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libbm3d_amd.so")
PLUGIN = os.path.join(OUT_DIR, "libbm3d_vs.so")   # the VapourSynth plugin shim over LIB
CXX = os.environ.get("CXX", "g++")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("BM3D_OFFLOAD_ARCH", "gfx950")

SOURCES = ["bm3d_kernels.hip", "bm3d_agg.hip", "bm3d_vkernels.hip", "lbm_kernels.hip", "bm3d_engine.cpp", "bm3d_run.cpp", "lbm.cpp",
           "bm3d_batch.cpp",
           "bm3d_noalloc.cpp"]

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def write_dct_header(path: str, nmax: int = 32) -> None:
    """Emits float(2 cos(pi k (2j+1) / (2N))) for N = 1..nmax as exact hex
    literals, computed with the C library's cos (Python's math.cos), so the
    kernels can use them as instruction constants.  The library checks them
    against the host's std::cos at bm3d_init."""
    import math
    import struct

    def f32(v: float) -> float:
        return struct.unpack("<f", struct.pack("<f", v))[0]

    lines = ["// Generated by build.py -- do not edit.", "#pragma once", "namespace bm3d {",
             "template <int N> struct DctC;"]
    for N in range(1, nmax + 1):
        vals = []
        for k in range(N):
            for j in range(N):
                vals.append(f32(2.0 * math.cos(math.pi * float(k) * float(2 * j + 1) / (2.0 * float(N)))))
        lit = ", ".join(float.hex(v) + "f" for v in vals)
        lines.append(f"template <> struct DctC<{N}> {{ static constexpr float v[{N * N}] = {{{lit}}}; }};")
    lines.append(f"constexpr int kDctCMax = {nmax};")
    lines.append("}  // namespace bm3d")
    tmp = path + ".tmp"
    with open(tmp, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    os.replace(tmp, path)


def build(verbose: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    write_dct_header(os.path.join(OUT_DIR, "bm3d_dct_coefs.h"))
    objs, cmds = [], []
    for src in SOURCES:
        obj = os.path.join(OUT_DIR, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".cpp") else []
        cmds.append([HIPCC, *FLAGS, "-I", OUT_DIR, *lang, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    # translation units compile in parallel (at most 4 hipcc at a time)
    jobs = max(1, min(4, os.cpu_count() or 1))
    pending = list(cmds)
    running = []
    failed = None
    while pending or running:
        while pending and len(running) < jobs and failed is None:
            cmd = pending.pop(0)
            if verbose:
                print(" ".join(shlex.quote(c) for c in cmd), flush=True)
            running.append((cmd, subprocess.Popen(cmd)))
        if failed is not None:
            pending.clear()
        cmd, proc = running.pop(0)
        if proc.wait() != 0 and failed is None:
            failed = subprocess.CalledProcessError(proc.returncode, cmd)
    if failed is not None:
        raise failed
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"]
    if verbose:
        print(" ".join(shlex.quote(c) for c in cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_plugin(verbose)
    return LIB


def build_plugin(verbose: bool = False) -> str:
    """The VapourSynth API v4 plugin (host code only, no device code): loads
    libbm3d_amd.so from its own directory."""
    tmp = PLUGIN + ".tmp"
    cmd = [CXX, "-O2", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden", "-Wall", "-Wextra",
           "-Wno-unused-parameter", os.path.join(HERE, "vs", "bm3d_vsplugin.cpp"), "-o", tmp, "-L", OUT_DIR,
           "-lbm3d_amd", "-Wl,-rpath,$ORIGIN", "-lpthread"]
    if verbose:
        print(" ".join(shlex.quote(c) for c in cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, PLUGIN)
    return PLUGIN


if __name__ == "__main__":
    if "--plugin-only" in sys.argv:
        print(build_plugin(verbose="-v" in sys.argv))
    else:
        print(build(verbose="-v" in sys.argv))
