"""Builds libptamd.so in-tree (pathtracingtutorial_amd/libptamd.so).

Host stages (scene loading, generators, light tables) are plain C++ built
with g++; kernels and the C ABI are HIP built with hipcc for gfx950 only.
The objects go to pathtracingtutorial_amd/_build/.  Also builds the `pt`
command-line renderer (pathtracingtutorial_amd/pt) on top of the C ABI.

Usage: python -m pathtracingtutorial_amd.build [-j N] [--quiet]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libptamd.so")
CLI = os.path.join(PKG, "pt")
INCLUDE = os.path.join(ROOT, "include")

ARCH = "gfx950"
HOST_SOURCES = ["host/json.cpp", "host/scene_io.cpp", "host/generators.cpp", "host/light_tables.cpp",
                "host/image_codec.cpp", "host/gltf.cpp", "host/bvh_wide.cpp"]
HIP_SOURCES = ["capi.hip", "device/bvh_build.hip", "device/render.hip", "device/wavefront.hip", "device/adjoint.hip",
               "device/light_build.hip", "device/bvh_wide_build.hip"]
CLI_SOURCES = ["cli/pt_main.cpp"]

CXXFLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-function"]
HIPFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # no packed-FP32 instructions (v_pk_*_f32) in any kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build libptamd)")


def _run(cmd: list[str], quiet: bool) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if proc.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    if out and not quiet:
        print(out)


def build(jobs: int | None = None, quiet: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    hipcc = _hipcc()
    incs = [f"-I{INCLUDE}", f"-I{CSRC}"]
    tasks = []
    objs = []
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, src.replace("/", "_") + ".o")
        tasks.append(["g++", *CXXFLAGS, *incs, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src.replace("/", "_") + ".o")
        tasks.append([hipcc, *HIPFLAGS, *incs, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for fut in [ex.submit(_run, t, quiet) for t in tasks]:
            fut.result()
    tmp = LIB + ".tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs], quiet)
    os.replace(tmp, LIB)
    cli_objs = []
    for src in CLI_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(BUILD, src.replace("/", "_") + ".o")
        _run(["g++", *CXXFLAGS, *incs, "-c", path, "-o", obj], quiet)
        cli_objs.append(obj)
    if cli_objs:
        _run(["g++", "-o", CLI, *cli_objs, f"-L{PKG}", "-lptamd", f"-Wl,-rpath,$ORIGIN"], quiet)
    return LIB


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("--quiet", action="store_true")
    a = ap.parse_args(argv)
    path = build(a.j, quiet=a.quiet)
    print(path)
    return 0


if __name__ == "__main__":
    sys.exit(main())
