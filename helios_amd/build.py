"""Build the native libraries in-tree.

* ``helios_amd/libhelios_rt.so``   -- gfx950 HIP kernels + the C ABI of include/helios_rt.h
* ``helios_amd/libhelios_scenes.so`` -- seeded synthetic scene generators (plain C, host only)
* ``helios_amd/libhelios_host.so`` -- C++ RadiationModel-style interface over the C ABI (host only)

Every HIP translation unit is built with ``-ffp-contract=off`` (the FP32 ray arithmetic must not be
contracted into FMAs, so rays are reproducible bit for bit) and with packed-FP32 instruction
selection disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_SOURCES = ["rt_scene.hip", "rt_trace.hip", "rt_solve.hip", "rt_api.hip", "rt_server.hip", "rt_backend.hip",
               "rt_autotune.hip", "rt_grid.hip"]
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]

RT_LIB = os.path.join(HERE, "libhelios_rt.so")
SCENE_LIB = os.path.join(HERE, "libhelios_scenes.so")
HOST_LIB = os.path.join(HERE, "libhelios_host.so")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd[:3]) + " ...")
    return r.stdout


def build_rt(verbose: bool = False) -> str:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    srcs = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]

    def one(src):
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        _run([HIPCC, *HIP_FLAGS, "-I", CSRC, "-I", os.path.join(HERE, "..", "include"), "-c",
              os.path.join(CSRC, src), "-o", obj])
        return obj

    with ThreadPoolExecutor(max_workers=min(6, len(srcs))) as ex:
        objs = list(ex.map(one, srcs))
    tmp = RT_LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, RT_LIB)
    if verbose:
        print("built", RT_LIB)
    return RT_LIB


def build_scenes(verbose: bool = False) -> str:
    tmp = SCENE_LIB + ".tmp"
    _run(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
          os.path.join(CSRC, "scene_gen.c"), "-o", tmp, "-lm"])
    os.replace(tmp, SCENE_LIB)
    if verbose:
        print("built", SCENE_LIB)
    return SCENE_LIB


def build_host(verbose: bool = False) -> str:
    tmp = HOST_LIB + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", os.path.join(HERE, "host", "radiation_model.cpp"),
          "-o", tmp, "-L" + HERE, "-lhelios_rt", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, HOST_LIB)
    if verbose:
        print("built", HOST_LIB)
    return HOST_LIB


def build_all(verbose: bool = False) -> None:
    build_scenes(verbose)
    build_rt(verbose)
    build_host(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
