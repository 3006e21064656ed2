"""Builds the HIP library ``openmps_amd/libmps_hip.so`` for gfx950 in-tree.

Every translation unit is compiled with ``hipcc --offload-arch=gfx950``:
* ``-ffp-contract=off``: no FMA contraction, so FP64 expressions round as the
  reference's host code does (bit-exact neighbour sums);
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32 ops.
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(CSRC, "_build")
LIB = os.path.join(HERE, "libmps_hip.so")
SOURCES = ["mps_api.cpp", "mps_batch.cpp", "mps_search.hip", "mps_particles.hip", "mps_ppe.hip", "mps_cg.hip",
           "mps_mlp.hip"]
ARCH = os.environ.get("MPS_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def flags() -> list[str]:
    return [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wno-unused-command-line-argument", "-I", os.path.join(HERE, "..", "include")]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj] + flags()
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    """Compiles every source and links the shared library; returns its path."""
    os.makedirs(BUILD, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def build_facade_driver(verbose: bool = False) -> str:
    """Compiles tests/cpp/facade_main.cpp (the C++ Computer facade driven like
    the reference's Main.cpp) against the library with the host compiler."""
    root = os.path.dirname(HERE)
    src = os.path.join(root, "tests", "cpp", "facade_main.cpp")
    out = os.path.join(root, "tests", "cpp", "facade_main")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-I", os.path.join(root, "include"), src,
           "-L", HERE, "-lmps_hip", "-Wl,-rpath,$ORIGIN/../../openmps_amd", "-o", out + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"facade driver build failed:\n{r.stdout}\n{r.stderr}")
    os.replace(out + ".tmp", out)
    if verbose:
        print(f"built {out}")
    return out


if __name__ == "__main__":
    build(verbose=True)
    build_facade_driver(verbose=True)
