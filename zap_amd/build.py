"""Builds libzaptex.so (the gfx950 kernels + C ABI) in-tree with hipcc.

Every translation unit is compiled with FMA contraction off, correctly rounded fp32 division and fp32
denormals kept, so the kernels reproduce the reference encoder's float operations bit for bit.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "_lib", "libzaptex.so")
PROBE = os.path.join(HERE, "_lib", "libztprobe.so")   # measurement kernels for the bench's roofline denominators
TOOL = os.path.join(HERE, "_lib", "zaptex-tex")   # the `zap tex` command on libzaptex (csrc/zaptex_tex.c)
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("ZAPTEX_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]
# zt_batch.hip includes zt_mip.hip, zt_bc14.hip and zt_bc7.hip: one translation unit for those kernels and the
# persistent batch kernel that runs their bodies
SOURCES = ["zt_batch.hip", "zt_bc7_full.hip", "zt_mlp.hip", "zt_bc6h.hip", "zt_astc.hip", "zt_decode.hip", "zt_rdo.hip", "zt_capi.cpp"]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    lang = ["-x", "hip"] if src.endswith(".cpp") else []
    cmd = [HIPCC, *FLAGS, *lang, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs: int | None = None) -> str:
    """Compile every source and link zap_amd/_lib/libzaptex.so; returns its path."""
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_tool()
    build_probe()
    return LIB


def build_probe() -> str:
    cmd = [HIPCC, *FLAGS, "-shared", os.path.join(CSRC, "zt_probe.hip"), "-o", PROBE + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"libztprobe build failed:\n{r.stdout}\n{r.stderr}")
    os.replace(PROBE + ".tmp", PROBE)
    return PROBE


def build_tool() -> str:
    """zap_amd/_lib/zaptex-tex: plain C11 against include/zaptex.h, linked to the library beside it."""
    cmd = ["gcc", "-std=c11", "-O2", "-Wall", "-Wextra", "-Werror", "-I", INCLUDE, os.path.join(CSRC, "zaptex_tex.c"),
           "-L", os.path.dirname(LIB), "-lzaptex", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib", "-o", TOOL + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"zaptex-tex build failed:\n{r.stdout}\n{r.stderr}")
    os.replace(TOOL + ".tmp", TOOL)
    return TOOL


if __name__ == "__main__":
    print(build(int(sys.argv[1]) if len(sys.argv) > 1 else None))
