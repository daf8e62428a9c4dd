"""Build recipe of the MI355X (gfx950) device library, its CLI and the oracle.

    python -m dg_hdg_euler_cpp_amd.build        # or __graft_entry__.build()

Outputs (all in-tree, git-ignored):
    dg_hdg_euler_cpp_amd/libdgh_device.so   HIP kernels + C ABI (include/dgh/dgh_device.h)
    dg_hdg_euler_cpp_amd/bin/dg_hdg_2d       reference-compatible CLI on the device path
    dg_hdg_euler_cpp_amd/bin/dg_hdg_1d       the 1D CLI (method dg) on the device path
    oracle/libdgh_oracle.so                 CPU FP64 oracle (test infrastructure)
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libdgh_device.so")
CLI = os.path.join(PKG, "bin", "dg_hdg_2d")
CLI1D = os.path.join(PKG, "bin", "dg_hdg_1d")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
ARCH = os.environ.get("DGH_OFFLOAD_ARCH", "gfx950")
# Packed FP32 instructions are disabled for every HIP translation unit.
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-Wno-unused-command-line-argument"]
SOURCES = ["residual.hip", "linalg.hip", "hdg.hip", "ilu.hip", "api.hip", "config.cpp", "basis.cpp", "oned.hip"]


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build_device(jobs: int = 8) -> str:
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    os.makedirs(os.path.dirname(CLI), exist_ok=True)
    tmp = CLI + ".tmp"
    _run([HIPCC, "-O2", "-std=c++17", "-o", tmp, os.path.join(PKG, "cli", "main.cpp"), "-L" + PKG, "-ldgh_device",
          "-Wl,-rpath,$ORIGIN/.."])
    os.replace(tmp, CLI)
    tmp = CLI1D + ".tmp"
    _run([HIPCC, "-O2", "-std=c++17", "-o", tmp, os.path.join(PKG, "cli", "main1d.cpp"), "-L" + PKG, "-ldgh_device",
          "-Wl,-rpath,$ORIGIN/.."])
    os.replace(tmp, CLI1D)
    return LIB


def build_oracle() -> str:
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    return os.path.join(ROOT, "oracle", "libdgh_oracle.so")


def main():
    lib = build_device()
    print("built", lib)
    print("built", build_oracle())


if __name__ == "__main__":
    sys.exit(main())
