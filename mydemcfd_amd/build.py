"""Compile the gfx950 HIP library mydemcfd_amd/lib/libsmac_hip.so with hipcc.

Flags:
  -ffp-contract=off   no FMA contraction (FP64 parity with the CPU restatement;
                      the reference builds with -fmad=false, makefile_CFD:5)
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed FP32 instructions in generated code
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
OUT = os.path.join(LIBDIR, "libsmac_hip.so")
SOURCES = ["smac_kernels.hip", "smac_mg.hip", "smac_refgmg.hip", "smac_adjoint.hip", "smac_post.hip", "smac_vof.hip", "smac_capi.hip"]
ARCH = os.environ.get("SMAC_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def flags():
    return ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument"]


def build(verbose: bool = False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    objs = []
    for src in SOURCES:
        obj = os.path.join(LIBDIR, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc()] + flags() + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC", "-Wl,-Bsymbolic", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    build_facade_driver(verbose)
    return OUT


ROOT = os.path.dirname(HERE)
DRIVER_SRC = os.path.join(ROOT, "tests", "cpp", "cavity_facade.cpp")
DRIVER = os.path.join(ROOT, "tests", "cpp", "build", "cavity_facade")
CONFIG_DRIVER_SRC = os.path.join(ROOT, "tests", "cpp", "config_driver.cpp")
CONFIG_DRIVER = os.path.join(ROOT, "tests", "cpp", "build", "config_driver")
CONFIG_TOOL_SRC = os.path.join(ROOT, "tests", "cpp", "config_tool.cpp")
CONFIG_TOOL = os.path.join(ROOT, "tests", "cpp", "build", "config_tool")


def build_facade_driver(verbose: bool = False) -> str:
    """The C++ drivers of tests/cpp: reference-shaped loops over the C++ facade
    (linked to the library) and the host-only configuration tool."""
    if not os.path.exists(DRIVER_SRC):
        return ""
    os.makedirs(os.path.dirname(DRIVER), exist_ok=True)
    inc = ["-I", os.path.join(ROOT, "include")]
    link = ["-L", LIBDIR, "-lsmac_hip", "-Wl,-rpath," + LIBDIR]
    cmds = [["g++", "-O2", "-std=c++17"] + inc + [DRIVER_SRC, "-o", DRIVER] + link,
            ["g++", "-O2", "-std=c++17"] + inc + [CONFIG_DRIVER_SRC, "-o", CONFIG_DRIVER] + link,
            ["g++", "-O2", "-std=c++17"] + inc + [CONFIG_TOOL_SRC, "-o", CONFIG_TOOL]]
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
    return DRIVER


def build_config_tool() -> str:
    """The host-only configuration tool alone (no HIP library needed)."""
    os.makedirs(os.path.dirname(CONFIG_TOOL), exist_ok=True)
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), CONFIG_TOOL_SRC, "-o",
                    CONFIG_TOOL], check=True)
    return CONFIG_TOOL


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
