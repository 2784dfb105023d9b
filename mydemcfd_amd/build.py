"""Compile the gfx950 HIP library mydemcfd_amd/lib/libsmac_hip.so with hipcc.

Flags:
  -ffp-contract=off   no FMA contraction (FP64 parity with the CPU restatement;
                      the reference builds with -fmad=false, makefile_CFD:5)
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed FP32 instructions in generated code

The translation units compile in parallel (one hipcc process each, at most
SMAC_BUILD_JOBS, default 8), then link into one shared library.
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
OUT = os.path.join(LIBDIR, "libsmac_hip.so")
SOURCES = ["smac_kernels.hip", "smac_mg.hip", "smac_refgmg.hip", "smac_adjoint.hip", "smac_adjoint_vof.hip",
           "smac_post.hip", "smac_vof.hip", "smac_capi.hip"]
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


def _compile(src: str, verbose: bool) -> str:
    obj = os.path.join(LIBDIR, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc()] + flags() + ["-c", os.path.join(CSRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    out = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass does not know the packed-fp32 feature; that warning is expected
    noise = "'-packed-fp32-ops' is not a recognized feature"
    msg = "\n".join(ln for ln in (out.stdout + out.stderr).splitlines() if noise not in ln)
    if msg.strip():
        print(msg, file=sys.stderr, flush=True)
    if out.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src} (status {out.returncode})")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    jobs = max(1, int(os.environ.get("SMAC_BUILD_JOBS", "8")))
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(jobs, len(SOURCES))) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), SOURCES))
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC", "-Wl,-Bsymbolic", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
