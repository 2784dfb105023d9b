"""Builds the C++ drivers of tests/cpp (test infrastructure, not the product):
reference-shaped loops over the C++ facade linked to the HIP library, and the
host-only configuration tool. Each is rebuilt when missing or older than its
sources."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
INCLUDE = os.path.join(ROOT, "include")
LIBDIR = os.path.join(ROOT, "mydemcfd_amd", "lib")
BUILD = os.path.join(HERE, "build")

DRIVER_SRC = os.path.join(HERE, "cavity_facade.cpp")
DRIVER = os.path.join(BUILD, "cavity_facade")
CONFIG_DRIVER_SRC = os.path.join(HERE, "config_driver.cpp")
CONFIG_DRIVER = os.path.join(BUILD, "config_driver")
CONFIG_TOOL_SRC = os.path.join(HERE, "config_tool.cpp")
CONFIG_TOOL = os.path.join(BUILD, "config_tool")

_HEADERS = [os.path.join(INCLUDE, h) for h in ("smac_c_api.h", "smac_facade.hpp", "smac_config.hpp")]


def _stale(out: str, srcs) -> bool:
    if not os.path.exists(out):
        return True
    t = os.path.getmtime(out)
    return any(os.path.exists(s) and os.path.getmtime(s) > t for s in srcs)


def _build(src: str, out: str, link_lib: bool) -> str:
    if not _stale(out, [src] + _HEADERS):
        return out
    os.makedirs(BUILD, exist_ok=True)
    cmd = ["g++", "-O2", "-std=c++17", "-I", INCLUDE, src, "-o", out]
    if link_lib:
        cmd += ["-L", LIBDIR, "-lsmac_hip", "-Wl,-rpath," + LIBDIR]
    subprocess.run(cmd, check=True)
    return out


def facade_driver() -> str:
    """The cavity loop of CFD/main.cu over the C++ facade (needs the HIP library)."""
    return _build(DRIVER_SRC, DRIVER, True)


def config_driver() -> str:
    """The config-driven loop (needs the HIP library)."""
    return _build(CONFIG_DRIVER_SRC, CONFIG_DRIVER, True)


def config_tool() -> str:
    """The host-only configuration tool (no HIP library)."""
    return _build(CONFIG_TOOL_SRC, CONFIG_TOOL, False)


def build_all() -> None:
    if os.path.exists(os.path.join(LIBDIR, "libsmac_hip.so")):
        facade_driver()
        config_driver()
    config_tool()


if __name__ == "__main__":
    build_all()
