"""MI355X (gfx950) SMAC projection step.

`load()` returns the HIP C-ABI library (mydemcfd_amd/lib/libsmac_hip.so) and raises
if it is missing or no device is present; there is no CPU fallback.
`SMACContext` is a context on that library.
"""
from __future__ import annotations

import ctypes as _C
import os as _os

from . import _abi
from ._abi import Context as _Context, SmacError  # noqa: F401

HERE = _os.path.dirname(_os.path.abspath(__file__))
LIB_PATH = _os.path.join(HERE, "lib", "libsmac_hip.so")

_lib = None


def load():
    """Load the HIP library. Raises RuntimeError when it has not been built.

    When PyTorch is installed it is imported first: its HIP runtime is then the
    one the library binds to (same soname), so one process holds a single HIP
    runtime and torch tensors and library calls share the device. Loading the
    library first would bring in a second runtime, after which torch cannot
    initialize the GPU."""
    global _lib
    if _lib is None:
        if not _os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        _lib = _abi.bind(_C.CDLL(LIB_PATH, mode=_C.RTLD_LOCAL))
    return _lib


class SMACContext(_Context):
    """A smac_ctx on the HIP library (device arrays live on the current GPU)."""

    def __init__(self, *a, **kw):
        super().__init__(load(), *a, **kw)
