"""Loader for the CPU FP64 oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this.
The oracle library exports the same C-ABI as the HIP product, plus a few
smac_oracle_* hooks for operator property tests.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from mydemcfd_amd import _abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "build", "libsmac_oracle.so")

_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    if force or not os.path.exists(LIB):
        subprocess.run(["make", "-C", HERE, "-B" if force else "all"], check=True,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    return LIB


def _bind_hooks(l):
    _abi.bind(l)
    l.smac_oracle_apply_A.restype = C.c_int
    l.smac_oracle_apply_A.argtypes = [C.c_void_p, _abi._D, _abi._D]
    l.smac_oracle_apply_mg.restype = C.c_int
    l.smac_oracle_apply_mg.argtypes = [C.c_void_p, C.POINTER(_abi.SolverOpts), _abi._D, _abi._D]
    l.smac_oracle_csum.restype = C.c_double
    l.smac_oracle_csum.argtypes = [C.c_void_p, _abi._D]
    l.smac_oracle_rg_beta.restype = C.c_int
    l.smac_oracle_rg_beta.argtypes = [C.c_void_p, C.POINTER(_abi.SolverOpts), C.c_int, C.c_int,
                                      C.POINTER(C.c_int * 3), _abi._D, C.c_size_t]
    l.smac_oracle_vofmath.restype = C.c_int
    l.smac_oracle_vofmath.argtypes = [C.c_int, _abi._D, _abi._D, C.c_size_t]
    l.smac_oracle_threads.restype = C.c_int
    l.smac_oracle_threads.argtypes = []
    return l


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = _bind_hooks(C.CDLL(LIB, mode=C.RTLD_LOCAL))
    return _lib


NATIVE_FLAGS = ["-O3", "-march=native", "-std=c11", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fno-fast-math"]


def native_lib(outdir: str):
    """The oracle built for the host it runs on (-O3 -march=native, no FMA
    contraction) into outdir: the CPU baseline's build. It is compiled where it
    runs, since -march=native code need not run on another host."""
    out = os.path.join(outdir, "libsmac_oracle_native.so")
    subprocess.run(["gcc"] + NATIVE_FLAGS + ["-shared", "-Wl,-Bsymbolic", "-o", out,
                                             os.path.join(HERE, "smac_oracle.c"), "-lm"], check=True)
    return _bind_hooks(C.CDLL(out, mode=C.RTLD_LOCAL))


def mutant_lib(mutant: int, outdir: str):
    """The oracle built with planted mistake `mutant` (SMAC_ORACLE_MUTANT in
    smac_oracle.c) into outdir, loaded and bound. Only the tests that show the
    closed-form pins catch such mistakes use it."""
    out = os.path.join(outdir, f"libsmac_oracle_mutant{int(mutant)}.so")
    if not os.path.exists(out):
        subprocess.run(["gcc", "-O1", "-std=c11", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fno-fast-math",
                        f"-DSMAC_ORACLE_MUTANT={int(mutant)}", "-shared", "-Wl,-Bsymbolic", "-o", out,
                        os.path.join(HERE, "smac_oracle.c"), "-lm"], check=True)
    return _bind_hooks(C.CDLL(out, mode=C.RTLD_LOCAL))


class OracleContext(_abi.Context):
    """A smac_ctx of the CPU oracle (`lib_`: another build of it, e.g. a
    mutant_lib)."""

    def __init__(self, *a, lib_=None, **kw):
        super().__init__(lib_ if lib_ is not None else lib(), *a, **kw)

    def apply_A(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        self._check(self.lib.smac_oracle_apply_A(self.h, x.ctypes.data_as(_abi._D), y.ctypes.data_as(_abi._D)),
                    "apply_A")
        return y

    def apply_mg(self, r, **kw):
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.zeros_like(r)
        o = self.opts(**kw)
        self._check(self.lib.smac_oracle_apply_mg(self.h, C.byref(o), r.ctypes.data_as(_abi._D),
                                                  z.ctypes.data_as(_abi._D)), "apply_mg")
        return z

    def rg_beta(self, level, axis, **kw):
        """Face coefficients of one axis on a level of the reference GMG, as a
        (n2+3, n1+3, n0+3) array (index = the reference's face/cell index)."""
        o = self.opts(precond="ref_gmg", **kw)
        dims = (C.c_int * 3)()
        L = self.lib.smac_oracle_rg_beta(self.h, C.byref(o), level, axis, C.byref(dims), None, 0)
        if L < 0:
            self._check(L, "rg_beta")
        n = [dims[0], dims[1], dims[2]]
        out = np.zeros((n[2] + 3, n[1] + 3, n[0] + 3))
        rc = self.lib.smac_oracle_rg_beta(self.h, C.byref(o), level, axis, C.byref(dims),
                                          out.ctypes.data_as(_abi._D), out.size)
        if rc < 0:
            self._check(rc, "rg_beta")
        return L, tuple(n), out

    def csum(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        return self.lib.smac_oracle_csum(self.h, x.ctypes.data_as(_abi._D))


def threads(l=None) -> int:
    return (l or lib()).smac_oracle_threads()


def vofmath(fn, x):
    """The VoF transport's exp (fn 0), log (1), cosh (2) and acos on [0, 1] (3)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    lib().smac_oracle_vofmath(int(fn), x.ctypes.data_as(_abi._D), y.ctypes.data_as(_abi._D), x.size)
    return y
