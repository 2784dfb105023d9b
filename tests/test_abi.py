"""The C-ABI boundary (no GPU needed): every entry point declared in
include/smac_c_api.h is exported by the HIP library and by the CPU oracle, and
the Python ctypes table mirrors the header."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "smac_c_api.h")


def header_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(smac_[a-z_]+)\s*\(", src)))


def test_header_parses():
    names = header_functions()
    assert "smac_create" in names and "smac_pressure_solve" in names and len(names) >= 20


def test_ctypes_table_matches_header():
    from mydemcfd_amd import _abi

    assert sorted(_abi.SIGNATURES) == header_functions()


def _exports(path):
    lib = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
    missing = [n for n in header_functions() if not hasattr(lib, n)]
    return lib, missing


def test_hip_library_exports_every_symbol():
    import mydemcfd_amd

    if not os.path.exists(mydemcfd_amd.LIB_PATH):
        pytest.skip("HIP library not built (run __graft_entry__.build())")
    _, missing = _exports(mydemcfd_amd.LIB_PATH)
    assert not missing, missing


def test_oracle_exports_every_symbol(oracle_mod):
    _, missing = _exports(oracle_mod.LIB)
    assert not missing, missing


def test_hip_library_without_device():
    """On a host without a GPU the library loads, reports its backend and
    refuses to create a context with SMAC_ERR_NO_DEVICE (no compute call)."""
    import mydemcfd_amd
    from mydemcfd_amd import _abi, cases

    if not os.path.exists(mydemcfd_amd.LIB_PATH):
        pytest.skip("HIP library not built")
    lib = mydemcfd_amd.load()
    assert lib.smac_backend_name() == b"hip-gfx950"
    assert lib.smac_strerror(_abi.SMAC_ERR_NO_DEVICE) == b"no device"
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(_abi.SmacError) as e:
        mydemcfd_amd.SMACContext(**cases.cavity_desc(8))
    assert e.value.status == _abi.SMAC_ERR_NO_DEVICE


def test_product_has_no_cpu_fallback():
    """The product loader only loads the HIP library and fails loudly when it is
    missing; it never imports the oracle."""
    import mydemcfd_amd

    src = open(os.path.join(ROOT, "mydemcfd_amd", "__init__.py")).read()
    assert "oracle" not in src
    for fn in os.listdir(os.path.join(ROOT, "mydemcfd_amd")):
        if fn.endswith(".py"):
            assert "import oracle" not in open(os.path.join(ROOT, "mydemcfd_amd", fn)).read()
            assert "from oracle" not in open(os.path.join(ROOT, "mydemcfd_amd", fn)).read()
    saved = mydemcfd_amd._lib, mydemcfd_amd.LIB_PATH
    try:
        mydemcfd_amd._lib = None
        mydemcfd_amd.LIB_PATH = os.path.join(ROOT, "does-not-exist.so")
        with pytest.raises(RuntimeError):
            mydemcfd_amd.load()
    finally:
        mydemcfd_amd._lib, mydemcfd_amd.LIB_PATH = saved
