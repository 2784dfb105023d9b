import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


def rel_err(a, b):
    """max|a-b| / max|b| (the parity metric of the FP64 form)."""
    den = float(np.max(np.abs(b)))
    return float(np.max(np.abs(a - b))) / (den if den > 0 else 1.0)


def rel_err_mod_mean(a, b, n):
    """Same, for a cell field compared modulo its interior mean."""
    ai = a[1:n[2] + 1, 1:n[1] + 1, 1:n[0] + 1]
    bi = b[1:n[2] + 1, 1:n[1] + 1, 1:n[0] + 1]
    return rel_err(ai - ai.mean(), bi - bi.mean())


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle

    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def hip():
    import mydemcfd_amd

    lib = mydemcfd_amd.load()
    return mydemcfd_amd
