"""Failure reporting and the solve statistics, on the CPU oracle.

The reference has no NaN check (`check_nan_all` is declared in
CFD/G_SMACSolver.h:88 and never defined) and its PCG stops silently at 2000
iterations (CFD/pressure_solver/G_PCGSolver.cu:635,693). The C-ABI reports both:
SMAC_ERR_NAN with `nan_detected` set, and `converged = 0`. `bytes_model` is the
canonical algorithmic byte count of the call; the bench derives its roofline
from the same model, and the two must agree.
"""
import numpy as np
import pytest

import bench
from mydemcfd_amd import _abi, cases


def _nan_u(ctx, where):
    u = np.zeros(ctx.shape("u"))
    u[where] = np.nan
    ctx.upload("u", u)


@pytest.mark.parametrize("precond", ["jacobi", "mg", "ref_gmg"])
def test_nan_in_u_stops_the_step(oracle_mod, precond):
    c = oracle_mod.OracleContext(**cases.cavity_desc(24))  # 24: three reference-GMG levels
    _nan_u(c, (3, 4, 5))
    with pytest.raises(_abi.SmacError) as e:
        c.step(1e-3, precond=precond, max_iter=2000)
    assert e.value.status == _abi.SMAC_ERR_NAN
    st = e.value.stats
    assert st.nan_detected == 1 and st.iters == 0 and st.converged == 0
    # the correction did not run: u keeps the NaN the caller put there and p is untouched
    assert np.isnan(c.download("u")[3, 4, 5])
    assert np.all(c.download("p") == 0.0)


@pytest.mark.parametrize("stop_rule", ["rel_l2", "ref_linf_dt"])
def test_nan_rhs_of_a_pressure_solve(oracle_mod, stop_rule):
    n = 8
    c = oracle_mod.OracleContext(**cases.periodic_desc(n))
    us = np.zeros(c.shape("ustar"))
    us[2, 2, 2] = np.inf
    c.upload("ustar", us)
    with pytest.raises(_abi.SmacError) as e:
        c.pressure_solve(1e-3, precond="mg", stop_rule=stop_rule)
    assert e.value.status == _abi.SMAC_ERR_NAN
    assert e.value.stats.nan_detected == 1 and e.value.stats.iters == 0


def test_finite_step_reports_no_nan(oracle_mod):
    c = oracle_mod.OracleContext(**cases.cavity_desc(8))
    st = c.step(1e-3, precond="mg", tol=1e-10)
    assert st.converged == 1 and st.nan_detected == 0


def test_bytes_model_matches_the_bench_model(oracle_mod):
    """One periodic FP64 MG step: bytes_model = B_step(N_it) x N^3 with the
    bench's B_step = 160 + (17.43 W + 32) + N_it (104 + 17.43 W)."""
    n = 16
    c = oracle_mod.OracleContext(**cases.periodic_desc(n))
    u, v, w = cases.tg_fields(n)
    for f, a in zip("uvw", (u, v, w)):
        c.upload(f, a)
    st = c.step(1e-3, precond="mg", tol=1e-8)
    assert st.iters > 0
    assert st.bytes_model == pytest.approx(bench.step_bytes(8.0, st.iters, False) * n ** 3, rel=1e-12)
    # a pressure solve alone: the step without the predictor (56) and the correction (72)
    c.upload("ustar", u)
    sp = c.pressure_solve(1e-3, precond="mg", tol=1e-8)
    assert sp.bytes_model == pytest.approx((bench.step_bytes(8.0, sp.iters, False) - 128.0) * n ** 3, rel=1e-12)


def test_bytes_model_variable_coefficients_and_jacobi(oracle_mod):
    """Two-phase grids add 48 B fixed, 24 B per iteration and (15 x 8/7) W per
    V-cycle; the Jacobi preconditioner costs 16 B (+24) per application."""
    n = (16, 8, 8)
    desc = cases.dam_break_desc(n, rho=(1.0, 20.0), mu=(1e-3, 5e-3))
    c = oracle_mod.OracleContext(**desc)
    c.upload("alpha", cases.dam_break_alpha(n))
    c.update_properties(True)
    st = c.step(1e-3, precond="mg", tol=1e-8)
    P = 17.43 * 8.0 + 15.0 * 8.0 / 7.0 * 8.0
    per_cell = 160.0 + 48.0 + (32.0 + P) + st.iters * (104.0 + P + 24.0)
    assert st.bytes_model == pytest.approx(per_cell * np.prod(n), rel=1e-12)
    c2 = oracle_mod.OracleContext(**cases.cavity_desc(8))
    sj = c2.step(1e-3, precond="jacobi", tol=1e-8, max_iter=500)
    assert sj.bytes_model == pytest.approx((160.0 + 48.0 + sj.iters * 120.0) * 8 ** 3, rel=1e-12)
