"""GPU: the four-rows-per-wave smoother (k_mg_rbgs_v2<T, VB, CPL, 4>, taken on
levels whose rows are 32 16-byte vectors long -- 64 cells in FP64, 128 in FP32,
256 in BF16 -- with ny % 16 == 0).

* FP64: steps bit-exact against the oracle on a two-phase 64x16x8 box
  (variable face coefficients on level 0) and a periodic 64x32x16 box.
* One V-cycle (the stationary iteration stopped after one iteration: dp = M b)
  in FP64 against the oracle, and in FP32 / BF16 against the GPU FP64 V-cycle.
  The mixed forms cannot be replayed on the CPU (their smoother divides by a
  hardware reciprocal), but a wrong neighbour, colour or lane edge in one
  sweep changes M b by O(1) at the cells it touches, where FP32 and BF16
  storage change it by ~1e-6 and ~1e-2 of its maximum.
"""
import numpy as np
import pytest

from conftest import rel_err, rel_err_mod_mean
from mydemcfd_amd import cases

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _desc(kind, n):
    if kind == "periodic":
        return cases.periodic_desc(n, length=1.0)
    length = (n[0] / 32.0, n[1] / 32.0, n[2] / 32.0)
    return cases.dam_break_desc(n, length=length, rho=(1.0, 20.0), mu=(1e-3, 5e-3))


def _alpha(n):
    length = (n[0] / 32.0, n[1] / 32.0, n[2] / 32.0)
    return cases.dam_break_alpha(n, length=length, x_front=0.3 * length[0], height=0.5 * length[2])


def _setup(ctx, kind, n, seed):
    rng = np.random.default_rng(seed)
    if kind == "two_phase":
        ctx.upload("alpha", _alpha(n))
        ctx.update_properties(True)
    for f in ("u", "v", "w"):
        ctx.upload(f, 0.1 * rng.standard_normal(ctx.shape(f)))


@pytest.mark.parametrize("kind,n", [("two_phase", (64, 16, 8)), ("periodic", (64, 32, 16))])
def test_rpw4_fp64_steps_match_oracle(hip, oracle_mod, kind, n):
    print(f"case: RPW=4 FP64 steps {kind} {n}", flush=True)
    desc = _desc(kind, n)
    g, o = hip.SMACContext(**desc), oracle_mod.OracleContext(**desc)
    for c in (g, o):
        _setup(c, kind, n, 0x5EED4)
    kw = dict(precond="mg", tol=1e-9, max_iter=500)
    for s in range(2):
        sg, so = g.step(1e-3, **kw), o.step(1e-3, **kw)
        assert sg.iters == so.iters and sg.converged == so.converged == 1, (s, sg.as_dict(), so.as_dict())
    for f in ("u", "v", "w"):
        assert rel_err(g.download(f), o.download(f)) <= TOL, f
    assert rel_err_mod_mean(g.download("p"), o.download("p"), n) <= TOL


def _one_vcycle(ctx, kind, n, precision):
    """dp = M b for the RHS of random u*: the stationary iteration, one step."""
    rng = np.random.default_rng(0x5EED5)
    if kind == "two_phase":
        ctx.upload("alpha", _alpha(n))
        ctx.update_properties(True)
    for f in ("ustar", "vstar", "wstar"):
        ctx.upload(f, rng.standard_normal(ctx.shape(f)))
    kw = dict(precond="mg", precision=precision, stationary=1, tol=0.0, max_iter=1)
    if precision != "fp64":
        kw["flexible"] = 1
    st = ctx.pressure_solve(1e-3, **kw)
    assert st.iters == 1
    return ctx.download("dp")


# (kind, n, precision): the levels that take the RPW=4 form
#   periodic 256x32x16: FP32 level 1 (128x16x8), BF16 level 0
#   two-phase 128x16x8: FP32 level 0 with face coefficients
#   two-phase 256x16x8: BF16 level 0 with face coefficients
#   two-phase 64x16x8 and periodic 64x32x16: FP64 level 0
MIXED = [("periodic", (256, 32, 16), "fp32", 5e-6), ("periodic", (256, 32, 16), "bf16", 5e-2),
         ("two_phase", (128, 16, 8), "fp32", 5e-6), ("two_phase", (256, 16, 8), "bf16", 5e-2)]


@pytest.mark.parametrize("kind,n", [("two_phase", (64, 16, 8)), ("periodic", (64, 32, 16)),
                                    ("periodic", (256, 32, 16)), ("two_phase", (128, 16, 8))])
def test_rpw4_fp64_vcycle_matches_oracle(hip, oracle_mod, kind, n):
    print(f"case: one FP64 V-cycle {kind} {n}", flush=True)
    desc = _desc(kind, n)
    dg = _one_vcycle(hip.SMACContext(**desc), kind, n, "fp64")
    do = _one_vcycle(oracle_mod.OracleContext(**desc), kind, n, "fp64")
    assert rel_err_mod_mean(dg, do, n) <= TOL


@pytest.mark.parametrize("kind,n,precision,tol", MIXED)
def test_rpw4_mixed_vcycle_vs_fp64(hip, kind, n, precision, tol):
    print(f"case: one {precision} V-cycle {kind} {n}", flush=True)
    desc = _desc(kind, n)
    d64 = _one_vcycle(hip.SMACContext(**desc), kind, n, "fp64")
    dmx = _one_vcycle(hip.SMACContext(**desc), kind, n, precision)
    err = rel_err_mod_mean(dmx, d64, n)
    print(f"  max|M_{precision} b - M_fp64 b| / max|M_fp64 b| = {err:.3e}", flush=True)
    assert err <= tol
