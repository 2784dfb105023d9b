"""GPU (HIP, gfx950) checks of the wall, slip and inflow boundaries.

* The closed-form pins of tests/test_oracle_boundary_pins.py (Couette flow
  between moving walls, slip channels, inflow plug flow, CFL closed forms) run
  on the GPU through the C-ABI.
* A box whose six sides carry every boundary kind the build accepts (inflow
  walls with a tangential velocity, a moving no-slip wall, slip walls, a wall
  at rest) steps with each preconditioner, bit for bit against the oracle.
  The inflow sides make the Poisson problem one that is not pure Neumann, so
  this runs the PCG branch without mean removal (G_PCGSolver.cu:448-586,
  619-770).
"""
import numpy as np
import pytest

from conftest import rel_err, rel_err_mod_mean
import test_oracle_boundary_pins as pins

from mydemcfd_amd import _abi

pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.mark.parametrize("a,b", pins.PAIRS)
def test_gpu_couette_and_slip_pins(hip, a, b):
    print(f"case: couette/slip pins comp {a} walls {b}", flush=True)
    pins.pin_couette(hip.SMACContext, a, b)
    pins.pin_slip_uniform(hip.SMACContext, a, b)
    pins.pin_slip_profile(hip.SMACContext, a, b)


@pytest.mark.parametrize("a", range(3))
def test_gpu_inflow_plug_pin(hip, a):
    print(f"case: inflow plug axis {a}", flush=True)
    pins.pin_inflow_plug(hip.SMACContext, a)


def test_gpu_cfl_pins(hip):
    print("case: cfl pins", flush=True)
    pins.pin_cfl_uniform(hip.SMACContext)
    pins.pin_cfl_spike(hip.SMACContext)
    pins.pin_cfl_two_phase(hip.SMACContext)


def mixed_bc_desc(n=(24, 16, 20), two_phase=False):
    """Every side a different boundary: inflow at both x sides (normal velocity
    0.5, tangential wall velocities differing between the sides), slip at
    y-min, a no-slip wall moving in x and z at y-max, no-slip at rest at
    z-min, slip at z-max."""
    table = ((_abi.BC_NOSLIP, 0.0, 0.0, 0.0),
             (_abi.BC_INFLOW, 0.5, 0.2, -0.1),
             (_abi.BC_INFLOW, 0.5, 0.1, 0.3),
             (_abi.BC_SLIP, 0.0, 0.0, 0.0),
             (_abi.BC_NOSLIP, 1.0, 0.0, 0.5),
             (_abi.BC_NOSLIP, 0.0, 0.0, 0.0))
    rho, mu, g = ((1.0, 1.0), (0.02, 0.02), (0.0, 0.0, 0.0))
    if two_phase:
        rho, mu, g = ((1.0, 50.0), (1e-3, 2e-2), (0.0, 0.0, -1.0))
    return dict(n=n, size=(1.2, 0.8, 1.0), origin=(0.0, 0.0, 0.0), periodic=(0, 0, 0),
                side_bcid=(1, 2, 3, 4, 5, 3), bc_table=table, rho=rho, mu=mu, g=g)


def mixed_bc_fields(ctx, seed=11):
    """Smooth random velocities with the normal boundary faces set to their
    boundary values (inflow 0.5 on the x walls, 0 on the others)."""
    rng = np.random.default_rng(seed)
    out = {}
    for a, name in enumerate(("u", "v", "w")):
        shape = ctx.shape(name)
        zz, yy, xx = np.meshgrid(*[np.linspace(0.0, 1.0, s) for s in shape], indexing="ij")
        f = np.zeros(shape)
        for _ in range(4):
            k = rng.integers(1, 4, size=3)
            f += 0.3 * rng.standard_normal() * np.sin(np.pi * (k[0] * xx + 0.3)) * np.cos(np.pi * k[1] * yy) * \
                np.sin(np.pi * (k[2] * zz + 0.1))
        if a == 0:
            f += 0.5
        ax = 2 - a  # numpy axis of the component's own direction
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[ax], hi[ax] = 1, shape[ax] - 2
        f[tuple(lo)] = 0.5 if a == 0 else 0.0
        f[tuple(hi)] = 0.5 if a == 0 else 0.0
        out[name] = f
    return out


@pytest.mark.parametrize("precond", ["jacobi", "mg", "ref_gmg"])
@pytest.mark.parametrize("two_phase", [False, True])
def test_mixed_boundaries_vs_oracle(hip, oracle_mod, precond, two_phase):
    from mydemcfd_amd import cases

    n = (24, 16, 20)
    print(f"case: mixed boundaries {n} two_phase={two_phase} {precond} x4", flush=True)
    desc = mixed_bc_desc(n, two_phase)
    g = hip.SMACContext(**desc)
    o = oracle_mod.OracleContext(**desc)
    fields = mixed_bc_fields(o)
    if two_phase:
        al = np.zeros(o.shape("alpha"))
        al[: n[2] // 2 + 1, :, : n[0] // 3 + 1] = 1.0
        fields["alpha"] = al
    for name, arr in fields.items():
        g.upload(name, arr)
        o.upload(name, arr)
    kw = dict(precond=precond, tol=1e-9, max_iter=3000)
    if precond == "ref_gmg":
        kw.update(stop_rule="ref_linf_dt", check_every=5, tol=1e-10)
    dt = 2e-3
    for s in range(4):
        sg = g.step(dt, **kw)
        so = o.step(dt, **kw)
        assert so.pure_neumann == 0 and sg.pure_neumann == 0
        assert sg.iters == so.iters and sg.converged == so.converged, (s, sg.as_dict(), so.as_dict())
        assert so.converged, so.as_dict()
    for f in ("u", "v", "w"):
        assert rel_err(g.download(f), o.download(f)) <= TOL, f
    assert rel_err_mod_mean(g.download("p"), o.download("p"), n) <= TOL
    assert g.calc_cfl(dt) == o.calc_cfl(dt)
    for c in (g, o):
        c.update_boundary_ghost(dt)
    for f in ("u", "v", "w", "p"):
        assert float(np.max(np.abs(g.download(f) - o.download(f)))) <= TOL * max(1.0, float(
            np.max(np.abs(o.download(f))))), ("ghosts", f)
    del cases
