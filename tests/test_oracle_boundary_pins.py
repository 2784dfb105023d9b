"""Closed-form pins of the oracle's wall getters and CFL numbers (no GPU).

The reference's tangential wall mirror (d_get_v*_*dir, CFD/G_BoundaryFunctions.h:100-367:
2*v_wall - v_in for no-slip and inflow walls, the mean of the wall velocities of
the two wall faces beside the mirrored face, v_in for slip and outlet walls),
its boundary face values (d_get_v*_*face, :9-97: inflow (1 - lambda)*v_bc with
lambda = 0 here) and its CFL numbers (k_calc_cfl, CFD/G_SMACSolver.cu:1365-1407)
are pinned here by flows whose discrete answer is known without the oracle's
formulas:

* plane Couette flow between two no-slip walls moving at different velocities,
  u_a = U_lo + (U_hi - U_lo) y_b / L: the mirror extends the linear profile
  exactly, so stress and advection vanish and u* = u, and a whole step keeps u
  (every wall axis b and tangential component a);
* a slip channel: a uniform tangential flow is steady (u* = u); a profile u(y)
  gets u* = u + dt nu D2_N u with D2_N the textbook 1-D Neumann second
  difference;
* plug flow between two inflow walls carrying the same velocity: u* = u on
  every face, and the step keeps it (this also runs the PCG's branch for grids
  that are not pure Neumann);
* the CFL of a uniform flow, dt sum_d |U_d|/h_d and (mu/rho) dt sum_d 1/h_d^2,
  of a single-face spike, and the viscous CFL of a two-phase box.

`test_pins_catch_planted_mistakes` builds the oracle with each planted mistake
of smac_oracle.c (SMAC_ORACLE_MUTANT) and checks that its pin fails.
"""
import itertools
import math

import numpy as np
import pytest

from mydemcfd_amd import _abi

COMP = ("u", "v", "w")
FIELDS = (_abi.FIELD_U, _abi.FIELD_V, _abi.FIELD_W)


def _desc(n, size, periodic, side_bcid, bc_table, rho=1.0, mu=0.1):
    return dict(n=n, size=size, origin=(0.0, 0.0, 0.0), periodic=periodic, side_bcid=side_bcid,
                bc_table=bc_table, rho=(rho, rho), mu=(mu, mu), g=(0.0, 0.0, 0.0))


def _coords(ctx, comp, axis):
    """Coordinate along `axis` of every index of a component-`comp` face array:
    face index f at (f-1)h on its own axis, cell index c at (c-1/2)h elsewhere."""
    nn = ctx.shape(FIELDS[comp])[::-1][axis]
    h = ctx.size[axis] / ctx.n[axis]
    idx = np.arange(nn, dtype=np.float64)
    return (idx - 1.0) * h if axis == comp else (idx - 0.5) * h


def _field(ctx, comp, fn_of_axis, axis):
    """Face array of component comp whose value depends on one axis only."""
    shape = ctx.shape(FIELDS[comp])
    x = fn_of_axis(_coords(ctx, comp, axis))
    bshape = [1, 1, 1]
    bshape[2 - axis] = x.size
    return np.broadcast_to(x.reshape(bshape), shape).copy()


def _interior_faces(ctx, comp):
    """Slices (z, y, x) of the faces of component comp that are interior
    (own axis 2..N on a wall axis, 1..N periodic; cells 1..N across)."""
    sl = []
    for ax in (2, 1, 0):
        n = ctx.n[ax]
        if ax == comp:
            sl.append(slice(1, n + 1) if ctx.periodic[ax] else slice(2, n + 1))
        else:
            sl.append(slice(1, n + 1))
    return tuple(sl)


def _wall_channel(make, b, bc_lo, bc_hi, n=(8, 6, 7), size=(1.0, 0.75, 0.875), mu=0.1, rho=1.0):
    periodic = tuple(0 if d == b else 1 for d in range(3))
    side = [0] * 6
    side[2 * b], side[2 * b + 1] = 1, 2
    table = ((_abi.BC_NOSLIP, 0.0, 0.0, 0.0), bc_lo, bc_hi)
    return make(**_desc(n, size, periodic, tuple(side), table, rho=rho, mu=mu))


def _bc(t, a, val):
    v = [0.0, 0.0, 0.0]
    v[a] = val
    return (t, v[0], v[1], v[2])


# ------------------------------------------------------------ pins (raise AssertionError)
def pin_couette(make, a, b, u_lo=0.5, u_hi=-0.25, dt=1e-2):
    c = _wall_channel(make, b, _bc(_abi.BC_NOSLIP, a, u_lo), _bc(_abi.BC_NOSLIP, a, u_hi))
    L = c.size[b]
    u = _field(c, a, lambda y: u_lo + (u_hi - u_lo) * y / L, b)
    c.upload(COMP[a], u)
    c.predictor(dt)
    sl = _interior_faces(c, a)
    us = c.download(COMP[a] + "star")
    assert np.max(np.abs(us[sl] - u[sl])) <= 1e-14, ("couette u*", a, b, float(np.max(np.abs(us[sl] - u[sl]))))
    for o in range(3):
        if o != a:
            assert np.max(np.abs(c.download(COMP[o] + "star"))) <= 1e-14, ("couette cross", a, b, o)
    c.upload(COMP[a], u)
    st = c.step(dt, precond="jacobi", stop_rule="rel_l2", tol=1e-10)
    un = c.download(COMP[a])
    assert np.max(np.abs(un[sl] - u[sl])) <= 1e-12, ("couette step", a, b, st.iters)


def pin_slip_uniform(make, a, b, U=0.6, dt=1e-2):
    c = _wall_channel(make, b, _bc(_abi.BC_SLIP, a, 0.0), _bc(_abi.BC_SLIP, a, 0.0))
    u = np.full(c.shape(FIELDS[a]), U)
    c.upload(COMP[a], u)
    c.predictor(dt)
    sl = _interior_faces(c, a)
    us = c.download(COMP[a] + "star")
    assert np.max(np.abs(us[sl] - U)) <= 1e-14, ("slip uniform", a, b, float(np.max(np.abs(us[sl] - U))))


def pin_slip_profile(make, a, b, dt=1e-2, mu=0.1):
    """u_a = f(y_b) between slip walls: advection vanishes (no flow along b or
    across a) and the stress is mu times the Neumann second difference."""
    c = _wall_channel(make, b, _bc(_abi.BC_SLIP, a, 0.0), _bc(_abi.BC_SLIP, a, 0.0), mu=mu)
    L = c.size[b]
    nb = c.n[b]
    h = L / nb
    y = (np.arange(nb) + 0.5) * h
    prof = np.cos(math.pi * y / L) + 0.3 * np.sin(2.0 * math.pi * y / L) + 0.1
    u = _field(c, a, lambda yy: np.interp(yy, y, prof), b)
    c.upload(COMP[a], u)
    c.predictor(dt)
    # textbook Neumann second difference (ghost = first interior value)
    D = np.zeros((nb, nb))
    for j in range(nb):
        if j > 0:
            D[j, j - 1] += 1.0
            D[j, j] -= 1.0
        if j < nb - 1:
            D[j, j + 1] += 1.0
            D[j, j] -= 1.0
    expect = prof + dt * mu * (D @ prof) / (h * h)
    us = c.download(COMP[a] + "star")
    line = [slice(1, 2)] * 3
    line[2 - b] = slice(1, nb + 1)
    if a != b:
        got = us[tuple(line)].ravel()
        assert np.max(np.abs(got - expect)) <= 1e-13, ("slip profile", a, b, float(np.max(np.abs(got - expect))))


def pin_inflow_plug(make, a, U=0.8, dt=1e-2):
    c = _wall_channel(make, a, _bc(_abi.BC_INFLOW, a, U), _bc(_abi.BC_INFLOW, a, U))
    u = np.full(c.shape(FIELDS[a]), U)
    c.upload(COMP[a], u)
    c.predictor(dt)
    us = c.download(COMP[a] + "star")
    sl = [slice(1, c.n[d] + 1) for d in (2, 1, 0)]
    sl[2 - a] = slice(1, c.n[a] + 2)  # interior and both inflow faces
    sl = tuple(sl)
    assert np.max(np.abs(us[sl] - U)) <= 1e-14, ("inflow u*", a, float(np.max(np.abs(us[sl] - U))))
    c.upload(COMP[a], u)
    st = c.step(dt, precond="mg", stop_rule="rel_l2", tol=1e-10)
    assert st.pure_neumann == 0
    un = c.download(COMP[a])
    assert np.max(np.abs(un[sl] - U)) <= 1e-12, ("inflow step", a, st.iters)


def pin_cfl_uniform(make, dt=3e-3):
    n = (8, 6, 10)
    size = (1.0, 0.9, 2.5)
    rho, mu = 2.0, 0.05
    c = make(**_desc(n, size, (1, 1, 1), (0,) * 6, ((_abi.BC_NOSLIP, 0.0, 0.0, 0.0),), rho=rho, mu=mu))
    U = (0.7, -1.3, 0.4)
    for a in range(3):
        c.upload(COMP[a], np.full(c.shape(FIELDS[a]), U[a]))
    cc, cv = c.calc_cfl(dt)
    h = [size[d] / n[d] for d in range(3)]
    want_c = dt * sum(abs(U[d]) / h[d] for d in range(3))
    want_v = (mu / rho) * dt * sum(1.0 / (h[d] * h[d]) for d in range(3))
    assert abs(cc - want_c) <= 1e-14 * want_c, ("cfl conv", cc, want_c)
    assert abs(cv - want_v) <= 1e-14 * want_v, ("cfl visc", cv, want_v)


def pin_cfl_spike(make, dt=1e-3):
    n = (8, 8, 8)
    c = make(**_desc(n, (1.0, 1.0, 1.0), (0, 0, 0), (0,) * 6, ((_abi.BC_NOSLIP, 0.0, 0.0, 0.0),), mu=0.0))
    s = -2.5
    for a in range(3):
        arr = np.zeros(c.shape(FIELDS[a]))
        if a == 1:
            arr[4, 5, 3] = s  # one interior y-face (z=4, y-face 5, x=3)
        c.upload(COMP[a], arr)
    cc, cv = c.calc_cfl(dt)
    want = dt * abs(s) / (1.0 / 8)
    assert abs(cc - want) <= 1e-15 * want, ("cfl spike", cc, want)
    assert cv == 0.0


def pin_cfl_two_phase(make, dt=1e-3):
    n = (8, 8, 8)
    rho, mu = (1.0, 1000.0), (1e-2, 1.0)
    c = make(n=n, size=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0), periodic=(0, 0, 0), side_bcid=(0,) * 6,
             bc_table=((_abi.BC_NOSLIP, 0.0, 0.0, 0.0),), rho=rho, mu=mu, g=(0.0, 0.0, 0.0))
    al = np.zeros(c.shape("alpha"))
    al[:, :, : n[0] // 2 + 1] = 1.0
    c.upload("alpha", al)
    c.update_properties(True)
    cc, cv = c.calc_cfl(dt)
    want = max(mu[0] / rho[0], mu[1] / rho[1]) * dt * 3 * 64.0
    assert cc == 0.0
    assert abs(cv - want) <= 1e-14 * want, ("cfl two-phase", cv, want)


PAIRS = [(a, b) for a, b in itertools.product(range(3), range(3)) if a != b]


# ------------------------------------------------------------ the pins on the oracle
@pytest.mark.parametrize("a,b", PAIRS)
def test_couette_moving_walls(oracle_mod, a, b):
    pin_couette(oracle_mod.OracleContext, a, b)


@pytest.mark.parametrize("a,b", PAIRS)
def test_slip_channel(oracle_mod, a, b):
    pin_slip_uniform(oracle_mod.OracleContext, a, b)
    pin_slip_profile(oracle_mod.OracleContext, a, b)


@pytest.mark.parametrize("a", range(3))
def test_inflow_plug_flow(oracle_mod, a):
    pin_inflow_plug(oracle_mod.OracleContext, a)


def test_cfl_closed_forms(oracle_mod):
    pin_cfl_uniform(oracle_mod.OracleContext)
    pin_cfl_spike(oracle_mod.OracleContext)
    pin_cfl_two_phase(oracle_mod.OracleContext)


# ------------------------------------------------------------ the pins catch planted mistakes
MUTANT_PINS = {
    1: lambda mk: pin_couette(mk, 0, 1),
    2: lambda mk: pin_couette(mk, 2, 0),
    3: lambda mk: pin_slip_uniform(mk, 0, 2),
    4: lambda mk: pin_inflow_plug(mk, 1),
    5: lambda mk: pin_cfl_spike(mk),
    6: lambda mk: pin_cfl_uniform(mk),
}


@pytest.fixture(scope="module")
def mutant_dir(tmp_path_factory):
    return str(tmp_path_factory.mktemp("oracle_mutants"))


@pytest.mark.parametrize("mutant", sorted(MUTANT_PINS))
def test_pins_catch_planted_mistakes(oracle_mod, mutant_dir, mutant):
    lib = oracle_mod.mutant_lib(mutant, mutant_dir)

    def make(**kw):
        return oracle_mod.OracleContext(lib_=lib, **kw)

    with pytest.raises(AssertionError):
        MUTANT_PINS[mutant](make)
