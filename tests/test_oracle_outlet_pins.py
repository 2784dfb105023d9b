"""Closed-form pins of the oracle's outlet boundary (no GPU).

The reference's outlet (BC_OUTLET) rules, restated in oracle/smac_oracle.c:
* face coefficient f_b of an outlet face = 2 f_b of the adjacent interior face
  ("pressure is assumed to be zero at the boundary",
  k_update_{x,y,z}_face_boundary_properties, CFD/G_SMACSolver_VoF.cu:1798-1941),
  summed into the Poisson diagonal while the off-diagonal is dropped
  (k_build_vof_poisson_A, CFD/pressure_solver/G_PCGSolver.cu:126-212): a
  Dirichlet p = 0 at the face, so the operator is regular and the PCG runs
  without mean removal (SimulationConfig::is_pure_neumann);
* the star value of an outlet face copies the adjacent interior face
  (d_get_v*star_*face, CFD/G_BoundaryFunctions.h:369-486);
* the predictor's face getter returns the stored face value, the tangential
  mirror the inner value (d_get_v*_*face / *dir, :9-367);
* after the step the outlet face takes u* + s f_b dt dp(c)/h, s = +1 on the
  high side and -1 on the low side, c the inner cell
  (k_update_v*_outlet, CFD/G_BoundaryFunctions.cu:143-255);
* a velocity ghost whose bc row is an outlet is 2 v(row 1) - v(row 2)
  (d_get_velocity_ghost_value, :257-275).

Pins, with answers known without the oracle's formulas:
* hydrostatic column under an outlet: a fluid at rest with gravity pointing
  away from the outlet side. The exact discrete solution of the first step has
  zero interior velocity, dp = rho_f(inner face) G h / 2 in the cell next to
  the outlet and dp(k-1) = dp(k) + rho_f(k) G h away from it, with rho_f the
  face mean of the cell densities: one phase and a smoothly layered
  two-phase column, outlet on either side of every axis;
* the same solve followed by the outlet update before the correction: the
  outlet face gets u* + s 2 beta_f dt dp/h = 0 exactly (the Dirichlet
  correction of that face);
* uniform plug flow from an inflow side to the opposite outlet: u stays U
  everywhere and p stays exactly 0 (no constant is free);
* a velocity ghost of outlet kind is the linear extrapolation of a linear
  profile.

`test_outlet_pins_catch_planted_mistakes` builds the oracle with each planted
outlet mistake (SMAC_ORACLE_MUTANT 7-13) and checks that its pin fails.
"""
import numpy as np
import pytest

from mydemcfd_amd import _abi

COMP = ("u", "v", "w")
FIELDS = (_abi.FIELD_U, _abi.FIELD_V, _abi.FIELD_W)
G = 9.81


def _box(make, a, outlet_hi, n=(6, 5, 8), size=(0.6, 0.5, 1.0), rho=(1.5, 1.5), mu=(0.01, 0.01), grav=None,
         lo_bc=None):
    """Periodic in the two axes other than a; along a, the outlet on the high
    (outlet_hi) or low side and lo_bc (default a no-slip wall at rest) on the
    other. Gravity points away from the outlet unless given."""
    n = tuple(n[(d - 2 + a) % 3] for d in range(3)) if a != 2 else tuple(n)
    size = tuple(size[(d - 2 + a) % 3] for d in range(3)) if a != 2 else tuple(size)
    periodic = tuple(0 if d == a else 1 for d in range(3))
    side = [0] * 6
    side[2 * a + (1 if outlet_hi else 0)] = 1
    side[2 * a + (0 if outlet_hi else 1)] = 2
    other = lo_bc if lo_bc is not None else (_abi.BC_NOSLIP, 0.0, 0.0, 0.0)
    table = ((_abi.BC_NOSLIP, 0.0, 0.0, 0.0), (_abi.BC_OUTLET, 0.0, 0.0, 0.0), other)
    g = [0.0, 0.0, 0.0]
    if grav is None:
        g[a] = -G if outlet_hi else G
    else:
        g[a] = grav
    return make(n=n, size=size, origin=(0.0, 0.0, 0.0), periodic=periodic, side_bcid=tuple(side),
                bc_table=table, rho=rho, mu=mu, g=tuple(g))


def _line(c, a, arr, lo, hi):
    """arr along axis a at the first cell of the other axes, indices lo..hi."""
    sl = [1, 1, 1]
    sl[2 - a] = slice(lo, hi + 1)
    return arr[tuple(sl)]


def _layer_alpha(c, a):
    """alpha varying smoothly in (0.1, 0.9) along a (a different density in
    every cell of the column)."""
    al = np.zeros(c.shape("alpha"))
    n = c.n[a]
    x = (np.arange(n + 2) - 0.5) / n
    prof = 0.5 + 0.4 * np.cos(2.6 * x + 0.3)
    bshape = [1, 1, 1]
    bshape[2 - a] = n + 2
    return al + prof.reshape(bshape)


def _hydro_expect(c, a, outlet_hi, rho_cells):
    """The closed-form dp along the column (cells 1..n)."""
    n = c.n[a]
    h = c.size[a] / n
    rf = 0.5 * (rho_cells[1:] + rho_cells[:-1])  # rf[m-2] = face m between cells m-1, m (m = 2..n)
    dp = np.zeros(n)
    if outlet_hi:
        dp[n - 1] = rf[n - 2] * G * h / 2.0
        for k in range(n - 1, 0, -1):  # cell k (1-based) from cell k+1 across face k+1
            dp[k - 1] = dp[k] + rf[k - 1] * G * h
    else:
        dp[0] = rf[0] * G * h / 2.0
        for k in range(2, n + 1):
            dp[k - 1] = dp[k - 2] + rf[k - 2] * G * h
    return dp


def _setup_hydro(make, a, outlet_hi, two_phase, n=(6, 5, 8)):
    rho = (1.0, 800.0) if two_phase else (1.5, 1.5)
    mu = (1e-3, 0.1) if two_phase else (0.01, 0.01)
    c = _box(make, a, outlet_hi, n=n, rho=rho, mu=mu)
    n = c.n[a]
    if two_phase:
        al = _layer_alpha(c, a)
        c.upload("alpha", al)
        a_line = _line(c, a, al, 1, n)
        rho_cells = (1.0 - a_line) * rho[0] + a_line * rho[1]
    else:
        rho_cells = np.full(n, rho[0])
    c.update_properties(True)
    return c, rho_cells


def pin_outlet_hydrostatic(make, a, outlet_hi, two_phase, precond="mg", dt=1e-3, n=(6, 5, 8)):
    c, rho_cells = _setup_hydro(make, a, outlet_hi, two_phase, n=n)
    n = c.n[a]
    kw = dict(precond=precond, stop_rule="rel_l2", tol=1e-13, max_iter=500)
    if precond == "ref_gmg":
        kw.update(mg_levels=2)
    st = c.step(dt, **kw)
    assert st.pure_neumann == 0 and st.converged, st.as_dict()
    want = _hydro_expect(c, a, outlet_hi, rho_cells)
    p = c.download("p")
    got = _line(c, a, p, 1, n)
    err = float(np.max(np.abs(got - want)) / np.max(np.abs(want)))
    assert err <= 1e-9, ("hydrostatic p", a, outlet_hi, two_phase, precond, err)
    # the same along every line of the column (periodic across)
    inner = p[1:c.n[2] + 1, 1:c.n[1] + 1, 1:c.n[0] + 1]
    assert float(np.max(np.abs(inner - inner.mean(axis=tuple(d for d in range(3) if d != 2 - a), keepdims=True)))) \
        <= 1e-9 * float(np.max(np.abs(want))), "hydrostatic p not uniform across"
    # interior faces of the axis at rest
    w = c.download(COMP[a])
    wi = _line(c, a, w, 2, n)
    assert float(np.max(np.abs(wi))) <= 1e-9 * G * dt, ("hydrostatic w", float(np.max(np.abs(wi))))


def pin_outlet_update(make, a, outlet_hi, two_phase, dt=1e-3):
    """predictor + solve, then the post-step update before the correction:
    the outlet face is corrected to zero."""
    c, _ = _setup_hydro(make, a, outlet_hi, two_phase)
    n = c.n[a]
    c.predictor(dt)
    st = c.pressure_solve(dt, precond="mg", stop_rule="rel_l2", tol=1e-13, max_iter=500)
    assert st.converged
    c.update_boundary_ghost(dt)
    w = c.download(COMP[a])
    f = n + 1 if outlet_hi else 1
    wo = _line(c, a, w, f, f)
    assert float(np.max(np.abs(wo))) <= 1e-9 * G * dt, ("outlet update", a, outlet_hi, float(np.max(np.abs(wo))))
    # after the correction (which clears dp) the update leaves u* = -g dt there
    c.correct(dt)
    c.update_boundary_ghost(dt)
    w = c.download(COMP[a])
    wo = _line(c, a, w, f, f)
    g = -G if outlet_hi else G
    assert float(np.max(np.abs(wo - g * dt))) <= 1e-15, ("outlet after correction", float(np.max(np.abs(wo - g * dt))))


def pin_outlet_plug(make, a, outlet_hi, U=0.7, dt=1e-2):
    s = 1.0 if outlet_hi else -1.0
    inflow = [0.0, 0.0, 0.0]
    inflow[a] = s * U
    c = _box(make, a, outlet_hi, grav=0.0, lo_bc=(_abi.BC_INFLOW, inflow[0], inflow[1], inflow[2]))
    u = np.full(c.shape(FIELDS[a]), s * U)
    c.upload(COMP[a], u)
    st = c.step(dt, precond="mg", stop_rule="rel_l2", tol=1e-12)
    assert st.pure_neumann == 0
    n = c.n[a]
    un = c.download(COMP[a])
    got = _line(c, a, un, 1, n + 1)
    assert float(np.max(np.abs(got - s * U))) <= 1e-13, ("plug u", a, outlet_hi, float(np.max(np.abs(got - s * U))))
    p = c.download("p")[1:c.n[2] + 1, 1:c.n[1] + 1, 1:c.n[0] + 1]
    assert float(np.max(np.abs(p))) <= 1e-12, ("plug p", float(np.max(np.abs(p))))
    us = c.download(COMP[a] + "star")
    assert float(np.max(np.abs(_line(c, a, us, 1, n + 1) - s * U))) <= 1e-13, "plug u*"
    c.update_boundary_ghost(dt)
    assert float(np.max(np.abs(_line(c, a, c.download(COMP[a]), 1, n + 1) - s * U))) <= 1e-13, "plug ghosts"


def pin_outlet_ghost(make):
    """bc row 0 of outlet kind decides the tangential ghosts off the walls'
    own faces: u linear in y extrapolates exactly across the y walls."""
    n = (6, 7, 5)
    size = (0.6, 0.7, 0.5)
    table = ((_abi.BC_OUTLET, 0.0, 0.0, 0.0), (_abi.BC_NOSLIP, 0.0, 0.0, 0.0))
    c = make(n=n, size=size, origin=(0.0, 0.0, 0.0), periodic=(0, 0, 0), side_bcid=(1,) * 6, bc_table=table,
             rho=(1.0, 1.0), mu=(0.0, 0.0), g=(0.0, 0.0, 0.0))
    h = size[1] / n[1]
    y = (np.arange(n[1] + 2) - 0.5) * h  # cell-centred along y for u
    u = np.broadcast_to((0.3 + 1.7 * y).reshape(1, -1, 1), c.shape("u")).copy()
    c.upload("u", u)
    c.update_boundary_ghost(1e-3)
    got = c.download("u")
    # ghosts across y at interior z and interior x-faces 2..n (the own faces 1
    # and n+1 take the wall's own rule)
    for j in (0, n[1] + 1):
        blk = got[1:n[2] + 1, j, 2:n[0] + 1]
        assert float(np.max(np.abs(blk - (0.3 + 1.7 * y[j])))) <= 1e-14, ("outlet ghost", j)


AXES = [(a, hi) for a in range(3) for hi in (True, False)]


@pytest.mark.parametrize("a,hi", AXES)
@pytest.mark.parametrize("two_phase", [False, True])
def test_outlet_hydrostatic(oracle_mod, a, hi, two_phase):
    pin_outlet_hydrostatic(oracle_mod.OracleContext, a, hi, two_phase)


@pytest.mark.parametrize("precond", ["jacobi", "ref_gmg"])
def test_outlet_hydrostatic_other_preconditioners(oracle_mod, precond):
    pin_outlet_hydrostatic(oracle_mod.OracleContext, 2, True, True, precond=precond, n=(12, 12, 24))


@pytest.mark.parametrize("a,hi", AXES)
def test_outlet_update(oracle_mod, a, hi):
    pin_outlet_update(oracle_mod.OracleContext, a, hi, a == 1)


@pytest.mark.parametrize("a,hi", AXES)
def test_outlet_plug_flow(oracle_mod, a, hi):
    pin_outlet_plug(oracle_mod.OracleContext, a, hi)


def test_outlet_ghost(oracle_mod):
    pin_outlet_ghost(oracle_mod.OracleContext)


MUTANT_PINS = {
    7: lambda mk: pin_outlet_hydrostatic(mk, 2, True, False),
    8: lambda mk: pin_outlet_hydrostatic(mk, 0, False, True),
    9: lambda mk: pin_outlet_hydrostatic(mk, 1, True, False),
    10: lambda mk: pin_outlet_plug(mk, 0, True),
    11: lambda mk: pin_outlet_ghost(mk),
    12: lambda mk: pin_outlet_hydrostatic(mk, 2, False, False),
    13: lambda mk: pin_outlet_update(mk, 2, True, False),
}


@pytest.fixture(scope="module")
def mutant_dir(tmp_path_factory):
    return str(tmp_path_factory.mktemp("oracle_outlet_mutants"))


@pytest.mark.parametrize("mutant", sorted(MUTANT_PINS))
def test_outlet_pins_catch_planted_mistakes(oracle_mod, mutant_dir, mutant):
    lib = oracle_mod.mutant_lib(mutant, mutant_dir)

    def make(**kw):
        return oracle_mod.OracleContext(lib_=lib, **kw)

    with pytest.raises(AssertionError):
        MUTANT_PINS[mutant](make)
