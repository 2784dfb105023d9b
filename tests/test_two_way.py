"""Pins of the oracle's two-way CFD-DEM variants (no GPU): the volume-fraction
predictor k_get_vof_vstar_rhouu_consistent_{x,y,z}_two_way
(CFD/G_SMACSolver.cu:1538-2203) with its mass flux
k_compute_mass_flux_from_alpha_flux_two_way (CFD/G_SMACSolver_VoF.cu:2472-2557),
the Poisson coefficients of update_poisson_beta_two_way
(Coupling/G_CFDDEMCoupling.cu:419-459) and the general Poisson rhs
(CFD/pressure_solver/G_PressureSolverBase.cu:49-118).

What fixes the oracle here is independent of its formulas:
  * a constant volume fraction c cancels from the step: every term of the
    two-way predictor carries one factor c (stresses, mass flux, old pressure
    gradient, the old momentum) and u* divides by c again, so u* equals the
    one-fluid predictor's to rounding -- for c = 1 and c = 0.6, walls and
    periodic, one and two phases. A dropped or doubled eps on any term breaks it;
  * the particle impulse J enters as (beta/eps_f) J/V: u*(J) - u*(0) is that,
    face by face;
  * uniform flow U along x in a periodic box with a varying volume fraction:
    no stress and no pressure, so u* = (beta/en_f)(eo_f rho U - dt (M+ - M-)/dx)
    with M = U times the face-averaged mass flux rho eps_half_f U -- restated
    here in numpy;
  * the Poisson operator becomes the 7-point operator with face coefficients
    eps_f/rho_f (assembled here in numpy) and returns to 1/rho_f after the next
    property update;
  * the rhs is -(eps - eps_old)/dt^2 for u* = 0 and c times the one-fluid rhs
    for eps = eps_old = c;
  * after a full step the discrete continuity of the volume-fraction form holds:
    D(eps_f u^{n+1}) = -(eps - eps_old)/dt to the solver tolerance.
"""
import numpy as np
import pytest

from mydemcfd_amd import _abi, cases


def _box(kind, n, two_phase=False):
    if kind == "periodic":
        desc = cases.periodic_desc(n, re=200.0, length=1.0)
    else:
        desc = cases.cavity_desc(n, re=100.0)
    if two_phase:
        desc.update(rho=(1.0, 5.0), mu=(0.01, 0.03), g=(0.0, -1.0, 0.0))
    return desc


def _fields(c, seed, scale=0.3):
    rng = np.random.default_rng(seed)
    f = {}
    for name in ("u", "v", "w", "p"):
        f[name] = scale * rng.standard_normal(c.shape(name))
    return f


def _setup(oracle_mod, kind, n, two_phase, seed):
    desc = _box(kind, n, two_phase)
    c = oracle_mod.OracleContext(**desc)
    for name, a in _fields(c, seed).items():
        c.upload(name, a)
    if two_phase:
        rng = np.random.default_rng(seed + 1)
        c.upload("alpha", (rng.uniform(size=c.shape("alpha")) > 0.5).astype(np.float64))
    c.update_properties(1)
    return c, desc


def _eps_all(c, desc, values):
    for name, v in zip(("eps", "eps_old", "eps_half"), values):
        a = v if isinstance(v, np.ndarray) else np.full(c.shape(name), float(v))
        c.upload(name, a)


@pytest.mark.parametrize("kind,two_phase", [("walls", False), ("walls", True), ("periodic", False)])
@pytest.mark.parametrize("cval", [1.0, 0.6])
def test_constant_volume_fraction_cancels(oracle_mod, kind, two_phase, cval):
    n = (10, 8, 6)
    dt = 1e-3
    one, desc = _setup(oracle_mod, kind, n, two_phase, 11)
    two, _ = _setup(oracle_mod, kind, n, two_phase, 11)
    _eps_all(two, desc, (cval, cval, cval))
    one.predictor(dt)
    two.predictor_two_way(dt)
    for ax, f in enumerate(("ustar", "vstar", "wstar")):
        a, b = two.download(f), one.download(f)
        if kind == "periodic":
            # the two-way mass flux is 0 off the faces 1..N+1 of an axis
            # (k_compute_mass_flux_from_alpha_flux_two_way writes 0 there), so
            # face 1 of a periodic axis sees a zero flux on its minus side: the
            # seam faces are left out of the comparison
            sl = [slice(None)] * 3
            sl[2 - ax] = slice(2, -2)
            a, b = a[tuple(sl)], b[tuple(sl)]
        assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(b)), f


def test_impulse_enters_linearly(oracle_mod):
    n = (10, 8, 6)
    dt = 2e-3
    ctxs = []
    for with_j in (False, True):
        c, desc = _setup(oracle_mod, "walls", n, True, 5)
        eps = cases.void_fraction(n, desc["periodic"], 7)
        _eps_all(c, desc, (eps, cases.void_fraction(n, desc["periodic"], 8),
                           cases.void_fraction(n, desc["periodic"], 9)))
        if with_j:
            rng = np.random.default_rng(3)
            for f in ("impx", "impy", "impz"):
                c.upload(f, 1e-4 * rng.standard_normal(c.shape(f)))
        c.predictor_two_way(dt)
        ctxs.append(c)
    c0, c1 = ctxs
    h = 1.0 / np.array(n, dtype=np.float64)
    inv_vol = (1.0 / h[0]) * (1.0 / h[1]) * (1.0 / h[2])
    rho = _cell_rho(c1)
    for a, (fs, fi) in enumerate((("ustar", "impx"), ("vstar", "impy"), ("wstar", "impz"))):
        d = c1.download(fs) - c0.download(fs)
        J = c1.download(fi)
        ax = 2 - a  # numpy axis of the face direction
        lo = [slice(1, -1)] * 3
        lo[ax] = slice(2, -2)  # interior faces 2..N
        mi = [slice(1, -1)] * 3
        mi[ax] = slice(1, -2)  # the minus cells 1..N-1
        pl = [slice(1, -1)] * 3
        pl[ax] = slice(2, -1)  # the plus cells 2..N (the face's own index)
        en_f = 0.5 * (eps[tuple(mi)] + eps[tuple(pl)])
        beta = 1.0 / (0.5 * (rho[tuple(mi)] + rho[tuple(pl)]))
        expect = (beta / en_f) * J[tuple(lo)] * inv_vol
        got = d[tuple(lo)]
        assert np.max(np.abs(got - expect)) <= 1e-12 * np.max(np.abs(expect)), fs


def _cell_rho(c):
    a = c.download("alpha")
    return (1.0 - a) * 1.0 + a * 5.0


def test_uniform_flow_closed_form(oracle_mod):
    n = (12, 6, 5)
    dt, U = 1e-2, 0.7
    desc = cases.periodic_desc(n, re=100.0, length=1.0)
    c = oracle_mod.OracleContext(**desc)
    c.upload("u", np.full(c.shape("u"), U))
    per = desc["periodic"]
    eps = cases.void_fraction(n, per, 21)
    eo = cases.void_fraction(n, per, 22)
    eh = cases.void_fraction(n, per, 23)
    _eps_all(c, desc, (eps, eo, eh))
    c.predictor_two_way(dt)
    us = c.download("ustar")
    rho, hx = 1.0, 1.0 / n[0]
    beta = 1.0 / rho
    # faces 3..N-1 along x: their mass-flux neighbours are the faces 2..N,
    # where eps_half_f is the two-cell mean (faces 1 and N+1 take the inner
    # cell alone)
    for i in range(3, n[0]):
        # mass flux on x-face f: rho0 * 0.5*(eh[f-1] + eh[f]) * U
        def mf(f):
            return rho * (0.5 * (eh[1:-1, 1:-1, f - 1] + eh[1:-1, 1:-1, f]) * U)
        mp = 0.5 * (mf(i) + mf(i + 1))
        mm = 0.5 * (mf(i - 1) + mf(i))
        conv = (mp * U - mm * U) * (1.0 / hx)
        eo_f = 0.5 * (eo[1:-1, 1:-1, i - 1] + eo[1:-1, 1:-1, i])
        en_f = 0.5 * (eps[1:-1, 1:-1, i - 1] + eps[1:-1, 1:-1, i])
        expect = (beta / en_f) * (eo_f * rho * U - dt * conv)
        got = us[1:-1, 1:-1, i]
        assert np.max(np.abs(got - expect)) <= 1e-13, i
    for f in ("vstar", "wstar"):
        assert np.max(np.abs(c.download(f)[1:-1, 1:-1, 1:-1])) == 0.0, f


def _operator_numpy(eps, n, h, beta0, periodic):
    """y = A x with face coefficients eps_f*beta0 (0 across walls), A = diag - offdiag."""
    nx, ny, nz = n

    def apply(x):
        y = np.zeros((nz, ny, nx))
        xi = x[1:-1, 1:-1, 1:-1]
        e = eps
        for ax, hd, per in ((2, h[0], periodic[0]), (1, h[1], periodic[1]), (0, h[2], periodic[2])):
            N = xi.shape[ax]
            for s in (+1, -1):
                # neighbour values and the face between
                sl_n = [slice(1, -1)] * 3
                sl_f = [slice(1, -1)] * 3
                idx = np.arange(1, N + 1) + s
                if per:
                    idx_w = (idx - 1) % N + 1
                else:
                    idx_w = idx
                xn = np.take(x, idx_w, axis=ax)[tuple(slice(1, -1) if a != ax else slice(None) for a in range(3))]
                en = np.take(e, idx, axis=ax)[tuple(slice(1, -1) if a != ax else slice(None) for a in range(3))]
                ec = e[1:-1, 1:-1, 1:-1]
                coef = 0.5 * (en + ec) * beta0 / hd ** 2
                if not per:
                    wall = (idx < 1) | (idx > N)
                    shape = [1, 1, 1]
                    shape[ax] = N
                    coef = np.where(wall.reshape(shape), 0.0, coef)
                y += coef * (xi - xn)
                del sl_n, sl_f
        return y

    return apply


@pytest.mark.parametrize("kind", ["walls", "periodic"])
def test_poisson_coefficients(oracle_mod, kind):
    n = (8, 6, 5)
    desc = _box(kind, n)
    c = oracle_mod.OracleContext(**desc)
    c.update_properties(1)
    per = desc["periodic"]
    eps = cases.void_fraction(n, per, 31)
    c.upload("eps", eps)
    rng = np.random.default_rng(4)
    x = rng.standard_normal(c.shape("p"))
    if per[0]:
        x = cases._embed_periodic(x[1:-1, 1:-1, 1:-1], _abi.FIELD_P)
    y_one = c.apply_A(x)
    c.update_poisson_beta_two_way()
    y_two = c.apply_A(x)
    h = (1.0 / n[0], 1.0 / n[1], 1.0 / n[2])
    ref = _operator_numpy(eps, n, h, 1.0, per)(x)
    got = y_two[1:-1, 1:-1, 1:-1]
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
    c.update_properties(0)
    y_back = c.apply_A(x)
    assert np.array_equal(y_back, y_one)


def test_rhs_closed_forms(oracle_mod):
    n = (8, 6, 5)
    dt = 1e-2
    desc = _box("walls", n)
    c = oracle_mod.OracleContext(**desc)
    per = desc["periodic"]
    eps = cases.void_fraction(n, per, 41)
    eo = cases.void_fraction(n, per, 42)
    c.upload("eps", eps)
    c.upload("eps_old", eo)
    c.upload("ustar", np.zeros(c.shape("ustar")))
    c.pressure_solve(dt, precond="jacobi", tol=1e-6)
    b = c.download("rhs")[1:-1, 1:-1, 1:-1]
    expect = -((eps - eo)[1:-1, 1:-1, 1:-1] * (1.0 / dt) * (1.0 / dt))
    assert np.array_equal(b, expect)
    # eps = eps_old = c: c times the one-fluid rhs
    rng = np.random.default_rng(2)
    us = {f: rng.standard_normal(c.shape(f)) for f in ("ustar", "vstar", "wstar")}
    one = oracle_mod.OracleContext(**desc)
    for ctx in (c, one):
        for f, a in us.items():
            ctx.upload(f, a)
    c.upload("eps", np.full(c.shape("eps"), 0.7))
    c.upload("eps_old", np.full(c.shape("eps"), 0.7))
    c.pressure_solve(dt, precond="jacobi", tol=1e-6)
    one.pressure_solve(dt, precond="jacobi", tol=1e-6)
    b1, b0 = c.download("rhs"), one.download("rhs")
    assert np.max(np.abs(b1 - 0.7 * b0)) <= 1e-13 * np.max(np.abs(b0))


@pytest.mark.parametrize("two_phase", [False, True])
def test_volume_fraction_continuity_after_step(oracle_mod, two_phase):
    n = (12, 10, 8)
    dt = 2e-3
    c, desc = _setup(oracle_mod, "walls", n, two_phase, 17)
    per = desc["periodic"]
    eo = cases.void_fraction(n, per, 51)
    # eps with the same interior total as eps_old (a closed box conserves it)
    eps = cases.void_fraction(n, per, 52)
    eps[1:-1, 1:-1, 1:-1] += (eo[1:-1, 1:-1, 1:-1].sum() - eps[1:-1, 1:-1, 1:-1].sum()) / np.prod(n)
    _eps_all(c, desc, (eps, eo, 0.5 * (eps + eo)))
    c.update_poisson_beta_two_way()
    c.predictor_two_way(dt)
    st = c.pressure_solve(dt, precond="mg", tol=1e-12, max_iter=200)
    assert st.converged, st.as_dict()
    c.correct(dt)
    u, v, w = c.download("u"), c.download("v"), c.download("w")
    nx, ny, nz = n
    h = 1.0 / np.array(n, dtype=np.float64)
    e = eps
    ex = 0.5 * (e[1:-1, 1:-1, :-1] + e[1:-1, 1:-1, 1:])   # x-faces 1..N+1 (ghost means at the walls)
    ey = 0.5 * (e[1:-1, :-1, 1:-1] + e[1:-1, 1:, 1:-1])
    ez = 0.5 * (e[:-1, 1:-1, 1:-1] + e[1:, 1:-1, 1:-1])
    fx = ex * u[1:-1, 1:-1, 1:-1]
    fy = ey * v[1:-1, 1:-1, 1:-1]
    fz = ez * w[1:-1, 1:-1, 1:-1]
    # the correction leaves the wall faces alone; their flux is that of u* (0)
    fx[:, :, [0, -1]] = 0.0
    fy[:, [0, -1], :] = 0.0
    fz[[0, -1], :, :] = 0.0
    div = ((fx[:, :, 1:] - fx[:, :, :-1]) / h[0] + (fy[:, 1:, :] - fy[:, :-1, :]) / h[1]
           + (fz[1:, :, :] - fz[:-1, :, :]) / h[2])
    deps = (eps - eo)[1:-1, 1:-1, 1:-1] / dt
    resid = div + deps
    assert np.max(np.abs(resid)) <= 1e-8 * np.max(np.abs(deps)), np.max(np.abs(resid))
