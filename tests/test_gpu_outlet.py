"""GPU (HIP, gfx950) checks of the outlet boundary.

* The closed-form outlet pins of tests/test_oracle_outlet_pins.py (hydrostatic
  column under an outlet on either side of every axis, one phase and layered
  two phase; the outlet face update before and after the correction; plug
  flow from an inflow side to the outlet; the outlet ghost rule) run on the
  GPU through the C-ABI.
* An inflow -> outlet channel (inflow carrying a tangential velocity at
  x-min, outlet at x-max, a no-slip wall at rest and a moving one across y,
  slip walls across z; one phase and a two-phase blob under gravity) steps
  with each preconditioner, the post-step update (outlet faces, ghosts) after
  every step as the reference loop does (CFD/main.cu:298-317), bit for bit
  against the oracle. The outlet faces carry the Dirichlet coefficient, so the
  solve is the regular (not pure-Neumann) PCG of G_PCGSolver.cu:448-586,619-770.
* The same channel in the mixed forms against the FP64 oracle.
"""
import numpy as np
import pytest

from conftest import rel_err
import test_oracle_outlet_pins as pins

from mydemcfd_amd import _abi

pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.mark.parametrize("a,hi", pins.AXES)
@pytest.mark.parametrize("two_phase", [False, True])
def test_gpu_outlet_hydrostatic_pin(hip, a, hi, two_phase):
    print(f"case: outlet hydrostatic axis {a} high {hi} two_phase {two_phase}", flush=True)
    pins.pin_outlet_hydrostatic(hip.SMACContext, a, hi, two_phase)


@pytest.mark.parametrize("precond", ["jacobi", "ref_gmg"])
def test_gpu_outlet_hydrostatic_other_preconditioners(hip, precond):
    print(f"case: outlet hydrostatic {precond}", flush=True)
    pins.pin_outlet_hydrostatic(hip.SMACContext, 2, True, True, precond=precond, n=(12, 12, 24))


@pytest.mark.parametrize("a,hi", pins.AXES)
def test_gpu_outlet_update_and_plug_pins(hip, a, hi):
    print(f"case: outlet update + plug axis {a} high {hi}", flush=True)
    pins.pin_outlet_update(hip.SMACContext, a, hi, a == 1)
    pins.pin_outlet_plug(hip.SMACContext, a, hi)


def test_gpu_outlet_ghost_pin(hip):
    print("case: outlet ghost pin", flush=True)
    pins.pin_outlet_ghost(hip.SMACContext)


def channel_desc(n=(24, 16, 20), two_phase=False):
    table = ((_abi.BC_NOSLIP, 0.0, 0.0, 0.0),
             (_abi.BC_INFLOW, 0.5, 0.1, -0.05),
             (_abi.BC_OUTLET, 0.0, 0.0, 0.0),
             (_abi.BC_NOSLIP, 0.0, 0.0, 0.0),
             (_abi.BC_NOSLIP, 0.3, 0.0, 0.1),
             (_abi.BC_SLIP, 0.0, 0.0, 0.0))
    rho, mu, g = (1.0, 1.0), (0.02, 0.02), (0.0, 0.0, 0.0)
    if two_phase:
        rho, mu, g = (1.0, 50.0), (1e-3, 2e-2), (0.0, -0.5, -1.0)
    return dict(n=n, size=(1.2, 0.8, 1.0), origin=(0.0, 0.0, 0.0), periodic=(0, 0, 0),
                side_bcid=(1, 2, 3, 4, 5, 5), bc_table=table, rho=rho, mu=mu, g=g)


def channel_fields(ctx, seed=5):
    """Smooth random velocities plus a mean flow 0.5 along x; the inflow face
    holds its boundary value, the outlet face a smooth value of its own, the y
    and z boundary faces zero."""
    rng = np.random.default_rng(seed)
    out = {}
    for a, name in enumerate(("u", "v", "w")):
        shape = ctx.shape(name)
        zz, yy, xx = np.meshgrid(*[np.linspace(0.0, 1.0, s) for s in shape], indexing="ij")
        f = np.zeros(shape)
        for _ in range(4):
            k = rng.integers(1, 4, size=3)
            f += 0.2 * rng.standard_normal() * np.sin(np.pi * (k[0] * xx + 0.2)) * np.cos(np.pi * k[1] * yy) * \
                np.sin(np.pi * (k[2] * zz + 0.3))
        ax = 2 - a
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[ax], hi[ax] = 1, shape[ax] - 2
        if a == 0:
            f += 0.5
            f[tuple(lo)] = 0.5
        else:
            f[tuple(lo)] = 0.0
            f[tuple(hi)] = 0.0
        out[name] = f
    return out


def _channel_pair(hip, oracle_mod, n, two_phase):
    desc = channel_desc(n, two_phase)
    g = hip.SMACContext(**desc)
    o = oracle_mod.OracleContext(**desc)
    fields = channel_fields(o)
    if two_phase:
        al = np.zeros(o.shape("alpha"))
        zz, yy, xx = np.meshgrid(*[np.arange(s) for s in al.shape], indexing="ij")
        blob = ((xx - n[0] * 0.4) / (0.25 * n[0])) ** 2 + ((yy - n[1] * 0.5) / (0.3 * n[1])) ** 2 + \
            ((zz - n[2] * 0.4) / (0.3 * n[2])) ** 2
        al[blob < 1.0] = 1.0
        fields["alpha"] = al
    for name, arr in fields.items():
        g.upload(name, arr)
        o.upload(name, arr)
    return g, o


@pytest.mark.parametrize("precond", ["jacobi", "mg", "ref_gmg"])
@pytest.mark.parametrize("two_phase", [False, True])
def test_outlet_channel_vs_oracle(hip, oracle_mod, precond, two_phase):
    n = (24, 16, 20)
    print(f"case: outlet channel {n} two_phase={two_phase} {precond} x4", flush=True)
    g, o = _channel_pair(hip, oracle_mod, n, two_phase)
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
        for c in (g, o):
            c.update_boundary_ghost(dt)
    # the Dirichlet outlet fixes the pressure constant: p is compared as it is
    for f in ("u", "v", "w", "p"):
        e = rel_err(g.download(f), o.download(f))
        assert e <= TOL, (f, e)
    for f in ("ustar", "vstar", "wstar"):
        assert rel_err(g.download(f), o.download(f)) <= TOL, f
    assert g.calc_cfl(dt) == o.calc_cfl(dt)


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_outlet_channel_mixed_vs_fp64_oracle(hip, oracle_mod, precision):
    n = (32, 16, 24)
    print(f"case: outlet channel {n} mixed {precision} x4", flush=True)
    g, o = _channel_pair(hip, oracle_mod, n, True)
    dt = 2e-3
    for s in range(4):
        sg = g.step(dt, precond="mg", precision=precision, flexible=1, tol=1e-8, max_iter=500)
        so = o.step(dt, precond="mg", tol=1e-8, max_iter=500)
        assert sg.converged and sg.rel_res < 1e-8, sg.as_dict()
        assert sg.iters <= so.iters + 6, (sg.iters, so.iters)
        for c in (g, o):
            c.update_boundary_ghost(dt)
    vel = {f: (g.download(f), o.download(f)) for f in ("u", "v", "w")}
    vscale = max(float(np.max(np.abs(b))) for _, b in vel.values())
    errs = {f: float(np.max(np.abs(a - b))) / vscale for f, (a, b) in vel.items()}
    errs["p"] = rel_err(g.download("p"), o.download("p"))
    print(f"outlet channel mixed {precision}: error vs FP64 oracle {errs}, iters {sg.iters} vs {so.iters}",
          flush=True)
    for f, e in errs.items():
        assert e < 1e-6, (f, e)


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16"])
def test_outlet_channel_wide_rows(hip, oracle_mod, precision):
    """Rows of 256 cells: the first SpMV of every solve forms the PCG start's
    d = z in its own pass, here without a mean shift (the outlet makes the
    operator regular). Isotropic cells, 3 steps, against the FP64 oracle."""
    n = (256, 8, 12)
    print(f"case: outlet channel {n} {precision} x3", flush=True)
    desc = channel_desc(n)
    desc["size"] = (2.56, 0.08, 0.12)
    g = hip.SMACContext(**desc)
    o = oracle_mod.OracleContext(**desc)
    for name, arr in channel_fields(o).items():
        g.upload(name, arr)
        o.upload(name, arr)
    dt = 5e-4
    for s in range(3):
        if precision == "fp64":
            sg = g.step(dt, precond="mg", tol=1e-9, max_iter=500)
        else:
            sg = g.step(dt, precond="mg", precision=precision, flexible=1, tol=1e-9, max_iter=500)
        so = o.step(dt, precond="mg", tol=1e-9, max_iter=500)
        assert so.pure_neumann == 0 and sg.pure_neumann == 0
        assert so.converged and sg.converged, (sg.as_dict(), so.as_dict())
        if precision == "fp64":
            assert sg.iters == so.iters, (s, sg.iters, so.iters)
        else:
            assert sg.iters <= so.iters + 6, (s, sg.iters, so.iters)
        for c in (g, o):
            c.update_boundary_ghost(dt)
    tol = TOL if precision == "fp64" else 1e-6
    for f in ("u", "v", "w", "p"):
        e = rel_err(g.download(f), o.download(f))
        assert e <= tol, (f, e)


def test_outlet_refusals(hip):
    """The adjoint refuses an outlet context; an outflow side (no case in the
    reference) is refused at create."""
    print("case: outlet refusals", flush=True)
    desc = channel_desc((16, 8, 8))
    desc_bad = dict(desc)
    desc_bad["bc_table"] = tuple((_abi.BC_OUTFLOW, 0.0, 0.0, 0.0) if i == 2 else e
                                 for i, e in enumerate(desc["bc_table"]))
    with pytest.raises(_abi.SmacError) as e:
        hip.SMACContext(**desc_bad)
    assert e.value.status == _abi.SMAC_ERR_UNSUPPORTED
    torch = pytest.importorskip("torch")
    c = hip.SMACContext(**desc)
    u_n = [torch.zeros(c.shape(f), dtype=torch.float64, device="cuda") for f in ("u", "v", "w")]
    ubar = [torch.zeros_like(t) for t in u_n]
    pbar = torch.zeros(c.shape("p"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    with pytest.raises(_abi.SmacError) as e:
        c.adjoint_step(1e-3, [t.data_ptr() for t in u_n], [t.data_ptr() for t in ubar], pbar.data_ptr())
    assert e.value.status == _abi.SMAC_ERR_UNSUPPORTED
