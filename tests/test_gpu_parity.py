"""GPU (HIP, gfx950) vs CPU FP64 oracle: the FP64 form of the SMAC step.

Bar: max|a-b|/max|b| <= 1e-10 on u, v, w and p after N
steps on identical seeded inputs (p compared modulo its mean). Both sides
share the arithmetic contract (no FMA contraction, canonical reduction
order), so in practice the fields agree exactly.
"""
import numpy as np
import pytest

from conftest import rel_err, rel_err_mod_mean
from cpp import drivers as cpp_drivers

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _pair(hip, oracle_mod, desc):
    return hip.SMACContext(**desc), oracle_mod.OracleContext(**desc)


def _upload_both(a, b, fields):
    for name, arr in fields.items():
        a.upload(name, arr)
        b.upload(name, arr)


def _compare(a, b, n, fields=("u", "v", "w"), tol=TOL):
    out = {}
    for f in fields:
        out[f] = rel_err(a.download(f), b.download(f))
    out["p"] = rel_err_mod_mean(a.download("p"), b.download("p"), n)
    for f, e in out.items():
        assert e <= tol, (f, e, out)
    return out


def test_predictor_cavity(hip, oracle_mod):
    from mydemcfd_amd import cases

    print("case: predictor cavity 12x10x8", flush=True)
    desc = cases.cavity_desc(8)
    desc["n"] = (12, 10, 8)
    g, o = _pair(hip, oracle_mod, desc)
    rng = np.random.default_rng(1)
    fields = {f: rng.standard_normal(g.shape(f)) for f in ("u", "v", "w", "p")}
    _upload_both(g, o, fields)
    g.predictor(2e-3)
    o.predictor(2e-3)
    for f in ("ustar", "vstar", "wstar"):
        assert rel_err(g.download(f), o.download(f)) == 0.0, f


@pytest.mark.parametrize("n", [(16, 8, 12), (128, 8, 40), (64, 12, 33), (64, 4, 4)])
def test_predictor_periodic(hip, oracle_mod, n):
    """Triply periodic predictor: the cell-parallel kernel (16x8x12) and the
    z-marching tile kernel (nx % 64 == 0, ny % 4 == 0): two and three z
    chunks of 32 planes (the last one partial, down to one plane), several
    y tiles, and a box of four planes (the z wrap reaches k-2 and k+3)."""
    from mydemcfd_amd import cases

    print(f"case: predictor periodic {n}", flush=True)
    desc = cases.periodic_desc(n)
    g, o = _pair(hip, oracle_mod, desc)
    rng = np.random.default_rng(2)
    fields = {f: rng.standard_normal(g.shape(f)) for f in ("u", "v", "w", "p")}
    _upload_both(g, o, fields)
    g.predictor(1e-3)
    o.predictor(1e-3)
    for f in ("ustar", "vstar", "wstar"):
        assert rel_err(g.download(f), o.download(f)) == 0.0, f


@pytest.mark.parametrize("precond,stop", [("jacobi", "rel_l2"), ("jacobi", "ref_linf_dt"), ("mg", "rel_l2")])
def test_cavity_steps(hip, oracle_mod, precond, stop):
    from mydemcfd_amd import cases

    print(f"case: cavity 16^3 x5 {precond} {stop}", flush=True)
    desc = cases.cavity_desc(16)
    g, o = _pair(hip, oracle_mod, desc)
    tol = 1e-8 if stop == "rel_l2" else 1e-9
    ce = 1 if precond == "mg" else 10
    for s in range(5):
        sg = g.step(1e-3, precond=precond, stop_rule=stop, tol=tol, check_every=ce)
        so = o.step(1e-3, precond=precond, stop_rule=stop, tol=tol, check_every=ce)
        assert sg.iters == so.iters and sg.converged == so.converged, (s, sg.as_dict(), so.as_dict())
    _compare(g, o, desc["n"])


def test_periodic_tg_mg(hip, oracle_mod):
    from mydemcfd_amd import cases

    print("case: TG periodic 32^3 x4 mg", flush=True)
    n = 32
    desc = cases.periodic_desc(n)
    g, o = _pair(hip, oracle_mod, desc)
    u, v, w = cases.tg_fields(n, pert="white")
    _upload_both(g, o, {"u": u, "v": v, "w": w})
    dt = cases.cfl_dt(o, 0.3)
    assert abs(cases.cfl_dt(g, 0.3) - dt) == 0.0
    for s in range(4):
        sg = g.step(dt, precond="mg", tol=1e-8)
        so = o.step(dt, precond="mg", tol=1e-8)
        assert sg.iters == so.iters, (s, sg.as_dict(), so.as_dict())
        assert sg.converged
    _compare(g, o, (n, n, n))


@pytest.mark.parametrize("n", [(512, 8, 8), (128, 8, 8)])
@pytest.mark.parametrize("flexible", [0, 1])
def test_wide_rows_mg_fp64(hip, oracle_mod, flexible, n):
    """Rows of 512 cells (the vector smoother) and of 128 cells with ny % 8 == 0
    (two rows per wave): the V-cycle's last fine half sweep then also forms the
    PCG's r.z (and z.Ad) row partials; FP64 steps exact against the oracle,
    iteration counts equal."""
    from mydemcfd_amd import cases

    print(f"case: periodic {n} mg fp64 flexible={flexible} x3", flush=True)
    desc = cases.periodic_desc(n)
    g, o = _pair(hip, oracle_mod, desc)
    rng = np.random.default_rng(11)
    _upload_both(g, o, {f: rng.standard_normal(g.shape(f)) for f in ("u", "v", "w")})
    for s in range(3):
        sg = g.step(1e-3, precond="mg", tol=1e-10, max_iter=300, flexible=flexible)
        so = o.step(1e-3, precond="mg", tol=1e-10, max_iter=300, flexible=flexible)
        assert sg.iters == so.iters and sg.converged, (s, sg.as_dict(), so.as_dict())
    _compare(g, o, n, tol=0.0)


def test_cfl(hip, oracle_mod):
    from mydemcfd_amd import cases

    print("case: cfl periodic 16^3", flush=True)
    desc = cases.periodic_desc(16, re=50.0)
    g, o = _pair(hip, oracle_mod, desc)
    u, v, w = cases.tg_fields(16, pert="white", pert_rms=1e-2)
    _upload_both(g, o, {"u": u, "v": v, "w": w})
    assert g.calc_cfl(1e-2) == o.calc_cfl(1e-2)


@pytest.mark.timeout(600)
def test_c1_cavity_64_100_steps(hip, oracle_mod):
    """Config 1: 64^3 lid-driven cavity, Re=100, dt=1e-3, 100 SMAC steps,
    Jacobi-PCG (pure Neumann) to 1e-8 relative residual."""
    from mydemcfd_amd import cases

    print("case: C1 cavity 64^3 x100 jacobi rel_l2 1e-8", flush=True)
    desc = cases.cavity_desc(64, re=100.0)
    g, o = _pair(hip, oracle_mod, desc)
    iters = []
    for s in range(100):
        sg = g.step(1e-3, precond="jacobi", stop_rule="rel_l2", tol=1e-8)
        so = o.step(1e-3, precond="jacobi", stop_rule="rel_l2", tol=1e-8)
        assert sg.iters == so.iters, (s, sg.iters, so.iters)
        iters.append(sg.iters)
        if s % 10 == 0:
            print(f"C1 step {s} iters {sg.iters}", flush=True)
    errs = _compare(g, o, desc["n"])
    print("C1 errors", errs, "iters first/last", iters[0], iters[-1], flush=True)


# ------------------------------------------------------------ mixed precision
@pytest.mark.parametrize("precision", ["fp32", "bf16"])
@pytest.mark.parametrize("kind", ["periodic", "cavity", "periodic_wide", "cavity_wide"])
def test_mixed_precision_vs_fp64_oracle(hip, oracle_mod, precision, kind):
    """Mixed forms (V-cycle in FP32, or BF16 storage with FP32 arithmetic,
    inside a flexible FP64 PCG; FP64 residual, dots and coarsest solve) solve to
    the same 1e-8 relative residual; their fields differ from the FP64 oracle
    only at that tolerance. The error is printed and bounded. The *_wide boxes
    (256-cell rows) take the vector (several cells per lane) smoother."""
    from mydemcfd_amd import cases

    n = 32
    if kind == "periodic":
        desc = cases.periodic_desc(n)
    elif kind == "cavity":
        desc = cases.cavity_desc(n, re=100.0)
    elif kind == "periodic_wide":
        desc = cases.periodic_desc((256, 16, 8), length=1.0)
    else:
        desc = cases.cavity_desc((256, 8, 8), re=100.0)
    print(f"case: mixed {precision} {kind} x4", flush=True)
    g, o = _pair(hip, oracle_mod, desc)
    if kind == "periodic":
        u, v, w = cases.tg_fields(n, pert="white")
        _upload_both(g, o, {"u": u, "v": v, "w": w})
        dt = cases.cfl_dt(o, 0.3)
    elif kind == "periodic_wide":
        rng = np.random.default_rng(9)
        _upload_both(g, o, {f: 0.1 * rng.standard_normal(g.shape(f)) for f in ("u", "v", "w")})
        dt = 2e-3
    else:
        dt = 2e-3
    for s in range(4):
        sg = g.step(dt, precond="mg", precision=precision, flexible=1, tol=1e-8)
        so = o.step(dt, precond="mg", tol=1e-8)
        assert sg.converged and sg.rel_res < 1e-8, sg.as_dict()
        # BF16 smoothing (8-bit mantissa) costs iterations on the 32:1
        # anisotropic cells of cavity_wide (153 against 108 in FP64)
        slack = max(6, so.iters // 2) if precision == "bf16" else 6
        assert sg.iters <= so.iters + slack, (sg.iters, so.iters)
    # velocity errors against the velocity scale of the whole field (a component
    # that is nearly zero, w in the 256x8x8 cavity, would inflate its own ratio)
    vel = {f: (g.download(f), o.download(f)) for f in ("u", "v", "w")}
    vscale = max(float(np.max(np.abs(b))) for _, b in vel.values())
    errs = {f: float(np.max(np.abs(a - b))) / vscale for f, (a, b) in vel.items()}
    errs["p"] = rel_err_mod_mean(g.download("p"), o.download("p"), desc["n"])
    print(f"mixed {precision} {kind}: error vs FP64 oracle {errs}, iters {sg.iters} vs {so.iters}", flush=True)
    for f, e in errs.items():
        assert e < 1e-6, (f, e)


def test_gpu_failure_handling(hip):
    from mydemcfd_amd import _abi, cases

    print("case: failure handling", flush=True)
    c = hip.SMACContext(**cases.cavity_desc(8))
    st = c.step(1e-3, precond="jacobi", tol=0.0, max_iter=3)
    assert st.converged == 0 and st.iters == 3
    with pytest.raises(_abi.SmacError) as e:
        c.step(1e-3, precond="mg", mg_levels=5)
    assert e.value.status == _abi.SMAC_ERR_INVALID_GRID
    with pytest.raises(_abi.SmacError) as e:
        c.step(1e-3, precond="jacobi", precision="fp32")
    assert e.value.status == _abi.SMAC_ERR_UNSUPPORTED
    with pytest.raises(_abi.SmacError) as e:
        hip.SMACContext(**cases.periodic_desc(3))
    assert e.value.status == _abi.SMAC_ERR_INVALID_GRID


@pytest.mark.parametrize("mode", ["jacobi", "gmg", "rbgs"])
def test_cpp_facade_driver(hip, oracle_mod, tmp_path, mode):
    """A C++ driver shaped like CFD/main.cu:225-317 (tests/cpp/cavity_facade.cpp),
    built against include/smac_facade.hpp, runs the cavity through the C-ABI
    with the reference's stopping rule (max|r|*dt < tol_, every 10 / 5
    iterations); its fields equal the oracle's."""
    import os
    import subprocess

    from mydemcfd_amd import cases

    driver = cpp_drivers.facade_driver()
    n = 24
    print(f"case: C++ facade cavity {n}^3 x5 {mode}", flush=True)
    args = [driver, str(n), "5", str(tmp_path), mode]
    out = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    o = oracle_mod.OracleContext(**cases.cavity_desc(n, re=100.0))
    precond = {"jacobi": "jacobi", "gmg": "ref_gmg", "rbgs": "mg"}[mode]
    for _ in range(5):
        o.step(1e-3, precond=precond, stop_rule="ref_linf_dt", tol=1e-8,
               check_every=10 if mode == "jacobi" else 5, mg_levels=0 if mode == "jacobi" else 4)
        o.update_boundary_ghost(1e-3)
        o.make_face_gradp()
    for f in ("u", "v", "w", "p", "gradpx", "gradpy", "gradpz"):
        a = np.fromfile(os.path.join(tmp_path, f + ".f64"), dtype=np.float64).reshape(o.shape(f))
        b = o.download(f)
        err = rel_err_mod_mean(a, b, (n, n, n)) if f == "p" else rel_err(a, b)
        assert err <= TOL, (f, err)


@pytest.mark.parametrize("n", [(256, 16, 8), (512, 8, 8), (384, 8, 8)])
def test_periodic_mg_wide_rows(hip, oracle_mod, n):
    """Anisotropic periodic boxes with 256- and 512-cell rows: the shuffle/LDS
    paths of the smoother, residual, SpMV and prolongation at four and eight
    cells per lane, the pair-order PCG kernels with two and four chunks per row
    (the x wrap across chunk edges), and the tiled predictor; 384-cell rows
    take the pair-order kernels with three chunks and the general SpMV."""
    from mydemcfd_amd import cases

    print(f"case: TG-like periodic {n} x3 mg", flush=True)
    desc = cases.periodic_desc(n)
    g, o = _pair(hip, oracle_mod, desc)
    rng = np.random.default_rng(21)
    fields = {f: 0.1 * rng.standard_normal(g.shape(f)) for f in ("u", "v", "w")}
    _upload_both(g, o, fields)
    for s in range(3):
        sg = g.step(2e-3, precond="mg", tol=1e-8)
        so = o.step(2e-3, precond="mg", tol=1e-8)
        assert sg.iters == so.iters, (s, sg.as_dict(), so.as_dict())
    _compare(g, o, n)


def test_cavity_mg_tiled(hip, oracle_mod):
    """Walled box of 64-cell rows: every level takes the load-per-neighbour
    row kernels (the shuffle/LDS paths need rows of 128 cells or more)."""
    from mydemcfd_amd import cases

    print("case: cavity 64x32x24 x3 mg", flush=True)
    n = (64, 32, 24)
    desc = cases.cavity_desc(n, re=100.0)
    g, o = _pair(hip, oracle_mod, desc)
    for s in range(3):
        sg = g.step(2e-3, precond="mg", tol=1e-8)
        so = o.step(2e-3, precond="mg", tol=1e-8)
        assert sg.iters == so.iters and sg.mg_levels == so.mg_levels, (s, sg.as_dict(), so.as_dict())
    _compare(g, o, n)


@pytest.mark.parametrize("n", [(128, 16, 12), (256, 8, 8), (512, 8, 8)])
def test_cavity_mg_wide_rows(hip, oracle_mod, n):
    """Walled box whose fine level takes the shuffle/LDS paths of the
    smoother, red residual, SpMV and prolongation (rows of 128, 256 and 512
    cells, ny % 4 == 0): the wall ghosts of the x shuffles and the LDS halo
    rows."""
    from mydemcfd_amd import cases

    print(f"case: cavity {n} x3 mg", flush=True)
    desc = cases.cavity_desc(n, re=100.0)
    g, o = _pair(hip, oracle_mod, desc)
    for s in range(3):
        sg = g.step(2e-3, precond="mg", tol=1e-8)
        so = o.step(2e-3, precond="mg", tol=1e-8)
        assert sg.iters == so.iters and sg.mg_levels == so.mg_levels, (s, sg.as_dict(), so.as_dict())
    _compare(g, o, n)


def test_two_phase_dam_break_wide_rows(hip, oracle_mod):
    """The two-phase dam break on 128-cell rows: the variable-coefficient
    shuffle/LDS paths of the smoother, residual and SpMV."""
    from mydemcfd_amd import cases

    n = (128, 8, 12)
    print(f"case: dam break {n} x2 mg", flush=True)
    desc = cases.dam_break_desc(n, length=(4.0, 0.25, 0.375))
    g, o = _pair(hip, oracle_mod, desc)
    _upload_both(g, o, {"alpha": cases.dam_break_alpha(n, length=(4.0, 0.25, 0.375), x_front=1.0, height=0.2)})
    kw = dict(precond="mg", tol=1e-8, max_iter=3000)
    for s in range(2):
        sg = g.step(1e-3, **kw)
        so = o.step(1e-3, **kw)
        assert sg.iters == so.iters and sg.converged == so.converged, (s, sg.as_dict(), so.as_dict())
    _compare(g, o, n)


@pytest.mark.parametrize("case", ["cavity24", "walls48x24x12", "periodic32"])
def test_ref_gmg_steps(hip, oracle_mod, case):
    """The reference's GMG-PCG (weighted-Jacobi V-cycle, mean restriction,
    constant prolongation, r mean removed each iteration, max|r|*dt stop
    every 5 iterations): iterations and fields equal the oracle's."""
    from mydemcfd_amd import cases

    print(f"case: ref_gmg {case} x4", flush=True)
    if case == "periodic32":
        n = (32, 32, 32)
        desc = cases.periodic_desc(n)
    else:
        n = (24, 24, 24) if case == "cavity24" else (48, 24, 12)
        desc = cases.cavity_desc(n, re=100.0)
    g, o = _pair(hip, oracle_mod, desc)
    if case == "periodic32":
        u, v, w = cases.tg_fields(32, pert="white")
        _upload_both(g, o, {"u": u, "v": v, "w": w})
        dt = cases.cfl_dt(o, 0.3)
    else:
        dt = 1e-3
    kw = dict(precond="ref_gmg", stop_rule="ref_linf_dt", tol=1e-8, check_every=5)
    for s in range(4):
        sg = g.step(dt, **kw)
        so = o.step(dt, **kw)
        assert sg.iters == so.iters and sg.mg_levels == so.mg_levels, (s, sg.as_dict(), so.as_dict())
        assert sg.converged
    _compare(g, o, n)


def test_ref_gmg_rejects(hip):
    from mydemcfd_amd import _abi, cases

    print("case: ref_gmg invalid grids and precision", flush=True)
    c = hip.SMACContext(**cases.cavity_desc(16))
    with pytest.raises(_abi.SmacError) as e:
        c.step(1e-3, precond="ref_gmg")
    assert e.value.status == _abi.SMAC_ERR_INVALID_GRID
    c = hip.SMACContext(**cases.cavity_desc(24))
    with pytest.raises(_abi.SmacError) as e:
        c.step(1e-3, precond="ref_gmg", precision="fp32")
    assert e.value.status == _abi.SMAC_ERR_UNSUPPORTED


def _bubble_alpha(shape, n, radius=0.25):
    """alpha = 1 inside a sphere at the box centre (cell centres, periodic box
    of unit cells), reference cell layout with periodic ghosts."""
    z, y, x = np.meshgrid(*[(np.arange(m + 2) - 0.5) / m for m in n[::-1]], indexing="ij")
    a = (((x % 1.0) - 0.5) ** 2 + ((y % 1.0) - 0.5) ** 2 + ((z % 1.0) - 0.5) ** 2 < radius ** 2)
    return a.astype(np.float64).reshape(shape)


@pytest.mark.parametrize("precond", ["jacobi", "mg", "ref_gmg"])
def test_two_phase_dam_break(hip, oracle_mod, precond):
    """Two-phase column (rho1/rho0 = 1000, mu1/mu0 = 100, gravity -z) in a
    closed box: properties from alpha on the GPU, variable-coefficient
    predictor, operator, preconditioners and correction equal the oracle."""
    from mydemcfd_amd import cases

    n = (32, 16, 16)
    print(f"case: dam break {n} x3 {precond}", flush=True)
    desc = cases.dam_break_desc(n)
    g, o = _pair(hip, oracle_mod, desc)
    _upload_both(g, o, {"alpha": cases.dam_break_alpha(n)})
    kw = dict(precond=precond, tol=1e-8, max_iter=3000)
    if precond == "ref_gmg":
        kw.update(stop_rule="ref_linf_dt", check_every=5)
    for s in range(3):
        sg = g.step(1e-3, **kw)
        so = o.step(1e-3, **kw)
        assert sg.iters == so.iters and sg.converged == so.converged, (s, sg.as_dict(), so.as_dict())
    _compare(g, o, n)
    cg, co = g.calc_cfl(1e-3), o.calc_cfl(1e-3)
    assert cg == co


def test_two_phase_periodic_bubble_mg(hip, oracle_mod):
    from mydemcfd_amd import cases

    n = (32, 32, 32)
    print("case: periodic heavy bubble 32^3 x3 mg", flush=True)
    desc = cases.periodic_desc(n, length=1.0)
    desc.update(rho=(1.0, 10.0), mu=(1e-3, 1e-2), g=(0.0, 0.0, -1.0))
    g, o = _pair(hip, oracle_mod, desc)
    _upload_both(g, o, {"alpha": _bubble_alpha(g.shape("alpha"), n)})
    for s in range(3):
        sg = g.step(2e-3, precond="mg", tol=1e-8)
        so = o.step(2e-3, precond="mg", tol=1e-8)
        assert sg.iters == so.iters, (s, sg.as_dict(), so.as_dict())
    _compare(g, o, n)


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_two_phase_mixed_vs_fp64_oracle(hip, oracle_mod, precision):
    from mydemcfd_amd import cases

    n = (32, 16, 16)
    print(f"case: dam break {n} x3 mg {precision}", flush=True)
    desc = cases.dam_break_desc(n)
    g, o = _pair(hip, oracle_mod, desc)
    _upload_both(g, o, {"alpha": cases.dam_break_alpha(n)})
    for s in range(3):
        sg = g.step(1e-3, precond="mg", precision=precision, flexible=1, tol=1e-8, max_iter=3000)
        so = o.step(1e-3, precond="mg", tol=1e-8, max_iter=3000)
        # BF16 smoothing is weaker across the 1000:1 density jump: more iterations, same tolerance
        assert sg.converged and sg.iters <= 2 * so.iters, (s, sg.as_dict(), so.as_dict())
    _compare(g, o, n, tol=1e-6)


@pytest.mark.parametrize("case", ["walls_mixed", "periodic_xz", "two_phase"])
def test_post_step_ghosts_and_gradp(hip, oracle_mod, case):
    """update_boundary_ghost + make_face_gradp from identical random state
    (ghosts included): every entry of u, v, w, p and of the three face
    gradients equals the oracle's."""
    from mydemcfd_amd import _abi, cases

    print(f"case: post-step ghosts and face gradient, {case}", flush=True)
    rng = np.random.default_rng(31)
    if case == "walls_mixed":
        n = (12, 10, 8)
        desc = dict(n=n, size=(1.2, 1.0, 0.8), origin=(0.0, 0.0, 0.0), periodic=(0, 0, 0),
                    side_bcid=(2, 1, 3, 1, 0, 1),
                    bc_table=((_abi.BC_SLIP, 0, 0, 0), (_abi.BC_NOSLIP, 0.3, -0.2, 0.1),
                              (_abi.BC_INFLOW, 0.5, 0.0, 0.0), (_abi.BC_SLIP, 0, 0, 0)),
                    rho=(2.0, 2.0), mu=(0.01, 0.01), g=(0.5, -9.81, 1.25))
    elif case == "periodic_xz":
        n = (16, 8, 12)
        desc = dict(n=n, size=(1.0, 0.5, 0.75), origin=(0.0, 0.0, 0.0), periodic=(1, 0, 1),
                    side_bcid=(0, 0, 1, 1, 0, 0),
                    bc_table=((_abi.BC_NOSLIP, 0, 0, 0), (_abi.BC_NOSLIP, 1.0, 0.0, 0.0)),
                    rho=(1.0, 1.0), mu=(0.01, 0.01), g=(0.0, -1.0, 0.0))
    else:
        n = (16, 8, 8)
        desc = cases.dam_break_desc(n, length=(2.0, 1.0, 1.0))
    g, o = _pair(hip, oracle_mod, desc)
    fields = {f: rng.standard_normal(g.shape(f)) for f in ("u", "v", "w", "p")}
    if case == "two_phase":
        fields["alpha"] = cases.dam_break_alpha(n)
    _upload_both(g, o, fields)
    for c in (g, o):
        c.update_boundary_ghost(1e-3)
        c.make_face_gradp()
    for f in ("u", "v", "w", "p", "gradpx", "gradpy", "gradpz"):
        a, b = g.download(f), o.download(f)
        diff = float(np.max(np.abs(a - b)))
        assert diff == 0.0, (f, diff)


@pytest.mark.parametrize("case", ["dam_break", "periodic_sphere"])
def test_vof_transport_and_step(hip, oracle_mod, case):
    """Split THINC/WLIC transport (2 substeps, the direction order rotating
    with the step), property update and a full step with the alpha flux in the
    mass flux, four times. The transport's exp/log/cosh/acos are explicit
    operation sequences shared bit for bit by the device and the CPU
    restatement, so every field is identical (40 MG-PCG iterations, tol 0)."""
    from mydemcfd_amd import _abi, cases

    print(f"case: VoF transport + step, {case}", flush=True)
    if case == "dam_break":
        n = (32, 16, 16)
        desc = cases.dam_break_desc(n, length=(2.0, 1.0, 1.0))
        alpha = cases.dam_break_alpha(n)
        fields = {}
        dt = 2e-3
    else:
        n = (16, 16, 16)
        desc = dict(n=n, size=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0), periodic=(1, 1, 1), side_bcid=(0,) * 6,
                    bc_table=((_abi.BC_NOSLIP, 0.0, 0.0, 0.0),), rho=(1.0, 1000.0), mu=(1e-3, 0.05),
                    g=(0.0, 0.0, 0.0))
        h = 1.0 / 16
        x = (np.arange(16) + 0.5) * h
        X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
        d = np.sqrt((X - 0.5) ** 2 + (Y - 0.45) ** 2 + (Z - 0.55) ** 2)
        inside = np.clip((0.3 - d) / h + 0.5, 0.0, 1.0)
        alpha = cases._embed_periodic(np.transpose(inside, (2, 1, 0)), _abi.FIELD_ALPHA)
        u, v, w = cases.white_solenoidal(16, seed=5, rms=0.5, length=1.0)
        fields = {f: cases._embed_periodic(a, _abi.FIELD_NAMES[f]) for f, a in zip("uvw", (u, v, w))}
        dt = 5e-3
    g, o = _pair(hip, oracle_mod, desc)
    fields["alpha"] = alpha
    _upload_both(g, o, fields)
    for c in (g, o):
        c.update_properties(1)
    opts = dict(precond="mg", stop_rule="rel_l2", tol=0.0, max_iter=40)
    for s in range(4):
        for c in (g, o):
            c.transport_alpha(dt, 2, s)
            c.step(dt, **opts)
        for f in ("alpha", "fluxx", "fluxy", "fluxz", "u", "v", "w", "p"):
            diff = float(np.max(np.abs(g.download(f) - o.download(f))))
            assert diff == 0.0, (s, f, diff)


@pytest.mark.parametrize("case", ["dam_break", "periodic", "cavity_wide"])
def test_two_way_steps(hip, oracle_mod, case):
    """CFD-DEM two-way steps in the coupled loop's order (Coupling/main.cu:517-549):
    (two-phase case) VoF transport and properties, new volume fractions and
    particle impulses, the Poisson coefficients eps_f/rho_f, the two-way
    predictor with its eps-weighted mass flux (and the alpha flux), the
    volume-fraction rhs, MG-PCG and the correction -- three steps, every field
    identical to the oracle (40 MG-PCG iterations, tol 0). cavity_wide runs the
    variable-coefficient shuffle/LDS smoother, residual and SpMV on 128-cell rows."""
    from mydemcfd_amd import cases

    print(f"case: two-way steps, {case}", flush=True)
    alpha = None
    if case == "dam_break":
        n = (16, 8, 8)
        desc = cases.dam_break_desc(n, length=(2.0, 1.0, 1.0), rho=(1.0, 50.0), mu=(1e-3, 1e-2))
        alpha = cases.dam_break_alpha(n)
        dt = 1e-3
    elif case == "periodic":
        n = (16, 16, 16)
        desc = cases.periodic_desc(n, re=400.0, length=1.0)
        dt = 2e-3
    else:
        n = (128, 8, 12)
        desc = cases.cavity_desc(n, re=100.0)
        dt = 1e-3
    per = desc["periodic"]
    g, o = _pair(hip, oracle_mod, desc)
    rng = np.random.default_rng(61)
    fields = {f: 0.2 * rng.standard_normal(g.shape(f)) for f in ("u", "v", "w")}
    if alpha is not None:
        fields["alpha"] = alpha
    _upload_both(g, o, fields)
    for c in (g, o):
        c.update_properties(1)
    opts = dict(precond="mg", stop_rule="rel_l2", tol=0.0, max_iter=40)
    eps_old = cases.void_fraction(n, per, 100)
    for s in range(3):
        eps = cases.void_fraction(n, per, 101 + s)
        tw = {"eps": eps, "eps_old": eps_old, "eps_half": 0.5 * (eps + eps_old)}
        for f in ("impx", "impy", "impz"):
            tw[f] = 1e-6 * rng.standard_normal(g.shape(f))
        for c in (g, o):
            if alpha is not None:
                c.transport_alpha(dt, 1, s)
                c.update_properties(0)
            for f, a in tw.items():
                c.upload(f, a)
            c.update_poisson_beta_two_way()
            c.predictor_two_way(dt)
        sg = g.pressure_solve(dt, **opts)
        so = o.pressure_solve(dt, **opts)
        assert sg.iters == so.iters, (s, sg.as_dict(), so.as_dict())
        for c in (g, o):
            c.correct(dt)
        for f in ("ustar", "vstar", "wstar", "rhs", "u", "v", "w", "p"):
            diff = float(np.max(np.abs(g.download(f) - o.download(f))))
            assert diff == 0.0, (s, f, diff)
        eps_old = eps


def test_config_driver_dam_break(hip, oracle_mod, tmp_path):
    """The reference's time loop (CFD/main.cu:258-346) over the C++ facade,
    driven by a JSON configuration (tests/golden/dam_break_config.json):
    set-up, CFL time steps, VoF substeps, properties, predictor, reference-GMG
    PCG, correction, ghosts, face gradient, VTI output. The same loop on the
    oracle, from initial fields computed independently in Python, gives the
    same time steps and fields (1e-8 relative: the VoF transport calls
    exp/log/cosh/acos), and the final VTI holds the oracle's cell data."""
    import json
    import os
    import subprocess

    import config_ref
    from mydemcfd_amd import cfdtime

    driver = cpp_drivers.config_driver()
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "dam_break_config.json")
    steps = 8
    print(f"case: config driver dam break 32x16x16 x{steps}", flush=True)
    out = subprocess.run([driver, path, str(tmp_path), str(steps)], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    rows = [ln.split() for ln in out.stdout.splitlines() if ln.startswith("step ")]
    assert len(rows) == steps
    with open(path) as f:
        cfg = json.load(f)
    desc = config_ref.desc(cfg)
    o = oracle_mod.OracleContext(**desc)
    alpha0, v0 = config_ref.initial_fields(cfg)
    o.upload("alpha", alpha0)
    o.upload("v", v0)
    o.update_properties(1)
    t = cfg["time"]
    ct = cfdtime.CFDTime(t["initial_dt"], t["maximum_dt"], t["interval"], t["end_time"], t["cfl"])
    substeps = max(1, int(np.ceil(t["cfl"] / t["vof_cfl"])))
    levels = cfg["pressure_solver"]["gmg_levels"]
    dt_prev = 0.001  # the facade's (and the reference's) G_SMACSolver::dt_ before the first step
    for s in range(steps):
        tol = 1e-11 if s == 0 else cfg["pressure_solver"]["tol"]
        cc, cv = o.calc_cfl(dt_prev)
        dt = ct.update_time(cc if cc > cv else cv)
        dt_prev = dt
        assert abs(dt - float(rows[s][3])) <= 1e-9 * dt, (s, dt, rows[s])
        o.clear_alpha_flux_accum()
        for _ in range(substeps):
            o.alpha_flux_split(dt / substeps, s)
        o.update_properties(0)
        o.predictor(dt)
        o.pressure_solve(dt, precond="ref_gmg", stop_rule="ref_linf_dt", tol=tol, check_every=5, max_iter=2000,
                         mg_levels=levels, ref_sweeps=4, ref_coarse_sweeps=30)
        o.correct(dt)
        o.update_boundary_ghost(dt)
        o.make_face_gradp()
    n = desc["n"]
    for f in ("u", "v", "w", "alpha"):
        a = np.fromfile(os.path.join(tmp_path, f + ".f64")).reshape(o.shape(f))
        err = rel_err(a, o.download(f))
        assert err <= 1e-8, (f, err)
    a = np.fromfile(os.path.join(tmp_path, "p.f64")).reshape(o.shape("p"))
    assert rel_err_mod_mean(a, o.download("p"), n) <= 1e-8
    attrs, arrs = config_ref.read_vti(os.path.join(tmp_path, "result_999999.vti"))
    p_o, al_o, vel_o = o.cell_output()
    assert attrs["extent"] == [0, n[0], 0, n[1], 0, n[2]]
    for name, ref in (("alpha", al_o), ("velocity", vel_o)):
        got = arrs[name][0]
        err = float(np.max(np.abs(got - ref.ravel()))) / max(float(np.max(np.abs(ref))), 1e-30)
        assert err <= 1e-6, (name, err)
    assert os.path.exists(os.path.join(tmp_path, "result_000000.vti"))
