"""The benchmark configurations that run at their own size on the GPU.

* C2: triply periodic 256^3 Taylor-Green plus the discrete curl of a seeded
  white-noise edge potential (seed 0x5EED0002, RMS 1e-3), Re 1600, the time
  step renewed every step from CFL 0.3, 20 steps, FP64 RBGS-MG-PCG V(2,2) with
  6 levels (256 -> 8) to a relative residual of 1e-8: every step's iteration
  count and dt, and u, v, w, p after 20 steps, equal the CPU FP64 oracle's
  (bar 1e-10 relative; p modulo its mean).
* C5: the VoF dam break on a 256x128x128 grid ([0,2]x[0,1]x[0,1], h = 1/128;
  heavy column x < 0.25, z < 0.5; density ratio 1000:1, viscosity ratio
  100:1, gravity -9.81 z; no-slip walls; dt = 1e-3, two VoF substeps per
  step). Five FP64 steps (transport + step, MG-PCG to 1e-8) equal the oracle
  (alpha, u, v, w, p). Then 50 steps in each mixed form (FP32 and BF16
  V-cycles in the flexible FP64 PCG) next to 50 FP64 steps on the GPU: every
  step converges, the projection identity ||D u||/||D u*|| = rel_res holds
  after the last step, the heavy phase's volume changes by less than 1e-9 of
  itself, and the mixed fields stay within 1e-6 of the FP64 fields (relative
  to each field's scale).
"""
import numpy as np
import pytest

from conftest import rel_err, rel_err_mod_mean

pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.mark.timeout(900)
def test_c2_taylor_green_256_20_steps(hip, oracle_mod):
    from mydemcfd_amd import cases

    n = 256
    print(f"case: C2 TG 256^3 + white curl 1e-3, CFL 0.3, 20 steps, FP64 MG-PCG (oracle threads "
          f"{oracle_mod.threads()})", flush=True)
    desc = cases.periodic_desc(n, re=1600.0)
    g = hip.SMACContext(**desc)
    o = oracle_mod.OracleContext(**desc)
    u, v, w = cases.tg_fields(n, seed=0x5EED0002, pert="white", pert_rms=1e-3)
    for c in (g, o):
        c.upload("u", u)
        c.upload("v", v)
        c.upload("w", w)
    del u, v, w
    kw = dict(precond="mg", stop_rule="rel_l2", tol=1e-8)
    for s in range(20):
        dt = cases.cfl_dt(o, 0.3)
        assert cases.cfl_dt(g, 0.3) == dt, s
        sg = g.step(dt, **kw)
        so = o.step(dt, **kw)
        assert sg.mg_levels == 6 and so.mg_levels == 6
        assert sg.iters == so.iters and sg.converged and so.converged, (s, sg.as_dict(), so.as_dict())
        print(f"C2 step {s} dt {dt:.6e} iters {sg.iters} rel_res {sg.rel_res:.3e}", flush=True)
    errs = {f: rel_err(g.download(f), o.download(f)) for f in ("u", "v", "w")}
    errs["p"] = rel_err_mod_mean(g.download("p"), o.download("p"), (n, n, n))
    print("C2 errors", errs, flush=True)
    for f, e in errs.items():
        assert e <= TOL, (f, e)


C5_N = (256, 128, 128)
C5_DT = 1e-3
C5_SUB = 2


def _c5_pair(hip, oracle_mod):
    from mydemcfd_amd import cases

    desc = cases.dam_break_desc(C5_N)
    g = hip.SMACContext(**desc)
    o = oracle_mod.OracleContext(**desc)
    al = cases.dam_break_alpha(C5_N)
    for c in (g, o):
        c.upload("alpha", al)
        c.update_properties(1)
    return g, o


def _div_norm_walls(torch, ctx, names, h):
    """||D u||_2 over the interior cells of a closed box (faces 1..N+1 of each
    axis, wall faces included)."""
    nx, ny, nz = ctx.n
    acc = None
    for axis, f in enumerate(names):
        t = torch.empty(ctx.shape(f), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        ctx.download_device(f, t.data_ptr(), t.numel())
        sl = [slice(1, nz + 1), slice(1, ny + 1), slice(1, nx + 1)]
        hi = list(sl)
        dim = 2 - axis
        hi[dim] = slice(2, (nz, ny, nx)[dim] + 2)
        d = (t[tuple(hi)] - t[tuple(sl)]) / h[axis]
        acc = d if acc is None else acc + d
        del t
    return float(torch.linalg.vector_norm(acc))


def _heavy_volume(ctx):
    nx, ny, nz = ctx.n
    return float(np.sum(ctx.download("alpha")[1:nz + 1, 1:ny + 1, 1:nx + 1]))


@pytest.mark.timeout(900)
def test_c5_dam_break_fp64_vs_oracle(hip, oracle_mod):
    print(f"case: C5 dam break {C5_N}, 5 FP64 steps vs oracle", flush=True)
    g, o = _c5_pair(hip, oracle_mod)
    kw = dict(precond="mg", stop_rule="rel_l2", tol=1e-8, max_iter=3000)
    for s in range(5):
        for c in (g, o):
            c.transport_alpha(C5_DT, C5_SUB, s)
        sg = g.step(C5_DT, **kw)
        so = o.step(C5_DT, **kw)
        assert sg.iters == so.iters and sg.converged and so.converged, (s, sg.as_dict(), so.as_dict())
        print(f"C5 step {s} iters {sg.iters} rel_res {sg.rel_res:.3e}", flush=True)
    errs = {f: rel_err(g.download(f), o.download(f)) for f in ("alpha", "u", "v", "w")}
    errs["p"] = rel_err_mod_mean(g.download("p"), o.download("p"), C5_N)
    print("C5 FP64 errors", errs, flush=True)
    for f, e in errs.items():
        assert e <= TOL, (f, e)


@pytest.mark.timeout(900)
def test_c5_dam_break_mixed_50_steps(hip):
    torch = pytest.importorskip("torch")
    from mydemcfd_amd import cases

    print(f"case: C5 dam break {C5_N}, 50 steps FP64 / mixed FP32 / mixed BF16 on the GPU", flush=True)
    desc = cases.dam_break_desc(C5_N)
    al = cases.dam_break_alpha(C5_N)
    h = [desc["size"][d] / C5_N[d] for d in range(3)]
    ref = {}
    for form, prec, flex in (("fp64", "fp64", 0), ("mixed_fp32", "fp32", 1), ("mixed_bf16", "bf16", 1)):
        ctx = hip.SMACContext(**desc)
        ctx.upload("alpha", al)
        ctx.update_properties(1)
        v0 = _heavy_volume(ctx)
        opts = ctx.opts(precond="mg", precision=prec, stop_rule="rel_l2", tol=1e-8, flexible=flex, max_iter=3000)
        iters = []
        for s in range(50):
            ctx.transport_alpha(C5_DT, C5_SUB, s)
            st = ctx.step(C5_DT, opts)
            assert st.converged, (form, s, st.as_dict())
            iters.append(st.iters)
        ns = _div_norm_walls(torch, ctx, ("ustar", "vstar", "wstar"), h)
        nu = _div_norm_walls(torch, ctx, ("u", "v", "w"), h)
        ratio = nu / ns
        dv = abs(_heavy_volume(ctx) - v0) / v0
        print(f"  {form}: iters {iters[0]}..{iters[-1]} (mean {sum(iters) / len(iters):.1f}) rel_res "
              f"{st.rel_res:.3e} |D u|/|D u*| {ratio:.3e} volume drift {dv:.2e}", flush=True)
        assert ratio <= 1.05 * st.rel_res + 1e-12 and ratio >= 0.5 * st.rel_res, (form, ratio, st.rel_res)
        assert dv <= 1e-9, (form, dv)
        fields = {f: ctx.download(f) for f in ("u", "v", "w", "p", "alpha")}
        if form == "fp64":
            ref = fields
        else:
            errs = {}
            for f in ("u", "v", "w", "alpha"):
                scale = 1.0 if f == "alpha" else max(float(np.max(np.abs(ref[x]))) for x in ("u", "v", "w"))
                errs[f] = float(np.max(np.abs(fields[f] - ref[f]))) / scale
            errs["p"] = rel_err_mod_mean(fields["p"], ref["p"], C5_N)
            print(f"  {form}: error vs FP64 {errs}", flush=True)
            for f, e in errs.items():
                assert e <= 1e-6, (form, f, e)
        ctx.close()
        del ctx
