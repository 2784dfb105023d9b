"""Adjoint of the VoF loop's step (smac_adjoint_vof_step, autograd.vof_rollout):
the split THINC/WLIC transport of alpha, the property map alpha -> (rho, mu,
beta), the alpha-flux mass flux, the alpha-dependent Poisson operator and the
correction.

The reference has no adjoint; the checks are the mathematics, as for the
one-fluid adjoint (tests/test_gpu_adjoint.py, whose forward conventions are
kept: the forward of the difference quotients solves with 40 stationary
multigrid cycles, a fixed linear map converged to rounding, and the density
ratio is 20:1, where that iteration converges):
  * the transpose identity <w, J d> = <J^T w, d> of one step (transport with
    two substeps, then the step), J d by central differences, for
    alpha -> alpha, alpha -> the whole output, u -> alpha and the full state;
  * the gradient (autograd.vof_rollout) of <w, state after 4 steps> against
    central differences along smooth directions of alpha and of u;
  * the mixed-precision gradient against the FP64 one.
The step is piecewise smooth. The transport's upwind donor, THINC/upwind
switch (|Delta alpha| < 1e-6, alpha within EPS of 0 or 1) and clip are
branches: the checks take a smooth alpha strictly inside (0.15, 0.85) and
monotone along every axis (a tilted tanh layer), so Delta alpha stays far above
the switch on every face and no cell is clipped. (A sharp column, as in the dam
break, sits on the clip's kink in its bulk cells: there the adjoint returns the
derivative of the branch the forward took.) The predictor's MUSCL-minmod slopes
and upwind sides are branches too, one set per face: with ~10^4 of them the
nearest switch along a random direction can lie within ~1e-7 of the base
point, and a central difference that straddles it is off by up to ~1e-5
(measured on a 16x8x8 box: the quotient moves from 8e-6 off the adjoint at
eps = 1e-6 to 8e-8 at eps = 1e-8, the kink on one face). Smaller steps do not
help the transport: the velocity enters the THINC flux through
lambda = u dt/h ~ 1e-4 in differences of O(1) logarithms, so its rounding over
eps * lambda swamps the quotient below eps ~ 1e-6. The adjoint is the
derivative of the branch the forward takes at the base point, so it is
compared with three estimates of that derivative at e = 1e-6, each exact to
O(e^2) when its stencil meets no switch: the central difference and the
second-order one-sided differences (4 (F(e) - F(0)) - (F(2e) - F(0)))/(2e) on
either side. A switch near the base point spoils the central difference and
the one-sided one on its side, the other side stays clean; the one-sided
stencils carry about five times the central one's rounding. The adjoint must
match the closest of the three to 1e-6 (all three are printed). Differences
of the outputs are taken before the cotangent is applied.
"""
import math

import numpy as np
import pytest

from test_gpu_adjoint import TIGHT, TWO_PHASE_RHO, _dot, smooth_field

pytestmark = pytest.mark.gpu

N = (16, 8, 12)
LENGTH = (2.0, 1.0, 1.5)
DT = 1e-3
SUB = 2


def _torch():
    import torch

    return torch


def _desc():
    from mydemcfd_amd import cases

    return cases.dam_break_desc(N, length=LENGTH, rho=TWO_PHASE_RHO, mu=(1e-3, 5e-3))


def smooth_alpha(n=N, length=LENGTH, w=0.6):
    """0.5 + 0.35 tanh(s/w), s = x - 0.45 Lx + 0.3 (y - Ly/2) + 0.2 (z - Lz/2)
    at the cell centres (reference layout with ghosts): in (0.15, 0.85),
    increasing along x, y and z."""
    c = [(np.arange(n[d] + 2) - 0.5) * (length[d] / n[d]) for d in range(3)]
    zz, yy, xx = np.meshgrid(c[2], c[1], c[0], indexing="ij")
    s = xx - 0.45 * length[0] + 0.3 * (yy - 0.5 * length[1]) + 0.2 * (zz - 0.5 * length[2])
    return 0.5 + 0.35 * np.tanh(s / w)


def _masks(desc, torch):
    from mydemcfd_amd import cases

    m = {f: torch.tensor(cases.state_mask(desc["n"], desc["periodic"], f), device="cuda")
         for f in ("u", "v", "w", "p")}
    m["alpha"] = m["p"]
    return m


def _state0(ctx, rng):
    torch = _torch()
    st = {"alpha": smooth_alpha()}
    for f in ("u", "v", "w"):
        st[f] = 0.05 * smooth_field(ctx.shape(f), rng)
    st["p"] = np.zeros(ctx.shape("p"))
    return {f: torch.tensor(a, dtype=torch.float64, device="cuda").contiguous() for f, a in st.items()}


def _forward(ctx, st, opts, nsteps=1):
    """nsteps of transport + step from st; the new state."""
    torch = _torch()
    torch.cuda.synchronize()
    for f, t in st.items():
        ctx.upload_device(f, t.data_ptr(), t.numel())
    ctx.update_properties(True)
    for s in range(nsteps):
        ctx.transport_alpha(DT, SUB, s)
        r = ctx.step(DT, opts)
        assert r.rel_res < (1e-12 if opts.stationary else 1e-9), r.as_dict()
    out = {f: torch.empty_like(t) for f, t in st.items()}
    for f, t in out.items():
        ctx.download_device(f, t.data_ptr(), t.numel())
    torch.cuda.synchronize()
    return out


def _adjoint(ctx, st, wv, opts):
    torch = _torch()
    g = {f: wv[f].clone() for f in wv}
    torch.cuda.synchronize()
    s = ctx.adjoint_vof_step(DT, SUB, 0, st["alpha"].data_ptr(), [st[f].data_ptr() for f in ("u", "v", "w")],
                             st["p"].data_ptr(), g["alpha"].data_ptr(), [g[f].data_ptr() for f in ("u", "v", "w")],
                             g["p"].data_ptr(), opts)
    assert s.rel_res < (1e-12 if opts.stationary else 1e-9), s.as_dict()
    torch.cuda.synchronize()
    return g


def _one_sided(f, ad, eps=1e-6):
    """f(e) = <w, out(x + e d) - out(x)>: the relative distances of the
    adjoint value ad to the second-order one-sided differences on both sides
    and to the central difference."""
    fp1, fp2, fm1, fm2 = f(eps), f(2 * eps), f(-eps), f(-2 * eps)
    plus = (4.0 * fp1 - fp2) / (2 * eps)
    minus = -(4.0 * fm1 - fm2) / (2 * eps)
    central = (fp1 - fm1) / (2 * eps)
    rel = lambda v: abs(v - ad) / abs(ad)
    return rel(plus), rel(minus), rel(central)


def test_vof_adjoint_transpose_identity(hip):
    torch = _torch()
    print(f"case: VoF adjoint transpose identity {N}, {SUB} substeps", flush=True)
    desc = _desc()
    ctx = hip.SMACContext(**desc)
    masks = _masks(desc, torch)
    opts = ctx.opts(**TIGHT)
    rng = np.random.default_rng(21)
    st0 = _state0(ctx, rng)
    d = {f: torch.tensor(smooth_field(t.shape, rng), device="cuda") * masks[f] for f, t in st0.items()}
    gen = torch.Generator(device="cuda")
    gen.manual_seed(21)
    wv = {f: torch.randn(t.shape, generator=gen, device="cuda", dtype=torch.float64) * masks[f]
          for f, t in st0.items()}
    for t in (d["p"], wv["p"]):
        t[masks["p"]] -= t[masks["p"]].mean()
    zero = {f: torch.zeros_like(t) for f, t in st0.items()}
    cases = {
        "alpha->alpha": (dict(zero, alpha=d["alpha"]), dict(zero, alpha=wv["alpha"])),
        "alpha->all": (dict(zero, alpha=d["alpha"]), wv),
        "u->alpha": (dict(zero, u=d["u"], v=d["v"], w=d["w"]), dict(zero, alpha=wv["alpha"])),
        "full": (d, wv),
    }
    base = _forward(ctx, st0, opts)
    worst = 0.0
    for name, (dd, ww) in cases.items():
        def f(e):
            out = _forward(ctx, {k: st0[k] + e * dd[k] for k in st0}, opts)
            return _dot(ww, {k: out[k] - base[k] for k in out}, masks)

        g = _adjoint(ctx, st0, ww, opts)
        for k in g:  # entries that are not state are zero on return
            assert float(g[k][~masks[k]].abs().max()) == 0.0, (name, k)
        rhs = _dot(g, dd, masks)
        rp, rm, rc = _one_sided(f, rhs)
        worst = max(worst, min(rp, rm, rc))
        print(f"  {name}: <J^T w, d> = {rhs:.12e}; one-sided rel {rp:.2e} (+) {rm:.2e} (-), central {rc:.2e}",
              flush=True)
        assert min(rp, rm, rc) <= 1e-6, (name, rp, rm, rc)
    print(f"  worst relative difference {worst:.2e}", flush=True)


def _objective(torch, out, wv):
    """<w, (alpha, u, v, w, p)^N> with fixed random cotangents w (the pressure
    one mean-free)."""
    return sum((o * wv[f]).sum() for o, f in zip(out, ("alpha", "u", "v", "w", "p")))


def _cotangents(torch, st0, masks, seed):
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    wv = {f: torch.randn(t.shape, generator=gen, device="cuda", dtype=torch.float64) * masks[f]
          for f, t in st0.items()}
    wv["p"][masks["p"]] -= wv["p"][masks["p"]].mean()
    return wv


def _rollout_grad(hip, opts_kw, nsteps=4):
    torch = _torch()
    from mydemcfd_amd import autograd

    desc = _desc()
    ctx = hip.SMACContext(**desc)
    masks = _masks(desc, torch)
    opts = ctx.opts(**opts_kw)
    rng = np.random.default_rng(33)
    st0 = _state0(ctx, rng)
    wv = _cotangents(torch, st0, masks, 55)
    leaves = {f: t.clone().requires_grad_() for f, t in st0.items()}
    out = autograd.vof_rollout(ctx, leaves["alpha"], leaves["u"], leaves["v"], leaves["w"], leaves["p"], DT, nsteps,
                               SUB, opts)
    J = _objective(torch, out, wv)
    J.backward()
    grads = {f: leaves[f].grad.detach() for f in leaves}
    for f in grads:  # entries that are not state get no gradient
        assert float(grads[f][~masks[f]].abs().max()) == 0.0, f
    return ctx, desc, masks, opts, st0, grads, float(J.detach()), wv


def test_vof_rollout_gradient_vs_finite_differences(hip):
    torch = _torch()
    nsteps = 4
    print(f"case: VoF rollout gradient {N}, {nsteps} steps x {SUB} substeps vs central differences", flush=True)
    ctx, desc, masks, opts, st0, grads, J0, wv = _rollout_grad(hip, TIGHT, nsteps)
    rng = np.random.default_rng(44)
    base = _forward(ctx, st0, opts, nsteps)
    worst = 0.0
    for r in range(6):
        d = {f: torch.zeros_like(t) for f, t in st0.items()}
        if r < 4:  # alpha directions
            d["alpha"] = torch.tensor(smooth_field(st0["alpha"].shape, rng), device="cuda") * masks["alpha"]
        else:  # velocity directions
            for f in ("u", "v", "w"):
                d[f] = torch.tensor(smooth_field(st0[f].shape, rng), device="cuda") * masks[f]

        def fdiff(e):
            out = _forward(ctx, {k: st0[k] + e * d[k] for k in st0}, opts, nsteps)
            return _dot(wv, {k: out[k] - base[k] for k in out}, masks)

        ad = _dot(grads, d, masks)
        rp, rm, rc = _one_sided(fdiff, ad)
        worst = max(worst, min(rp, rm, rc))
        print(f"  direction {r} ({'alpha' if r < 4 else 'u'}): adjoint {ad:.12e}; one-sided rel {rp:.2e} (+) "
              f"{rm:.2e} (-), central {rc:.2e}", flush=True)
        assert min(rp, rm, rc) <= 1e-6, (r, rp, rm, rc)
    print(f"  <w, state_N> = {J0:.12e}; worst relative difference {worst:.2e}", flush=True)


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_vof_mixed_adjoint_vs_fp64(hip, precision):
    torch = _torch()
    print(f"case: VoF rollout gradient, mixed {precision} against FP64", flush=True)
    base = dict(precond="mg", stop_rule="rel_l2", tol=1e-11, max_iter=200)
    *_, g64, J64, _ = _rollout_grad(hip, dict(base, precision="fp64"), 3)
    *_, gmx, Jmx, _ = _rollout_grad(hip, dict(base, precision=precision, flexible=1, tol=1e-9), 3)
    errs = {f: float((gmx[f] - g64[f]).abs().max() / g64[f].abs().max()) for f in ("alpha", "u", "v", "w")}
    print(f"  J {J64:.12e} / {Jmx:.12e}; gradient error vs FP64 (max rel.) {errs}", flush=True)
    for f, e in errs.items():
        assert e <= 1e-6, (f, e)


@pytest.mark.timeout(900)
def test_c5_column_front_gradient(hip):
    """The C5 configuration: VoF dam break on 256x128x128 (1000:1 density,
    100:1 viscosity, gravity -z, two VoF substeps, dt = 1e-3), 50 steps of the
    mixed FP32 MG-PCG to 1e-8 and the same in FP64; the gradient of the final
    column-front position on the floor, x_f = (h_x / ny) sum alpha over the
    bottom cell layer, with respect to the initial VoF field. The mixed
    gradient is reported against the FP64 one. (Finite differences are not
    the check here: the sharp initial column's bulk cells sit on the clip's
    kink, where the adjoint gives the derivative of the branch taken.)"""
    torch = _torch()
    from mydemcfd_amd import autograd, cases

    n, nsteps = (256, 128, 128), 50
    desc = cases.dam_break_desc(n)
    hx = desc["size"][0] / n[0]
    res = {}
    for prec in ("fp64", "fp32"):
        print(f"case: C5 column-front gradient {n}, {nsteps} steps, {prec}", flush=True)
        ctx = hip.SMACContext(**desc)
        opts = ctx.opts(precond="mg", precision=prec, flexible=1 if prec != "fp64" else 0, stop_rule="rel_l2",
                        tol=1e-8, max_iter=3000)
        st = {"alpha": cases.dam_break_alpha(n)}
        for f in ("u", "v", "w", "p"):
            st[f] = np.zeros(ctx.shape(f))
        leaves = {f: torch.tensor(a, dtype=torch.float64, device="cuda").requires_grad_() for f, a in st.items()}
        import time

        t0 = time.perf_counter()
        out = autograd.vof_rollout(ctx, leaves["alpha"], leaves["u"], leaves["v"], leaves["w"], leaves["p"], DT,
                                   nsteps, SUB, opts)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        front = out[0][1, 1:n[1] + 1, 1:n[0] + 1].sum() * (hx / n[1])
        front.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        g = leaves["alpha"].grad.detach()[1:n[2] + 1, 1:n[1] + 1, 1:n[0] + 1]
        assert bool(torch.isfinite(g).all())
        res[prec] = (float(front.detach()), g)
        print(f"  x_f = {float(front.detach()):.12e}; forward {t1 - t0:.1f} s, backward {t2 - t1:.1f} s; "
              f"|grad|_max {float(g.abs().max()):.4e}, |grad|_2 {float(torch.linalg.vector_norm(g)):.4e}", flush=True)
        del out, leaves, ctx
        torch.cuda.empty_cache()
    f64, g64 = res["fp64"]
    f32, g32 = res["fp32"]
    ef = abs(f32 - f64) / abs(f64)
    eg = float((g32 - g64).abs().max() / g64.abs().max())
    e2 = float(torch.linalg.vector_norm(g32 - g64) / torch.linalg.vector_norm(g64))
    print(f"  mixed FP32 vs FP64: x_f {ef:.2e}, gradient max-norm {eg:.2e}, 2-norm {e2:.2e}", flush=True)
    assert ef <= 1e-9
    assert e2 <= 1e-5 and eg <= 1e-4
