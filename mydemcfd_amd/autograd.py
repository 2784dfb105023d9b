"""torch.autograd binding of SMAC rollouts on the HIP library.

`rollout(ctx, u, v, w, p, dt, nsteps, opts)` advances the state (float64 CUDA
tensors in the reference layout) by `nsteps` SMAC steps through the C-ABI and
returns the new (u, v, w, p). Its backward runs smac_adjoint_step once per
step in reverse order from per-step checkpoints of the velocity (the pressure
enters the step linearly, so it needs no checkpoint). Entries that are not
state (ghosts, periodic images) receive zero gradient.

The forward and backward run on the context's HIP stream; both synchronize
with torch's current stream at entry and exit.
"""
from __future__ import annotations

import torch


def _sync():
    torch.cuda.synchronize()


class SMACRollout(torch.autograd.Function):
    @staticmethod
    def forward(fctx, u, v, w, p, smac, dt, nsteps, opts):
        for t in (u, v, w, p):
            if t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous():
                raise ValueError("rollout state must be contiguous float64 CUDA tensors")
        _sync()
        for name, t in zip(("u", "v", "w", "p"), (u, v, w, p)):
            smac.upload_device(name, t.data_ptr(), t.numel())
        ckpts = []
        stats = []
        for _ in range(nsteps):
            cu = [torch.empty_like(t) for t in (u, v, w)]
            for name, t in zip(("u", "v", "w"), cu):
                smac.download_device(name, t.data_ptr(), t.numel())
            ckpts.append(cu)
            st = smac.step(dt, opts)
            if opts.tol > 0 and not st.converged:  # tol = 0: a fixed iteration count
                raise RuntimeError(f"forward step did not converge: {st.as_dict()}")
            stats.append(st.as_dict())
        outs = [torch.empty_like(t) for t in (u, v, w, p)]
        for name, t in zip(("u", "v", "w", "p"), outs):
            smac.download_device(name, t.data_ptr(), t.numel())
        fctx.smac, fctx.dt, fctx.opts, fctx.ckpts = smac, dt, opts, ckpts
        fctx.forward_stats = stats
        return tuple(outs)

    @staticmethod
    def backward(fctx, gu, gv, gw, gp):
        ref = fctx.ckpts[0] if fctx.ckpts else None
        grads = []
        for g, like in zip((gu, gv, gw), ref if ref else (None, None, None)):
            grads.append(torch.zeros_like(like) if g is None else g.detach().to(torch.float64).contiguous().clone())
        pb = gp.detach().to(torch.float64).contiguous().clone() if gp is not None else None
        if pb is None:
            raise RuntimeError("rollout backward needs the pressure gradient tensor")
        _sync()
        for cu in reversed(fctx.ckpts):
            st = fctx.smac.adjoint_step(fctx.dt, [t.data_ptr() for t in cu], [t.data_ptr() for t in grads],
                                        pb.data_ptr(), fctx.opts)
            if fctx.opts.tol > 0 and not st.converged:
                raise RuntimeError(f"adjoint solve did not converge: {st.as_dict()}")
        _sync()
        return grads[0], grads[1], grads[2], pb, None, None, None, None


def rollout(smac, u, v, w, p, dt, nsteps, opts):
    """(u, v, w, p) after nsteps steps, differentiable with respect to the
    initial (u, v, w, p)."""
    return SMACRollout.apply(u, v, w, p, smac, dt, nsteps, opts)


class VoFRollout(torch.autograd.Function):
    """nsteps of the VoF loop -- smac_transport_alpha(dt, substeps, step)
    then smac_step(dt, opts), CFD/main.cu:283-317 -- from (alpha, u, v, w, p);
    its backward runs smac_adjoint_vof_step once per step in reverse order from
    per-step checkpoints of (alpha, u, v, w, p)."""

    @staticmethod
    def forward(fctx, alpha, u, v, w, p, smac, dt, nsteps, substeps, opts, step0):
        for t in (alpha, u, v, w, p):
            if t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous():
                raise ValueError("rollout state must be contiguous float64 CUDA tensors")
        _sync()
        for name, t in zip(("alpha", "u", "v", "w", "p"), (alpha, u, v, w, p)):
            smac.upload_device(name, t.data_ptr(), t.numel())
        smac.update_properties(True)  # rho_old of the first step = rho(alpha^0)
        names = ("alpha", "u", "v", "w", "p")
        ckpts = []
        stats = []
        for s in range(nsteps):
            ck = [torch.empty_like(t) for t in (alpha, u, v, w, p)]
            for name, t in zip(names, ck):
                smac.download_device(name, t.data_ptr(), t.numel())
            ckpts.append(ck)
            smac.transport_alpha(dt, substeps, step0 + s)
            st = smac.step(dt, opts)
            if opts.tol > 0 and not st.converged:
                raise RuntimeError(f"forward step did not converge: {st.as_dict()}")
            stats.append(st.as_dict())
        outs = [torch.empty_like(t) for t in (alpha, u, v, w, p)]
        for name, t in zip(names, outs):
            smac.download_device(name, t.data_ptr(), t.numel())
        fctx.smac, fctx.dt, fctx.opts, fctx.ckpts = smac, dt, opts, ckpts
        fctx.substeps, fctx.step0 = substeps, step0
        fctx.forward_stats = stats
        return tuple(outs)

    @staticmethod
    def backward(fctx, ga, gu, gv, gw, gp):
        like = fctx.ckpts[0]
        grads = [torch.zeros_like(l) if g is None else g.detach().to(torch.float64).contiguous().clone()
                 for g, l in zip((ga, gu, gv, gw, gp), like)]
        _sync()
        for s in range(len(fctx.ckpts) - 1, -1, -1):
            ca, cu, cv, cw, cp = fctx.ckpts[s]
            st = fctx.smac.adjoint_vof_step(fctx.dt, fctx.substeps, fctx.step0 + s, ca.data_ptr(),
                                            [cu.data_ptr(), cv.data_ptr(), cw.data_ptr()], cp.data_ptr(),
                                            grads[0].data_ptr(), [t.data_ptr() for t in grads[1:4]],
                                            grads[4].data_ptr(), fctx.opts)
            if fctx.opts.tol > 0 and not st.converged:
                raise RuntimeError(f"adjoint solve did not converge: {st.as_dict()}")
        _sync()
        return tuple(grads) + (None,) * 6


def vof_rollout(smac, alpha, u, v, w, p, dt, nsteps, substeps, opts, step0=0):
    """(alpha, u, v, w, p) after nsteps steps of the VoF loop (transport with
    `substeps` substeps, then the step; the sweep order of step s is
    (step0 + s) % 3), differentiable with respect to the initial state."""
    return VoFRollout.apply(alpha, u, v, w, p, smac, dt, nsteps, substeps, opts, step0)
