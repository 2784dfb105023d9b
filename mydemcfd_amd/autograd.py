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
