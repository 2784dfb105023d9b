"""GPU: NaN reporting, the byte model and zero-copy binding, against the oracle
and against the copying entry points.

* NaN injected into u returns SMAC_ERR_NAN with `nan_detected` set before any
  PCG iteration (the oracle does the same), leaves u and p untouched.
* `bytes_model` equals the oracle's for the same call (the same iteration
  counts: the FP64 form is bit-identical).
* smac_bind_device: a step on torch tensors bound in the device layout gives
  bit-identical fields to the same step on uploaded copies; unbinding hands the
  contents back; a bound alpha receives the VoF transport's result.
"""
import ctypes as C

import numpy as np
import pytest

from mydemcfd_amd import _abi, cases

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    assert torch.cuda.is_available()
    return torch


@pytest.mark.parametrize("precond,precision", [("jacobi", "fp64"), ("mg", "fp64"), ("mg", "fp32"),
                                               ("ref_gmg", "fp64")])
def test_nan_in_u_matches_oracle(hip, oracle_mod, precond, precision):
    print(f"case: NaN in u, cavity 24^3, {precond} {precision}", flush=True)
    u = np.zeros((26, 26, 27))
    u[3, 4, 5] = np.nan
    kw = dict(precond=precond, precision=precision, max_iter=2000)
    if precision != "fp64":
        kw["flexible"] = 1
    out = {}
    for name, ctx in (("gpu", hip.SMACContext(**cases.cavity_desc(24))),
                      ("oracle", oracle_mod.OracleContext(**cases.cavity_desc(24)))):
        ctx.upload("u", u)
        if name == "oracle" and precision != "fp64":
            kw_o = dict(kw, precision="fp64", flexible=0)
        else:
            kw_o = kw
        with pytest.raises(_abi.SmacError) as e:
            ctx.step(1e-3, **kw_o)
        assert e.value.status == _abi.SMAC_ERR_NAN, name
        st = e.value.stats
        assert st.nan_detected == 1 and st.iters == 0 and st.converged == 0, name
        assert np.isnan(ctx.download("u")[3, 4, 5]) and np.all(ctx.download("p") == 0.0), name
        out[name] = st
    # a NaN right-hand side never reaches the iteration on either side
    assert out["gpu"].iters == out["oracle"].iters == 0


def _tg(n):
    return cases.tg_fields(n)


@pytest.mark.parametrize("case", ["periodic_mg", "cavity_jacobi", "two_phase_mg", "cavity_ref_gmg"])
def test_bytes_model_matches_oracle(hip, oracle_mod, case):
    print(f"case: bytes_model {case}", flush=True)
    if case == "periodic_mg":
        desc, kw, fields = cases.periodic_desc(16), dict(precond="mg", tol=1e-8), dict(zip("uvw", _tg(16)))
    elif case == "cavity_jacobi":
        desc, kw, fields = cases.cavity_desc(16), dict(precond="jacobi", tol=1e-8, max_iter=500), {}
    elif case == "cavity_ref_gmg":
        desc, kw, fields = cases.cavity_desc(24), dict(precond="ref_gmg", tol=1e-8, stop_rule="ref_linf_dt"), {}
    else:
        n = (16, 8, 8)
        desc = cases.dam_break_desc(n, rho=(1.0, 20.0), mu=(1e-3, 5e-3))
        kw, fields = dict(precond="mg", tol=1e-8), {"alpha": cases.dam_break_alpha(n)}
    res = {}
    for name, ctx in (("gpu", hip.SMACContext(**desc)), ("oracle", oracle_mod.OracleContext(**desc))):
        for f, a in fields.items():
            ctx.upload(f, a)
        if "alpha" in fields:
            ctx.update_properties(True)
        res[name] = ctx.step(1e-3, **kw)
    g, o = res["gpu"], res["oracle"]
    assert g.iters == o.iters > 0
    assert g.bytes_model == pytest.approx(o.bytes_model, rel=1e-12)
    assert g.nan_detected == 0


def _bound(torch, ctx, field, a):
    """A torch buffer in the device layout holding the reference-layout array a."""
    planes, py, px = ctx.device_layout()
    t = torch.zeros(planes * py * px, dtype=torch.float64, device="cuda")
    sz, sy, sx = a.shape
    t.view(planes, py, px)[:sz, :sy, :sx].copy_(torch.from_numpy(a))
    torch.cuda.synchronize()
    ctx.bind_device(field, t.data_ptr(), t.numel())
    return t


def _view(ctx, t, field):
    planes, py, px = ctx.device_layout()
    sz, sy, sx = ctx.shape(field)
    return t.view(planes, py, px)[:sz, :sy, :sx].cpu().numpy()


def _interior(a, n, axis=None):
    hi = [n[0] + 1, n[1] + 1, n[2] + 1]
    if axis is not None:
        hi[axis] += 1
    return a[1:hi[2], 1:hi[1], 1:hi[0]]


def test_bind_device_step_matches_copies(hip):
    torch = _torch()
    n = 32
    print(f"case: bind u,v,w,p periodic {n}^3, two MG FP64 steps", flush=True)
    fields = dict(zip("uvw", _tg(n)))
    fields["p"] = np.zeros((n + 2, n + 2, n + 2))
    a = hip.SMACContext(**cases.periodic_desc(n))
    b = hip.SMACContext(**cases.periodic_desc(n))
    for f, arr in fields.items():
        a.upload(f, arr)
    bufs = {f: _bound(torch, b, f, arr) for f, arr in fields.items()}
    kw = dict(precond="mg", tol=1e-10)
    for _ in range(2):
        sa, sb = a.step(2e-3, **kw), b.step(2e-3, **kw)
        assert sa.iters == sb.iters
    b.sync()
    # faces 1..n: on a periodic axis face n+1 is the image of face 1, which a
    # host download fills and the device array does not keep
    for f in "uvw":
        assert np.array_equal(_interior(_view(b, bufs[f], f), (n, n, n)), _interior(a.download(f), (n, n, n))), f
    assert np.array_equal(_interior(_view(b, bufs["p"], "p"), (n, n, n)), _interior(a.download("p"), (n, n, n)))
    # unbind: the context copies the contents into its own arrays and goes on
    for f in fields:
        b.bind_device(f, None, 0)
    for f in fields:
        assert np.array_equal(b.download(f), a.download(f)), f
    a.step(2e-3, **kw)
    b.step(2e-3, **kw)
    for f in fields:
        assert np.array_equal(b.download(f), a.download(f)), f


def test_bind_alpha_receives_the_transport(hip):
    torch = _torch()
    n = (16, 8, 12)
    print(f"case: bind alpha, dam break {n}, transport (2 substeps) + step x2", flush=True)
    desc = cases.dam_break_desc(n, length=(2.0, 1.0, 1.5), rho=(1.0, 20.0), mu=(1e-3, 5e-3))
    alpha = cases.dam_break_alpha(n, length=(2.0, 1.0, 1.5))
    a = hip.SMACContext(**desc)
    b = hip.SMACContext(**desc)
    # both grids take the property arrays: a by a device upload of alpha, b by the binding
    ta = torch.from_numpy(alpha).cuda()
    torch.cuda.synchronize()
    a.upload_device("alpha", ta.data_ptr(), ta.numel())
    tb = _bound(torch, b, "alpha", alpha)
    kw = dict(precond="mg", tol=1e-10)
    for ctx in (a, b):
        ctx.update_properties(True)
    for s in range(2):
        for ctx in (a, b):
            ctx.transport_alpha(1e-3, 2, s)
            ctx.step(1e-3, **kw)
    b.sync()
    assert np.array_equal(_interior(_view(b, tb, "alpha"), n), _interior(a.download("alpha"), n))
    for f in "uvwp":
        assert np.array_equal(b.download(f), a.download(f)), f


def test_bind_device_rejects(hip):
    torch = _torch()
    print("case: bind_device argument checks", flush=True)
    c = hip.SMACContext(**cases.cavity_desc(8))
    planes, py, px = c.device_layout()
    cnt = planes * py * px
    t = torch.zeros(cnt + 64, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    for args, status in (((("u", t.data_ptr(), cnt - 1)), _abi.SMAC_ERR_INVALID_ARG),  # wrong count
                         ((("u", t.data_ptr() + 8, cnt)), _abi.SMAC_ERR_INVALID_ARG),  # not 256-byte aligned
                         ((("rhs", t.data_ptr(), cnt)), _abi.SMAC_ERR_UNSUPPORTED)):
        with pytest.raises(_abi.SmacError) as e:
            c.bind_device(*args)
        assert e.value.status == status, args
    host = np.zeros(cnt + 64)
    ptr = (host.ctypes.data + 255) // 256 * 256
    with pytest.raises(_abi.SmacError) as e:
        c.bind_device("u", ptr, cnt)  # host memory
    assert e.value.status == _abi.SMAC_ERR_INVALID_ARG
    # nothing was bound: the context still steps on its own arrays
    c.step(1e-3, precond="mg", tol=1e-8)
    c.bind_device("u", None, 0)  # unbinding an unbound field is a no-op
