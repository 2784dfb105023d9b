"""Workload definitions and seeded initial conditions.

Every array is float64 in the reference padded layout (numpy shape (z, y, x),
x fastest; see _abi.field_shape). Face index f along its own axis sits at
coordinate origin + (f-1)*h; cell index c at origin + (c-1/2)*h, as in
StaggeredGrid::get_cell_coord (CFD/StaggeredGrid.h:121-142).

Workloads:
  cavity   lid-driven cavity, unit cube, all walls no-slip, lid u=1 on y-max
           (the reference's gravity convention is -y, CFD/config.json), Re=UL/nu.
  tg       triply periodic [0,2pi)^3 Taylor-Green, u=sin x cos y cos z,
           v=-cos x sin y cos z, w=0, plus a seeded solenoidal perturbation.
"""
from __future__ import annotations

import math

import numpy as np

from . import _abi


def cavity_desc(n=64, re=100.0, lid=1.0):
    """Context kwargs for the lid-driven cavity: bc id 1 = no-slip at rest on
    five sides, bc id 2 = no-slip moving with u=lid on the y-max side."""
    nu = lid * 1.0 / re
    nn = tuple(n) if isinstance(n, (tuple, list)) else (n, n, n)
    return dict(n=nn, size=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0), periodic=(0, 0, 0),
                side_bcid=(1, 1, 1, 2, 1, 1),
                bc_table=((_abi.BC_NOSLIP, 0.0, 0.0, 0.0), (_abi.BC_NOSLIP, 0.0, 0.0, 0.0),
                          (_abi.BC_NOSLIP, lid, 0.0, 0.0)),
                rho=(1.0, 1.0), mu=(nu, nu), g=(0.0, 0.0, 0.0))


def periodic_desc(n=256, re=1600.0, length=2.0 * math.pi):
    nu = 1.0 / re
    nn = n if isinstance(n, (tuple, list)) else (n, n, n)
    return dict(n=tuple(nn), size=(length, length, length), origin=(0.0, 0.0, 0.0), periodic=(1, 1, 1),
                side_bcid=(0,) * 6, bc_table=((_abi.BC_NOSLIP, 0.0, 0.0, 0.0),),
                rho=(1.0, 1.0), mu=(nu, nu), g=(0.0, 0.0, 0.0))


def _embed_periodic(a0, field):
    """Place a periodic (z, y, x) array of interior values (face/cell index 1..N)
    into the reference padded layout, filling ghosts and duplicate faces."""
    nz, ny, nx = a0.shape
    shape = _abi.field_shape((nx, ny, nz), field)
    out = np.empty(shape, dtype=np.float64)
    iz = (np.arange(shape[0]) - 1) % nz
    iy = (np.arange(shape[1]) - 1) % ny
    ix = (np.arange(shape[2]) - 1) % nx
    out[...] = a0[np.ix_(iz, iy, ix)]
    return out


def taylor_green(n, length=2.0 * math.pi):
    """(u, v, w) interior arrays (z, y, x), N^3 each, of the Taylor-Green field."""
    h = length / n
    xf = np.arange(n) * h            # face coordinate of face index 1..N
    xc = (np.arange(n) + 0.5) * h    # cell centre
    u = np.sin(xf)[None, None, :] * np.cos(xc)[None, :, None] * np.cos(xc)[:, None, None]
    v = -np.cos(xc)[None, None, :] * np.sin(xf)[None, :, None] * np.cos(xc)[:, None, None]
    w = np.zeros((n, n, n))
    return u, v, w


def discrete_curl(Ax, Ay, Az, h):
    """MAC-grid curl of an edge-centred vector potential on a periodic box:
    exactly (to rounding) divergence free under the staggered divergence."""
    r = np.roll
    u = (r(Az, -1, 1) - Az) / h - (r(Ay, -1, 0) - Ay) / h
    v = (r(Ax, -1, 0) - Ax) / h - (r(Az, -1, 2) - Az) / h
    w = (r(Ay, -1, 2) - Ay) / h - (r(Ax, -1, 1) - Ax) / h
    return u, v, w


def white_solenoidal(n, seed, rms, length=2.0 * math.pi):
    """Curl of an i.i.d. N(0,1) edge potential, scaled to the given velocity RMS."""
    rng = np.random.default_rng(seed)
    A = [rng.standard_normal((n, n, n)) for _ in range(3)]
    u, v, w = discrete_curl(*A, length / n)
    s = math.sqrt((np.mean(u * u) + np.mean(v * v) + np.mean(w * w)) / 3.0)
    return u * (rms / s), v * (rms / s), w * (rms / s)


def spectral_solenoidal(n, seed, kmax=32, slope=-5.0 / 3.0, length=2.0 * math.pi):
    """Random-phase solenoidal field with shell energy E(k) ~ k^slope for
    1 <= |k| <= kmax: potential amplitude |A(k)| ~ k^((slope-4)/2), curled
    discretely so the field is divergence free on the MAC grid. Unit RMS."""
    rng = np.random.default_rng(seed)
    k1 = np.fft.fftfreq(n, d=1.0 / n)
    kz = k1[:, None, None]
    ky = k1[None, :, None]
    kx = np.fft.rfftfreq(n, d=1.0 / n)[None, None, :]
    kk = np.sqrt(kx * kx + ky * ky + kz * kz)
    amp = np.where((kk >= 1.0) & (kk <= kmax), np.power(np.maximum(kk, 1.0), (slope - 4.0) / 2.0), 0.0)
    A = []
    for _ in range(3):
        ph = rng.uniform(0.0, 2.0 * math.pi, size=amp.shape)
        mag = rng.standard_normal(amp.shape)
        spec = amp * mag * np.exp(1j * ph)
        A.append(np.fft.irfftn(spec, s=(n, n, n)).real)
        del spec, ph, mag
    u, v, w = discrete_curl(*A, length / n)
    s = math.sqrt((np.mean(u * u) + np.mean(v * v) + np.mean(w * w)) / 3.0)
    return u / s, v / s, w / s


def tg_fields(n, seed=0x5EED0002, pert="white", pert_rms=1e-3, energy_frac=0.05):
    """Taylor-Green plus perturbation, as reference-layout (u, v, w).
    pert="white": curl of white noise scaled to RMS pert_rms (config 2);
    pert="k53": k^-5/3 spectrum, k<=32, scaled to energy_frac of the total
    kinetic energy (config 3)."""
    u, v, w = taylor_green(n)
    if pert == "white":
        pu, pv, pw = white_solenoidal(n, seed, pert_rms)
    elif pert == "k53":
        pu, pv, pw = spectral_solenoidal(n, seed)
        e_tg = 0.5 * (np.mean(u * u) + np.mean(v * v) + np.mean(w * w))
        e_p = 0.5 * (np.mean(pu * pu) + np.mean(pv * pv) + np.mean(pw * pw))
        scale = math.sqrt(energy_frac / (1.0 - energy_frac) * e_tg / e_p)
        pu, pv, pw = pu * scale, pv * scale, pw * scale
    elif pert is None:
        pu = pv = pw = 0.0
    else:
        raise ValueError(pert)
    return (_embed_periodic(u + pu, _abi.FIELD_U), _embed_periodic(v + pv, _abi.FIELD_V),
            _embed_periodic(w + pw, _abi.FIELD_W))


def cfl_dt(ctx, cfl_target, dt_probe=1.0, eps=1e-12):
    """dt with max(CFL_conv, CFL_visc) = cfl_target, from the CFL of a probe dt
    (both CFL numbers are linear in dt); the rule of CFDTime::updateTime
    (CFD/CFDTime.cpp:22-48) with the clamps removed."""
    cc, cv = ctx.calc_cfl(dt_probe)
    return dt_probe * cfl_target / (max(cc, cv) + eps)


# ----------------------------------------------------------------- torch (GPU) generators
def tg_fields_torch(n, seed=0x5EED0003, kmax=32, slope=-5.0 / 3.0, energy_frac=0.05, device="cuda",
                    length=2.0 * math.pi):
    """Config-3 field on the GPU: Taylor-Green plus a random-phase solenoidal
    k^slope perturbation (|k| <= kmax) holding energy_frac of the total kinetic
    energy, generated with torch.fft on `device`. Returns reference-layout
    float64 tensors (u, v, w), contiguous, on `device`."""
    import torch

    f64 = torch.float64
    gen = torch.Generator(device=device)
    gen.manual_seed(int(seed))
    k1 = torch.fft.fftfreq(n, d=1.0 / n, device=device, dtype=f64)
    kr = torch.fft.rfftfreq(n, d=1.0 / n, device=device, dtype=f64)
    kk = torch.sqrt(kr[None, None, :] ** 2 + k1[None, :, None] ** 2 + k1[:, None, None] ** 2)
    amp = torch.where((kk >= 1.0) & (kk <= kmax), torch.clamp(kk, min=1.0) ** ((slope - 4.0) / 2.0),
                      torch.zeros_like(kk))
    del kk
    A = []
    for _ in range(3):
        ph = torch.rand(amp.shape, generator=gen, device=device, dtype=f64) * (2.0 * math.pi)
        mag = torch.randn(amp.shape, generator=gen, device=device, dtype=f64) * amp
        spec = torch.complex(mag * torch.cos(ph), mag * torch.sin(ph))
        del ph, mag
        A.append(torch.fft.irfftn(spec, s=(n, n, n)))
        del spec
    del amp
    h = length / n
    r = torch.roll
    Ax, Ay, Az = A
    pu = (r(Az, -1, 1) - Az) / h - (r(Ay, -1, 0) - Ay) / h
    pv = (r(Ax, -1, 0) - Ax) / h - (r(Az, -1, 2) - Az) / h
    pw = (r(Ay, -1, 2) - Ay) / h - (r(Ax, -1, 1) - Ax) / h
    del A, Ax, Ay, Az
    xf = torch.arange(n, device=device, dtype=f64) * h
    xc = (torch.arange(n, device=device, dtype=f64) + 0.5) * h
    u = torch.sin(xf)[None, None, :] * torch.cos(xc)[None, :, None] * torch.cos(xc)[:, None, None]
    v = -torch.cos(xc)[None, None, :] * torch.sin(xf)[None, :, None] * torch.cos(xc)[:, None, None]
    e_tg = 0.5 * float(torch.mean(u * u) + torch.mean(v * v))
    e_p = 0.5 * float(torch.mean(pu * pu) + torch.mean(pv * pv) + torch.mean(pw * pw))
    s = math.sqrt(energy_frac / (1.0 - energy_frac) * e_tg / e_p)
    u = u + s * pu
    v = v + s * pv
    w = s * pw
    del pu, pv, pw
    out = []
    for a0, fid in ((u, _abi.FIELD_U), (v, _abi.FIELD_V), (w, _abi.FIELD_W)):
        shape = _abi.field_shape((n, n, n), fid)
        iz = (torch.arange(shape[0], device=device) - 1) % n
        iy = (torch.arange(shape[1], device=device) - 1) % n
        ix = (torch.arange(shape[2], device=device) - 1) % n
        out.append(a0.index_select(0, iz).index_select(1, iy).index_select(2, ix).contiguous())
    return tuple(out)


def dam_break_desc(n=(256, 128, 128), length=(2.0, 1.0, 1.0), rho=(1.0, 1000.0), mu=(1e-5, 1e-3),
                   g=(0.0, 0.0, -9.81)):
    """Context kwargs for the two-phase dam break: a closed box of no-slip
    walls, phase 0 (light, alpha = 0) and phase 1 (heavy, alpha = 1) with
    densities rho and viscosities mu, gravity along -z. The extents default to
    [0,2]x[0,1]x[0,1] with h = 1/128 at 256x128x128."""
    return dict(n=tuple(n), size=tuple(float(v) for v in length), origin=(0.0, 0.0, 0.0), periodic=(0, 0, 0),
                side_bcid=(1, 1, 1, 1, 1, 1), bc_table=((_abi.BC_NOSLIP, 0.0, 0.0, 0.0), (_abi.BC_NOSLIP, 0.0, 0.0, 0.0)),
                rho=tuple(rho), mu=tuple(mu), g=tuple(g))


def dam_break_alpha(n=(256, 128, 128), length=(2.0, 1.0, 1.0), x_front=0.25, height=0.5):
    """alpha (reference cell layout, (z, y, x) with ghosts) = 1 in the column
    x < x_front, z < height (cell centres), 0 elsewhere; ghosts copy the
    nearest interior cell."""
    nx, ny, nz = n
    hx, hz = length[0] / nx, length[2] / nz
    xc = (np.arange(nx + 2) - 0.5) * hx
    zc = (np.arange(nz + 2) - 0.5) * hz
    xc[0], xc[-1] = xc[1], xc[-2]
    zc[0], zc[-1] = zc[1], zc[-2]
    col = (xc[None, :] < x_front) & (zc[:, None] < height)  # (z, x)
    a = np.repeat(col[:, None, :], ny + 2, axis=1).astype(np.float64)
    return a


def adjoint_fields(n, seed=0x5EED0004, kmax=32):
    """Random-phase solenoidal k^-5/3 field with unit RMS on the periodic
    [0,2pi)^3 box (|k| <= kmax), reference-layout (u, v, w): the initial
    velocity of the gradient checks (Re = 500, dt = 2e-3)."""
    u, v, w = spectral_solenoidal(n, seed, kmax=min(kmax, n // 2))
    return (_embed_periodic(u, _abi.FIELD_U), _embed_periodic(v, _abi.FIELD_V), _embed_periodic(w, _abi.FIELD_W))


def state_mask(n, periodic, field):
    """Boolean mask (reference layout) of the entries of `field` that are state:
    interior cells, faces 1..N+1 along a wall axis of their own direction and
    1..N along a periodic one; ghosts and periodic images are not state."""
    fid = _abi.FIELD_NAMES[field] if isinstance(field, str) else field
    shape = _abi.field_shape(n, fid)
    axis = {_abi.FIELD_U: 0, _abi.FIELD_V: 1, _abi.FIELD_W: 2, _abi.FIELD_USTAR: 0, _abi.FIELD_VSTAR: 1,
            _abi.FIELD_WSTAR: 2}.get(fid, -1)
    sl = []
    for d in (2, 1, 0):  # numpy axes z, y, x
        hi = n[d] + (1 if (d == axis and not periodic[d]) else 0)
        sl.append(slice(1, hi + 1))
    m = np.zeros(shape, dtype=bool)
    m[tuple(sl)] = True
    return m


def void_fraction(n, periodic, seed, lo=0.45, hi=0.95, kmax=2, modes=5):
    """A smooth fluid volume fraction on cells in the reference layout ((z, y, x)
    with ghosts), values in [lo, hi]: a sum of random cosine modes over the
    interior, ghost cells the periodic image (periodic axes) or the nearest
    interior cell (walls), as a coupling fills them."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = n
    idx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    dims = (nz, ny, nx)
    f = np.zeros(dims)
    for _ in range(modes):
        k = rng.integers(-kmax, kmax + 1, size=3)
        ph = rng.uniform(0.0, 2.0 * math.pi)
        f += rng.standard_normal() * np.cos(sum(2.0 * math.pi * k[d] * idx[d] / dims[d] for d in range(3)) + ph)
    f = (f - f.min()) / max(f.max() - f.min(), 1e-30)
    f = lo + (hi - lo) * f
    out = np.zeros((nz + 2, ny + 2, nx + 2))
    out[1:-1, 1:-1, 1:-1] = f
    per = (periodic[2], periodic[1], periodic[0])  # (z, y, x) order
    for ax in range(3):
        src_lo = -2 if per[ax] else 1
        src_hi = 1 if per[ax] else -2
        sl_lo = [slice(None)] * 3
        sl_hi = [slice(None)] * 3
        s_lo = [slice(None)] * 3
        s_hi = [slice(None)] * 3
        sl_lo[ax], sl_hi[ax] = 0, -1
        s_lo[ax], s_hi[ax] = src_lo, src_hi
        out[tuple(sl_lo)] = out[tuple(s_lo)]
        out[tuple(sl_hi)] = out[tuple(s_hi)]
    return out
