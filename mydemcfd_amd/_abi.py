"""ctypes mirror of include/smac_c_api.h.

The structures and function signatures here are plain data describing the
C-ABI. `bind(lib)` attaches them to a loaded shared library; `Context` is a thin
Python handle over one `smac_ctx`. The product (mydemcfd_amd.smac) binds it to the
HIP library only.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

# status codes
SMAC_OK = 0
SMAC_ERR_INVALID_ARG = -1
SMAC_ERR_INVALID_GRID = -2
SMAC_ERR_HIP = -3
SMAC_ERR_NO_DEVICE = -4
SMAC_ERR_OOM = -5
SMAC_ERR_NAN = -6
SMAC_ERR_UNSUPPORTED = -7

# enum smac_bc_type (CFD/Enums.h:10-16)
BC_NOSLIP, BC_SLIP, BC_INFLOW, BC_OUTFLOW, BC_OUTLET = range(5)
XMIN, XMAX, YMIN, YMAX, ZMIN, ZMAX = range(6)

(FIELD_U, FIELD_V, FIELD_W, FIELD_USTAR, FIELD_VSTAR, FIELD_WSTAR, FIELD_P, FIELD_DP, FIELD_RHS, FIELD_ALPHA,
 FIELD_GRADPX, FIELD_GRADPY, FIELD_GRADPZ, FIELD_FLUXX, FIELD_FLUXY, FIELD_FLUXZ,
 FIELD_EPS, FIELD_EPS_OLD, FIELD_EPS_HALF, FIELD_IMPX, FIELD_IMPY, FIELD_IMPZ) = range(22)
FIELD_NAMES = {"u": FIELD_U, "v": FIELD_V, "w": FIELD_W, "ustar": FIELD_USTAR, "vstar": FIELD_VSTAR,
               "wstar": FIELD_WSTAR, "p": FIELD_P, "dp": FIELD_DP, "rhs": FIELD_RHS, "alpha": FIELD_ALPHA,
               "gradpx": FIELD_GRADPX, "gradpy": FIELD_GRADPY, "gradpz": FIELD_GRADPZ,
               "fluxx": FIELD_FLUXX, "fluxy": FIELD_FLUXY, "fluxz": FIELD_FLUXZ,
               "eps": FIELD_EPS, "eps_old": FIELD_EPS_OLD, "eps_half": FIELD_EPS_HALF,
               "impx": FIELD_IMPX, "impy": FIELD_IMPY, "impz": FIELD_IMPZ}

PRECOND_JACOBI, PRECOND_MG, PRECOND_REF_GMG = 0, 1, 2
PREC_FP64, PREC_MIXED_FP32, PREC_MIXED_BF16 = 0, 1, 2
STOP_REF_LINF_DT, STOP_REL_L2 = 0, 1


class BcEntry(C.Structure):
    _fields_ = [("type", C.c_int32), ("vx", C.c_double), ("vy", C.c_double), ("vz", C.c_double)]


class GridDesc(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("sizex", C.c_double), ("sizey", C.c_double), ("sizez", C.c_double),
                ("originx", C.c_double), ("originy", C.c_double), ("originz", C.c_double),
                ("periodic", C.c_int32 * 3), ("side_bcid", C.c_uint8 * 6),
                ("num_bc", C.c_int32), ("bc_table", C.POINTER(BcEntry))]


class FluidDesc(C.Structure):
    _fields_ = [("rho0", C.c_double), ("rho1", C.c_double), ("mu0", C.c_double), ("mu1", C.c_double),
                ("gx", C.c_double), ("gy", C.c_double), ("gz", C.c_double)]


class SolverOpts(C.Structure):
    _fields_ = [("precond", C.c_int32), ("precision", C.c_int32), ("stop_rule", C.c_int32),
                ("tol", C.c_double), ("check_every", C.c_int32), ("max_iter", C.c_int32),
                ("mg_levels", C.c_int32), ("nu1", C.c_int32), ("nu2", C.c_int32), ("flexible", C.c_int32),
                ("ref_sweeps", C.c_int32), ("ref_coarse_sweeps", C.c_int32),
                ("stationary", C.c_int32)]


class SolveStats(C.Structure):
    _fields_ = [("iters", C.c_int32), ("converged", C.c_int32), ("pure_neumann", C.c_int32),
                ("mg_levels", C.c_int32), ("rel_res", C.c_double), ("linf_res_dt", C.c_double),
                ("bnorm", C.c_double), ("ms_predictor", C.c_double), ("ms_solve", C.c_double),
                ("ms_correct", C.c_double), ("nan_detected", C.c_int32), ("reserved", C.c_int32),
                ("bytes_model", C.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


_P = C.c_void_p
_D = C.POINTER(C.c_double)

# name -> (restype, argtypes)
SIGNATURES = {
    "smac_create": (C.c_int, [C.POINTER(GridDesc), C.POINTER(FluidDesc), C.POINTER(_P)]),
    "smac_destroy": (None, [_P]),
    "smac_set_fluid": (C.c_int, [_P, C.POINTER(FluidDesc)]),
    "smac_upload": (C.c_int, [_P, C.c_int32, _D, C.c_size_t]),
    "smac_download": (C.c_int, [_P, C.c_int32, _D, C.c_size_t]),
    "smac_upload_device": (C.c_int, [_P, C.c_int32, _P, C.c_size_t]),
    "smac_download_device": (C.c_int, [_P, C.c_int32, _P, C.c_size_t]),
    "smac_field_count": (C.c_size_t, [_P, C.c_int32]),
    "smac_device_layout": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "smac_bind_device": (C.c_int, [_P, C.c_int32, _P, C.c_size_t]),
    "smac_update_properties": (C.c_int, [_P, C.c_int32]),
    "smac_compute_mass_flux": (C.c_int, [_P, C.c_double]),
    "smac_transport_alpha": (C.c_int, [_P, C.c_double, C.c_int32, C.c_int32]),
    "smac_clear_alpha_flux_accum": (C.c_int, [_P]),
    "smac_alpha_flux_split": (C.c_int, [_P, C.c_double, C.c_int32]),
    "smac_predictor": (C.c_int, [_P, C.c_double]),
    "smac_predictor_two_way": (C.c_int, [_P, C.c_double]),
    "smac_update_poisson_beta_two_way": (C.c_int, [_P]),
    "smac_pressure_solve": (C.c_int, [_P, C.c_double, C.POINTER(SolverOpts), C.POINTER(SolveStats)]),
    "smac_correct": (C.c_int, [_P, C.c_double]),
    "smac_step": (C.c_int, [_P, C.c_double, C.POINTER(SolverOpts), C.POINTER(SolveStats)]),
    "smac_calc_cfl": (C.c_int, [_P, C.c_double, _D, _D]),
    "smac_calc_cfl_noalloc": (C.c_int, [_P, C.c_double, _D, _D]),
    "smac_update_boundary_ghost": (C.c_int, [_P, C.c_double]),
    "smac_make_face_gradp": (C.c_int, [_P]),
    "smac_cell_output": (C.c_int, [_P, C.POINTER(C.c_float), C.c_size_t]),
    "smac_adjoint_step": (C.c_int, [_P, C.c_double, C.POINTER(SolverOpts), C.POINTER(_P * 3), C.POINTER(_P * 3), _P,
                                    C.POINTER(SolveStats)]),
    "smac_adjoint_vof_step": (C.c_int, [_P, C.c_double, C.c_int32, C.c_int32, C.POINTER(SolverOpts), _P,
                                        C.POINTER(_P * 3), _P, _P, C.POINTER(_P * 3), _P, C.POINTER(SolveStats)]),
    "smac_default_opts": (None, [C.POINTER(SolverOpts)]),
    "smac_set_profiling": (C.c_int, [_P, C.c_int32]),
    "smac_kernel_time": (C.c_int, [_P, C.c_char_p, C.POINTER(C.c_int64), _D]),
    "smac_kernel_names": (C.c_int, [_P, C.c_char_p, C.c_size_t]),
    "smac_reset_kernel_times": (C.c_int, [_P]),
    "smac_sync": (C.c_int, [_P]),
    "smac_stream": (_P, [_P]),
    "smac_strerror": (C.c_char_p, [C.c_int]),
    "smac_backend_name": (C.c_char_p, []),
}


def bind(lib: C.CDLL) -> C.CDLL:
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


class SmacError(RuntimeError):
    def __init__(self, lib, status, what):
        msg = lib.smac_strerror(status).decode() if lib is not None else str(status)
        super().__init__(f"{what} failed: {msg} ({status})")
        self.status = status
        self.stats = None  # the SolveStats of a failed solve or step (e.g. nan_detected)


def default_opts(lib, **kw) -> SolverOpts:
    o = SolverOpts()
    lib.smac_default_opts(C.byref(o))
    for k, v in kw.items():
        if k == "precond" and isinstance(v, str):
            v = {"jacobi": PRECOND_JACOBI, "mg": PRECOND_MG, "ref_gmg": PRECOND_REF_GMG}[v]
        if k == "precision" and isinstance(v, str):
            v = {"fp64": PREC_FP64, "fp32": PREC_MIXED_FP32, "bf16": PREC_MIXED_BF16}[v]
        if k == "stop_rule" and isinstance(v, str):
            v = {"ref_linf_dt": STOP_REF_LINF_DT, "rel_l2": STOP_REL_L2}[v]
        setattr(o, k, v)
    return o


def field_shape(n, field: int):
    """(z, y, x) numpy shape of a field in the reference padded layout."""
    nx, ny, nz = n
    sh = [nx + 2, ny + 2, nz + 2]
    if field in (FIELD_U, FIELD_USTAR, FIELD_GRADPX, FIELD_FLUXX, FIELD_IMPX):
        sh[0] += 1
    elif field in (FIELD_V, FIELD_VSTAR, FIELD_GRADPY, FIELD_FLUXY, FIELD_IMPY):
        sh[1] += 1
    elif field in (FIELD_W, FIELD_WSTAR, FIELD_GRADPZ, FIELD_FLUXZ, FIELD_IMPZ):
        sh[2] += 1
    return (sh[2], sh[1], sh[0])


class Context:
    """One smac_ctx on `lib`. Arrays are numpy float64 in the reference layout,
    shape (z, y, x) (x fastest), see field_shape()."""

    def __init__(self, lib, n, size=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0), periodic=(0, 0, 0),
                 side_bcid=(0, 0, 0, 0, 0, 0), bc_table=((BC_NOSLIP, 0.0, 0.0, 0.0),),
                 rho=(1.0, 1.0), mu=(0.0, 0.0), g=(0.0, 0.0, 0.0)):
        self.lib = lib
        self.n = tuple(int(v) for v in n)
        self._bc = (BcEntry * len(bc_table))(*[BcEntry(int(t), float(a), float(b), float(c))
                                               for (t, a, b, c) in bc_table])
        gd = GridDesc()
        gd.nx, gd.ny, gd.nz = self.n
        gd.sizex, gd.sizey, gd.sizez = size
        gd.originx, gd.originy, gd.originz = origin
        for d in range(3):
            gd.periodic[d] = int(periodic[d])
        for s in range(6):
            gd.side_bcid[s] = int(side_bcid[s])
        gd.num_bc = len(bc_table)
        gd.bc_table = C.cast(self._bc, C.POINTER(BcEntry))
        self.fluid = FluidDesc(rho[0], rho[1], mu[0], mu[1], g[0], g[1], g[2])
        h = _P()
        self._check(lib.smac_create(C.byref(gd), C.byref(self.fluid), C.byref(h)), "smac_create")
        self.h = h
        self.size = tuple(size)
        self.periodic = tuple(int(v) for v in periodic)

    def _check(self, rc, what, stats=None):
        if rc != SMAC_OK:
            e = SmacError(self.lib, rc, what)
            e.stats = stats
            raise e

    def close(self):
        if getattr(self, "h", None):
            self.lib.smac_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _field(f):
        return FIELD_NAMES[f] if isinstance(f, str) else int(f)

    def shape(self, f):
        return field_shape(self.n, self._field(f))

    def upload(self, f, arr):
        fid = self._field(f)
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.shape != self.shape(fid):
            raise ValueError(f"field {f}: expected shape {self.shape(fid)}, got {a.shape}")
        self._check(self.lib.smac_upload(self.h, fid, a.ctypes.data_as(_D), a.size), "smac_upload")

    def download(self, f):
        fid = self._field(f)
        a = np.empty(self.shape(fid), dtype=np.float64)
        self._check(self.lib.smac_download(self.h, fid, a.ctypes.data_as(_D), a.size), "smac_download")
        return a

    def upload_device(self, f, ptr: int, count: int):
        self._check(self.lib.smac_upload_device(self.h, self._field(f), ptr, count), "smac_upload_device")

    def download_device(self, f, ptr: int, count: int):
        self._check(self.lib.smac_download_device(self.h, self._field(f), ptr, count), "smac_download_device")

    def device_layout(self):
        """(planes, pitch_y, pitch_x): the device box every field lives in; the
        reference-layout field is the sub-box [:sz, :sy, :sx] of shape(f)."""
        px, py, pz = C.c_int64(), C.c_int64(), C.c_int64()
        self._check(self.lib.smac_device_layout(self.h, C.byref(px), C.byref(py), C.byref(pz)), "smac_device_layout")
        return pz.value, py.value, px.value

    def bind_device(self, f, ptr, count: int):
        """Zero-copy: the context uses the device buffer at `ptr` (device_layout
        box, `count` doubles) as field f (u, v, w, p or alpha) until
        bind_device(f, None, 0)."""
        self._check(self.lib.smac_bind_device(self.h, self._field(f), ptr, count), "smac_bind_device")

    def set_fluid(self, rho=(1.0, 1.0), mu=(0.0, 0.0), g=(0.0, 0.0, 0.0)):
        self.fluid = FluidDesc(rho[0], rho[1], mu[0], mu[1], g[0], g[1], g[2])
        self._check(self.lib.smac_set_fluid(self.h, C.byref(self.fluid)), "smac_set_fluid")

    def update_properties(self, initial=True):
        self._check(self.lib.smac_update_properties(self.h, int(initial)), "smac_update_properties")

    def predictor(self, dt):
        self._check(self.lib.smac_predictor(self.h, dt), "smac_predictor")

    def predictor_two_way(self, dt):
        """G_SMACSolver::get_vof_vstar_rhouu_consistent_two_way (with its mass flux
        and the u* boundary update); needs eps, eps_old and eps_half uploaded."""
        self._check(self.lib.smac_predictor_two_way(self.h, dt), "smac_predictor_two_way")

    def update_poisson_beta_two_way(self):
        """G_CFDDEMCoupling::update_poisson_beta_two_way: Poisson coefficients
        eps_f/rho_f until the next update_properties."""
        self._check(self.lib.smac_update_poisson_beta_two_way(self.h), "smac_update_poisson_beta_two_way")

    def pressure_solve(self, dt, opts=None, **kw):
        o = opts if opts is not None else default_opts(self.lib, **kw)
        st = SolveStats()
        self._check(self.lib.smac_pressure_solve(self.h, dt, C.byref(o), C.byref(st)), "smac_pressure_solve", st)
        return st

    def correct(self, dt):
        self._check(self.lib.smac_correct(self.h, dt), "smac_correct")

    def step(self, dt, opts=None, **kw):
        o = opts if opts is not None else default_opts(self.lib, **kw)
        st = SolveStats()
        self._check(self.lib.smac_step(self.h, dt, C.byref(o), C.byref(st)), "smac_step", st)
        return st

    def adjoint_step(self, dt, u_n, ubar, pbar, opts=None, **kw):
        """Adjoint of one step from the state u_n (three device pointers, the
        reference layout): ubar (three device pointers) and pbar (one) hold the
        cotangents of the step's output on entry and of its input on return."""
        o = opts if opts is not None else default_opts(self.lib, **kw)
        st = SolveStats()
        un = (_P * 3)(*[int(x) for x in u_n])
        ub = (_P * 3)(*[int(x) for x in ubar])
        self._check(self.lib.smac_adjoint_step(self.h, dt, C.byref(o), C.byref(un), C.byref(ub), int(pbar),
                                               C.byref(st)), "smac_adjoint_step")
        return st

    def adjoint_vof_step(self, dt, substeps, step, alpha_n, u_n, p_n, alpha_bar, ubar, pbar, opts=None, **kw):
        """Adjoint of transport_alpha(dt, substeps, step) + step(dt) from the
        state (alpha_n, u_n, p_n) (device pointers, the reference layout):
        alpha_bar, ubar (three) and pbar hold the cotangents of the output on
        entry and of the input on return."""
        o = opts if opts is not None else default_opts(self.lib, **kw)
        st = SolveStats()
        un = (_P * 3)(*[int(x) for x in u_n])
        ub = (_P * 3)(*[int(x) for x in ubar])
        self._check(self.lib.smac_adjoint_vof_step(self.h, dt, int(substeps), int(step), C.byref(o), int(alpha_n),
                                                   C.byref(un), int(p_n), int(alpha_bar), C.byref(ub), int(pbar),
                                                   C.byref(st)), "smac_adjoint_vof_step")
        return st

    def calc_cfl(self, dt):
        a, b = C.c_double(), C.c_double()
        self._check(self.lib.smac_calc_cfl(self.h, dt, C.byref(a), C.byref(b)), "smac_calc_cfl")
        return a.value, b.value

    def transport_alpha(self, dt, substeps=1, step=0):
        """VoF transport of alpha over one time step (split THINC/WLIC); the
        accumulated flux enters the next predictor's mass flux."""
        self._check(self.lib.smac_transport_alpha(self.h, dt, int(substeps), int(step)), "smac_transport_alpha")

    def clear_alpha_flux_accum(self):
        self._check(self.lib.smac_clear_alpha_flux_accum(self.h), "smac_clear_alpha_flux_accum")

    def alpha_flux_split(self, sub_dt, step):
        """One VoF substep (split THINC/WLIC), its flux added to the accumulator."""
        self._check(self.lib.smac_alpha_flux_split(self.h, sub_dt, int(step)), "smac_alpha_flux_split")

    def update_boundary_ghost(self, dt):
        """Ghost pressure and tangential velocity ghosts after a step
        (G_SMACSolver::update_boundary_ghost)."""
        self._check(self.lib.smac_update_boundary_ghost(self.h, dt), "smac_update_boundary_ghost")

    def make_face_gradp(self):
        """Face pressure gradient into gradpx/gradpy/gradpz (G_SMACSolver::make_face_gradp)."""
        self._check(self.lib.smac_make_face_gradp(self.h), "smac_make_face_gradp")

    def cell_output(self):
        """(p, alpha, velocity) cell arrays of the VTI output as float32:
        shapes (nz, ny, nx) and (nz, ny, nx, 3)."""
        nx, ny, nz = self.n
        n = nx * ny * nz
        buf = np.empty(5 * n, dtype=np.float32)
        self._check(self.lib.smac_cell_output(self.h, buf.ctypes.data_as(C.POINTER(C.c_float)), buf.size),
                    "smac_cell_output")
        return (buf[:n].reshape(nz, ny, nx), buf[n:2 * n].reshape(nz, ny, nx), buf[2 * n:].reshape(nz, ny, nx, 3))

    def set_profiling(self, on=True):
        self._check(self.lib.smac_set_profiling(self.h, int(on)), "smac_set_profiling")

    def kernel_time(self, name):
        n, ms = C.c_int64(), C.c_double()
        self._check(self.lib.smac_kernel_time(self.h, name.encode(), C.byref(n), C.byref(ms)), "smac_kernel_time")
        return n.value, ms.value

    def kernel_names(self):
        buf = C.create_string_buffer(8192)
        self._check(self.lib.smac_kernel_names(self.h, buf, len(buf)), "smac_kernel_names")
        v = buf.value.decode()
        return [x for x in v.split(",") if x]

    def reset_kernel_times(self):
        self._check(self.lib.smac_reset_kernel_times(self.h), "smac_reset_kernel_times")

    def sync(self):
        self._check(self.lib.smac_sync(self.h), "smac_sync")

    def opts(self, **kw):
        return default_opts(self.lib, **kw)
