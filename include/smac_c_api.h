/*
 * smac_c_api.h -- C-ABI of the MI355X (gfx950) SMAC projection step.
 *
 * One context owns every device array of one staggered (MAC) grid plus one HIP
 * stream. All entry points are plain C: opaque handle, plain pointers and sizes,
 * int status codes (0 = ok, negative = SMAC_ERR_*). Non-convergence of the
 * pressure solve is not an error; it is reported in smac_solve_stats.
 *
 * Each entry point below names the reference interface it replaces
 * (bi-rugaoishii/myDEMCFD, paths relative to the reference root).
 *
 * Layout of host buffers (smac_upload / smac_download) is the reference's
 * padded x-fastest layout, MyArray<T,3>::operator() (CFD/MyArray.h:35-57):
 *   cells   (Nx+2)(Ny+2)(Nz+2)            index (k*(Ny+2)+j)*(Nx+2)+i
 *   x-faces (Nx+3)(Ny+2)(Nz+2)            index (k*(Ny+2)+j)*(Nx+3)+i
 *   y-faces (Nx+2)(Ny+3)(Nz+2)            index (k*(Ny+3)+j)*(Nx+2)+i
 *   z-faces (Nx+2)(Ny+2)(Nz+3)            index (k*(Ny+2)+j)*(Nx+2)+i
 * (CFD/memberList/gridMembers.def:14-128). Cell i in 1..Nx is interior; x-face
 * i lies between cells i-1 and i; x-faces 1 and Nx+1 are the walls.
 */
#ifndef SMAC_C_API_H
#define SMAC_C_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SMAC_API_VERSION 1

/* ---- status codes -------------------------------------------------------- */
enum {
    SMAC_OK = 0,
    SMAC_ERR_INVALID_ARG = -1,   /* null pointer, bad enum, bad count        */
    SMAC_ERR_INVALID_GRID = -2,  /* sizes the multigrid hierarchy cannot use  */
    SMAC_ERR_HIP = -3,           /* a HIP runtime call failed                 */
    SMAC_ERR_NO_DEVICE = -4,     /* no gfx950 device visible                  */
    SMAC_ERR_OOM = -5,           /* device allocation failed                  */
    SMAC_ERR_NAN = -6,           /* NaN/Inf met in the solve                  */
    SMAC_ERR_UNSUPPORTED = -7    /* option combination not built              */
};

/* ---- enums mirrored from CFD/Enums.h:3-36 -------------------------------- */
enum smac_bc_type {              /* BoundaryType (CFD/Enums.h:10-16)          */
    SMAC_BC_NOSLIP = 0,
    SMAC_BC_SLIP = 1,
    SMAC_BC_INFLOW = 2,
    SMAC_BC_OUTFLOW = 3,
    SMAC_BC_OUTLET = 4
};

enum smac_side { SMAC_XMIN = 0, SMAC_XMAX, SMAC_YMIN, SMAC_YMAX, SMAC_ZMIN, SMAC_ZMAX };

/* Fields addressable through smac_upload/smac_download. Names follow the
 * reference members of CFD/memberList/gridMembers.def. */
enum smac_field {
    SMAC_FIELD_U = 0,        /* f_vx_       x-faces */
    SMAC_FIELD_V = 1,        /* f_vy_       y-faces */
    SMAC_FIELD_W = 2,        /* f_vz_       z-faces */
    SMAC_FIELD_USTAR = 3,    /* f_vx_star_  x-faces */
    SMAC_FIELD_VSTAR = 4,    /* f_vy_star_  y-faces */
    SMAC_FIELD_WSTAR = 5,    /* f_vz_star_  z-faces */
    SMAC_FIELD_P = 6,        /* p_          cells   */
    SMAC_FIELD_DP = 7,       /* p_delta_    cells   */
    SMAC_FIELD_RHS = 8,      /* rhs_        cells   */
    SMAC_FIELD_ALPHA = 9,    /* alpha_      cells   (VoF fraction, drives rho/mu) */
    SMAC_FIELD_GRADPX = 10,  /* f_gradp_x_  x-faces (written by smac_make_face_gradp) */
    SMAC_FIELD_GRADPY = 11,  /* f_gradp_y_  y-faces */
    SMAC_FIELD_GRADPZ = 12,  /* f_gradp_z_  z-faces */
    SMAC_FIELD_FLUXX = 13,   /* f_Fx_accum_ x-faces (VoF alpha flux of the last transport) */
    SMAC_FIELD_FLUXY = 14,   /* f_Fy_accum_ y-faces */
    SMAC_FIELD_FLUXZ = 15,   /* f_Fz_accum_ z-faces */
    /* CFD-DEM (two-way) inputs, written by the coupling (CFD/memberList/gridMembers.def:144-152) */
    SMAC_FIELD_EPS = 16,      /* void_fraction_        cells   (fluid volume fraction, t^{n+1}) */
    SMAC_FIELD_EPS_OLD = 17,  /* void_fraction_old_    cells   (t^n) */
    SMAC_FIELD_EPS_HALF = 18, /* void_fraction_half_   cells   (t^{n+1/2}, in the two-way mass flux) */
    SMAC_FIELD_IMPX = 19,     /* f_coupling_impulse_x_ x-faces (particle momentum impulse over the step) */
    SMAC_FIELD_IMPY = 20,     /* f_coupling_impulse_y_ y-faces */
    SMAC_FIELD_IMPZ = 21,     /* f_coupling_impulse_z_ z-faces */
    SMAC_FIELD_COUNT = 22
};

/* PCG_TYPES (CFD/Enums.h:28-33) plus the red-black Gauss-Seidel multigrid. */
enum smac_precond {
    SMAC_PRECOND_JACOBI = 0,     /* STANDARD_PCG / PURENEUMANN_STANDARD_PCG  */
    SMAC_PRECOND_MG = 1,         /* RBGS V-cycle geometric multigrid          */
    SMAC_PRECOND_REF_GMG = 2     /* GMG_PCG / PURENEUMANN_GMG_PCG: the reference's
                                    weighted-Jacobi V-cycle with mean restriction
                                    and constant prolongation
                                    (pressure_solver/G_GMGSolver.cu:10-147,925-998) */
};

enum smac_precision {
    SMAC_PREC_FP64 = 0,          /* everything in FP64                        */
    SMAC_PREC_MIXED_FP32 = 1,    /* V-cycle storage+arithmetic FP32           */
    SMAC_PREC_MIXED_BF16 = 2     /* V-cycle storage BF16, FP32 arithmetic     */
};

enum smac_stop_rule {
    SMAC_STOP_REF_LINF_DT = 0,   /* max|r|*dt < tol  (G_PCGSolver.cu:884-903) */
    SMAC_STOP_REL_L2 = 1         /* ||r||_2 < tol*||b||_2                     */
};

/* ---- descriptors --------------------------------------------------------- */
typedef struct smac_bc_entry {   /* one row of the bc table,
                                    boundaryConditionMembers.def:1-4          */
    int32_t type;                /* enum smac_bc_type */
    double vx, vy, vz;           /* wall / inflow velocity */
} smac_bc_entry;

typedef struct smac_grid_desc {
    int32_t nx, ny, nz;          /* interior cells per direction              */
    double sizex, sizey, sizez;  /* domain extents (SMACSolver::set_calc_properties) */
    double originx, originy, originz;
    int32_t periodic[3];         /* 1 = periodic in that direction (extension:
                                    the reference has no periodic BC)         */
    uint8_t side_bcid[6];        /* bc id of each wall side, indexed by smac_side
                                    (StaggeredGrid::set_boundary_id,
                                    StaggeredGrid.h:75-115); unused if periodic */
    int32_t num_bc;              /* rows of bc_table (<= 256)                 */
    const smac_bc_entry* bc_table;
} smac_grid_desc;

typedef struct smac_fluid_desc { /* G_SMACSolver::rho0_.. gx_ (G_SMACSolver.h:11-16) */
    double rho0, rho1;           /* phase densities                           */
    double mu0, mu1;             /* phase viscosities                         */
    double gx, gy, gz;           /* gravity                                   */
} smac_fluid_desc;

typedef struct smac_solver_opts {
    int32_t precond;             /* enum smac_precond                         */
    int32_t precision;           /* enum smac_precision                       */
    int32_t stop_rule;           /* enum smac_stop_rule                       */
    double tol;                  /* G_PressureSolverBase::tol_                */
    int32_t check_every;         /* convergence check cadence (ref: 10 Jacobi, 5 GMG) */
    int32_t max_iter;            /* ref: 2000                                 */
    int32_t mg_levels;           /* MG: 0 = automatic. REF_GMG: the requested
                                    num_levels of G_GMGSolver::initialize, 0 = 4
                                    (CFD/config.json "gmg_levels")            */
    int32_t nu1, nu2;            /* pre/post RBGS sweeps per level            */
    int32_t flexible;            /* 1 = Polak-Ribiere beta (auto-on for mixed) */
    int32_t ref_sweeps;          /* REF_GMG weighted-Jacobi sweeps before and
                                    after the coarse correction, num_iter_fine_ (4) */
    int32_t ref_coarse_sweeps;   /* REF_GMG sweeps on the coarsest level,
                                    num_iter_coarse_ (30, G_GMGSolver.h:21-22) */
    int32_t stationary;          /* 0 = preconditioned CG (default); 1 = the
                                    stationary iteration x += M r with the same
                                    preconditioner M: with tol = 0 a fixed linear
                                    map of the right-hand side (finite-difference
                                    checks of the adjoint). Extension. */
} smac_solver_opts;

typedef struct smac_solve_stats {
    int32_t iters;               /* PCG iterations (G_PCGSolver.cu:921)       */
    int32_t converged;
    int32_t pure_neumann;
    int32_t mg_levels;
    double rel_res;              /* ||r||_2/||b||_2 at exit                   */
    double linf_res_dt;          /* max|r|*dt at exit                         */
    double bnorm;                /* ||b - mean b||_2                          */
    double ms_predictor, ms_solve, ms_correct; /* device time, when profiling */
    int32_t nan_detected;        /* 1: ||b||^2 or a checked residual was NaN/Inf;
                                    the solve stopped there and the call returned
                                    SMAC_ERR_NAN (the reference has no check:
                                    check_nan_all is declared, never defined,
                                    CFD/G_SMACSolver.h) */
    int32_t reserved;
    double bytes_model;          /* algorithmic HBM bytes of the call (whole
                                    arrays read or written once per pass, per
                                    interior cell x cells): RHS 32 (8 for a
                                    given RHS, +16 with the volume fraction),
                                    PCG start 32 + P, per iteration
                                    104 + P (+16 flexible, +24 face
                                    coefficients); P = one V-cycle, 17.43 W_s
                                    (+17.14 W_s face coefficients), W_s the
                                    smoother width 8/4/2; Jacobi P = 16 (+24);
                                    the reference GMG: 8 sweeps x 3 words +
                                    4.25 words transfers per level, 90 words
                                    coarsest, geometric 1/8 per level. smac_step
                                    adds predictor 56 and correction 72 (+24
                                    each with face coefficients). */
} smac_solve_stats;

typedef struct smac_ctx smac_ctx;

/* ---- lifecycle ------------------------------------------------------------
 * Replaces G_SMACSolver::set_calc_properties + solver_malloc + set_block_grid
 * (CFD/G_SMACSolver.h:25,84,90-100; CFD/G_SMACSolver.cu:1451-1496) and the host
 * setup SMACSolver::set_face_type / set_cell_type / set_face_internal_direction
 * (CFD/SMACSolver.cpp:37-248). All fields start at zero (the reference
 * calloc's its host arrays, CFD/SMACSolver.cpp:423-440). Side kinds: no-slip,
 * slip, inflow and outlet; an outlet side is a Dirichlet face of the pressure
 * (its face coefficient 2 f_b of the inner face, update_boundary_faces,
 * CFD/G_SMACSolver_VoF.cu:1798-1950), so such a grid is not pure Neumann.
 * SMAC_BC_OUTFLOW, which no reference kernel handles, returns
 * SMAC_ERR_UNSUPPORTED. */
int smac_create(const smac_grid_desc* grid, const smac_fluid_desc* fluid, smac_ctx** out);
/* Replaces G_SMACSolver::solver_free (CFD/G_SMACSolver.cu:1433-1449). */
void smac_destroy(smac_ctx* ctx);
/* Replaces the rho/mu/g members written by the driver (CFD/main.cu:130-133,203-206). */
int smac_set_fluid(smac_ctx* ctx, const smac_fluid_desc* fluid);

/* ---- data movement ----------------------------------------------------------
 * Replaces G_SMACSolver::cpuTogpu / gpuTocpu (CFD/G_SMACSolver.h:82-83,
 * CFD/G_SMACSolver.cu:1507-1535) for one field. count = element count of the
 * field in the reference layout (above). Synchronous. */
int smac_upload(smac_ctx* ctx, int32_t field, const double* host, size_t count);
int smac_download(smac_ctx* ctx, int32_t field, double* host, size_t count);
/* Same, from/to a device buffer (e.g. a torch tensor) in the reference layout.
 * The copy runs on the context's non-blocking stream, which is not ordered with
 * any other stream: finish the other streams' work on the buffer first (for a
 * torch tensor, torch.cuda.synchronize(); the caching allocator hands a freed
 * block to a new tensor while kernels queued on its stream may still read it).
 * A device upload of SMAC_FIELD_ALPHA is not scanned for zeros: the grid then
 * keeps its properties in arrays, as a two-phase grid. */
int smac_upload_device(smac_ctx* ctx, int32_t field, const double* dev, size_t count);
int smac_download_device(smac_ctx* ctx, int32_t field, double* dev, size_t count);
/* Element count of a field in the reference layout. */
size_t smac_field_count(const smac_ctx* ctx, int32_t field);
/* Device layout of every field: a box of planes x pitch_y x pitch_x doubles,
 * element (i, j, k) of the reference index at (k*pitch_y + j)*pitch_x + i
 * (pitch_x >= Nx+3, rows 128-byte aligned; pitch_y = Ny+3; planes = Nz+3). The
 * reference-layout extent of a field is the sub-box from the origin; the rest is
 * padding. Extension (the reference keeps MyArray's dense layout). */
int smac_device_layout(const smac_ctx* ctx, int64_t* pitch_x, int64_t* pitch_y, int64_t* planes);
/* Zero-copy binding (e.g. of a torch tensor): from this call the context reads
 * and writes `dev` -- device memory of the context's GPU, 256-byte aligned,
 * count = pitch_x*pitch_y*planes doubles in the device layout above -- as
 * field U, V, W, P or ALPHA in place of its own array, and takes its contents
 * as the field's values. Borrowed: the context never frees it; the caller keeps
 * it alive and orders its own streams' work on it with the context stream, as
 * for smac_upload_device. dev = NULL (count 0) unbinds: the context copies the
 * bound contents into its own array and uses that again. smac_destroy drops
 * bindings without copying. Binding ALPHA marks the grid two-phase (as a device
 * upload of ALPHA); the VoF transport writes its result into the bound array.
 * Periodic images and ghost entries are the context's working values (a host
 * download fills periodic images; the bound array is not rewritten for them).
 * Returns SMAC_ERR_UNSUPPORTED for other fields, SMAC_ERR_INVALID_ARG for a
 * wrong count, misaligned or non-device pointer. Extension: the reference's
 * cpuTogpu/gpuTocpu (CFD/G_SMACSolver.cu:1507-1535) always copy. */
int smac_bind_device(smac_ctx* ctx, int32_t field, double* dev, size_t count);


/* ---- the step ---------------------------------------------------------------
 * Property assembly from alpha: G_SMACSolver::update_properties_by_alpha +
 * update_boundary_faces (CFD/G_SMACSolver_VoF.cu:1544-1566,1944-1950), the t=0
 * host form SMACSolver::update_properties_by_alpha_initial
 * (CFD/SMACSolver.cpp:283-421) when initial != 0 (rho_old = rho). */
int smac_update_properties(smac_ctx* ctx, int32_t initial);
/* VoF transport of alpha over one time step dt by the split THINC/WLIC scheme
 * (CFD/main.cu:283-295): G_SMACSolver::clear_alpha_flux_accum, then for each of
 * `substeps` substeps clear_alpha_flux, alpha_flux_thincwlic_split(dt/substeps, step)
 * and alpha_flux_accum (CFD/G_SMACSolver_VoF.cu:451-1190,1575-1706). step (the
 * time-step counter) picks the direction order, step % 3 -> xyz, yzx, zxy. The
 * accumulated face fluxes (SMAC_FIELD_FLUX{X,Y,Z}) enter the mass flux of the next
 * predictor. Call smac_update_properties(ctx, 1) once before the first
 * transport, as the reference's t = 0 setup does. */
int smac_transport_alpha(smac_ctx* ctx, double dt, int32_t substeps, int32_t step);
/* The two halves of smac_transport_alpha, for callers that keep the reference's
 * substep loop: G_SMACSolver::clear_alpha_flux_accum
 * (CFD/G_SMACSolver_VoF.cu:1688-1696), and one substep = clear_alpha_flux +
 * alpha_flux_thincwlic_split(sub_dt, step) + alpha_flux_accum
 * (CFD/G_SMACSolver_VoF.cu:468-505,1575-1706). */
int smac_clear_alpha_flux_accum(smac_ctx* ctx);
int smac_alpha_flux_split(smac_ctx* ctx, double sub_dt, int32_t step);
/* Mass flux G_SMACSolver::compute_mass_flux_from_alpha_flux
 * (CFD/G_SMACSolver_VoF.cu:1708-1795): mf = rho0*u_f + (rho1-rho0)*F_accum*h/dt
 * right after smac_transport_alpha, else mf = rho0*u_f (no alpha flux). The
 * predictor forms it itself; the transported flux serves one predictor. */
int smac_compute_mass_flux(smac_ctx* ctx, double dt);
/* G_SMACSolver::get_vof_vstar_rhouu_consistent + update_vstar_boundary
 * (CFD/G_SMACSolver.h:71,74; CFD/G_SMACSolver.cu:1044-1069). */
int smac_predictor(smac_ctx* ctx, double dt);
/* Two-way CFD-DEM variants (the coupled loop Coupling/main.cu:531-540,661-670).
 * G_SMACSolver::compute_mass_flux_from_alpha_flux_two_way +
 * get_vof_vstar_rhouu_consistent_two_way + update_vstar_boundary
 * (CFD/G_SMACSolver_VoF.cu:2472-2557; CFD/G_SMACSolver.cu:1538-2203,1044-1052):
 * the predictor with the stresses and the old pressure gradient weighted by the
 * t^n volume fraction, the mass flux rho0*eps_half_f*u_f (+ the alpha flux term
 * right after a VoF transport), and u* = (beta/eps_f)(eps^n_f rho^n_f u + dt*rhs
 * + J/V) + dt*g with J the impulse fields (zero when not uploaded). Needs
 * SMAC_FIELD_EPS, EPS_OLD and EPS_HALF uploaded (ghost cells included, as the
 * coupling fills them); else SMAC_ERR_INVALID_ARG. */
int smac_predictor_two_way(smac_ctx* ctx, double dt);
/* G_CFDDEMCoupling::update_poisson_beta_two_way (Coupling/G_CFDDEMCoupling.cu:419-459):
 * the Poisson face coefficients become eps_f/rho_f (eps_f the mean of the two cells'
 * SMAC_FIELD_EPS) until the next smac_update_properties, which restores 1/rho_f as
 * the reference's face-property update does. The correction keeps 1/rho_f. */
int smac_update_poisson_beta_two_way(smac_ctx* ctx);
/* G_SMACSolver::solve_poisson -> G_PCGSolver::solve (CFD/G_SMACSolver.cu:1499-1501,
 * CFD/pressure_solver/G_PCGSolver.cu:592-617): RHS assembly
 * (G_PressureSolverBase.cu:49-118), operator build (G_PCGSolver.cu:126-212) and the
 * preconditioned CG on A dp = b, dp starting from 0. Once SMAC_FIELD_EPS and
 * EPS_OLD are uploaded the RHS carries the volume fraction,
 * b = -(1/dt) D(eps_f u*) - (eps - eps_old)/dt^2 (the reference's general form;
 * without them eps = 1). */
int smac_pressure_solve(smac_ctx* ctx, double dt, const smac_solver_opts* opts,
                        smac_solve_stats* stats);
/* G_SMACSolver::correct_vof_velocity (CFD/G_SMACSolver.cu:1072-1142,1269-1275):
 * u = u* - beta*dt*grad(dp) on interior faces, p += dp, dp = 0. */
int smac_correct(smac_ctx* ctx, double dt);
/* One whole step: properties (if variable), mass flux, predictor, solve, correct,
 * in the order of the driver loop CFD/main.cu:298-317. */
int smac_step(smac_ctx* ctx, double dt, const smac_solver_opts* opts, smac_solve_stats* stats);
/* G_SMACSolver::calc_cfl (CFD/G_SMACSolver.cu:1365-1430): convective and viscous
 * CFL numbers for time step dt. */
int smac_calc_cfl(smac_ctx* ctx, double dt, double* cfl_conv, double* cfl_visc);
/* smac_calc_cfl that allocates no device memory. No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Same kernels
 * and the same result as smac_calc_cfl, bit for bit. Its scratch (two row
 * partials of ny*nz doubles, the lane sums of the two-stage reduction and the
 * two result slots, 8 MiB + 16.25 KiB) is one static device array per GPU,
 * resolved once per device by smac_create and held until the process ends.
 * Every call on a GPU shares that GPU's array: calls that overlap on one GPU
 * (other threads, other contexts' streams) race; the array keeps the last
 * call's contents and nothing in it needs to start at zero (every slot a call
 * reads is written first by that call). Returns SMAC_ERR_INVALID_GRID when
 * ny*nz > SMAC_NOALLOC_MAX_ROWS, and SMAC_ERR_UNSUPPORTED for a two-phase
 * context whose properties are not built yet (they would be allocated: call
 * smac_update_properties or take a step first). Synchronous. */
#define SMAC_NOALLOC_MAX_ROWS (1 << 19)
int smac_calc_cfl_noalloc(smac_ctx* ctx, double dt, double* cfl_conv, double* cfl_visc);

/* ---- after the step (CFD/main.cu:316-317) -----------------------------------
 * G_SMACSolver::update_boundary_ghost (CFD/G_SMACSolver.cu:1345-1362):
 *  - ghost pressure p(g) = p(s) + rho(s)*(r.g) with s the nearest interior cell
 *    and r = x(g) - x(s) (k_update_cell_ghost_pressure, CFD/G_SMACSolver.cu:1227-1267,
 *    which overwrites every ghost k_update_cell_boundary_pressure wrote);
 *  - tangential velocity ghosts by d_get_velocity_ghost_value with the bc id of the
 *    face one row inside, and edge ghosts as the mean of the two adjacent ghosts
 *    (k_update_v{x,y,z}_ghost[_corner], CFD/G_BoundaryFunctions.cu:257-540);
 *  - outlet faces first (k_update_v{x,y,z}_outlet, CFD/G_BoundaryFunctions.cu:143-255):
 *    u = u* + s beta_f dt dp(c)/h on the faces of an outlet side, c the inner
 *    cell, s = +1 on the high side and -1 on the low side (after smac_correct,
 *    which clears dp as the reference's memset does, the face keeps u*);
 * Periodic directions (the extension) get periodic images instead. dt is the
 * outlet update's time step. */
int smac_update_boundary_ghost(smac_ctx* ctx, double dt);
/* G_SMACSolver::make_face_gradp (CFD/G_SMACSolver.cu:1277-1343): face pressure
 * gradient (p(c) - p(c-1))/h of each axis into SMAC_FIELD_GRADP{X,Y,Z}, on the
 * faces whose three indices all lie in 1..N+1 (others stay 0). */
int smac_make_face_gradp(smac_ctx* ctx);
/* The cell data of FileInOut::output_vti_binary_cellData (CFD/FileInOut.cpp:145-255):
 * host_out receives count = 5*Nx*Ny*Nz floats, x fastest over the interior
 * cells: pressure[n], alpha[n], then velocity[3n] with (u, v, w) interleaved,
 * each component the mean of the cell's two faces. Packed on the device and
 * copied once; the VTI file itself is written by smac_config::write_vti
 * (include/smac_config.hpp). */
int smac_cell_output(smac_ctx* ctx, float* host_out, size_t count);

/* Adjoint (vector-Jacobian product) of one smac_step with time step dt taken
 * from the velocity u_n (three device arrays, x/y/z faces, reference layout):
 * on entry ubar[3] (faces) and pbar (cells), device arrays in the reference
 * layout, hold the cotangents of the step's output (u^{n+1}, p^{n+1}); on
 * return they hold the cotangents of its input (u^n, p^n). The predictor is
 * transposed at u_n (upwind side and minmod branch of the forward), the
 * Poisson solve is repeated with opts (the operator is self-adjoint) and its
 * statistics go to stats. Entries that are not state (ghosts, periodic images)
 * are ignored on entry and zero on return. The context's velocity is
 * overwritten with u_n; two-phase properties are those of the context.
 * SMAC_ERR_UNSUPPORTED while a VoF flux is pending or once volume fractions
 * (SMAC_FIELD_EPS) are in the context: the adjoint is that of the one-fluid step.
 * No counterpart in the reference (it has no adjoint); it serves gradient-based
 * callers through a torch.autograd.Function (mydemcfd_amd/autograd.py). */
int smac_adjoint_step(smac_ctx* ctx, double dt, const smac_solver_opts* opts, const double* const u_n[3],
                      double* const ubar[3], double* pbar, smac_solve_stats* stats);

/* Adjoint (vector-Jacobian product) of one step of the VoF loop -- the
 * transport smac_transport_alpha(ctx, dt, substeps, step) followed by
 * smac_step(ctx, dt, opts), the order of CFD/main.cu:283-317 -- taken from the
 * state (alpha_n, u_n, p_n) (device arrays, reference layout). On entry
 * alpha_bar (cells), ubar[3] (faces) and pbar (cells) hold the cotangents of
 * the step's output (alpha^{n+1}, u^{n+1}, p^{n+1}); on return those of its
 * input. Covers the split THINC/WLIC transport (the branch the forward took:
 * upwind donor, THINC/WLIC or first order, the clip), the property map
 * alpha -> (rho, mu, beta = 1/rho_f) of alpha^{n+1} and rho_old = rho(alpha^n),
 * the alpha-flux mass flux of the predictor, the alpha-dependent Poisson
 * operator (a second solve with opts) and the correction. The step is taken
 * again from the given state (the context's fields are overwritten: alpha
 * ends at alpha^{n+1}); statistics of the second solve go to stats. Entries
 * that are not state are ignored on entry and zero on return.
 * SMAC_ERR_UNSUPPORTED with volume fractions or outlet sides. No counterpart
 * in the reference (it has no adjoint); it serves gradient-based callers of the
 * VoF loop through mydemcfd_amd/autograd.py (vof_rollout). */
int smac_adjoint_vof_step(smac_ctx* ctx, double dt, int32_t substeps, int32_t step, const smac_solver_opts* opts,
                          const double* alpha_n, const double* const u_n[3], const double* p_n, double* alpha_bar,
                          double* const ubar[3], double* pbar, smac_solve_stats* stats);

/* Fill opts with the defaults (RBGS multigrid, FP64, REL_L2, 1e-8). */
void smac_default_opts(smac_solver_opts* opts);
/* Device-time profiling of the step stages and of named kernels (hipEvents). */
int smac_set_profiling(smac_ctx* ctx, int32_t on);
/* Accumulated count and device milliseconds of kernel `name` since profiling was
 * switched on. Returns SMAC_ERR_INVALID_ARG for an unknown name. */
int smac_kernel_time(smac_ctx* ctx, const char* name, int64_t* launches, double* ms);
/* Comma-separated names of the kernels timed so far (buf of len bytes). */
int smac_kernel_names(smac_ctx* ctx, char* buf, size_t len);
/* Zero the accumulated kernel times. */
int smac_reset_kernel_times(smac_ctx* ctx);
/* Synchronize the context stream. */
int smac_sync(smac_ctx* ctx);
/* HIP stream of the context (hipStream_t), for callers that order work with it. */
void* smac_stream(smac_ctx* ctx);
const char* smac_strerror(int status);
const char* smac_backend_name(void);

#ifdef __cplusplus
}
#endif
#endif /* SMAC_C_API_H */
