/*
 * smac_oracle.c -- CPU FP64 restatement of the reference SMAC projection step.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP product
 * (mydemcfd_amd/csrc). Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it. It exports the same C-ABI as the product
 * (include/smac_c_api.h) so both can be driven by identical test code.
 *
 * Reference: bi-rugaoishii/myDEMCFD. The reference CPU SMACSolver has no step
 * (CFD/SMACSolver.cpp holds setup only; CFD/main.cu:60-69 refuses CPU runs), so
 * this file restates the reference's GPU kernels, with eps (void fraction) = 1,
 * surface tension = 0, no IBM solid and no CFD-DEM coupling impulse:
 *   predictor          CFD/G_SMACSolver.cu:67-88 (minmod), 90-395 (x), 397-717 (y), 721-1012 (z)
 *   boundary getters   CFD/G_BoundaryFunctions.h:9-486, star BC G_BoundaryFunctions.cu:3-71
 *   properties         CFD/SMACSolver.cpp:283-421, CFD/G_SMACSolver_VoF.cu:1374-1566,1798-1950
 *   mass flux          CFD/G_SMACSolver_VoF.cu:1708-1790
 *   Poisson RHS        CFD/pressure_solver/G_PressureSolverBase.cu:49-118
 *   operator, SpMV     CFD/pressure_solver/G_PCGSolver.cu:87-245
 *   Jacobi PCG         CFD/pressure_solver/G_PCGSolver.cu:307-586
 *   mean removal       CFD/pressure_solver/G_PressureSolverBase.cu:142-155
 *   correction         CFD/G_SMACSolver.cu:1072-1142,1269-1275
 *   CFL                CFD/G_SMACSolver.cu:1365-1430
 * Extensions the reference lacks (periodic directions, the RBGS multigrid
 * V-cycle preconditioner, the relative-L2 stopping rule) follow the build's own
 * specification; the GPU product implements the same arithmetic.
 *
 * Determinism contract shared with the GPU product: no FMA contraction
 * (-ffp-contract=off), every expression evaluated in the order written here,
 * and every sum over interior cells taken in the canonical order of csum()
 * (row partials over 64 lanes, halving tree, then a 1024-wide final pass).
 */
#include "../include/smac_c_api.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { F_INTERIOR = 0, F_BOUNDARY = 1, F_GHOST = 2 };   /* CFD/Enums.h:3-8 */

/* Planted mistakes, for the tests that show the closed-form pins catch them
 * (tests/test_oracle_pins.py builds this file with -DSMAC_ORACLE_MUTANT=n into
 * a scratch directory; the default build, and everything that checks the
 * product, has n = 0):
 *   1  wall mirror drops the wall velocity (-v_in instead of 2 v_wall - v_in)
 *   2  the two wall faces' velocities summed, not averaged
 *   3  slip wall mirrored like a no-slip wall
 *   4  inflow face value taken as zero
 *   5  convective CFL from the smaller of a cell's two face speeds
 *   6  viscous CFL from mu instead of mu/rho
 *   7  outlet face coefficient beta of the inner face instead of 2 beta
 *   8  outlet face coefficient 2/rho of the inner cell instead of 2 beta of
 *      the inner face
 *   9  star value of an outlet face taken as zero (no copy of the inner face)
 *  10  outlet face value taken as zero in the predictor's getter
 *  11  outlet velocity ghost copies the inner row (bval) instead of 2 bval - inval
 *  12  an outlet side counted as pure Neumann (mean removal of the rhs)
 *  13  outlet face update with the sign of the dp term flipped */
#ifndef SMAC_ORACLE_MUTANT
#define SMAC_ORACLE_MUTANT 0
#endif

#define MAXL 12

typedef struct level {
    int n[3];            /* interior cells */
    int p[3];            /* pitches n+3 */
    size_t tot;
    double inv_d2[3];
    double *e, *f, *res; /* correction, rhs, residual (cells) */
    double *b[3];        /* face coefficients beta */
} level;

struct smac_ctx {
    int n[3], p[3];
    size_t tot;
    double d[3], inv_d[3], inv_d2[3];
    int per[3];
    uint8_t side_bcid[6];
    int nbc;
    int has_outlet;
    int bctype[256];
    double bcv[256][3];
    smac_fluid_desc fl;
    /* fields, all in the common (n+3)^3 layout */
    double *vel[3], *vst[3], *p_, *dp, *rhs, *alpha;
    double *rho, *rho_old, *inv_rho, *mu;
    double *beta[3], *inv_rhof[3], *muf[3], *mf[3], *rhof[3];
    double *gradp[3];    /* f_gradp_{x,y,z}_ (smac_make_face_gradp) */
    /* VoF transport (allocated on first use): alpha_new_, f_F{x,y,z}_,
     * f_F{x,y,z}_accum_, correct_coeff_; vof_armed = the accumulated fluxes
     * belong to the coming predictor */
    double *alpha_new, *vf[3], *vfacc[3];
    unsigned char* vcoeff;
    int vof_armed;
    /* CFD-DEM (two-way) inputs (allocated on upload): void_fraction_,
     * void_fraction_old_, void_fraction_half_ (cells), f_coupling_impulse_*
     * (faces); bpois = the Poisson face coefficients eps_f/rho_f in use from
     * smac_update_poisson_beta_two_way to the next smac_update_properties */
    double *eps, *eps_old, *eps_half, *imp[3], *bpois[3];
    int bpois_on;
    /* PCG */
    double *r, *z, *dvec, *Ad;
    double *Adiag, *Aoff[6];
    /* multigrid */
    int nlev;
    level lev[MAXL];
    double* cinv;        /* dense pseudo-inverse on the coarsest level */
    int cinv_n;
    int mg_ready;
    int props_ready;
    /* the reference's GMG (REF_GMG) */
    int nrg, rg_req;
    struct rglevel* rg;
};

/* ---------------------------------------------------------------- indexing */
static inline size_t IDX(const struct smac_ctx* c, int i, int j, int k) {
    return ((size_t)k * (size_t)c->p[1] + (size_t)j) * (size_t)c->p[0] + (size_t)i;
}
static inline int wrap(int i, int n) { return i < 1 ? i + n : (i > n ? i - n : i); }
/* resolve an index along dimension d: periodic directions wrap into 1..n */
static inline int W(const struct smac_ctx* c, int d, int i) { return c->per[d] ? wrap(i, c->n[d]) : i; }

typedef struct { int v[3]; } ivec;
static inline ivec iv(int i, int j, int k) { ivec r = {{i, j, k}}; return r; }
static inline ivec add(ivec a, int d, int s) { a.v[d] += s; return a; }
static inline ivec wr(const struct smac_ctx* c, ivec a) {
    for (int d = 0; d < 3; d++) a.v[d] = W(c, d, a.v[d]);
    return a;
}
static inline size_t IX(const struct smac_ctx* c, ivec a) { return IDX(c, a.v[0], a.v[1], a.v[2]); }

/* Face type of the a-face at (already wrapped) index q: analytic form of
 * SMACSolver::set_face_type (CFD/SMACSolver.cpp:104-202) for a box without
 * internal walls; periodic directions have no walls. */
static int face_type(const struct smac_ctx* c, int a, ivec q) {
    for (int d = 0; d < 3; d++) {
        if (d == a) continue;
        if (q.v[d] < 1 || q.v[d] > c->n[d]) return F_GHOST;
    }
    int i = q.v[a], n = c->n[a];
    if (c->per[a]) return (i >= 1 && i <= n) ? F_INTERIOR : F_GHOST;
    if (i == 1 || i == n + 1) return F_BOUNDARY;
    if (i >= 2 && i <= n) return F_INTERIOR;
    return F_GHOST;
}
/* bc id of an a-face; 0 off the walls as in the calloc'd reference array
 * (StaggeredGrid::set_boundary_id, CFD/StaggeredGrid.h:75-115). */
static int face_bcid(const struct smac_ctx* c, int a, ivec q) {
    int i = q.v[a];
    if (c->per[a]) return 0;
    for (int d = 0; d < 3; d++) {
        if (d == a) continue;
        if (q.v[d] < 1 || q.v[d] > c->n[d]) return 0;
    }
    if (i == 1) return c->side_bcid[2 * a];
    if (i == c->n[a] + 1) return c->side_bcid[2 * a + 1];
    return 0;
}
/* cell type: interior iff inside (SMACSolver::set_cell_type, CFD/SMACSolver.cpp:206-248) */
static inline int cell_interior(const struct smac_ctx* c, ivec q) {
    for (int d = 0; d < 3; d++)
        if (q.v[d] < 1 || q.v[d] > c->n[d]) return 0;
    return 1;
}

/* ------------------------------------------------------ boundary getters */
/* d_get_v{x,y,z}_{x,y,z}face (CFD/G_BoundaryFunctions.h:9-97) */
static double get_face(const struct smac_ctx* c, int a, ivec q) {
    q = wr(c, q);
    int t = face_type(c, a, q);
    if (t == F_INTERIOR) return c->vel[a][IX(c, q)];
    if (t == F_BOUNDARY) {
        int bid = face_bcid(c, a, q);
        int bt = c->bctype[bid];
        if (bt == SMAC_BC_INFLOW) return SMAC_ORACLE_MUTANT == 4 ? 0. : (1. - 0.0) * c->bcv[bid][a];
        if (bt == SMAC_BC_NOSLIP) return 0.;
        if (bt == SMAC_BC_OUTLET) return SMAC_ORACLE_MUTANT == 10 ? 0. : c->vel[a][IX(c, q)];
        if (bt == SMAC_BC_SLIP && a != 0) return c->vel[a][IX(c, q)];
    }
    return 0.;
}
/* d_get_v{a}_{b}dir (CFD/G_BoundaryFunctions.h:100-367): the a-velocity one
 * step along b, mirrored across a wall. q is an interior a-face. */
static double get_dir(const struct smac_ctx* c, int a, int b, ivec q, int s) {
    q = wr(c, q);
    double vin = c->vel[a][IX(c, q)];
    ivec q2 = wr(c, add(q, b, s));
    if (face_type(c, a, q2) == F_INTERIOR) return c->vel[a][IX(c, q2)];
    ivec qf = q;
    qf.v[b] = s > 0 ? q.v[b] + 1 : q.v[b];
    double vw = 0.0;
    int count = 0;
    if (face_type(c, b, qf) == F_BOUNDARY) {
        vw += c->bcv[face_bcid(c, b, qf)][a];
        count++;
    }
    ivec qf2 = wr(c, add(qf, a, -1));
    if (face_type(c, b, qf2) == F_BOUNDARY) {
        vw += c->bcv[face_bcid(c, b, qf2)][a];
        count++;
    }
    if (count > 0 && SMAC_ORACLE_MUTANT != 2) vw /= (double)count;
    int bt = c->bctype[face_bcid(c, b, qf)];
    if ((bt == SMAC_BC_SLIP && SMAC_ORACLE_MUTANT != 3) || bt == SMAC_BC_OUTLET) return vin;
    if (SMAC_ORACLE_MUTANT == 1) return -vin;
    return 2.0 * vw - vin;
}

/* d_minmod (CFD/G_SMACSolver.cu:67-88) */
static inline double minmod(double dp, double dm) {
    double prod = dp * dm;
    double s = 0.;
    if (prod > 0.) s = fabs(dp) < fabs(dm) ? dp : dm;
    return s;
}
/* sgn (CFD/G_SMACSolver.h:130-139) */
static inline int sgn(double a) { return a < 0 ? -1 : 1; }

/* ------------------------------------------------------------- predictor */
/* One component of k_get_vof_vstar_rhouu_consistent_{x,y,z}
 * (CFD/G_SMACSolver.cu:139-392 / 441-662 / 763-1003) at interior a-face q. */
/* Mass flux of k_compute_mass_flux_from_alpha_flux_two_way
 * (CFD/G_SMACSolver_VoF.cu:2472-2557) at the raw a-face index r: 0 off the
 * faces 1..N+1 of axis a; eps_f = the t^{n+1/2} volume fraction of the face,
 * the mean of its two cells, the inner cell on faces 1 and N+1
 * (d_get_void_fraction_half_*face, 1954-1967); the alpha-flux term after a
 * VoF transport. */
static double mf_two_way(const struct smac_ctx* c, int a, ivec r, int with_flux, double inv_dt) {
    int m = r.v[a];
    if (m < 1 || m > c->n[a] + 1) return 0.0;
    double ef;
    if (m == 1) ef = c->eps_half[IX(c, r)];
    else if (m == c->n[a] + 1) ef = c->eps_half[IX(c, add(r, a, -1))];
    else ef = 0.5 * (c->eps_half[IX(c, add(r, a, -1))] + c->eps_half[IX(c, r)]);
    double q = ef * get_face(c, a, r);
    if (!with_flux) return c->fl.rho0 * q;
    double hbydt = c->d[a] * inv_dt;
    return c->fl.rho0 * q + (c->fl.rho1 - c->fl.rho0) * (c->vfacc[a][IX(c, wr(c, r))] * hbydt);
}
/* mass flux seen by the predictor at a-face r: the precomputed mf (one-fluid
 * form) or the two-way form (tw: 1 = without, 2 = with the alpha flux) */
static inline double mf_at(const struct smac_ctx* c, int a, ivec r, int tw, double inv_dt) {
    if (tw) return mf_two_way(c, a, r, tw == 2, inv_dt);
    return c->mf[a][IX(c, wr(c, r))];
}

/* tw != 0: the two-way form k_get_vof_vstar_rhouu_consistent_{x,y,z}_two_way
 * (CFD/G_SMACSolver.cu:1538-2194): stresses weighted by the t^n volume
 * fraction (cells for the own direction, the mean of the four cells around an
 * edge, in memory order, for a cross direction), mass flux mf_two_way, the old
 * pressure gradient weighted by the face mean of eps^n, and
 * u* = (beta/eps^{n+1}_f)(eps^n_f rho^n_f u + dt tmp + J/V) + dt g; volume
 * fractions at raw cell indices (the caller fills the ghosts). */
static double predict_face(const struct smac_ctx* c, int a, ivec q, double dt, int tw) {
    const double* inv_d = c->inv_d;
    const double inv_dt = 1. / dt;
    double v0 = c->vel[a][IX(c, q)];
    double vp1 = get_face(c, a, add(q, a, +1));
    double vm1 = get_face(c, a, add(q, a, -1));
    ivec cr = add(q, a, -1);  /* minus cell, raw index */
    ivec cm = wr(c, cr);      /* cell on the minus side */
    const double* eo = c->eps_old;

    /* viscous, own direction */
    double tau_p = c->mu[IX(c, q)] * (vp1 - v0);
    double tau_m = c->mu[IX(c, cm)] * (v0 - vm1);
    double tmp = tw ? 2. * (eo[IX(c, q)] * tau_p - eo[IX(c, cr)] * tau_m) * c->inv_d2[a]
                    : 2. * (tau_p - tau_m) * c->inv_d2[a];

    /* viscous, cross directions in increasing axis order */
    for (int b = 0; b < 3; b++) {
        if (b == a) continue;
        double vbm = get_dir(c, a, b, q, -1);
        double vbp = get_dir(c, a, b, q, +1);
        double wpp = get_face(c, b, add(q, b, +1));
        double wmp = get_face(c, b, add(add(q, b, +1), a, -1));
        double wp0 = get_face(c, b, q);
        double wm0 = get_face(c, b, add(q, a, -1));
        ivec f1 = wr(c, add(q, b, +1)), f2 = wr(c, add(add(q, b, +1), a, -1));
        ivec f3 = wr(c, q), f4 = wr(c, add(q, a, -1));
        double mu_p = 0.5 * (c->muf[b][IX(c, f1)] + c->muf[b][IX(c, f2)]);
        double t_p = mu_p * ((wpp - wmp) * inv_d[a] + (vbp - v0) * inv_d[b]);
        double mu_m = 0.5 * (c->muf[b][IX(c, f3)] + c->muf[b][IX(c, f4)]);
        double t_m = mu_m * ((wp0 - wm0) * inv_d[a] + (v0 - vbm) * inv_d[b]);
        if (tw) {
            ivec qa = add(q, a, -1);
            double ep, em;
            if (b > a) {
                ep = 0.25 * (eo[IX(c, qa)] + eo[IX(c, q)] + eo[IX(c, add(qa, b, +1))] + eo[IX(c, add(q, b, +1))]);
                em = 0.25 * (eo[IX(c, add(qa, b, -1))] + eo[IX(c, add(q, b, -1))] + eo[IX(c, qa)] + eo[IX(c, q)]);
            } else {
                ep = 0.25 * (eo[IX(c, qa)] + eo[IX(c, add(qa, b, +1))] + eo[IX(c, q)] + eo[IX(c, add(q, b, +1))]);
                em = 0.25 * (eo[IX(c, add(qa, b, -1))] + eo[IX(c, qa)] + eo[IX(c, add(q, b, -1))] + eo[IX(c, q)]);
            }
            tmp += (ep * t_p - em * t_m) * inv_d[b];
        } else {
            tmp += (t_p - t_m) * inv_d[b];
        }
    }

    /* convection, MUSCL-minmod, directions x, y, z */
    for (int e = 0; e < 3; e++) {
        double mp, mm, up_p, up_m, Mp, Mm;
        if (e == a) {
            mp = 0.5 * (mf_at(c, a, q, tw, inv_dt) + mf_at(c, a, add(q, a, +1), tw, inv_dt));
            mm = 0.5 * (mf_at(c, a, add(q, a, -1), tw, inv_dt) + mf_at(c, a, q, tw, inv_dt));
            /* + face */
            ivec up = mp > 0. ? q : add(q, a, +1);
            if (face_type(c, a, wr(c, up)) != F_INTERIOR) {
                up_p = mp > 0. ? v0 : vp1;
            } else {
                double dpv = get_face(c, a, add(up, a, +1)) - get_face(c, a, up);
                double dmv = get_face(c, a, up) - get_face(c, a, add(up, a, -1));
                double s = minmod(dpv, dmv);
                int dir = sgn(mp);
                double uw = mp > 0. ? v0 : vp1;
                up_p = uw + dir * 0.5 * s;
            }
            Mp = mp * up_p;
            /* - face */
            up = mm > 0. ? add(q, a, -1) : q;
            if (face_type(c, a, wr(c, up)) != F_INTERIOR) {
                up_m = mm > 0. ? vm1 : v0;
            } else {
                double dpv = get_face(c, a, add(up, a, +1)) - get_face(c, a, up);
                double dmv = get_face(c, a, up) - get_face(c, a, add(up, a, -1));
                double s = minmod(dpv, dmv);
                int dir = sgn(mm);
                double uw = mm > 0. ? vm1 : v0;
                up_m = uw + dir * 0.5 * s;
            }
            Mm = mm * up_m;
        } else {
            double vcm = get_dir(c, a, e, q, -1);
            double vcp = get_dir(c, a, e, q, +1);
            mp = 0.5 * (mf_at(c, e, add(add(q, e, +1), a, -1), tw, inv_dt) + mf_at(c, e, add(q, e, +1), tw, inv_dt));
            mm = 0.5 * (mf_at(c, e, q, tw, inv_dt) + mf_at(c, e, add(q, a, -1), tw, inv_dt));
            ivec up = mp > 0. ? q : add(q, e, +1);
            ivec upw = wr(c, up);
            if (face_type(c, a, upw) != F_INTERIOR) {
                up_p = mp > 0. ? v0 : vcp;
            } else {
                double vu = c->vel[a][IX(c, upw)];
                double dpv = get_dir(c, a, e, upw, +1) - vu;
                double dmv = vu - get_dir(c, a, e, upw, -1);
                double s = minmod(dpv, dmv);
                int dir = sgn(mp);
                double uw = mp > 0. ? v0 : vcp;
                up_p = uw + dir * 0.5 * s;
            }
            Mp = mp * up_p;
            up = mm > 0. ? add(q, e, -1) : q;
            upw = wr(c, up);
            if (face_type(c, a, upw) != F_INTERIOR) {
                up_m = mm > 0. ? vcm : v0;
            } else {
                double vu = c->vel[a][IX(c, upw)];
                double dpv = get_dir(c, a, e, upw, +1) - vu;
                double dmv = vu - get_dir(c, a, e, upw, -1);
                double s = minmod(dpv, dmv);
                int dir = sgn(mm);
                double uw = mm > 0. ? vcm : v0;
                up_m = uw + dir * 0.5 * s;
            }
            Mm = mm * up_m;
        }
        tmp -= (Mp - Mm) * inv_d[e];
    }

    double beta = c->beta[a][IX(c, q)];
    double rho_old_f = 0.5 * (c->rho_old[IX(c, q)] + c->rho_old[IX(c, cm)]);
    double g = a == 0 ? c->fl.gx : (a == 1 ? c->fl.gy : c->fl.gz);
    if (tw) {
        double eo_f = 0.5 * (eo[IX(c, cr)] + eo[IX(c, q)]);
        double en_f = 0.5 * (c->eps[IX(c, cr)] + c->eps[IX(c, q)]);
        tmp -= eo_f * (c->p_[IX(c, q)] - c->p_[IX(c, cm)]) * inv_d[a];
        double inv_vol = c->inv_d[0] * c->inv_d[1] * c->inv_d[2];
        double jd = c->imp[a] ? c->imp[a][IX(c, q)] * inv_vol : 0.0;
        double inv_eps_rho = beta / en_f;
        return inv_eps_rho * (eo_f * rho_old_f * v0 + dt * tmp + jd) + dt * g;
    }
    tmp -= (c->p_[IX(c, q)] - c->p_[IX(c, cm)]) * inv_d[a];
    return beta * (v0 * rho_old_f) + dt * (beta * tmp + g);
}

/* range of face indices along a that are F_INTERIOR */
static inline int face_lo(const struct smac_ctx* c, int a) { return c->per[a] ? 1 : 2; }
static inline int face_hi(const struct smac_ctx* c, int a) { return c->n[a]; }

/* d_get_v{x,y,z}star_{x,y,z}face (CFD/G_BoundaryFunctions.h:369-486) */
static double get_star_boundary(const struct smac_ctx* c, int a, ivec q) {
    int bid = face_bcid(c, a, q);
    int bt = c->bctype[bid];
    if (bt == SMAC_BC_INFLOW) return (1. - 0.0) * c->bcv[bid][a];
    if (bt == SMAC_BC_NOSLIP) return 0.;
    if (bt == SMAC_BC_OUTLET) {
        int internal = (q.v[a] == 1) ? 0 : -1; /* set_face_internal_direction */
        if (SMAC_ORACLE_MUTANT == 9) return 0.;
        return internal < 0 ? c->vst[a][IX(c, add(q, a, -1))] : c->vst[a][IX(c, add(q, a, +1))];
    }
    return 0.;
}

static void predictor_run(smac_ctx* c, double dt, int tw);
int smac_predictor(smac_ctx* c, double dt) {
    if (!c || !(dt > 0)) return SMAC_ERR_INVALID_ARG;
    if (!c->props_ready) smac_update_properties(c, 1);
    smac_compute_mass_flux(c, dt);
    c->vof_armed = 0; /* the accumulated alpha flux serves one predictor */
    predictor_run(c, dt, 0);
    return SMAC_OK;
}
/* two-way form (G_SMACSolver::compute_mass_flux_from_alpha_flux_two_way +
 * get_vof_vstar_rhouu_consistent_two_way + update_vstar_boundary,
 * CFD/G_SMACSolver_VoF.cu:2472-2557, CFD/G_SMACSolver.cu:1538-2203) */
int smac_predictor_two_way(smac_ctx* c, double dt) {
    if (!c || !(dt > 0)) return SMAC_ERR_INVALID_ARG;
    if (!c->eps || !c->eps_old || !c->eps_half) return SMAC_ERR_INVALID_ARG;
    if (!c->props_ready) smac_update_properties(c, 1);
    int tw = c->vof_armed ? 2 : 1;
    c->vof_armed = 0;
    predictor_run(c, dt, tw);
    return SMAC_OK;
}
static void predictor_run(smac_ctx* c, double dt, int tw) {
    for (int a = 0; a < 3; a++) {
        int lo[3], hi[3];
        for (int d = 0; d < 3; d++) { lo[d] = 1; hi[d] = c->n[d]; }
        lo[a] = face_lo(c, a);
        hi[a] = face_hi(c, a);
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = lo[2]; k <= hi[2]; k++)
            for (int j = lo[1]; j <= hi[1]; j++)
                for (int i = lo[0]; i <= hi[0]; i++) {
                    ivec q = iv(i, j, k);
                    c->vst[a][IX(c, q)] = predict_face(c, a, q, dt, tw);
                }
    }
    /* update_vstar_boundary: wall faces of each component */
    for (int a = 0; a < 3; a++) {
        if (c->per[a]) continue;
        int b1 = (a + 1) % 3, b2 = (a + 2) % 3;
        for (int side = 0; side < 2; side++) {
            int fi = side == 0 ? 1 : c->n[a] + 1;
            for (int t2 = 1; t2 <= c->n[b2]; t2++)
                for (int t1 = 1; t1 <= c->n[b1]; t1++) {
                    ivec q;
                    q.v[a] = fi; q.v[b1] = t1; q.v[b2] = t2;
                    c->vst[a][IX(c, q)] = get_star_boundary(c, a, q);
                }
        }
    }
}

/* -------------------------------------------------- canonical reductions */
/* Sum of f(i,j,k) over interior cells. Row (j,k): lane l accumulates
 * i = 1+l, 1+l+64, ... in that order; the 64 lanes are combined by a halving
 * tree (lane l += lane l+s, s = 32..1). Row partials, in row order
 * (k-1)*Ny+(j-1), are combined by 1024 lanes the same way. The GPU kernels
 * (wave64 shuffles) produce bit-identical results. */
typedef double (*cellfn)(const struct smac_ctx*, size_t, const void*);

static double tree(double* a, int n) {
    for (int s = n / 2; s >= 1; s /= 2)
        for (int l = 0; l < s; l++) a[l] = a[l] + a[l + s];
    return a[0];
}
static double final_reduce(const double* part, size_t nrows) {
    double lanes[1024];
    for (int t = 0; t < 1024; t++) {
        double acc = 0.0;
        for (size_t r = (size_t)t; r < nrows; r += 1024) acc += part[r];
        lanes[t] = acc;
    }
    return tree(lanes, 1024);
}

/* generic dot-like sum: sum_c x[c]*y[c] (y may be NULL: sum x; sq: x*x).
 * Natural modes: lane l accumulates cells i = 1+l, 1+l+64, ...
 * Pair modes (RED_*_S2): the order of a row held in the row-split layout
 * (even i first, then odd i) and read two positions per lane: lane l
 * accumulates split positions q = 1+2l, 2+2l, 1+2l+128, 2+2l+128, ...
 * (cell i = 2q for q <= nx/2, else 2(q - nx/2) - 1). */
enum { RED_SUM = 0, RED_DOT = 1, RED_SQ = 2, RED_DOT_S2 = 3, RED_SQ_S2 = 4 };
static double csum(const struct smac_ctx* c, int mode, const double* x, const double* y) {
    int nx = c->n[0], ny = c->n[1], nz = c->n[2];
    size_t nrows = (size_t)ny * (size_t)nz;
    double* part = (double*)malloc(nrows * sizeof(double));
    const int pair = mode == RED_DOT_S2 || mode == RED_SQ_S2;
    const int step = pair ? 128 : 64, width = pair ? 2 : 1;
#pragma omp parallel for schedule(static)
    for (long r = 0; r < (long)nrows; r++) {
        int j = (int)(r % ny) + 1, k = (int)(r / ny) + 1;
        double lanes[64];
        for (int l = 0; l < 64; l++) {
            double acc = 0.0;
            for (int q0 = 1 + width * l; q0 <= nx; q0 += step)
                for (int t = 0; t < width; t++) {
                    int q = q0 + t, i = q;
                    if (q > nx) break;
                    if (pair) i = q <= nx / 2 ? 2 * q : 2 * (q - nx / 2) - 1;
                    size_t id = IDX(c, i, j, k);
                    double v = mode == RED_SUM ? x[id]
                               : (mode == RED_SQ || mode == RED_SQ_S2 ? x[id] * x[id] : x[id] * y[id]);
                    acc += v;
                }
            lanes[l] = acc;
        }
        part[r] = tree(lanes, 64);
    }
    double s = final_reduce(part, nrows);
    free(part);
    return s;
}
static double cmaxabs(const struct smac_ctx* c, const double* x) {
    double m = 0.0;
    for (int k = 1; k <= c->n[2]; k++)
        for (int j = 1; j <= c->n[1]; j++)
            for (int i = 1; i <= c->n[0]; i++) {
                double v = fabs(x[IDX(c, i, j, k)]);
                if (v > m || v != v) m = v;
            }
    return m;
}

/* ----------------------------------------------------------- properties */
int smac_update_properties(smac_ctx* c, int32_t initial) {
    if (!c) return SMAC_ERR_INVALID_ARG;
    if (c->bpois_on) { c->bpois_on = 0; c->mg_ready = 0; }
    const smac_fluid_desc* f = &c->fl;
    size_t tot = c->tot;
    if (!initial) { /* rho_old <- rho (the swap of update_properties_by_alpha) */
        double* t = c->rho_old; c->rho_old = c->rho; c->rho = t;
    }
    for (int k = 1; k <= c->n[2]; k++)
        for (int j = 1; j <= c->n[1]; j++)
            for (int i = 1; i <= c->n[0]; i++) {
                size_t id = IDX(c, i, j, k);
                double al = c->alpha[id];
                c->rho[id] = (1. - al) * f->rho0 + al * f->rho1;
                c->inv_rho[id] = 1. / c->rho[id];
                c->mu[id] = (1. - al) * f->mu0 + al * f->mu1;
                if (initial) c->rho_old[id] = c->rho[id];
            }
    (void)tot;
    /* ghost cells: periodic images, or copies of the inner cell
     * (k_update_cell_boundary_properties_by_alpha, G_SMACSolver_VoF.cu:1244-1307) */
    for (int k = 0; k <= c->n[2] + 1; k++)
        for (int j = 0; j <= c->n[1] + 1; j++)
            for (int i = 0; i <= c->n[0] + 1; i++) {
                ivec q = iv(i, j, k);
                if (cell_interior(c, q)) continue;
                ivec s = q;
                for (int d = 0; d < 3; d++) {
                    if (c->per[d]) s.v[d] = wrap(s.v[d], c->n[d]);
                    else if (s.v[d] < 1) s.v[d] = 1;
                    else if (s.v[d] > c->n[d]) s.v[d] = c->n[d];
                }
                size_t id = IX(c, q), is = IX(c, s);
                c->rho[id] = c->rho[is];
                c->mu[id] = c->mu[is];
                c->inv_rho[id] = c->inv_rho[is];
                if (initial) c->rho_old[id] = c->rho_old[is];
            }
    /* faces (k_update_{x,y,z}_face_properties_by_alpha, G_SMACSolver_VoF.cu:1409-1535;
     * boundary faces: update_boundary_faces, 1798-1941) */
    for (int a = 0; a < 3; a++) {
        for (int k = 0; k < c->p[2]; k++)
            for (int j = 0; j < c->p[1]; j++)
                for (int i = 0; i < c->p[0]; i++) {
                    ivec q = iv(i, j, k);
                    size_t id = IX(c, q);
                    int t = face_type(c, a, q);
                    if (t == F_INTERIOR) {
                        ivec cm = wr(c, add(q, a, -1));
                        ivec cp = wr(c, q);
                        c->rhof[a][id] = 0.5 * (c->rho[IX(c, cp)] + c->rho[IX(c, cm)]);
                        c->beta[a][id] = 1. / c->rhof[a][id];
                        c->inv_rhof[a][id] = c->beta[a][id];
                        c->muf[a][id] = 0.5 * (c->mu[IX(c, cp)] + c->mu[IX(c, cm)]);
                    } else if (t == F_BOUNDARY) {
                        int inner = q.v[a] == 1 ? 0 : -1;
                        ivec ci = add(q, a, inner);
                        c->rhof[a][id] = c->rho[IX(c, ci)];
                        c->muf[a][id] = c->mu[IX(c, ci)];
                        /* an outlet face: 2 f_b of the adjacent interior face
                         * ("pressure is assumed to be zero at the boundary",
                         * k_update_{x,y,z}_face_boundary_properties); the
                         * interior face's value is formed here, as its own
                         * branch forms it (the loop may not have reached it) */
                        double b = 0.;
                        if (c->bctype[face_bcid(c, a, q)] == SMAC_BC_OUTLET) {
                            ivec qi = add(q, a, inner < 0 ? -1 : 1);
                            double rfi = 0.5 * (c->rho[IX(c, qi)] + c->rho[IX(c, add(qi, a, -1))]);
                            if (SMAC_ORACLE_MUTANT == 8) rfi = c->rho[IX(c, ci)];
                            b = (SMAC_ORACLE_MUTANT == 7 ? 1.0 : 2.0) * (1. / rfi);
                        }
                        c->beta[a][id] = b;
                        c->inv_rhof[a][id] = 0.;
                    } else {
                        c->beta[a][id] = 0.;
                        c->inv_rhof[a][id] = 0.;
                    }
                }
    }
    c->props_ready = 1;
    c->mg_ready = 0;
    return SMAC_OK;
}

/* k_compute_mass_flux_from_alpha_flux (CFD/G_SMACSolver_VoF.cu:1708-1790):
 * mf = rho0*u_f + (rho1 - rho0)*(F_accum*h/dt) after a VoF transport, else
 * mf = rho0*u_f (no alpha flux) */
int smac_compute_mass_flux(smac_ctx* c, double dt) {
    if (!c) return SMAC_ERR_INVALID_ARG;
    if (c->vof_armed) {
        if (!(dt > 0)) return SMAC_ERR_INVALID_ARG;
        double inv_dt = 1. / dt, drho = c->fl.rho1 - c->fl.rho0;
        for (int a = 0; a < 3; a++) {
            double hbydt = c->d[a] * inv_dt;
            for (int k = 0; k < c->p[2]; k++)
                for (int j = 0; j < c->p[1]; j++)
                    for (int i = 0; i < c->p[0]; i++) {
                        ivec q = iv(i, j, k);
                        int ok = 1;
                        for (int d = 0; d < 3; d++)
                            if (q.v[d] > c->n[d] + (d == a ? 2 : 1)) ok = 0;
                        if (!ok) continue;
                        size_t id = IX(c, q);
                        c->mf[a][id] = c->fl.rho0 * get_face(c, a, q) + drho * (c->vfacc[a][id] * hbydt);
                    }
        }
        return SMAC_OK;
    }
    for (int a = 0; a < 3; a++)
        for (int k = 0; k < c->p[2]; k++)
            for (int j = 0; j < c->p[1]; j++)
                for (int i = 0; i < c->p[0]; i++) {
                    ivec q = iv(i, j, k);
                    int ok = 1;
                    for (int d = 0; d < 3; d++)
                        if (q.v[d] > c->n[d] + (d == a ? 2 : 1)) ok = 0;
                    if (!ok) continue;
                    c->mf[a][IX(c, q)] = c->fl.rho0 * get_face(c, a, q);
                }
    return SMAC_OK;
}

/* ---------------------------------------------------------------- Poisson */
/* base_k_make_poisson_rhs with eps = 1 and no surface tension. */
static void make_rhs(struct smac_ctx* c, double inv_dt) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= c->n[2]; k++)
        for (int j = 1; j <= c->n[1]; j++)
            for (int i = 1; i <= c->n[0]; i++) {
                ivec q = iv(i, j, k);
                double div = (c->vst[0][IX(c, wr(c, add(q, 0, 1)))] - c->vst[0][IX(c, q)]) * c->inv_d[0] +
                             (c->vst[1][IX(c, wr(c, add(q, 1, 1)))] - c->vst[1][IX(c, q)]) * c->inv_d[1] +
                             (c->vst[2][IX(c, wr(c, add(q, 2, 1)))] - c->vst[2][IX(c, q)]) * c->inv_d[2];
                c->rhs[IX(c, q)] = -inv_dt * div;
            }
}
/* base_k_make_poisson_rhs with the volume fraction and no surface tension
 * (CFD/pressure_solver/G_PressureSolverBase.cu:49-118):
 * b = -(1/dt) sum_d (eps_{d+} u*_{d+} - eps_{d-} u*_{d-})/h_d - (eps - eps_old)/dt^2,
 * eps_{d+-} the face means of eps^{n+1} with the raw neighbour cells. */
static void make_rhs_eps(struct smac_ctx* c, double inv_dt) {
    const double* e = c->eps;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= c->n[2]; k++)
        for (int j = 1; j <= c->n[1]; j++)
            for (int i = 1; i <= c->n[0]; i++) {
                ivec q = iv(i, j, k);
                size_t id = IX(c, q);
                double exp_ = 0.5 * (e[id] + e[IX(c, add(q, 0, 1))]), exm = 0.5 * (e[IX(c, add(q, 0, -1))] + e[id]);
                double eyp = 0.5 * (e[id] + e[IX(c, add(q, 1, 1))]), eym = 0.5 * (e[IX(c, add(q, 1, -1))] + e[id]);
                double ezp = 0.5 * (e[id] + e[IX(c, add(q, 2, 1))]), ezm = 0.5 * (e[IX(c, add(q, 2, -1))] + e[id]);
                double div = (exp_ * c->vst[0][IX(c, wr(c, add(q, 0, 1)))] - exm * c->vst[0][id]) * c->inv_d[0] +
                             (eyp * c->vst[1][IX(c, wr(c, add(q, 1, 1)))] - eym * c->vst[1][id]) * c->inv_d[1] +
                             (ezp * c->vst[2][IX(c, wr(c, add(q, 2, 1)))] - ezm * c->vst[2][id]) * c->inv_d[2];
                double deps = e[id] - c->eps_old[id];
                c->rhs[id] = -inv_dt * div - deps * inv_dt * inv_dt;
            }
}
/* Poisson face coefficient array of axis d: eps_f/rho_f after
 * smac_update_poisson_beta_two_way, else 1/rho_f */
static inline const double* pois_b(const struct smac_ctx* c, int d) { return c->bpois_on ? c->bpois[d] : c->beta[d]; }
/* G_CFDDEMCoupling::update_poisson_beta_two_way
 * (Coupling/G_CFDDEMCoupling.cu:419-459): on the a-faces 1..N+1 with the
 * other indices in 1..N, bpois = eps_f * (1/rho_f) with eps_f the mean of the
 * two cells' eps^{n+1}; other faces keep 1/rho_f. */
int smac_update_poisson_beta_two_way(smac_ctx* c) {
    if (!c || !c->eps) return SMAC_ERR_INVALID_ARG;
    if (!c->props_ready) smac_update_properties(c, 1);
    for (int a = 0; a < 3; a++) {
        if (!c->bpois[a]) c->bpois[a] = (double*)calloc(c->tot, sizeof(double));
        if (!c->bpois[a]) return SMAC_ERR_OOM;
        for (int k = 0; k < c->p[2]; k++)
            for (int j = 0; j < c->p[1]; j++)
                for (int i = 0; i < c->p[0]; i++) {
                    ivec q = iv(i, j, k);
                    size_t id = IX(c, q);
                    int in = q.v[a] >= 1 && q.v[a] <= c->n[a] + 1;
                    for (int d = 0; d < 3; d++)
                        if (d != a) in = in && q.v[d] >= 1 && q.v[d] <= c->n[d];
                    if (in) {
                        double epsf = 0.5 * (c->eps[IX(c, add(q, a, -1))] + c->eps[id]);
                        c->bpois[a][id] = epsf * c->beta[a][id];
                    } else {
                        c->bpois[a][id] = c->beta[a][id];
                    }
                }
    }
    c->bpois_on = 1;
    c->mg_ready = 0;
    return SMAC_OK;
}

/* k_build_vof_poisson_A (CFD/pressure_solver/G_PCGSolver.cu:126-212).
 * Off-diagonal order: xp, xm, yp, ym, zp, zm. */
static void build_A(struct smac_ctx* c) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= c->n[2]; k++)
        for (int j = 1; j <= c->n[1]; j++)
            for (int i = 1; i <= c->n[0]; i++) {
                ivec q = iv(i, j, k);
                size_t id = IX(c, q);
                double a6[6];
                for (int d = 0; d < 3; d++) {
                    ivec fp = wr(c, add(q, d, 1));
                    a6[2 * d] = pois_b(c, d)[IX(c, fp)] * c->inv_d2[d];
                    a6[2 * d + 1] = pois_b(c, d)[id] * c->inv_d2[d];
                }
                c->Adiag[id] = a6[0] + a6[1] + a6[2] + a6[3] + a6[4] + a6[5];
                for (int d = 0; d < 3; d++) {
                    if (face_type(c, d, wr(c, add(q, d, 1))) == F_BOUNDARY) a6[2 * d] = 0.;
                    if (face_type(c, d, q) == F_BOUNDARY) a6[2 * d + 1] = 0.;
                }
                for (int m = 0; m < 6; m++) c->Aoff[m][id] = a6[m];
            }
}

/* (A x)_c = Adiag*x_c - (Axp*x_xp + Axm*x_xm + Ayp*x_yp + Aym*x_ym + Azp*x_zp + Azm*x_zm)
 * (k_make_pAp, G_PCGSolver.cu:87-124) */
static inline double apply_A(const struct smac_ctx* c, const double* x, int i, int j, int k) {
    size_t id = IDX(c, i, j, k);
    double xp = x[IDX(c, W(c, 0, i + 1), j, k)], xm = x[IDX(c, W(c, 0, i - 1), j, k)];
    double yp = x[IDX(c, i, W(c, 1, j + 1), k)], ym = x[IDX(c, i, W(c, 1, j - 1), k)];
    double zp = x[IDX(c, i, j, W(c, 2, k + 1))], zm = x[IDX(c, i, j, W(c, 2, k - 1))];
    double s = c->Aoff[0][id] * xp + c->Aoff[1][id] * xm + c->Aoff[2][id] * yp +
               c->Aoff[3][id] * ym + c->Aoff[4][id] * zp + c->Aoff[5][id] * zm;
    return c->Adiag[id] * x[id] - s;
}

/* --------------------------------------------------------------- multigrid */
static int mg_auto_levels(const struct smac_ctx* c) {
    int n[3] = {c->n[0], c->n[1], c->n[2]};
    int L = 1;
    while (L < MAXL) {
        long cells = (long)n[0] * n[1] * n[2];
        if (cells <= 512) break;
        if ((n[0] & 1) || (n[1] & 1) || (n[2] & 1)) break;
        if (n[0] < 4 || n[1] < 4 || n[2] < 4) break;
        n[0] /= 2; n[1] /= 2; n[2] /= 2;
        L++;
    }
    return L;
}
static int mg_check_levels(const struct smac_ctx* c, int L) {
    if (L < 1 || L > MAXL) return 0;
    int n[3] = {c->n[0], c->n[1], c->n[2]};
    for (int l = 1; l < L; l++) {
        for (int d = 0; d < 3; d++) {
            if (n[d] & 1) return 0;
            n[d] /= 2;
            if (n[d] < 2) return 0;
        }
    }
    if ((long)n[0] * n[1] * n[2] > 1024) return 0;
    return 1;
}

static inline size_t LIDX(const level* L, int i, int j, int k) {
    return ((size_t)k * (size_t)L->p[1] + (size_t)j) * (size_t)L->p[0] + (size_t)i;
}
static inline int LW(const struct smac_ctx* c, const level* L, int d, int i) {
    return c->per[d] ? wrap(i, L->n[d]) : i;
}
static int lface_boundary(const struct smac_ctx* c, const level* L, int d, int i) {
    return !c->per[d] && (i == 1 || i == L->n[d] + 1);
}

static void mg_free(struct smac_ctx* c) {
    for (int l = 0; l < MAXL; l++) {
        level* L = &c->lev[l];
        free(L->e); free(L->f); free(L->res);
        if (l > 0) { free(L->b[0]); free(L->b[1]); free(L->b[2]); }
        memset(L, 0, sizeof(*L));
    }
    free(c->cinv);
    c->cinv = NULL;
    c->nlev = 0;
    c->mg_ready = 0;
}

/* coefficient of face (I,J,K) of direction d on level l+1: mean of the four
 * fine faces covering it, ((b00 + b10) + (b01 + b11)) * 0.25 */
static void mg_coarsen_beta(const struct smac_ctx* c, const level* F, level* C) {
    for (int d = 0; d < 3; d++) {
        int t1 = (d + 1) % 3, t2 = (d + 2) % 3;
        for (int K = 0; K < C->p[2]; K++)
            for (int J = 0; J < C->p[1]; J++)
                for (int I = 0; I < C->p[0]; I++) {
                    int q[3] = {I, J, K};
                    int ok = 1;
                    for (int e = 0; e < 3; e++) {
                        int hi = C->n[e] + (e == d ? 1 : 0);
                        if (q[e] < 1 || q[e] > hi) ok = 0;
                    }
                    if (!ok) { C->b[d][LIDX(C, I, J, K)] = 0.; continue; }
                    /* layer sums S(m) of the 2x2 fine faces 2q-1+m along d (m = -1, 0, 1;
                     * periodic: wrapped); a face on a wall or outlet side takes its own
                     * layer, an interior face the weights 1/4, 1/2, 1/4 along the normal */
                    const int bnd = !c->per[d] && (q[d] == 1 || q[d] == C->n[d] + 1);
                    double S3[3] = {0., 0., 0.};
                    for (int m = -1; m <= 1; m++) {
                        if (bnd && m != 0) continue;
                        int fq[3];
                        fq[d] = 2 * q[d] - 1 + m;
                        if (m != 0 && c->per[d]) fq[d] = wrap(fq[d], F->n[d]);
                        double v[2][2];
                        for (int s2 = 0; s2 < 2; s2++)
                            for (int s1 = 0; s1 < 2; s1++) {
                                fq[t1] = 2 * q[t1] - 1 + s1;
                                fq[t2] = 2 * q[t2] - 1 + s2;
                                v[s2][s1] = F->b[d][LIDX(F, fq[0], fq[1], fq[2])];
                            }
                        S3[m + 1] = (v[0][0] + v[0][1]) + (v[1][0] + v[1][1]);
                    }
                    C->b[d][LIDX(C, I, J, K)] =
                        bnd ? S3[1] * 0.25 : ((0.25 * S3[0] + 0.5 * S3[1]) + 0.25 * S3[2]) * 0.25;
                }
    }
}

/* level operator: a6 = beta/h^2 per face (xp,xm,yp,ym,zp,zm), diag = sum of all six
 * (before zeroing walls), off-diagonals zero across walls */
static inline void lcoef(const struct smac_ctx* c, const level* L, int i, int j, int k, double a6[6], double* diag) {
    int q[3] = {i, j, k};
    for (int d = 0; d < 3; d++) {
        int qp[3] = {i, j, k};
        qp[d] = q[d] + 1;
        if (c->per[d]) qp[d] = wrap(qp[d], L->n[d]);
        a6[2 * d] = L->b[d][LIDX(L, qp[0], qp[1], qp[2])] * L->inv_d2[d];
        a6[2 * d + 1] = L->b[d][LIDX(L, i, j, k)] * L->inv_d2[d];
    }
    *diag = a6[0] + a6[1] + a6[2] + a6[3] + a6[4] + a6[5];
    for (int d = 0; d < 3; d++) {
        if (lface_boundary(c, L, d, q[d] + 1)) a6[2 * d] = 0.;
        if (lface_boundary(c, L, d, q[d])) a6[2 * d + 1] = 0.;
    }
}
static inline double lnbsum(const struct smac_ctx* c, const level* L, const double* x, int i, int j, int k,
                            const double a6[6]) {
    double xp = x[LIDX(L, LW(c, L, 0, i + 1), j, k)], xm = x[LIDX(L, LW(c, L, 0, i - 1), j, k)];
    double yp = x[LIDX(L, i, LW(c, L, 1, j + 1), k)], ym = x[LIDX(L, i, LW(c, L, 1, j - 1), k)];
    double zp = x[LIDX(L, i, j, LW(c, L, 2, k + 1))], zm = x[LIDX(L, i, j, LW(c, L, 2, k - 1))];
    return a6[0] * xp + a6[1] * xm + a6[2] * yp + a6[3] * ym + a6[4] * zp + a6[5] * zm;
}

static void dense_pinv(const struct smac_ctx* c, const level* L, double* out, int pure) {
    int n = L->n[0] * L->n[1] * L->n[2];
    double* A = (double*)calloc((size_t)n * n, sizeof(double));
    for (int k = 1; k <= L->n[2]; k++)
        for (int j = 1; j <= L->n[1]; j++)
            for (int i = 1; i <= L->n[0]; i++) {
                int row = ((k - 1) * L->n[1] + (j - 1)) * L->n[0] + (i - 1);
                double a6[6], diag;
                lcoef(c, L, i, j, k, a6, &diag);
                A[(size_t)row * n + row] += diag;
                int nb[6][3] = {{i + 1, j, k}, {i - 1, j, k}, {i, j + 1, k}, {i, j - 1, k}, {i, j, k + 1}, {i, j, k - 1}};
                for (int m = 0; m < 6; m++) {
                    if (a6[m] == 0.) continue;
                    int q[3] = {nb[m][0], nb[m][1], nb[m][2]};
                    for (int d = 0; d < 3; d++)
                        if (c->per[d]) q[d] = wrap(q[d], L->n[d]);
                    int col = ((q[2] - 1) * L->n[1] + (q[1] - 1)) * L->n[0] + (q[0] - 1);
                    A[(size_t)row * n + col] -= a6[m];
                }
            }
    double cshift = pure ? A[0] / (double)n : 0.0;
    if (pure)
        for (size_t t = 0; t < (size_t)n * n; t++) A[t] += cshift;
    /* Cholesky A = L L^T (lower, in place) */
    for (int j = 0; j < n; j++) {
        double s = A[(size_t)j * n + j];
        for (int k = 0; k < j; k++) s -= A[(size_t)j * n + k] * A[(size_t)j * n + k];
        double djj = sqrt(s);
        A[(size_t)j * n + j] = djj;
        for (int i = j + 1; i < n; i++) {
            double t = A[(size_t)i * n + j];
            for (int k = 0; k < j; k++) t -= A[(size_t)i * n + k] * A[(size_t)j * n + k];
            A[(size_t)i * n + j] = t / djj;
        }
    }
    double* y = (double*)malloc((size_t)n * sizeof(double));
    for (int col = 0; col < n; col++) {
        for (int i = 0; i < n; i++) {
            double t = (i == col) ? 1.0 : 0.0;
            for (int k = 0; k < i; k++) t -= A[(size_t)i * n + k] * y[k];
            y[i] = t / A[(size_t)i * n + i];
        }
        for (int i = n - 1; i >= 0; i--) {
            double t = y[i];
            for (int k = i + 1; k < n; k++) t -= A[(size_t)k * n + i] * y[k];
            y[i] = t / A[(size_t)i * n + i];
        }
        for (int i = 0; i < n; i++) out[(size_t)i * n + col] = y[i];
    }
    if (pure) {
        double corr = 1.0 / (cshift * (double)n * (double)n);
        for (size_t t = 0; t < (size_t)n * n; t++) out[t] -= corr;
    }
    free(y);
    free(A);
}

static int pure_neumann(const struct smac_ctx* c) {
    for (int d = 0; d < 3; d++) {
        if (c->per[d]) continue;
        for (int s = 0; s < 2; s++) {
            int bt = c->bctype[c->side_bcid[2 * d + s]];
            if ((bt == SMAC_BC_OUTLET && SMAC_ORACLE_MUTANT != 12) || bt == SMAC_BC_INFLOW) return 0;
        }
    }
    return 1;
}

static int mg_setup(struct smac_ctx* c, int L) {
    if (c->mg_ready && c->nlev == L) return SMAC_OK;
    mg_free(c);
    c->nlev = L;
    for (int l = 0; l < L; l++) {
        level* V = &c->lev[l];
        for (int d = 0; d < 3; d++) {
            V->n[d] = c->n[d] >> l;
            V->p[d] = V->n[d] + 3;
            double h = c->d[d] * (double)(1 << l);
            V->inv_d2[d] = 1.0 / (h * h);
        }
        V->tot = (size_t)V->p[0] * V->p[1] * V->p[2];
        V->e = (double*)calloc(V->tot, sizeof(double));
        V->f = (double*)calloc(V->tot, sizeof(double));
        V->res = (double*)calloc(V->tot, sizeof(double));
        if (l == 0) {
            for (int d = 0; d < 3; d++) V->b[d] = (double*)pois_b(c, d);
        } else {
            for (int d = 0; d < 3; d++) V->b[d] = (double*)calloc(V->tot, sizeof(double));
            mg_coarsen_beta(c, &c->lev[l - 1], V);
        }
    }
    level* C = &c->lev[L - 1];
    int n = C->n[0] * C->n[1] * C->n[2];
    c->cinv = (double*)malloc((size_t)n * n * sizeof(double));
    c->cinv_n = n;
    /* outlet faces (Dirichlet) make A regular; without them it is singular */
    dense_pinv(c, C, c->cinv, c->has_outlet ? 0 : 1);
    c->mg_ready = 1;
    return SMAC_OK;
}

/* RBGS half-sweep of colour col ((i+j+k)&1 == col) on level V: e = (f + S)/diag.
 * zero_start: e is known to be zero everywhere (first half-sweep). */
static void rbgs(const struct smac_ctx* c, level* V, int col) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= V->n[2]; k++)
        for (int j = 1; j <= V->n[1]; j++)
            for (int i = 1 + ((col + j + k + 1) & 1); i <= V->n[0]; i += 2) {
                double a6[6], diag;
                lcoef(c, V, i, j, k, a6, &diag);
                double S = lnbsum(c, V, V->e, i, j, k, a6);
                V->e[LIDX(V, i, j, k)] = (V->f[LIDX(V, i, j, k)] + S) / diag;
            }
}
/* red_only: the pre-smoothing ended with a black half sweep, which leaves the
 * black equations satisfied: the residual is 0 on black cells and formed on
 * red cells only */
static void residual(const struct smac_ctx* c, level* V, int red_only) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= V->n[2]; k++)
        for (int j = 1; j <= V->n[1]; j++)
            for (int i = 1; i <= V->n[0]; i++) {
                if (red_only && ((i + j + k) & 1)) {
                    V->res[LIDX(V, i, j, k)] = 0.0;
                    continue;
                }
                double a6[6], diag;
                lcoef(c, V, i, j, k, a6, &diag);
                double S = lnbsum(c, V, V->e, i, j, k, a6);
                size_t id = LIDX(V, i, j, k);
                V->res[id] = V->f[id] - (diag * V->e[id] - S);
            }
}
/* 1-D restriction weights of coarse index I: fine indices and weights, the
 * transpose of cell-centred linear prolongation (walls fold the mirror weight) */
static int rweights(const struct smac_ctx* c, const level* C, int d, int I, int idx[4], double w[4]) {
    int nc = C->n[d], nf = 2 * nc;
    int m = 0;
    int f0 = 2 * I - 2, f1 = 2 * I - 1, f2 = 2 * I, f3 = 2 * I + 1;
    if (c->per[d]) {
        idx[0] = wrap(f0, nf); w[0] = 0.25;
        idx[1] = f1; w[1] = 0.75;
        idx[2] = f2; w[2] = 0.75;
        idx[3] = wrap(f3, nf); w[3] = 0.25;
        return 4;
    }
    if (I > 1) { idx[m] = f0; w[m] = 0.25; m++; }
    idx[m] = f1; w[m] = (I == 1) ? 1.0 : 0.75; m++;
    idx[m] = f2; w[m] = (I == nc) ? 1.0 : 0.75; m++;
    if (I < nc) { idx[m] = f3; w[m] = 0.25; m++; }
    return m;
}
/* coarse f = (1/8) sum_z wz sum_y wy sum_x wx res over the folded 4x4x4 fine
 * neighbourhood (the transpose of trilinear prolongation, scaled by 1/8) */
static void restrict_res(const struct smac_ctx* c, const level* F, level* C) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int K = 1; K <= C->n[2]; K++)
        for (int J = 1; J <= C->n[1]; J++)
            for (int I = 1; I <= C->n[0]; I++) {
                int ix[4], iy[4], iz[4];
                double wx[4], wy[4], wz[4];
                int mx = rweights(c, C, 0, I, ix, wx);
                int my = rweights(c, C, 1, J, iy, wy);
                int mz = rweights(c, C, 2, K, iz, wz);
                /* separable: x innermost, then y, then z, then the 1/8 */
                double tz = 0.0;
                for (int a = 0; a < mz; a++) {
                    double ty = 0.0;
                    for (int b = 0; b < my; b++) {
                        double tx = 0.0;
                        for (int g = 0; g < mx; g++) tx += wx[g] * F->res[LIDX(F, ix[g], iy[b], iz[a])];
                        ty += wy[b] * tx;
                    }
                    tz += wz[a] * ty;
                }
                C->f[LIDX(C, I, J, K)] = 0.125 * tz;
            }
}
/* fine index i: parent (i+1)/2 weight 0.75, neighbour parent-1 (odd i) or
 * parent+1 (even i) weight 0.25; walls mirror the neighbour onto the parent */
static inline void pweights(const struct smac_ctx* c, const level* C, int d, int i, int idx[2], double w[2]) {
    int par = (i + 1) >> 1;
    int nb = (i & 1) ? par - 1 : par + 1;
    if (c->per[d]) nb = wrap(nb, C->n[d]);
    else if (nb < 1 || nb > C->n[d]) nb = par;
    idx[0] = par; w[0] = 0.75;
    idx[1] = nb; w[1] = 0.25;
}
static void prolong_add(const struct smac_ctx* c, const level* C, level* F) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= F->n[2]; k++)
        for (int j = 1; j <= F->n[1]; j++)
            for (int i = 1; i <= F->n[0]; i++) {
                int ix[2], iy[2], iz[2];
                double wx[2], wy[2], wz[2];
                pweights(c, C, 0, i, ix, wx);
                pweights(c, C, 1, j, iy, wy);
                pweights(c, C, 2, k, iz, wz);
                double acc = 0.0;
                for (int a = 0; a < 2; a++)
                    for (int b = 0; b < 2; b++)
                        for (int g = 0; g < 2; g++) {
                            double w = wz[a] * wy[b] * wx[g];
                            acc += w * C->e[LIDX(C, ix[g], iy[b], iz[a])];
                        }
                F->e[LIDX(F, i, j, k)] += acc;
            }
}
static void coarse_solve(const struct smac_ctx* c, level* V) {
    int n = c->cinv_n;
    double* rv = (double*)malloc((size_t)n * sizeof(double));
    for (int k = 1; k <= V->n[2]; k++)
        for (int j = 1; j <= V->n[1]; j++)
            for (int i = 1; i <= V->n[0]; i++)
                rv[((k - 1) * V->n[1] + (j - 1)) * V->n[0] + (i - 1)] = V->f[LIDX(V, i, j, k)];
    /* row sum in the canonical order: lane l of 64 accumulates columns
     * l, l+64, ... in order, then the halving tree */
#pragma omp parallel for schedule(static)
    for (int row = 0; row < n; row++) {
        double lanes[64];
        const double* a = c->cinv + (size_t)row * n;
        for (int l = 0; l < 64; l++) {
            double acc = 0.0;
            for (int col = l; col < n; col += 64) acc += a[col] * rv[col];
            lanes[l] = acc;
        }
        int i = row % V->n[0] + 1, j = (row / V->n[0]) % V->n[1] + 1, k = row / (V->n[0] * V->n[1]) + 1;
        V->e[LIDX(V, i, j, k)] = tree(lanes, 64);
    }
    free(rv);
}
static void zero_level(level* V) { memset(V->e, 0, V->tot * sizeof(double)); }

static void vcycle(const struct smac_ctx* c, int l, int nu1, int nu2) {
    level* V = (level*)&c->lev[l];
    if (l == c->nlev - 1) {
        coarse_solve(c, V);
        return;
    }
    zero_level(V);
    for (int s = 0; s < nu1; s++) { rbgs(c, V, 0); rbgs(c, V, 1); }
    residual(c, V, nu1 > 0);
    level* C = (level*)&c->lev[l + 1];
    restrict_res(c, V, C);
    vcycle(c, l + 1, nu1, nu2);
    prolong_add(c, C, V);
    for (int s = 0; s < nu2; s++) { rbgs(c, V, 1); rbgs(c, V, 0); }
}

/* ------------------------------------------------ the reference's GMG
 * G_GMGSolver (CFD/pressure_solver/G_GMGSolver.cu, G_GMGSolver.h): weighted
 * Jacobi (omega = 0.6) V-cycle, mean restriction over the 2x2x2 children,
 * constant prolongation, coarse face coefficients = mean of the fine face
 * coefficients of the 2x2x2 block (two faces along the face's own axis),
 * coarse ghosts held at zero. Level 0 is already coarse (N/2). Directions
 * whose size halved would drop below min_size_ = 6 stop coarsening. Periodic
 * directions (absent in the reference) wrap face and cell indices. */
#define RG_MIN_SIZE 6        /* G_GMGSolver.h:20 */
#define RG_MAX_LEVELS 16     /* G_GMGSolver.h:2  */
#define RG_EPS 1.E-12        /* CFD/hardCodedParameters.h:7 */

typedef struct rglevel {
    int n[3], p[3], coarse[3];
    size_t tot;
    double inv_d2[3];
    double *q, *qt, *rhs, *res, *b[3];
} rglevel;

static inline size_t RIDX(const rglevel* L, int i, int j, int k) {
    return ((size_t)k * (size_t)L->p[1] + (size_t)j) * (size_t)L->p[0] + (size_t)i;
}

static void rg_free(struct smac_ctx* c) {
    if (c->rg) {
        for (int l = 0; l < c->nrg; l++) {
            rglevel* L = &c->rg[l];
            free(L->q); free(L->qt); free(L->rhs); free(L->res);
            free(L->b[0]); free(L->b[1]); free(L->b[2]);
        }
        free(c->rg);
    }
    c->rg = NULL;
    c->nrg = 0;
    c->rg_req = 0;
}

/* G_GMGSolver::initialize (G_GMGSolver.cu:10-147). Returns the level count,
 * or a negative status: odd size in a coarsened direction (the reference
 * aborts), or a single level (the reference then prolongs through
 * levels_[-1], G_GMGSolver.cu:971). */
static int rg_setup(struct smac_ctx* c, int req) {
    if (req < 1 || req > RG_MAX_LEVELS) return SMAC_ERR_INVALID_ARG;
    if (c->rg && c->rg_req == req) return c->nrg;
    rg_free(c);
    rglevel tmp[RG_MAX_LEVELS];
    memset(tmp, 0, sizeof(tmp));
    int prev[3] = {c->n[0], c->n[1], c->n[2]};
    double pinv[3] = {c->inv_d2[0], c->inv_d2[1], c->inv_d2[2]};
    int L = 0;
    for (int l = 0; l < req; l++) {
        rglevel* V = &tmp[l];
        int any = 0;
        for (int d = 0; d < 3; d++) {
            V->coarse[d] = !(prev[d] / 2 < RG_MIN_SIZE);
            any |= V->coarse[d];
        }
        if (l > 0 && !any) break;
        for (int d = 0; d < 3; d++)
            if (V->coarse[d] && (prev[d] & 1)) return SMAC_ERR_INVALID_GRID;
        for (int d = 0; d < 3; d++) {
            V->n[d] = V->coarse[d] ? prev[d] / 2 : prev[d];
            V->inv_d2[d] = V->coarse[d] ? pinv[d] * 0.25 : pinv[d];
            V->p[d] = V->n[d] + 3;
            prev[d] = V->n[d];
            pinv[d] = V->inv_d2[d];
        }
        L++;
    }
    if (L < 2) return SMAC_ERR_INVALID_GRID;
    c->rg = (rglevel*)calloc((size_t)L, sizeof(rglevel));
    for (int l = 0; l < L; l++) {
        rglevel* V = &c->rg[l];
        *V = tmp[l];
        V->tot = (size_t)V->p[0] * V->p[1] * V->p[2];
        V->q = (double*)calloc(V->tot, sizeof(double));
        V->qt = (double*)calloc(V->tot, sizeof(double));
        V->rhs = (double*)calloc(V->tot, sizeof(double));
        V->res = (double*)calloc(V->tot, sizeof(double));
        for (int d = 0; d < 3; d++) V->b[d] = (double*)calloc(V->tot, sizeof(double));
    }
    c->nrg = L;
    c->rg_req = req;
    return L;
}

/* face coefficient of axis d at face index (i,j,k) of the level below
 * (lv < 0: the grid's f_b*). Indices past the last face of a wall direction
 * are ghost faces and read 0 (F_GHOST, SMACSolver.cpp:358-360); periodic
 * directions wrap. */
static double rg_fine_beta(const struct smac_ctx* c, int lv, int d, int q[3]) {
    int n[3];
    for (int e = 0; e < 3; e++) n[e] = lv < 0 ? c->n[e] : c->rg[lv].n[e];
    int w[3] = {q[0], q[1], q[2]};
    if (c->per[d]) w[d] = wrap(w[d], n[d]);
    else if (w[d] > n[d] + 1) return 0.0;
    if (lv < 0) return pois_b(c, d)[IDX(c, w[0], w[1], w[2])];
    return c->rg[lv].b[d][RIDX(&c->rg[lv], w[0], w[1], w[2])];
}

/* k_create_level0_coeffs / k_create_level_coeffs (G_GMGSolver.cu:499-644) */
static void rg_coeffs(struct smac_ctx* c) {
    for (int l = 0; l < c->nrg; l++) {
        rglevel* V = &c->rg[l];
        for (int d = 0; d < 3; d++) {
            int hi[3] = {V->n[0], V->n[1], V->n[2]};
            hi[d] += 1;
            for (int k = 1; k <= hi[2]; k++)
                for (int j = 1; j <= hi[1]; j++)
                    for (int i = 1; i <= hi[0]; i++) {
                        int q[3] = {i, j, k}, B[3], R[3];
                        for (int e = 0; e < 3; e++) {
                            B[e] = V->coarse[e] ? 2 * q[e] - 1 : q[e];
                            R[e] = V->coarse[e] ? 2 : 1;
                        }
                        double count = 0., tmp = 0.;
                        for (int iz = 0; iz < R[2]; iz++)
                            for (int iy = 0; iy < R[1]; iy++)
                                for (int ix = 0; ix < R[0]; ix++) {
                                    int f[3] = {B[0] + ix, B[1] + iy, B[2] + iz};
                                    tmp += rg_fine_beta(c, l - 1, d, f);
                                    count += 1.0;
                                }
                        V->b[d][RIDX(V, i, j, k)] = tmp / count;
                    }
        }
    }
}

/* Level coefficients of k_create_level_A (G_GMGSolver.cu:646-685). */
static inline void rg_coef(const rglevel* V, int i, int j, int k, double a6[6], double* diag, double* inv) {
    a6[0] = V->b[0][RIDX(V, i + 1, j, k)] * V->inv_d2[0];
    a6[1] = V->b[0][RIDX(V, i, j, k)] * V->inv_d2[0];
    a6[2] = V->b[1][RIDX(V, i, j + 1, k)] * V->inv_d2[1];
    a6[3] = V->b[1][RIDX(V, i, j, k)] * V->inv_d2[1];
    a6[4] = V->b[2][RIDX(V, i, j, k + 1)] * V->inv_d2[2];
    a6[5] = V->b[2][RIDX(V, i, j, k)] * V->inv_d2[2];
    *diag = (a6[0] + a6[1] + a6[2] + a6[3] + a6[4] + a6[5]);
    *inv = fabs(*diag) < RG_EPS ? 0. : 1. / *diag;
}
/* neighbour of cell (i,j,k) along d by s on a level: ghosts are zero on
 * walls, periodic directions wrap */
static inline double rg_nb(const struct smac_ctx* c, const rglevel* V, const double* x, int i, int j, int k, int d,
                           int s) {
    int q[3] = {i, j, k};
    q[d] += s;
    if (c->per[d]) q[d] = wrap(q[d], V->n[d]);
    return x[RIDX(V, q[0], q[1], q[2])];
}
static inline double rg_nbsum(const struct smac_ctx* c, const rglevel* V, const double* x, int i, int j, int k,
                              const double a6[6]) {
    return a6[0] * rg_nb(c, V, x, i, j, k, 0, 1) + a6[1] * rg_nb(c, V, x, i, j, k, 0, -1) +
           a6[2] * rg_nb(c, V, x, i, j, k, 1, 1) + a6[3] * rg_nb(c, V, x, i, j, k, 1, -1) +
           a6[4] * rg_nb(c, V, x, i, j, k, 2, 1) + a6[5] * rg_nb(c, V, x, i, j, k, 2, -1);
}

/* k_jacobi_level_iteration (G_GMGSolver.cu:692-717) + the pointer swap */
static void rg_jacobi(const struct smac_ctx* c, rglevel* V) {
    const double omega = 0.6;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= V->n[2]; k++)
        for (int j = 1; j <= V->n[1]; j++)
            for (int i = 1; i <= V->n[0]; i++) {
                double a6[6], diag, inv;
                rg_coef(V, i, j, k, a6, &diag, &inv);
                double tmp_p = rg_nbsum(c, V, V->q, i, j, k, a6);
                size_t id = RIDX(V, i, j, k);
                V->qt[id] = (1.0 - omega) * V->q[id] + omega * (tmp_p + V->rhs[id]) * inv;
            }
    double* t = V->q; V->q = V->qt; V->qt = t;
}
/* k_get_res_levels (G_GMGSolver.cu:152-188) */
static void rg_residual(const struct smac_ctx* c, rglevel* V) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= V->n[2]; k++)
        for (int j = 1; j <= V->n[1]; j++)
            for (int i = 1; i <= V->n[0]; i++) {
                double a6[6], diag, inv;
                rg_coef(V, i, j, k, a6, &diag, &inv);
                size_t id = RIDX(V, i, j, k);
                double Ap = V->q[id] * diag - rg_nbsum(c, V, V->q, i, j, k, a6);
                V->res[id] = V->rhs[id] - Ap;
            }
}
/* k_restrict_grid_to_level0 / k_restrict_level_to_level (G_GMGSolver.cu:416-493):
 * C->rhs = mean of the children of `res` (fine box of pitch fp) */
static void rg_restrict(const rglevel* C, const double* res, const int fp[3], int fine_is_grid,
                        const struct smac_ctx* c) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= C->n[2]; k++)
        for (int j = 1; j <= C->n[1]; j++)
            for (int i = 1; i <= C->n[0]; i++) {
                int q[3] = {i, j, k}, B[3], R[3];
                for (int e = 0; e < 3; e++) {
                    B[e] = C->coarse[e] ? 2 * q[e] - 1 : q[e];
                    R[e] = C->coarse[e] ? 2 : 1;
                }
                double count = 0., tmp = 0.;
                for (int iz = 0; iz < R[2]; iz++)
                    for (int iy = 0; iy < R[1]; iy++)
                        for (int ix = 0; ix < R[0]; ix++) {
                            size_t id = fine_is_grid ? IDX(c, B[0] + ix, B[1] + iy, B[2] + iz)
                                                     : ((size_t)(B[2] + iz) * fp[1] + (B[1] + iy)) * fp[0] + (B[0] + ix);
                            tmp += res[id];
                            count += 1.0;
                        }
                C->rhs[RIDX(C, i, j, k)] = tmp / count;
            }
}
/* k_propagate_level_to_level / k_propagate_level0_to_array (G_GMGSolver.cu:285-389) */
static void rg_prolong(const rglevel* C, double* fq, const int fp[3], int fine_is_grid, const struct smac_ctx* c) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= C->n[2]; k++)
        for (int j = 1; j <= C->n[1]; j++)
            for (int i = 1; i <= C->n[0]; i++) {
                int q[3] = {i, j, k}, B[3], R[3];
                for (int e = 0; e < 3; e++) {
                    B[e] = C->coarse[e] ? 2 * q[e] - 1 : q[e];
                    R[e] = C->coarse[e] ? 2 : 1;
                }
                double prop = C->q[RIDX(C, i, j, k)];
                for (int iz = 0; iz < R[2]; iz++)
                    for (int iy = 0; iy < R[1]; iy++)
                        for (int ix = 0; ix < R[0]; ix++) {
                            size_t id = fine_is_grid ? IDX(c, B[0] + ix, B[1] + iy, B[2] + iz)
                                                     : ((size_t)(B[2] + iz) * fp[1] + (B[1] + iy)) * fp[0] + (B[0] + ix);
                            fq[id] += prop;
                        }
            }
}
/* k_jacobi_iteration_general (G_GMGSolver.cu:719-746) on the grid's operator:
 * out = (1-omega) x + omega (sum a_nb x_nb + r) / Adiag */
static void rg_fine_jacobi(const struct smac_ctx* c, const double* x, double* out, const double* r) {
    const double omega = 0.6;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= c->n[2]; k++)
        for (int j = 1; j <= c->n[1]; j++)
            for (int i = 1; i <= c->n[0]; i++) {
                size_t id = IDX(c, i, j, k);
                double xp = x[IDX(c, W(c, 0, i + 1), j, k)], xm = x[IDX(c, W(c, 0, i - 1), j, k)];
                double yp = x[IDX(c, i, W(c, 1, j + 1), k)], ym = x[IDX(c, i, W(c, 1, j - 1), k)];
                double zp = x[IDX(c, i, j, W(c, 2, k + 1))], zm = x[IDX(c, i, j, W(c, 2, k - 1))];
                double tmp_p = c->Aoff[0][id] * xp + c->Aoff[1][id] * xm + c->Aoff[2][id] * yp +
                               c->Aoff[3][id] * ym + c->Aoff[4][id] * zp + c->Aoff[5][id] * zm;
                out[id] = (1.0 - omega) * x[id] + omega * (tmp_p + r[id]) * (1. / c->Adiag[id]);
            }
}

/* v_cycle_as_preconditioner (G_GMGSolver.cu:925-998): z = M^{-1} r */
static void rg_vcycle(struct smac_ctx* c, const smac_solver_opts* o) {
    int nf = o->ref_sweeps, nc = o->ref_coarse_sweeps, L = c->nrg;
    double* buf = (double*)calloc(c->tot, sizeof(double));
    double* t = buf;
    double* x = c->z;
    memset(x, 0, c->tot * sizeof(double));
    for (int l = 0; l < L; l++) memset(c->rg[l].q, 0, c->rg[l].tot * sizeof(double));
    for (int s = 0; s < nf; s++) { rg_fine_jacobi(c, x, t, c->r); double* u = x; x = t; t = u; }
    /* k_get_res_general: residual of the grid's operator, into t */
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= c->n[2]; k++)
        for (int j = 1; j <= c->n[1]; j++)
            for (int i = 1; i <= c->n[0]; i++) t[IDX(c, i, j, k)] = c->r[IDX(c, i, j, k)] - apply_A(c, x, i, j, k);
    rg_restrict(&c->rg[0], t, c->p, 1, c);
    for (int l = 0; l < L - 1; l++) {
        rglevel* V = &c->rg[l];
        for (int s = 0; s < nf; s++) rg_jacobi(c, V);
        rg_residual(c, V);
        rg_restrict(&c->rg[l + 1], V->res, V->p, 0, c);
    }
    for (int s = 0; s < nc; s++) rg_jacobi(c, &c->rg[L - 1]);
    rg_prolong(&c->rg[L - 1], c->rg[L - 2].q, c->rg[L - 2].p, 0, c);
    for (int l = L - 2; l >= 0; l--) {
        rglevel* V = &c->rg[l];
        for (int s = 0; s < nf; s++) rg_jacobi(c, V);
        if (l != 0) rg_prolong(V, c->rg[l - 1].q, c->rg[l - 1].p, 0, c);
        else rg_prolong(V, x, c->p, 1, c);
    }
    for (int s = 0; s < nf; s++) { rg_fine_jacobi(c, x, t, c->r); double* u = x; x = t; t = u; }
    if (x != c->z) memcpy(c->z, x, c->tot * sizeof(double));
    free(buf);
}

/* z = M^{-1} r */
static void precond(struct smac_ctx* c, const smac_solver_opts* o) {
    int nx = c->n[0], ny = c->n[1], nz = c->n[2];
    if (o->precond == SMAC_PRECOND_JACOBI) {
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = 1; k <= nz; k++)
            for (int j = 1; j <= ny; j++)
                for (int i = 1; i <= nx; i++) {
                    size_t id = IDX(c, i, j, k);
                    c->z[id] = (1. / c->Adiag[id]) * c->r[id];
                }
        return;
    }
    if (o->precond == SMAC_PRECOND_REF_GMG) {
        rg_vcycle(c, o);
        return;
    }
    level* V = &c->lev[0];
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) V->f[LIDX(V, i, j, k)] = c->r[IDX(c, i, j, k)];
    vcycle(c, 0, o->nu1, o->nu2);
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) c->z[IDX(c, i, j, k)] = V->e[LIDX(V, i, j, k)];
}

/* --------------------------------------------------------------- the PCG */
static void add_scalar_interior(struct smac_ctx* c, double* q, double a, double sum) {
    for (int k = 1; k <= c->n[2]; k++)
        for (int j = 1; j <= c->n[1]; j++)
            for (int i = 1; i <= c->n[0]; i++) {
                size_t id = IDX(c, i, j, k);
                q[id] += a * sum;
            }
}

void smac_default_opts(smac_solver_opts* o) {
    if (!o) return;
    o->precond = SMAC_PRECOND_MG;
    o->precision = SMAC_PREC_FP64;
    o->stop_rule = SMAC_STOP_REL_L2;
    o->tol = 1e-8;
    o->check_every = 1;
    o->max_iter = 2000;
    o->mg_levels = 0;
    o->nu1 = 2;
    o->nu2 = 2;
    o->flexible = 0;
    o->ref_sweeps = 4;
    o->ref_coarse_sweeps = 30;
    o->stationary = 0;
}

/* Algorithmic HBM bytes per interior cell of one solve, the model the HIP
 * library reports in smac_solve_stats.bytes_model (each array read or written
 * once per pass; W = smoother width). */
static int var_props(const smac_ctx* c) {
    if (c->fl.rho0 != c->fl.rho1 || c->fl.mu0 != c->fl.mu1 || c->has_outlet) return 1;
    for (size_t t = 0; t < c->tot; t++)
        if (c->alpha[t] != 0.0) return 1;
    return 0;
}
static double solve_bytes_model(const smac_ctx* c, const smac_solver_opts* o, int L, int iters) {
    int var = c->bpois_on || var_props(c);
    double W = o->precision == SMAC_PREC_MIXED_FP32 ? 4.0 : o->precision == SMAC_PREC_MIXED_BF16 ? 2.0 : 8.0;
    double P;
    if (o->precond == SMAC_PRECOND_MG) {
        P = 17.43 * W + (var ? (15.0 * 8.0 / 7.0) * W : 0.0);
    } else if (o->precond == SMAC_PRECOND_JACOBI) {
        P = 16.0 + (var ? 24.0 : 0.0);
    } else {
        double words = 0.0, f = 1.0;
        for (int l = 0; l < L - 1; l++, f /= 8.0) words += f * ((8.0 * 3.0 + 4.25) + (var ? 8.0 * 3.0 + 3.0 : 0.0));
        words += f * 30.0 * (var ? 6.0 : 3.0);
        P = words * 8.0;
    }
    double rhs = 32.0 + ((c->eps && c->eps_old) ? 16.0 : 0.0);
    double it = 104.0 + P + (o->flexible ? 16.0 : 0.0) + (var ? 24.0 : 0.0);
    return rhs + (32.0 + P) + iters * it;
}

int smac_pressure_solve(smac_ctx* c, double dt, const smac_solver_opts* opts, smac_solve_stats* st) {
    if (!c || !(dt > 0)) return SMAC_ERR_INVALID_ARG;
    smac_solver_opts o;
    if (opts) o = *opts; else smac_default_opts(&o);
    if (o.check_every < 1) o.check_every = 1;
    if (o.precision != SMAC_PREC_FP64) return SMAC_ERR_UNSUPPORTED; /* the oracle is FP64 */
    if (!c->props_ready) smac_update_properties(c, 1);
    int nx = c->n[0], ny = c->n[1], nz = c->n[2];
    double Ncell = (double)nx * (double)ny * (double)nz;
    int pure = pure_neumann(c);
    double inv_dt = 1. / dt;

    if (c->eps && c->eps_old) make_rhs_eps(c, inv_dt);
    else make_rhs(c, inv_dt);
    build_A(c);
    int L = 1;
    if (o.precond == SMAC_PRECOND_MG) {
        L = o.mg_levels > 0 ? o.mg_levels : mg_auto_levels(c);
        if (!mg_check_levels(c, L)) return SMAC_ERR_INVALID_GRID;
        mg_setup(c, L);
    } else if (o.precond == SMAC_PRECOND_REF_GMG) {
        if (o.ref_sweeps < 0 || o.ref_coarse_sweeps < 0) return SMAC_ERR_INVALID_ARG;
        L = rg_setup(c, o.mg_levels > 0 ? o.mg_levels : 4);
        if (L < 0) return L;
        rg_coeffs(c); /* create_coeffs on every solve (G_PCGSolver.cu:805) */
    } else if (o.precond != SMAC_PRECOND_JACOBI) {
        return SMAC_ERR_INVALID_ARG;
    }
    /* dp = 0 on entry (correct_vof_velocity clears it), so A dp = 0 and
     * r = b - mean(b) (pure Neumann) or r = b. */
    double mean_b = 0.0;
    if (pure) mean_b = csum(c, RED_SUM, c->rhs, NULL) / Ncell;
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                size_t id = IDX(c, i, j, k);
                c->dp[id] = 0.0;
                c->r[id] = c->rhs[id] - mean_b;
            }
    double bb = csum(c, RED_SQ, c->r, NULL);
    double bnorm = sqrt(bb);
    double size_inv = 1. / Ncell;
    if (pure) add_scalar_interior(c, c->r, -1. * size_inv, csum(c, RED_SUM, c->r, NULL));
    precond(c, &o);
    if (pure) add_scalar_interior(c, c->z, -1. * size_inv, csum(c, RED_SUM, c->z, NULL));
    memcpy(c->dvec, c->z, c->tot * sizeof(double));
    /* pair order, the order of every later r.z */
    double rz = csum(c, RED_DOT_S2, c->r, c->z);

    int iter = 0, conv = 0;
    double rr = bb, linf = 0.0;
    if (bb == 0.0) conv = 1;
    /* a NaN/Inf right-hand side stops the solve at once (the reference has no
     * check: check_nan_all is declared, never defined, CFD/G_SMACSolver.h:88) */
    int nan = !isfinite(bb);
    /* |r|^2 after an update: pair order from the update itself, natural order
     * where the mean of r is removed after it */
    const int rr_mode = pure && o.precond == SMAC_PRECOND_REF_GMG ? RED_SQ : RED_SQ_S2;
    for (iter = 0; iter < o.max_iter && !conv && !nan; ++iter) {
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = 1; k <= nz; k++)
            for (int j = 1; j <= ny; j++)
                for (int i = 1; i <= nx; i++) c->Ad[IDX(c, i, j, k)] = apply_A(c, c->dvec, i, j, k);
        double pAp = csum(c, RED_DOT_S2, c->dvec, c->Ad);
        /* stationary: x += z, r -= A z (a fixed linear map of b for tol = 0) */
        double alpha = o.stationary ? 1.0 : rz / pAp;
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = 1; k <= nz; k++)
            for (int j = 1; j <= ny; j++)
                for (int i = 1; i <= nx; i++) {
                    size_t id = IDX(c, i, j, k);
                    c->dp[id] += alpha * c->dvec[id];
                    c->r[id] -= alpha * c->Ad[id];
                }
        /* solve_pureneumann_gmgpcg removes the mean of r after every update
         * (G_PCGSolver.cu:866) */
        if (pure && o.precond == SMAC_PRECOND_REF_GMG)
            add_scalar_interior(c, c->r, -1. * size_inv, csum(c, RED_SUM, c->r, NULL));
        precond(c, &o);
        double rz_new = csum(c, RED_DOT_S2, c->r, c->z);
        double zAd = 0.0;
        if (o.flexible) zAd = csum(c, RED_DOT_S2, c->z, c->Ad);
        if (iter % o.check_every == 0) {
            if (!isfinite(csum(c, rr_mode, c->r, NULL))) { iter++; nan = 1; break; }
            if (o.stop_rule == SMAC_STOP_REF_LINF_DT) {
                linf = cmaxabs(c, c->r) * dt;
                if (linf < o.tol) { iter++; conv = 1; break; }
            } else {
                rr = csum(c, rr_mode, c->r, NULL);
                if (sqrt(rr) < o.tol * bnorm) { iter++; conv = 1; break; }
            }
        }
        double beta = o.stationary ? 0.0 : (o.flexible ? (-alpha * zAd) / rz : rz_new / rz);
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = 1; k <= nz; k++)
            for (int j = 1; j <= ny; j++)
                for (int i = 1; i <= nx; i++) {
                    size_t id = IDX(c, i, j, k);
                    c->dvec[id] = c->z[id] + beta * c->dvec[id];
                }
        rz = rz_new;
    }
    if (st) {
        memset(st, 0, sizeof(*st));
        st->iters = iter;
        st->converged = conv;
        st->pure_neumann = pure;
        st->mg_levels = o.precond != SMAC_PRECOND_JACOBI ? L : 0;
        if (iter > 0) rr = csum(c, rr_mode, c->r, NULL);
        st->bnorm = bnorm;
        st->rel_res = bnorm > 0 ? sqrt(rr) / bnorm : 0.0;
        st->linf_res_dt = cmaxabs(c, c->r) * dt;
        st->nan_detected = nan || !(rr == rr);
        st->bytes_model = solve_bytes_model(c, &o, L, iter) * Ncell;
    }
    if (nan || !(bb == bb) || !(rr == rr)) return SMAC_ERR_NAN;
    return SMAC_OK;
}

/* k_correct_vof_velocity + memset dp (CFD/G_SMACSolver.cu:1072-1142,1269-1275) */
int smac_correct(smac_ctx* c, double dt) {
    if (!c) return SMAC_ERR_INVALID_ARG;
    for (int a = 0; a < 3; a++) {
        int lo[3], hi[3];
        for (int d = 0; d < 3; d++) { lo[d] = 1; hi[d] = c->n[d]; }
        lo[a] = face_lo(c, a);
        hi[a] = face_hi(c, a);
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = lo[2]; k <= hi[2]; k++)
            for (int j = lo[1]; j <= hi[1]; j++)
                for (int i = lo[0]; i <= hi[0]; i++) {
                    ivec q = iv(i, j, k);
                    size_t id = IX(c, q);
                    ivec cm = wr(c, add(q, a, -1));
                    double g = (c->dp[id] - c->dp[IX(c, cm)]) * c->inv_d[a];
                    c->vel[a][id] = c->vst[a][id] + c->inv_rhof[a][id] * dt * (0.0 - g);
                }
    }
    for (int k = 1; k <= c->n[2]; k++)
        for (int j = 1; j <= c->n[1]; j++)
            for (int i = 1; i <= c->n[0]; i++) {
                size_t id = IDX(c, i, j, k);
                c->p_[id] += c->dp[id];
            }
    memset(c->dp, 0, c->tot * sizeof(double));
    return SMAC_OK;
}

int smac_step(smac_ctx* c, double dt, const smac_solver_opts* o, smac_solve_stats* st) {
    int rc;
    if (!c) return SMAC_ERR_INVALID_ARG;
    if (!c->props_ready) smac_update_properties(c, 1);
    else if (c->fl.rho0 != c->fl.rho1 || c->fl.mu0 != c->fl.mu1) smac_update_properties(c, 0);
    if ((rc = smac_predictor(c, dt))) return rc;
    if ((rc = smac_pressure_solve(c, dt, o, st))) return rc;
    /* predictor 56 and correction 72 bytes per cell (+24 each with property arrays) */
    if (st) st->bytes_model += (128.0 + (var_props(c) ? 48.0 : 0.0)) * (double)c->n[0] * c->n[1] * c->n[2];
    return smac_correct(c, dt);
}

/* The oracle restates the reference, which has no adjoint: the adjoint is
 * pinned by central finite differences of the forward step instead. */
int smac_adjoint_step(smac_ctx* c, double dt, const smac_solver_opts* o, const double* const u_n[3],
                      double* const ubar[3], double* pbar, smac_solve_stats* st) {
    (void)c; (void)dt; (void)o; (void)u_n; (void)ubar; (void)pbar; (void)st;
    return SMAC_ERR_UNSUPPORTED;
}

/* Likewise the adjoint of the VoF loop's step: checked against central
 * differences of the forward transport + step. */
int smac_adjoint_vof_step(smac_ctx* c, double dt, int32_t S, int32_t step, const smac_solver_opts* o,
                          const double* alpha_n, const double* const u_n[3], const double* p_n, double* alpha_bar,
                          double* const ubar[3], double* pbar, smac_solve_stats* st) {
    (void)c; (void)dt; (void)S; (void)step; (void)o; (void)alpha_n; (void)u_n; (void)p_n; (void)alpha_bar;
    (void)ubar; (void)pbar; (void)st;
    return SMAC_ERR_UNSUPPORTED;
}

/* k_calc_cfl + calc_cfl (CFD/G_SMACSolver.cu:1365-1430) */
int smac_calc_cfl(smac_ctx* c, double dt, double* cc, double* cv) {
    if (!c) return SMAC_ERR_INVALID_ARG;
    if (!c->props_ready) smac_update_properties(c, 1);
    double mc = 0.0, mv = 0.0;
    for (int k = 1; k <= c->n[2]; k++)
        for (int j = 1; j <= c->n[1]; j++)
            for (int i = 1; i <= c->n[0]; i++) {
                ivec q = iv(i, j, k);
                double m[3];
                for (int a = 0; a < 3; a++) {
                    double vp = fabs(c->vel[a][IX(c, wr(c, add(q, a, 1)))]);
                    double vm = fabs(c->vel[a][IX(c, q)]);
                    m[a] = (SMAC_ORACLE_MUTANT == 5) ? (vp < vm ? vp : vm) : (vp > vm ? vp : vm);
                }
                double cf = dt * (m[0] * c->inv_d[0] + m[1] * c->inv_d[1] + m[2] * c->inv_d[2]);
                size_t id = IX(c, q);
                double cvv = (SMAC_ORACLE_MUTANT == 6 ? 1.0 : c->inv_rho[id]) * c->mu[id] * dt *
                             (c->inv_d2[0] + c->inv_d2[1] + c->inv_d2[2]);
                if (cf > mc) mc = cf;
                if (cvv > mv) mv = cvv;
            }
    if (cc) *cc = mc;
    if (cv) *cv = mv;
    return SMAC_OK;
}

/* The no-allocation variant of the C-ABI (no counterpart in the reference):
 * on the CPU it is smac_calc_cfl behind the same argument checks. */
int smac_calc_cfl_noalloc(smac_ctx* c, double dt, double* cc, double* cv) {
    if (!c) return SMAC_ERR_INVALID_ARG;
    if ((long)c->n[1] * (long)c->n[2] > SMAC_NOALLOC_MAX_ROWS) return SMAC_ERR_INVALID_GRID;
    if (!c->props_ready && (c->fl.rho0 != c->fl.rho1 || c->fl.mu0 != c->fl.mu1)) return SMAC_ERR_UNSUPPORTED;
    return smac_calc_cfl(c, dt, cc, cv);
}

/* ------------------------------------------------ VoF transport */
#define THINC_BETA 2.0             /* BETA, CFD/hardCodedParameters.h:6 */
#define VOF_EPS 1.E-12              /* EPS, CFD/hardCodedParameters.h:7 */
#define VOF_PI 3.14159265358979323846

/* ------------------------------------------------ VoF transcendentals
 * exp, log, cosh and acos of the THINC/WLIC flux as a fixed sequence of IEEE
 * double operations, so that this file and the device kernels
 * (mydemcfd_amd/csrc/smac_vof.hip, the same operations in the same order)
 * give the same bits: exp by Cody-Waite reduction (k = round(x/ln2),
 * ln2 split into a 33-bit head and a tail) and the Taylor polynomial of degree
 * 13 on |r| <= ln2/2; log by frexp to m in [sqrt(1/2), sqrt(2)) and the atanh
 * series 2 atanh((m-1)/(m+1)) to degree 21; acos on [0, 1] by the asin series
 * to degree 53 on [0, 1/2] and acos x = 2 asin(sqrt((1-x)/2)) above 1/2.
 * Coefficients are the nearest doubles of the exact rationals. Accuracy a few
 * ulp (checked against libm in tests/test_vof.py). */
static double vm_exp(double x) {
    if (x > 709.0) return HUGE_VAL;
    if (x < -745.0) return 0.0;
    double kd = floor(x * 0x1.71547652b82fep+0 + 0.5);
    double r = (x - kd * 0x1.62e42fee00000p-1) - kd * 0x1.a39ef35793c76p-33;
    double p = 0x1.6124613a86d09p-33;
    p = p * r + 0x1.1eed8eff8d898p-29;
    p = p * r + 0x1.ae64567f544e4p-26;
    p = p * r + 0x1.27e4fb7789f5cp-22;
    p = p * r + 0x1.71de3a556c734p-19;
    p = p * r + 0x1.a01a01a01a01ap-16;
    p = p * r + 0x1.a01a01a01a01ap-13;
    p = p * r + 0x1.6c16c16c16c17p-10;
    p = p * r + 0x1.1111111111111p-7;
    p = p * r + 0x1.5555555555555p-5;
    p = p * r + 0x1.5555555555555p-3;
    p = p * r + 0x1.0000000000000p-1;
    p = p * r + 0x1.0000000000000p+0;
    p = p * r + 0x1.0000000000000p+0;
    return ldexp(p, (int)kd);
}
static double vm_log(double x) {
    int e;
    double m = frexp(x, &e);
    if (m < 0x1.6a09e667f3bcdp-1) {
        m = m * 2.0;
        e = e - 1;
    }
    double f = (m - 1.0) / (m + 1.0);
    double f2 = f * f;
    double p = 0x1.8618618618618p-5;
    p = p * f2 + 0x1.af286bca1af28p-5;
    p = p * f2 + 0x1.e1e1e1e1e1e1ep-5;
    p = p * f2 + 0x1.1111111111111p-4;
    p = p * f2 + 0x1.3b13b13b13b14p-4;
    p = p * f2 + 0x1.745d1745d1746p-4;
    p = p * f2 + 0x1.c71c71c71c71cp-4;
    p = p * f2 + 0x1.2492492492492p-3;
    p = p * f2 + 0x1.999999999999ap-3;
    p = p * f2 + 0x1.5555555555555p-2;
    p = p * f2 + 0x1.0000000000000p+0;
    double lm = 2.0 * f * p;
    double ed = (double)e;
    return ed * 0x1.62e42fee00000p-1 + (ed * 0x1.a39ef35793c76p-33 + lm);
}
static double vm_cosh(double x) { return 0.5 * (vm_exp(x) + vm_exp(-x)); }
/* asin on [0, 1/2] */
static double vm_asin05(double y) {
    double y2 = y * y;
    double p = 0x1.1052bc5fa960ap-9;
    p = p * y2 + 0x1.208d3570ae5a6p-9;
    p = p * y2 + 0x1.3275586c5f2f0p-9;
    p = p * y2 + 0x1.464c0950f7d47p-9;
    p = p * y2 + 0x1.5c5f56efaaaabp-9;
    p = p * y2 + 0x1.750de64d7d05fp-9;
    p = p * y2 + 0x1.90cb77f60c7cep-9;
    p = p * y2 + 0x1.b026f57b13b14p-9;
    p = p * y2 + 0x1.d3d2a8e0dd67dp-9;
    p = p * y2 + 0x1.fcaf8fb6db6dbp-9;
    p = p * y2 + 0x1.15ee9d45d1746p-8;
    p = p * y2 + 0x1.31683bdef7bdfp-8;
    p = p * y2 + 0x1.51ba308d3dcb1p-8;
    p = p * y2 + 0x1.782dda12f684cp-8;
    p = p * y2 + 0x1.a6863d70a3d71p-8;
    p = p * y2 + 0x1.df3bd37a6f4dfp-8;
    p = p * y2 + 0x1.12ef3cf3cf3cfp-7;
    p = p * y2 + 0x1.3fde50d79435ep-7;
    p = p * y2 + 0x1.7a87878787878p-7;
    p = p * y2 + 0x1.c99999999999ap-7;
    p = p * y2 + 0x1.1c4ec4ec4ec4fp-6;
    p = p * y2 + 0x1.6e8ba2e8ba2e9p-6;
    p = p * y2 + 0x1.f1c71c71c71c7p-6;
    p = p * y2 + 0x1.6db6db6db6db7p-5;
    p = p * y2 + 0x1.3333333333333p-4;
    p = p * y2 + 0x1.5555555555555p-3;
    p = p * y2 + 0x1.0000000000000p+0;
    return y * p;
}
/* acos on [0, 1] (NaN above 1, as acos) */
static double vm_acos01(double x) {
    if (x <= 0.5) return (0x1.921fb54442d18p+0 - vm_asin05(x)) + 0x1.1a62633145c07p-54;
    double y = sqrt((1.0 - x) * 0.5);
    return 2.0 * vm_asin05(y);
}

/* find_xi0_analytic (CFD/G_SMACSolver_VoF.cu:459-465) */
static double thinc_xi0(double alpha, double gamma) {
    double A = vm_exp(THINC_BETA * gamma * (2. * alpha - 1.));
    return 1. / (2. * THINC_BETA) * vm_log((vm_exp(THINC_BETA) - A) / (A - vm_exp(-THINC_BETA)));
}
/* integrate_thinc (CFD/G_SMACSolver_VoF.cu:451-455) */
static double thinc_integral(double a, double b, double gamma, double xi) {
    return 0.5 * (b - a) + gamma / (2. * THINC_BETA) *
                               (vm_log(vm_cosh(THINC_BETA * (b - xi))) - vm_log(vm_cosh(THINC_BETA * (a - xi))));
}

/* test hook: f = 0 exp, 1 log, 2 cosh, 3 acos on [0, 1] */
int smac_oracle_vofmath(int fn, const double* x, double* y, size_t n) {
    for (size_t t = 0; t < n; t++) {
        double v = x[t];
        y[t] = fn == 0 ? vm_exp(v) : (fn == 1 ? vm_log(v) : (fn == 2 ? vm_cosh(v) : vm_acos01(v)));
    }
    return SMAC_OK;
}
static inline double sgnd(double a) { return a < 0 ? -1. : 1.; }

/* one tangential component of the donor cell's normal (-grad alpha):
 * central, one-sided where a neighbour is not interior, else 0 */
static double vof_normal_t(const struct smac_ctx* c, ivec dn, int t) {
    ivec qp = add(dn, t, +1), qm = add(dn, t, -1);
    int ip = c->per[t] || (qp.v[t] >= 1 && qp.v[t] <= c->n[t]);
    int im = c->per[t] || (qm.v[t] >= 1 && qm.v[t] <= c->n[t]);
    const double* al = c->alpha;
    if (ip && im) return -(al[IX(c, wr(c, qp))] - al[IX(c, wr(c, qm))]) * (1. / (2. * c->d[t]));
    if (!ip && im) return -(al[IX(c, dn)] - al[IX(c, wr(c, qm))]) * c->inv_d[t];
    if (ip && !im) return -(al[IX(c, wr(c, qp))] - al[IX(c, dn)]) * c->inv_d[t];
    return 0.;
}

/* THINC/WLIC flux of the a-face q: k_alpha_flux_thincwlic_{x,y,z}
 * (CFD/G_SMACSolver_VoF.cu:507-666, 668-827, 829-984). q is a computed face
 * (2..N along a wall axis, 1..N along a periodic one), so its donor cell is
 * interior and the reference's F_BOUNDARY branch has no case. */
static double vof_face_flux(const struct smac_ctx* c, int a, ivec q, double dt) {
    const double* al = c->alpha;
    double v = c->vel[a][IX(c, q)];
    ivec dn = wr(c, v > 0.0 ? add(q, a, -1) : q);
    double ad = al[IX(c, dn)];
    double dtbyd = dt * c->inv_d[a];
    ivec dp = add(dn, a, +1), dm = add(dn, a, -1);
    int ip = c->per[a] || dp.v[a] <= c->n[a];
    int im = c->per[a] || dm.v[a] >= 1;
    if (!ip || !im) return v * ad * dtbyd;
    double gam = al[IX(c, wr(c, dp))] - al[IX(c, wr(c, dm))];
    if (ad < VOF_EPS || ad > 1.0 - VOF_EPS || fabs(gam) < 1e-6) return v * ad * dtbyd;
    double nrm[3];
    nrm[a] = -gam * (1. / (2. * c->d[a]));
    for (int t = 0; t < 3; t++)
        if (t != a) nrm[t] = vof_normal_t(c, dn, t);
    double s_sq = nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2];
    double sn = sqrt(s_sq);
    double inv_s = 1.0 / (sn + VOF_EPS);
    double theta = vm_acos01(fabs(nrm[a]) * inv_s);
    double wc = 2. * theta / VOF_PI;
    double w = 1.0 - wc * wc;
    double gamma = sgnd(gam);
    double xi0 = thinc_xi0(ad, gamma);
    double lambda = v * dtbyd;
    double ft = v > 0.0 ? thinc_integral(1.0 - lambda, 1.0, gamma, xi0) : -thinc_integral(0.0, -lambda, gamma, xi0);
    double fu = lambda * ad;
    return w * ft + (1. - w) * fu;
}

/* k_alpha_flux_thincwlic_{x,y,z} over the computed faces, then
 * k_transport_alpha_{x,y,z} (CFD/G_SMACSolver_VoF.cu:1011-1152) into
 * alpha_new and the swap; the accumulation of k_alpha_flux_accum
 * (468-505) is taken here, per computed face (other faces hold 0) */
static void vof_direction(struct smac_ctx* c, int a, double dt) {
    int lo[3] = {1, 1, 1}, hi[3] = {c->n[0], c->n[1], c->n[2]};
    lo[a] = c->per[a] ? 1 : 2;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = lo[2]; k <= hi[2]; k++)
        for (int j = lo[1]; j <= hi[1]; j++)
            for (int i = lo[0]; i <= hi[0]; i++) {
                ivec q = iv(i, j, k);
                size_t id = IX(c, q);
                double F = vof_face_flux(c, a, q, dt);
                c->vf[a][id] = F;
                c->vfacc[a][id] += F;
            }
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= c->n[2]; k++)
        for (int j = 1; j <= c->n[1]; j++)
            for (int i = 1; i <= c->n[0]; i++) {
                ivec q = iv(i, j, k);
                size_t id = IX(c, q);
                double flux = 0.;
                flux += c->vf[a][IX(c, wr(c, add(q, a, +1)))] - c->vf[a][id];
                double up = get_face(c, a, add(q, a, +1));
                double um = get_face(c, a, q);
                double coeff = (double)c->vcoeff[id];
                flux -= coeff * (up - um) * dt * c->inv_d[a];
                c->alpha_new[id] = c->alpha[id] - flux;
            }
    double* t = c->alpha;
    c->alpha = c->alpha_new;
    c->alpha_new = t;
}

/* G_SMACSolver::alpha_flux_thincwlic_split (CFD/G_SMACSolver_VoF.cu:1575-1674):
 * correct_coeff from alpha (k_set_correct_coeff, 986-1009), the three
 * directions in the order of step % 3 (xyz, yzx, zxy), then k_clip_alpha
 * (1173-1190) */
static void vof_split(struct smac_ctx* c, double dt, int step) {
    for (int k = 1; k <= c->n[2]; k++)
        for (int j = 1; j <= c->n[1]; j++)
            for (int i = 1; i <= c->n[0]; i++) {
                size_t id = IDX(c, i, j, k);
                c->vcoeff[id] = c->alpha[id] > 0.5 ? 1 : 0;
            }
    int first = step % 3;
    for (int s = 0; s < 3; s++) vof_direction(c, (first + s) % 3, dt);
    for (int k = 1; k <= c->n[2]; k++)
        for (int j = 1; j <= c->n[1]; j++)
            for (int i = 1; i <= c->n[0]; i++) {
                size_t id = IDX(c, i, j, k);
                if (c->alpha[id] < VOF_EPS) c->alpha[id] = 0.;
                else if (c->alpha[id] > 1.) c->alpha[id] = 1.;
            }
}

static int vof_alloc(struct smac_ctx* c) {
    if (c->alpha_new) return SMAC_OK;
    c->alpha_new = (double*)calloc(c->tot, sizeof(double));
    c->vcoeff = (unsigned char*)calloc(c->tot, 1);
    for (int a = 0; a < 3; a++) {
        c->vf[a] = (double*)calloc(c->tot, sizeof(double));
        c->vfacc[a] = (double*)calloc(c->tot, sizeof(double));
    }
    if (!c->alpha_new || !c->vcoeff || !c->vf[0] || !c->vf[1] || !c->vf[2] || !c->vfacc[0] || !c->vfacc[1] ||
        !c->vfacc[2])
        return SMAC_ERR_OOM;
    return SMAC_OK;
}

/* G_SMACSolver::clear_alpha_flux_accum (CFD/G_SMACSolver_VoF.cu:1688-1696) */
int smac_clear_alpha_flux_accum(smac_ctx* c) {
    if (!c) return SMAC_ERR_INVALID_ARG;
    if (vof_alloc(c)) return SMAC_ERR_OOM;
    for (int a = 0; a < 3; a++) memset(c->vfacc[a], 0, c->tot * sizeof(double));
    c->vof_armed = 0;
    return SMAC_OK;
}

/* one substep: clear_alpha_flux, alpha_flux_thincwlic_split(sub_dt, step),
 * alpha_flux_accum (CFD/main.cu:289-292) */
int smac_alpha_flux_split(smac_ctx* c, double sub_dt, int32_t step) {
    if (!c || !(sub_dt > 0) || step < 0) return SMAC_ERR_INVALID_ARG;
    if (vof_alloc(c)) return SMAC_ERR_OOM;
    vof_split(c, sub_dt, step);
    c->vof_armed = 1;
    return SMAC_OK;
}

/* One time step of VoF transport (CFD/main.cu:283-295) */
int smac_transport_alpha(smac_ctx* c, double dt, int32_t substeps, int32_t step) {
    if (!c || !(dt > 0) || substeps < 1 || step < 0) return SMAC_ERR_INVALID_ARG;
    int rc = smac_clear_alpha_flux_accum(c);
    double sub_dt = dt / (double)substeps;
    for (int s = 0; s < substeps && !rc; s++) rc = smac_alpha_flux_split(c, sub_dt, step);
    return rc;
}

/* ------------------------------------------------ after the step */
/* d_get_velocity_ghost_value (CFD/G_BoundaryFunctions.cu:257-275) with the IBM
 * solid fraction lambda = 0 (no immersed solid): bval is the face one row
 * inside the wall, inval the face two rows inside. */
static double ghost_value(int bt, double bv, double bval, double inval) {
    if (bt == SMAC_BC_INFLOW) return 2.0 * (1.0 - 0.0) * bv - inval;
    if (bt == SMAC_BC_NOSLIP) return -inval;
    if (bt == SMAC_BC_OUTLET) return SMAC_ORACLE_MUTANT == 11 ? bval : 2.0 * bval - inval;
    if (bt == SMAC_BC_SLIP) return inval;
    return bval;
}

/* ghost of component a across the wall of tangential axis t at q (q.v[t] is 0
 * or n+1): the body of k_update_v{x,y,z}_ghost (CFD/G_BoundaryFunctions.cu:277-336,
 * 365-424, 454-513), bc id of the a-face one row inside */
static double wall_ghost(const struct smac_ctx* c, int a, int t, ivec q) {
    int lower = q.v[t] == 0;
    ivec qb = q, qi = q;
    qb.v[t] = lower ? 1 : c->n[t];
    qi.v[t] = lower ? 2 : c->n[t] - 1;
    int bid = face_bcid(c, a, qb);
    return ghost_value(c->bctype[bid], c->bcv[bid][a], c->vel[a][IX(c, qb)], c->vel[a][IX(c, qi)]);
}

/* value of velocity component a at position q of its face array after the
 * ghost update: periodic directions resolve to their image; a ghost across one
 * wall follows wall_ghost; an edge ghost across two walls is the mean of the
 * two adjacent wall ghosts, the one across the lower-numbered axis first
 * (k_update_v{x,y,z}_ghost_corner, CFD/G_BoundaryFunctions.cu:338-363,426-452,515-540) */
static double ghost_vel(const struct smac_ctx* c, int a, ivec q) {
    for (int d = 0; d < 3; d++) {
        if (!c->per[d]) continue;
        q.v[d] = wrap(q.v[d], c->n[d]);
        if (d == a && q.v[d] > c->n[d]) q.v[d] -= c->n[d];
    }
    int t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2;
    int g1 = q.v[t1] == 0 || q.v[t1] == c->n[t1] + 1;
    int g2 = q.v[t2] == 0 || q.v[t2] == c->n[t2] + 1;
    if (!g1 && !g2) return c->vel[a][IX(c, q)];
    if (g1 && !g2) return wall_ghost(c, a, t1, q);
    if (!g1 && g2) return wall_ghost(c, a, t2, q);
    ivec q1 = q, q2 = q;
    q1.v[t2] = q.v[t2] == 0 ? 1 : c->n[t2];
    q2.v[t1] = q.v[t1] == 0 ? 1 : c->n[t1];
    return 0.5 * (wall_ghost(c, a, t1, q1) + wall_ghost(c, a, t2, q2));
}

/* G_SMACSolver::update_boundary_ghost (CFD/G_SMACSolver.cu:1345-1362) */
int smac_update_boundary_ghost(smac_ctx* c, double dt) {
    if (!c) return SMAC_ERR_INVALID_ARG;
    if (!c->props_ready) smac_update_properties(c, 1);
    const double gv[3] = {c->fl.gx, c->fl.gy, c->fl.gz};
    /* k_update_cell_ghost_pressure (CFD/G_SMACSolver.cu:1227-1267) */
    for (int k = 0; k <= c->n[2] + 1; k++)
        for (int j = 0; j <= c->n[1] + 1; j++)
            for (int i = 0; i <= c->n[0] + 1; i++) {
                ivec q = iv(i, j, k);
                if (cell_interior(c, q)) continue;
                ivec s = q;
                double r[3];
                for (int d = 0; d < 3; d++) {
                    if (c->per[d]) s.v[d] = wrap(s.v[d], c->n[d]);
                    else if (s.v[d] < 1) s.v[d] = 1;
                    else if (s.v[d] > c->n[d]) s.v[d] = c->n[d];
                    r[d] = c->per[d] ? 0.0 : (double)(q.v[d] - s.v[d]) * c->d[d];
                }
                size_t is = IX(c, s);
                c->p_[IX(c, q)] = c->p_[is] + c->rho[is] * (r[0] * gv[0] + r[1] * gv[1] + r[2] * gv[2]);
            }
    /* k_update_v{x,y,z}_outlet (CFD/G_BoundaryFunctions.cu:143-255), before
     * the velocity ghosts, which read these faces: on the boundary a-faces of
     * an outlet side, u = u* + s*f_b*dt*dp(c)/h with c the inner cell
     * (internal id 0 on the low side, -1 on the high side) and s = +1 when the
     * id is negative, else -1 */
    for (int a = 0; a < 3; a++) {
        if (c->per[a]) continue;
        int b1 = (a + 1) % 3, b2 = (a + 2) % 3;
        for (int side = 0; side < 2; side++) {
            if (c->bctype[c->side_bcid[2 * a + side]] != SMAC_BC_OUTLET) continue;
            int shift = side ? -1 : 0;
            int sgn = (shift < 0) == (SMAC_ORACLE_MUTANT != 13) ? 1 : -1;
            for (int t2 = 1; t2 <= c->n[b2]; t2++)
                for (int t1 = 1; t1 <= c->n[b1]; t1++) {
                    ivec q;
                    q.v[a] = side ? c->n[a] + 1 : 1;
                    q.v[b1] = t1;
                    q.v[b2] = t2;
                    size_t id = IX(c, q);
                    c->vel[a][id] = c->vst[a][id] + sgn * c->beta[a][id] * dt * c->dp[IX(c, add(q, a, shift))] *
                                                        c->inv_d[a];
                }
        }
    }
    /* velocity ghosts: every position of the face array with a ghost index
     * along a tangential axis, or along its own axis when that is periodic;
     * faces 0 and n+2 along a wall axis are not touched (the kernels run
     * the own index over 1..n+1). Values are computed from entries that no
     * write changes, then stored. */
    for (int a = 0; a < 3; a++) {
        int sh[3] = {c->n[0] + 2, c->n[1] + 2, c->n[2] + 2};
        sh[a]++;
        size_t cnt = (size_t)sh[0] * sh[1] * sh[2], m = 0;
        size_t* pos = (size_t*)malloc(cnt * sizeof(size_t));
        double* val = (double*)malloc(cnt * sizeof(double));
        if (!pos || !val) { free(pos); free(val); return SMAC_ERR_OOM; }
        for (int k = 0; k < sh[2]; k++)
            for (int j = 0; j < sh[1]; j++)
                for (int i = 0; i < sh[0]; i++) {
                    ivec q = iv(i, j, k);
                    int ghost = 0, skip = 0;
                    for (int d = 0; d < 3; d++) {
                        int x = q.v[d], n = c->n[d];
                        if (d == a) {
                            if (c->per[d]) ghost |= (x < 1 || x > n);
                            else skip |= (x < 1 || x > n + 1);
                        } else {
                            ghost |= (x < 1 || x > n);
                        }
                    }
                    if (!ghost || skip) continue;
                    pos[m] = IX(c, q);
                    val[m] = ghost_vel(c, a, q);
                    m++;
                }
        for (size_t t = 0; t < m; t++) c->vel[a][pos[t]] = val[t];
        free(pos);
        free(val);
    }
    return SMAC_OK;
}

/* G_SMACSolver::make_face_gradp (CFD/G_SMACSolver.cu:1277-1343) */
int smac_make_face_gradp(smac_ctx* c) {
    if (!c) return SMAC_ERR_INVALID_ARG;
    for (int a = 0; a < 3; a++)
        for (int k = 1; k <= c->n[2] + 1; k++)
            for (int j = 1; j <= c->n[1] + 1; j++)
                for (int i = 1; i <= c->n[0] + 1; i++) {
                    ivec q = iv(i, j, k);
                    c->gradp[a][IX(c, q)] = (c->p_[IX(c, q)] - c->p_[IX(c, add(q, a, -1))]) * c->inv_d[a];
                }
    return SMAC_OK;
}

/* FileInOut::output_vti_binary_cellData (CFD/FileInOut.cpp:204-236): cell
 * data as float, x fastest: p, alpha, face-mean velocity interleaved */
int smac_cell_output(smac_ctx* c, float* h, size_t count) {
    if (!c || !h) return SMAC_ERR_INVALID_ARG;
    size_t n = (size_t)c->n[0] * c->n[1] * c->n[2];
    if (count != 5 * n) return SMAC_ERR_INVALID_ARG;
    size_t m = 0;
    for (int k = 1; k <= c->n[2]; k++)
        for (int j = 1; j <= c->n[1]; j++)
            for (int i = 1; i <= c->n[0]; i++, m++) {
                size_t id = IDX(c, i, j, k);
                h[m] = (float)c->p_[id];
                h[n + m] = (float)c->alpha[id];
                h[2 * n + 3 * m + 0] = (float)(0.5 * (c->vel[0][id] + c->vel[0][IDX(c, i + 1, j, k)]));
                h[2 * n + 3 * m + 1] = (float)(0.5 * (c->vel[1][id] + c->vel[1][IDX(c, i, j + 1, k)]));
                h[2 * n + 3 * m + 2] = (float)(0.5 * (c->vel[2][id] + c->vel[2][IDX(c, i, j, k + 1)]));
            }
    return SMAC_OK;
}

/* --------------------------------------------------------- lifecycle, IO */
static int shape_of(const struct smac_ctx* c, int field, int sh[3]) {
    sh[0] = c->n[0] + 2; sh[1] = c->n[1] + 2; sh[2] = c->n[2] + 2;
    switch (field) {
        case SMAC_FIELD_U: case SMAC_FIELD_USTAR: case SMAC_FIELD_GRADPX: sh[0]++; return 0;
        case SMAC_FIELD_V: case SMAC_FIELD_VSTAR: case SMAC_FIELD_GRADPY: sh[1]++; return 1;
        case SMAC_FIELD_W: case SMAC_FIELD_WSTAR: case SMAC_FIELD_GRADPZ: sh[2]++; return 2;
        case SMAC_FIELD_FLUXX: case SMAC_FIELD_IMPX: sh[0]++; return 0;
        case SMAC_FIELD_FLUXY: case SMAC_FIELD_IMPY: sh[1]++; return 1;
        case SMAC_FIELD_FLUXZ: case SMAC_FIELD_IMPZ: sh[2]++; return 2;
        default: return -1;
    }
}
static double* field_ptr(struct smac_ctx* c, int field) {
    switch (field) {
        case SMAC_FIELD_U: return c->vel[0];
        case SMAC_FIELD_V: return c->vel[1];
        case SMAC_FIELD_W: return c->vel[2];
        case SMAC_FIELD_USTAR: return c->vst[0];
        case SMAC_FIELD_VSTAR: return c->vst[1];
        case SMAC_FIELD_WSTAR: return c->vst[2];
        case SMAC_FIELD_P: return c->p_;
        case SMAC_FIELD_DP: return c->dp;
        case SMAC_FIELD_RHS: return c->rhs;
        case SMAC_FIELD_ALPHA: return c->alpha;
        case SMAC_FIELD_GRADPX: return c->gradp[0];
        case SMAC_FIELD_GRADPY: return c->gradp[1];
        case SMAC_FIELD_GRADPZ: return c->gradp[2];
        case SMAC_FIELD_FLUXX: return c->vfacc[0];
        case SMAC_FIELD_FLUXY: return c->vfacc[1];
        case SMAC_FIELD_FLUXZ: return c->vfacc[2];
        case SMAC_FIELD_EPS: return c->eps;
        case SMAC_FIELD_EPS_OLD: return c->eps_old;
        case SMAC_FIELD_EPS_HALF: return c->eps_half;
        case SMAC_FIELD_IMPX: return c->imp[0];
        case SMAC_FIELD_IMPY: return c->imp[1];
        case SMAC_FIELD_IMPZ: return c->imp[2];
        default: return NULL;
    }
}
size_t smac_field_count(const smac_ctx* c, int32_t field) {
    if (!c || field < 0 || field >= SMAC_FIELD_COUNT) return 0;
    int sh[3];
    shape_of(c, field, sh);
    return (size_t)sh[0] * sh[1] * sh[2];
}

/* Interior entries are taken from the host buffer; on download, periodic
 * images (ghost cells, duplicate faces) are filled from their sources. */
int smac_upload(smac_ctx* c, int32_t field, const double* h, size_t count) {
    if (!c || !h || field < 0 || field >= SMAC_FIELD_COUNT) return SMAC_ERR_INVALID_ARG;
    if (count != smac_field_count(c, field)) return SMAC_ERR_INVALID_ARG;
    int sh[3];
    shape_of(c, field, sh);
    if (field >= SMAC_FIELD_FLUXX && field <= SMAC_FIELD_FLUXZ && vof_alloc(c)) return SMAC_ERR_OOM;
    if (field >= SMAC_FIELD_EPS) { /* CFD-DEM inputs, allocated on first upload */
        double** slot = field == SMAC_FIELD_EPS ? &c->eps
                      : field == SMAC_FIELD_EPS_OLD ? &c->eps_old
                      : field == SMAC_FIELD_EPS_HALF ? &c->eps_half : &c->imp[field - SMAC_FIELD_IMPX];
        if (!*slot) *slot = (double*)calloc(c->tot, sizeof(double));
        if (!*slot) return SMAC_ERR_OOM;
    }
    double* dst = field_ptr(c, field);
    for (int k = 0; k < sh[2]; k++)
        for (int j = 0; j < sh[1]; j++)
            for (int i = 0; i < sh[0]; i++)
                dst[IDX(c, i, j, k)] = h[((size_t)k * sh[1] + j) * sh[0] + i];
    if (field == SMAC_FIELD_ALPHA) c->props_ready = 0;
    return SMAC_OK;
}
int smac_download(smac_ctx* c, int32_t field, double* h, size_t count) {
    if (!c || !h || field < 0 || field >= SMAC_FIELD_COUNT) return SMAC_ERR_INVALID_ARG;
    if (count != smac_field_count(c, field)) return SMAC_ERR_INVALID_ARG;
    int sh[3];
    int a = shape_of(c, field, sh);
    const double* src = field_ptr(c, field);
    if (!src) {
        memset(h, 0, count * sizeof(double));
        return SMAC_OK;
    }
    for (int k = 0; k < sh[2]; k++)
        for (int j = 0; j < sh[1]; j++)
            for (int i = 0; i < sh[0]; i++) {
                int q[3] = {i, j, k};
                for (int d = 0; d < 3; d++) {
                    if (!c->per[d]) continue;
                    /* cells: 0 -> n, n+1 -> 1; faces along a: 0 -> n, n+1 -> 1, n+2 -> 2 */
                    q[d] = wrap(q[d], c->n[d]);
                    if (d == a && q[d] > c->n[d]) q[d] -= c->n[d];
                }
                h[((size_t)k * sh[1] + j) * sh[0] + i] = src[IDX(c, q[0], q[1], q[2])];
            }
    return SMAC_OK;
}
int smac_upload_device(smac_ctx* c, int32_t f, const double* d, size_t n) {
    (void)c; (void)f; (void)d; (void)n;
    return SMAC_ERR_UNSUPPORTED;
}
int smac_download_device(smac_ctx* c, int32_t f, double* d, size_t n) {
    (void)c; (void)f; (void)d; (void)n;
    return SMAC_ERR_UNSUPPORTED;
}
/* The oracle's arrays are host boxes of (Nx+3)(Ny+3)(Nz+3); nothing binds. */
int smac_device_layout(const smac_ctx* c, int64_t* px, int64_t* py, int64_t* planes) {
    if (!c || !px || !py || !planes) return SMAC_ERR_INVALID_ARG;
    *px = c->n[0] + 3;
    *py = c->n[1] + 3;
    *planes = c->n[2] + 3;
    return SMAC_OK;
}
int smac_bind_device(smac_ctx* c, int32_t f, double* d, size_t n) {
    (void)c; (void)f; (void)d; (void)n;
    return SMAC_ERR_UNSUPPORTED;
}

int smac_set_fluid(smac_ctx* c, const smac_fluid_desc* f) {
    if (!c || !f) return SMAC_ERR_INVALID_ARG;
    c->fl = *f;
    c->props_ready = 0;
    return SMAC_OK;
}

int smac_create(const smac_grid_desc* g, const smac_fluid_desc* f, smac_ctx** out) {
    if (!g || !out) return SMAC_ERR_INVALID_ARG;
    *out = NULL;
    int n[3] = {g->nx, g->ny, g->nz};
    double sz[3] = {g->sizex, g->sizey, g->sizez};
    for (int d = 0; d < 3; d++) {
        if (n[d] < 2 || !(sz[d] > 0)) return SMAC_ERR_INVALID_GRID;
        if (g->periodic[d] && n[d] < 4) return SMAC_ERR_INVALID_GRID;
    }
    if (g->num_bc < 0 || g->num_bc > 256 || (g->num_bc > 0 && !g->bc_table)) return SMAC_ERR_INVALID_ARG;
    struct smac_ctx* c = (struct smac_ctx*)calloc(1, sizeof(*c));
    for (int d = 0; d < 3; d++) {
        c->n[d] = n[d];
        c->p[d] = n[d] + 3;
        c->per[d] = g->periodic[d] ? 1 : 0;
        c->d[d] = sz[d] / (double)n[d];
        c->inv_d[d] = 1. / c->d[d];
        c->inv_d2[d] = 1. / (c->d[d] * c->d[d]);
    }
    memcpy(c->side_bcid, g->side_bcid, 6);
    c->nbc = g->num_bc;
    for (int b = 0; b < g->num_bc; b++) {
        c->bctype[b] = g->bc_table[b].type;
        c->bcv[b][0] = g->bc_table[b].vx;
        c->bcv[b][1] = g->bc_table[b].vy;
        c->bcv[b][2] = g->bc_table[b].vz;
    }
    /* BC_OUTFLOW has no case in the reference's kernels (and no config
     * spelling): refused. An outlet side makes the operator regular. */
    for (int s = 0; s < 6; s++) {
        if (c->per[s / 2]) continue;
        int bt = c->bctype[c->side_bcid[s]];
        if (bt == SMAC_BC_OUTFLOW || bt < SMAC_BC_NOSLIP || bt > SMAC_BC_OUTLET) {
            free(c);
            return SMAC_ERR_UNSUPPORTED;
        }
        if (bt == SMAC_BC_OUTLET) c->has_outlet = 1;
    }
    if (f) c->fl = *f;
    else { c->fl.rho0 = c->fl.rho1 = 1.0; c->fl.mu0 = c->fl.mu1 = 0.0; }
    c->tot = (size_t)c->p[0] * c->p[1] * c->p[2];
    double** arrs[] = {&c->vel[0], &c->vel[1], &c->vel[2], &c->vst[0], &c->vst[1], &c->vst[2], &c->p_, &c->dp,
                       &c->rhs, &c->alpha, &c->rho, &c->rho_old, &c->inv_rho, &c->mu, &c->beta[0], &c->beta[1],
                       &c->beta[2], &c->inv_rhof[0], &c->inv_rhof[1], &c->inv_rhof[2], &c->muf[0], &c->muf[1],
                       &c->muf[2], &c->mf[0], &c->mf[1], &c->mf[2], &c->rhof[0], &c->rhof[1], &c->rhof[2],
                       &c->r, &c->z, &c->dvec, &c->Ad, &c->Adiag, &c->Aoff[0], &c->Aoff[1], &c->Aoff[2],
                       &c->Aoff[3], &c->Aoff[4], &c->Aoff[5], &c->gradp[0], &c->gradp[1], &c->gradp[2]};
    for (size_t t = 0; t < sizeof(arrs) / sizeof(arrs[0]); t++) {
        *arrs[t] = (double*)calloc(c->tot, sizeof(double));
        if (!*arrs[t]) { smac_destroy(c); return SMAC_ERR_OOM; }
    }
    *out = c;
    return SMAC_OK;
}

void smac_destroy(smac_ctx* c) {
    if (!c) return;
    mg_free(c);
    rg_free(c);
    double* arrs[] = {c->vel[0], c->vel[1], c->vel[2], c->vst[0], c->vst[1], c->vst[2], c->p_, c->dp, c->rhs,
                      c->alpha, c->rho, c->rho_old, c->inv_rho, c->mu, c->beta[0], c->beta[1], c->beta[2],
                      c->inv_rhof[0], c->inv_rhof[1], c->inv_rhof[2], c->muf[0], c->muf[1], c->muf[2], c->mf[0],
                      c->mf[1], c->mf[2], c->rhof[0], c->rhof[1], c->rhof[2], c->r, c->z, c->dvec, c->Ad,
                      c->Adiag, c->Aoff[0], c->Aoff[1], c->Aoff[2], c->Aoff[3], c->Aoff[4], c->Aoff[5],
                      c->gradp[0], c->gradp[1], c->gradp[2], c->alpha_new, c->vf[0], c->vf[1], c->vf[2],
                      c->vfacc[0], c->vfacc[1], c->vfacc[2], c->eps, c->eps_old, c->eps_half, c->imp[0],
                      c->imp[1], c->imp[2], c->bpois[0], c->bpois[1], c->bpois[2]};
    for (size_t t = 0; t < sizeof(arrs) / sizeof(arrs[0]); t++) free(arrs[t]);
    free(c->vcoeff);
    free(c);
}

int smac_set_profiling(smac_ctx* c, int32_t on) { (void)c; (void)on; return SMAC_OK; }
int smac_kernel_time(smac_ctx* c, const char* name, int64_t* n, double* ms) {
    (void)c; (void)name;
    if (n) *n = 0;
    if (ms) *ms = 0;
    return SMAC_ERR_INVALID_ARG;
}
int smac_kernel_names(smac_ctx* c, char* buf, size_t len) {
    (void)c;
    if (!buf || len == 0) return SMAC_ERR_INVALID_ARG;
    buf[0] = 0;
    return SMAC_OK;
}
int smac_reset_kernel_times(smac_ctx* c) { (void)c; return SMAC_OK; }
int smac_sync(smac_ctx* c) { (void)c; return SMAC_OK; }
void* smac_stream(smac_ctx* c) { (void)c; return NULL; }
const char* smac_backend_name(void) { return "oracle-cpu-fp64"; }
const char* smac_strerror(int s) {
    switch (s) {
        case SMAC_OK: return "ok";
        case SMAC_ERR_INVALID_ARG: return "invalid argument";
        case SMAC_ERR_INVALID_GRID: return "grid sizes not supported by the multigrid hierarchy";
        case SMAC_ERR_HIP: return "HIP runtime error";
        case SMAC_ERR_NO_DEVICE: return "no device";
        case SMAC_ERR_OOM: return "out of memory";
        case SMAC_ERR_NAN: return "NaN in solve";
        case SMAC_ERR_UNSUPPORTED: return "unsupported option";
        default: return "unknown error";
    }
}

/* ------------------------------------------------------------ test hooks */
/* Apply the assembled Poisson operator to a cell field x (reference layout),
 * for operator property tests (symmetry, null space, eigenfunctions). */
int smac_oracle_apply_A(smac_ctx* c, const double* x_ref, double* y_ref) {
    if (!c) return SMAC_ERR_INVALID_ARG;
    if (!c->props_ready) smac_update_properties(c, 1);
    build_A(c);
    size_t cnt = smac_field_count(c, SMAC_FIELD_P);
    smac_upload(c, SMAC_FIELD_DP, x_ref, cnt);
    for (int k = 1; k <= c->n[2]; k++)
        for (int j = 1; j <= c->n[1]; j++)
            for (int i = 1; i <= c->n[0]; i++) c->Ad[IDX(c, i, j, k)] = apply_A(c, c->dp, i, j, k);
    double* save = c->rhs;
    c->rhs = c->Ad;
    smac_download(c, SMAC_FIELD_RHS, y_ref, cnt);
    c->rhs = save;
    memset(c->dp, 0, c->tot * sizeof(double));
    return SMAC_OK;
}
/* Apply the multigrid preconditioner z = M r (reference layout). */
int smac_oracle_apply_mg(smac_ctx* c, const smac_solver_opts* o, const double* r_ref, double* z_ref) {
    if (!c || !o) return SMAC_ERR_INVALID_ARG;
    if (!c->props_ready) smac_update_properties(c, 1);
    build_A(c);
    if (o->precond == SMAC_PRECOND_REF_GMG) {
        int L = rg_setup(c, o->mg_levels > 0 ? o->mg_levels : 4);
        if (L < 0) return L;
        rg_coeffs(c);
    } else {
        int L = o->mg_levels > 0 ? o->mg_levels : mg_auto_levels(c);
        if (!mg_check_levels(c, L)) return SMAC_ERR_INVALID_GRID;
        mg_setup(c, L);
    }
    size_t cnt = smac_field_count(c, SMAC_FIELD_P);
    double* save = c->rhs;
    c->rhs = c->r;
    smac_upload(c, SMAC_FIELD_RHS, r_ref, cnt);
    c->rhs = save;
    precond(c, o);
    save = c->rhs;
    c->rhs = c->z;
    smac_download(c, SMAC_FIELD_RHS, z_ref, cnt);
    c->rhs = save;
    return SMAC_OK;
}
/* Face coefficients of axis `axis` on level `level` of the reference GMG:
 * dims <- the level's cell counts; out (if cap is large enough) <- the
 * (n0+3)(n1+3)(n2+3) box, x fastest. Returns the level count or a status. */
int smac_oracle_rg_beta(smac_ctx* c, const smac_solver_opts* o, int level, int axis, int dims[3], double* out,
                        size_t cap) {
    if (!c || !o || axis < 0 || axis > 2) return SMAC_ERR_INVALID_ARG;
    if (!c->props_ready) smac_update_properties(c, 1);
    int L = rg_setup(c, o->mg_levels > 0 ? o->mg_levels : 4);
    if (L < 0) return L;
    if (level < 0 || level >= L) return SMAC_ERR_INVALID_ARG;
    rg_coeffs(c);
    const rglevel* V = &c->rg[level];
    for (int d = 0; d < 3; d++) dims[d] = V->n[d];
    if (out && cap >= V->tot) memcpy(out, V->b[axis], V->tot * sizeof(double));
    return L;
}
/* Canonical sum of an interior cell field (reference layout). */
double smac_oracle_csum(smac_ctx* c, const double* x_ref) {
    size_t cnt = smac_field_count(c, SMAC_FIELD_P);
    double* save = c->rhs;
    c->rhs = c->Ad;
    smac_upload(c, SMAC_FIELD_RHS, x_ref, cnt);
    c->rhs = save;
    return csum(c, RED_SUM, c->Ad, NULL);
}
int smac_oracle_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
