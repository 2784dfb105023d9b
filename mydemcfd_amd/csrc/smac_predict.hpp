// smac_predict.hpp -- the per-face predictor of the SMAC step as device
// templates, shared by the forward kernels (smac_kernels.hip) and the adjoint
// (smac_adjoint.hip, which re-evaluates a face's terms to differentiate them).
#pragma once

#include "smac_internal.hpp"

namespace smac {

// --------------------------------------------------------------- predictor
// Boundary policies of the predictor. WallBC applies the reference getters
// (face types, wall mirror, bc table); PeriodicBC is the triply periodic box
// where every face is interior and the getters reduce to wrapped loads.
struct WallBC {
  static __device__ __forceinline__ double face(const Geo& g, const Vel& V, int a, I3 q) {
    return get_face(g, V, a, q);
  }
  static __device__ __forceinline__ double dir(const Geo& g, const Vel& V, int a, int b, I3 q, int s) {
    return get_dir(g, V, a, b, q, s);
  }
  static __device__ __forceinline__ bool interior(const Geo& g, int a, I3 q) {
    return face_type(g, a, wr(g, q)) == F_INTERIOR;
  }
};
struct PeriodicBC {
  static __device__ __forceinline__ I3 w3(const Geo& g, I3 q) {
    q.v[0] = wrapi(q.v[0], g.n[0]);
    q.v[1] = wrapi(q.v[1], g.n[1]);
    q.v[2] = wrapi(q.v[2], g.n[2]);
    return q;
  }
  static __device__ __forceinline__ double face(const Geo& g, const Vel& V, int a, I3 q) {
    return V.c[a][IX(g, w3(g, q))];
  }
  static __device__ __forceinline__ double dir(const Geo& g, const Vel& V, int a, int b, I3 q, int s) {
    return V.c[a][IX(g, w3(g, sh(q, b, s)))];
  }
  static __device__ __forceinline__ bool interior(const Geo&, int, I3) { return true; }
};

// Cross-direction viscous term of component A along axis B (B != A).
// Edge viscosities are the mean of the two B-face viscosities around the edge.
template <class BC, class PR, int A, int B>
__device__ __forceinline__ double visc_cross(const Geo& g, const Props& P, const Vel& V, I3 q, double v0) {
  double vbm = BC::dir(g, V, A, B, q, -1);
  double vbp = BC::dir(g, V, A, B, q, +1);
  double wpp = BC::face(g, V, B, sh(q, B, +1));
  double wmp = BC::face(g, V, B, sh(sh(q, B, +1), A, -1));
  double wp0 = BC::face(g, V, B, q);
  double wm0 = BC::face(g, V, B, sh(q, A, -1));
  double mu_p = PR::mu_edge(g, P, B, wr(g, sh(q, B, +1)), wr(g, sh(sh(q, B, +1), A, -1)));
  double t_p = mu_p * ((wpp - wmp) * g.inv_d[A] + (vbp - v0) * g.inv_d[B]);
  double mu_m = PR::mu_edge(g, P, B, wr(g, q), wr(g, sh(q, A, -1)));
  double t_m = mu_m * ((wp0 - wm0) * g.inv_d[A] + (v0 - vbm) * g.inv_d[B]);
  return (t_p - t_m) * g.inv_d[B];
}

// Mass flux policies of the predictor, the face value mf of
// k_compute_mass_flux_from_alpha_flux (CFD/G_SMACSolver_VoF.cu:1708-1790):
// MfRho0 has no alpha flux (mf = rho0 u_f); MfVof adds (rho1-rho0) F h/dt with
// F the alpha flux accumulated by the VoF transport.
struct MfRho0 {
  template <int E>
  static __device__ __forceinline__ double mf(const Geo& g, const VofF&, I3, double fv) {
    return g.rho0 * fv;
  }
};
struct MfVof {
  template <int E>
  static __device__ __forceinline__ double mf(const Geo& g, const VofF& M, I3 r, double fv) {
    return g.rho0 * fv + M.drho * (M.F[E][IX(g, wr(g, r))] * M.hbydt[E]);
  }
};

// Convective flux difference (M+ - M-)/h_E of component A along axis E,
// MUSCL-minmod with first-order upwind where the stencil leaves the interior.
template <class BC, class MF, int A, int E>
__device__ __forceinline__ double conv_term(const Geo& g, const Vel& V, const VofF& M, I3 q, double v0, double vp1,
                                            double vm1) {
  double mp, mm, up_p, up_m;
  if (E == A) {
    const I3 qp = sh(q, A, +1), qm = sh(q, A, -1);
    mp = 0.5 * (MF::template mf<A>(g, M, q, BC::face(g, V, A, q)) + MF::template mf<A>(g, M, qp, BC::face(g, V, A, qp)));
    mm = 0.5 * (MF::template mf<A>(g, M, qm, BC::face(g, V, A, qm)) + MF::template mf<A>(g, M, q, BC::face(g, V, A, q)));
    I3 up = mp > 0. ? q : sh(q, A, +1);
    if (!BC::interior(g, A, up)) {
      up_p = mp > 0. ? v0 : vp1;
    } else {
      double dpv = BC::face(g, V, A, sh(up, A, +1)) - BC::face(g, V, A, up);
      double dmv = BC::face(g, V, A, up) - BC::face(g, V, A, sh(up, A, -1));
      double s = minmod(dpv, dmv);
      int dir = sgn(mp);
      double uw = mp > 0. ? v0 : vp1;
      up_p = uw + dir * 0.5 * s;
    }
    up = mm > 0. ? sh(q, A, -1) : q;
    if (!BC::interior(g, A, up)) {
      up_m = mm > 0. ? vm1 : v0;
    } else {
      double dpv = BC::face(g, V, A, sh(up, A, +1)) - BC::face(g, V, A, up);
      double dmv = BC::face(g, V, A, up) - BC::face(g, V, A, sh(up, A, -1));
      double s = minmod(dpv, dmv);
      int dir = sgn(mm);
      double uw = mm > 0. ? vm1 : v0;
      up_m = uw + dir * 0.5 * s;
    }
  } else {
    double vcm = BC::dir(g, V, A, E, q, -1);
    double vcp = BC::dir(g, V, A, E, q, +1);
    const I3 r1 = sh(sh(q, E, +1), A, -1), r2 = sh(q, E, +1), r3 = sh(q, A, -1);
    mp = 0.5 * (MF::template mf<E>(g, M, r1, BC::face(g, V, E, r1)) + MF::template mf<E>(g, M, r2, BC::face(g, V, E, r2)));
    mm = 0.5 * (MF::template mf<E>(g, M, q, BC::face(g, V, E, q)) + MF::template mf<E>(g, M, r3, BC::face(g, V, E, r3)));
    I3 up = mp > 0. ? q : sh(q, E, +1);
    I3 upw = wr(g, up);
    if (!BC::interior(g, A, upw)) {
      up_p = mp > 0. ? v0 : vcp;
    } else {
      double vu = V.c[A][IX(g, upw)];
      double dpv = BC::dir(g, V, A, E, upw, +1) - vu;
      double dmv = vu - BC::dir(g, V, A, E, upw, -1);
      double s = minmod(dpv, dmv);
      int dir = sgn(mp);
      double uw = mp > 0. ? v0 : vcp;
      up_p = uw + dir * 0.5 * s;
    }
    up = mm > 0. ? sh(q, E, -1) : q;
    upw = wr(g, up);
    if (!BC::interior(g, A, upw)) {
      up_m = mm > 0. ? vcm : v0;
    } else {
      double vu = V.c[A][IX(g, upw)];
      double dpv = BC::dir(g, V, A, E, upw, +1) - vu;
      double dmv = vu - BC::dir(g, V, A, E, upw, -1);
      double s = minmod(dpv, dmv);
      int dir = sgn(mm);
      double uw = mm > 0. ? vcm : v0;
      up_m = uw + dir * 0.5 * s;
    }
  }
  double Mp = mp * up_p;
  double Mm = mm * up_m;
  return (Mp - Mm) * g.inv_d[E];
}

// One component of k_get_vof_vstar_rhouu_consistent_{x,y,z}
// (CFD/G_SMACSolver.cu:139-392, 441-662, 763-1003). Mass flux from the
// policy MF; mu, face mu, beta and rho_old from the property policy PR. Terms
// are accumulated in the reference's order: own-direction stress, cross
// stresses in increasing axis order, convection along x, y, z, old pressure
// gradient.
// The bracket tmp of the predictor: stresses, convection and the old
// pressure gradient of face q of axis A
template <class BC, class PR, class MF, int A>
__device__ __forceinline__ double predict_tmp(const Geo& g, const Props& P, const Vel& V, const VofF& M,
                                              const double* __restrict__ p, I3 q) {
  constexpr int B1 = A == 0 ? 1 : 0;
  constexpr int B2 = A == 2 ? 1 : 2;
  double v0 = V.c[A][IX(g, q)];
  double vp1 = BC::face(g, V, A, sh(q, A, +1));
  double vm1 = BC::face(g, V, A, sh(q, A, -1));
  I3 cm = wr(g, sh(q, A, -1));

  double tau_p = PR::mu(g, P, q) * (vp1 - v0);
  double tau_m = PR::mu(g, P, cm) * (v0 - vm1);
  double tmp = 2. * (tau_p - tau_m) * g.inv_d2[A];
  tmp += visc_cross<BC, PR, A, B1>(g, P, V, q, v0);
  tmp += visc_cross<BC, PR, A, B2>(g, P, V, q, v0);
  tmp -= conv_term<BC, MF, A, 0>(g, V, M, q, v0, vp1, vm1);
  tmp -= conv_term<BC, MF, A, 1>(g, V, M, q, v0, vp1, vm1);
  tmp -= conv_term<BC, MF, A, 2>(g, V, M, q, v0, vp1, vm1);
  tmp -= (p[IX(g, q)] - p[IX(g, cm)]) * g.inv_d[A];
  return tmp;
}

template <class BC, class PR, class MF, int A>
__device__ double predict_face(const Geo& g, const Props& P, const Vel& V, const VofF& M, const double* __restrict__ p,
                               I3 q, double dt) {
  const double tmp = predict_tmp<BC, PR, MF, A>(g, P, V, M, p, q);
  const double v0 = V.c[A][IX(g, q)];
  const I3 cm = wr(g, sh(q, A, -1));
  double beta = PR::beta(g, P, A, q);
  double rho_old_f = PR::rho_old_f(g, P, q, cm);
  return beta * (v0 * rho_old_f) + dt * (beta * tmp + g.g[A]);
}

}  // namespace smac
