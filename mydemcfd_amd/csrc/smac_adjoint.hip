// smac_adjoint.hip -- adjoint (vector-Jacobian product) of one SMAC step.
//
// Forward step, from (u^n, p^n):
//   u*      = F(u^n) - dt beta G p^n           predictor (k_predict; F holds
//                                              the time, stress and MUSCL terms)
//   b       = -(1/dt) D u*                     RHS
//   dp      = A^+ (b - mean b)                 pressure solve
//   u^{n+1} = u* - dt beta G dp  (interior faces; wall faces keep u^n)
//   p^{n+1} = p^n + dp
// With cotangents (ub', pb') of (u^{n+1}, p^{n+1}) the cotangents of (u^n, p^n):
//   dpb  = pb' + dt D_i(beta ub')              D_i: divergence over interior faces
//   bb   = A^+ (dpb - mean dpb)                second (self-adjoint) Poisson solve
//   ub*  = ub' + (1/dt) G bb      on interior faces
//   pb   = pb' + dt D_i(beta ub*)
//   ub   = (dF/du)^T ub*  (+ ub' on wall faces)
// (dF/du)^T is formed face by face: each interior face re-evaluates its
// predictor stencil with tracked loads (value, source array, index, factor)
// and scatters the partial derivatives of the branch the forward took
// (upwind side, minmod slope) with FP64 atomics into ub. Sums over the
// scatter are therefore order-dependent in the last bits; every other stage is
// a deterministic gather.
#include "smac_internal.hpp"
#include "smac_predict.hpp"

namespace smac {

namespace {

constexpr dim3 CELLB_A(64, 4);

// A loaded value with its linear dependence on one input: v = f * u_c[i] + const
// (c < 0: constant).
struct TV {
  double v;
  int c;
  size_t i;
  double f;
};
__device__ __forceinline__ TV tconst(double v) { return TV{v, -1, 0, 0.0}; }
__device__ __forceinline__ TV tmem(const Vel& V, int c, size_t i) { return TV{V.c[c][i], c, i, 1.0}; }

// tracked forms of the boundary getters (same values as get_face / get_dir);
// Fwd is the forward policy with the same getters
struct WallT {
  using Fwd = WallBC;
  static __device__ __forceinline__ TV face(const Geo& g, const Vel& V, int a, I3 q) {
    q = wr(g, q);
    int t = face_type(g, a, q);
    if (t == F_INTERIOR) return tmem(V, a, IX(g, q));
    if (t == F_BOUNDARY) {
      int s = face_slot(g, a, q);
      int bt = btype_of(g, s);
      if (bt == BC_INFLOW) return tconst((1. - 0.0) * bv_of(g, s, a));
      if (bt == BC_NOSLIP) return tconst(0.);
      if (bt == BC_OUTLET) return tmem(V, a, IX(g, q));
      if (bt == BC_SLIP && a != 0) return tmem(V, a, IX(g, q));
    }
    return tconst(0.);
  }
  static __device__ __forceinline__ TV dir(const Geo& g, const Vel& V, int a, int b, I3 q, int s) {
    q = wr(g, q);
    const size_t iq = IX(g, q);
    double vin = V.c[a][iq];
    I3 q2 = wr(g, sh(q, b, s));
    if (face_type(g, a, q2) == F_INTERIOR) return tmem(V, a, IX(g, q2));
    I3 qf = q;
    qf.v[b] = s > 0 ? q.v[b] + 1 : q.v[b];
    double vw = 0.0;
    int count = 0;
    if (face_type(g, b, qf) == F_BOUNDARY) {
      vw += bv_of(g, face_slot(g, b, qf), a);
      count++;
    }
    I3 qf2 = wr(g, sh(qf, a, -1));
    if (face_type(g, b, qf2) == F_BOUNDARY) {
      vw += bv_of(g, face_slot(g, b, qf2), a);
      count++;
    }
    if (count > 0) vw /= (double)count;
    int bt = btype_of(g, face_slot(g, b, qf));
    if (bt == BC_SLIP || bt == BC_OUTLET) return TV{vin, a, iq, 1.0};
    return TV{2.0 * vw - vin, a, iq, -1.0};
  }
  static __device__ __forceinline__ bool interior(const Geo& g, int a, I3 q) {
    return face_type(g, a, wr(g, q)) == F_INTERIOR;
  }
};
struct PeriodicT {
  using Fwd = PeriodicBC;
  static __device__ __forceinline__ I3 w3(const Geo& g, I3 q) {
    q.v[0] = wrapi(q.v[0], g.n[0]);
    q.v[1] = wrapi(q.v[1], g.n[1]);
    q.v[2] = wrapi(q.v[2], g.n[2]);
    return q;
  }
  static __device__ __forceinline__ TV face(const Geo& g, const Vel& V, int a, I3 q) {
    return tmem(V, a, IX(g, w3(g, q)));
  }
  static __device__ __forceinline__ TV dir(const Geo& g, const Vel& V, int a, int b, I3 q, int s) {
    return tmem(V, a, IX(g, w3(g, sh(q, b, s))));
  }
  static __device__ __forceinline__ bool interior(const Geo&, int, I3) { return true; }
};

struct Acc {
  double* ub[3];
};
// Cotangent sinks of the VoF adjoint (the step after a transport): Fb = the
// accumulated alpha flux of each axis (faces), anp1 = alpha^{n+1} (the
// properties rho, mu, beta of the step), an = alpha^n (rho_old); the
// property derivatives d rho / d alpha = drho, d mu / d alpha = dmu.
struct AccV {
  double* Fb[3];
  double* anp1;
  double* an;
  double drho, dmu;
};
__device__ __forceinline__ void aadd(double* base, size_t i, double v) {
  if (v != 0.0) unsafeAtomicAdd(base + i, v);
}
__device__ __forceinline__ void scat(const Acc& U, const TV& t, double c) {
  if (t.c >= 0 && c != 0.0) {
    // select chain: a dynamic index into the by-value argument would force a scratch copy
    double* base = t.c == 0 ? U.ub[0] : (t.c == 1 ? U.ub[1] : U.ub[2]);
    unsafeAtomicAdd(base + t.i, t.f * c);
  }
}
// mass-flux face value of conv_term (MfRho0 or MfVof, the same expression)
template <bool VOF, int E>
__device__ __forceinline__ double mflux(const Geo& g, const VofF& M, I3 r, const TV& f) {
  if constexpr (VOF) return g.rho0 * f.v + M.drho * (M.F[E][IX(g, wr(g, r))] * M.hbydt[E]);
  return g.rho0 * f.v;
}
// cotangent c of a mass-flux face value: rho0 c to the velocity, (for VOF)
// drho h/dt c to the accumulated alpha flux at r
template <bool VOF, int E>
__device__ __forceinline__ void mflux_bar(const Geo& g, const VofF& M, const AccV& AV, const Acc& U, I3 r,
                                          const TV& f, double c) {
  scat(U, f, g.rho0 * c);
  if constexpr (VOF) aadd(AV.Fb[E], IX(g, wr(g, r)), M.drho * M.hbydt[E] * c);
}

// branch of minmod(dpv, dmv): 0 -> 0, 1 -> dpv, 2 -> dmv (d_minmod,
// CFD/G_SMACSolver.cu:67-88)
__device__ __forceinline__ int minmod_sel(double dp, double dm) {
  if (dp * dm > 0.) return fabs(dp) < fabs(dm) ? 1 : 2;
  return 0;
}

// adjoint of one side of the convective flux: M = m * (uw + sgn(m) 0.5 s),
// m = 0.5 (mf(f0) + mf(f1)) with the mass-flux faces at r0, r1; cM = cotangent of M
template <bool VOF, int E>
__device__ __forceinline__ void flux_vjp(const Geo& g, const VofF& M, const AccV& AV, const Acc& U, I3 r0,
                                         const TV& f0, I3 r1, const TV& f1, const TV& uw, bool slope, const TV& sa,
                                         const TV& sb, const TV& sc, double cM) {
  const double m = 0.5 * (mflux<VOF, E>(g, M, r0, f0) + mflux<VOF, E>(g, M, r1, f1));
  double s = 0.0;
  int sel = 0;
  if (slope) {
    const double dpv = sa.v - sb.v, dmv = sb.v - sc.v;
    sel = minmod_sel(dpv, dmv);
    s = sel == 1 ? dpv : (sel == 2 ? dmv : 0.0);
  }
  const int dir = sgn(m);
  const double up = uw.v + dir * 0.5 * s;
  // dM/dm = up; dm/dmf0 = dm/dmf1 = 0.5
  mflux_bar<VOF, E>(g, M, AV, U, r0, f0, 0.5 * up * cM);
  mflux_bar<VOF, E>(g, M, AV, U, r1, f1, 0.5 * up * cM);
  // dM/duw = m; dM/ds = m sgn(m) 0.5
  scat(U, uw, m * cM);
  const double cs = m * dir * 0.5 * cM;
  if (sel == 1) {  // s = sa - sb
    scat(U, sa, cs);
    scat(U, sb, -cs);
  } else if (sel == 2) {  // s = sb - sc
    scat(U, sb, cs);
    scat(U, sc, -cs);
  }
}

// adjoint of conv_term<BC, MF, A, E> scaled by cot (cotangent of the term)
template <class BC, bool VOF, int A, int E>
__device__ __forceinline__ void conv_vjp(const Geo& g, const Vel& V, const VofF& M, const AccV& AV, const Acc& U, I3 q,
                                         const TV& v0, const TV& vp1, const TV& vm1, double cot) {
  const double cE = cot * g.inv_d[E];  // d(term)/dMp = inv_d, d/dMm = -inv_d
  if (E == A) {
    const I3 qp = sh(q, A, +1), qm = sh(q, A, -1);
    const TV fa0 = BC::face(g, V, A, q), fap = BC::face(g, V, A, qp), fam = BC::face(g, V, A, qm);
    const double mp = 0.5 * (mflux<VOF, A>(g, M, q, fa0) + mflux<VOF, A>(g, M, qp, fap));
    const double mm = 0.5 * (mflux<VOF, A>(g, M, qm, fam) + mflux<VOF, A>(g, M, q, fa0));
    {
      const I3 up = mp > 0. ? q : qp;
      const bool in = BC::interior(g, A, up);
      const TV uw = mp > 0. ? v0 : vp1;
      TV sa = tconst(0.), sb = tconst(0.), sc = tconst(0.);
      if (in) {
        sa = BC::face(g, V, A, sh(up, A, +1));
        sb = BC::face(g, V, A, up);
        sc = BC::face(g, V, A, sh(up, A, -1));
      }
      flux_vjp<VOF, A>(g, M, AV, U, q, fa0, qp, fap, uw, in, sa, sb, sc, cE);
    }
    {
      const I3 up = mm > 0. ? qm : q;
      const bool in = BC::interior(g, A, up);
      const TV uw = mm > 0. ? vm1 : v0;
      TV sa = tconst(0.), sb = tconst(0.), sc = tconst(0.);
      if (in) {
        sa = BC::face(g, V, A, sh(up, A, +1));
        sb = BC::face(g, V, A, up);
        sc = BC::face(g, V, A, sh(up, A, -1));
      }
      flux_vjp<VOF, A>(g, M, AV, U, qm, fam, q, fa0, uw, in, sa, sb, sc, -cE);
    }
  } else {
    const TV vcm = BC::dir(g, V, A, E, q, -1), vcp = BC::dir(g, V, A, E, q, +1);
    const I3 r1 = sh(sh(q, E, +1), A, -1), r2 = sh(q, E, +1), r3 = sh(q, A, -1);
    const TV e1 = BC::face(g, V, E, r1), e2 = BC::face(g, V, E, r2);
    const TV e3 = BC::face(g, V, E, q), e4 = BC::face(g, V, E, r3);
    const double mp = 0.5 * (mflux<VOF, E>(g, M, r1, e1) + mflux<VOF, E>(g, M, r2, e2));
    const double mm = 0.5 * (mflux<VOF, E>(g, M, q, e3) + mflux<VOF, E>(g, M, r3, e4));
    {
      const I3 upw = wr(g, mp > 0. ? q : sh(q, E, +1));
      const bool in = BC::interior(g, A, upw);
      const TV uw = mp > 0. ? v0 : vcp;
      TV sa = tconst(0.), sb = tconst(0.), sc = tconst(0.);
      if (in) {
        sb = tmem(V, A, IX(g, upw));
        sa = BC::dir(g, V, A, E, upw, +1);
        sc = BC::dir(g, V, A, E, upw, -1);
      }
      flux_vjp<VOF, E>(g, M, AV, U, r1, e1, r2, e2, uw, in, sa, sb, sc, cE);
    }
    {
      const I3 upw = wr(g, mm > 0. ? sh(q, E, -1) : q);
      const bool in = BC::interior(g, A, upw);
      const TV uw = mm > 0. ? vcm : v0;
      TV sa = tconst(0.), sb = tconst(0.), sc = tconst(0.);
      if (in) {
        sb = tmem(V, A, IX(g, upw));
        sa = BC::dir(g, V, A, E, upw, +1);
        sc = BC::dir(g, V, A, E, upw, -1);
      }
      flux_vjp<VOF, E>(g, M, AV, U, q, e3, r3, e4, uw, in, sa, sb, sc, -cE);
    }
  }
}

// cotangent c of the face viscosity muf[B] at f (VOF): to the cells it is
// formed from (k_props_faces: the mean of its two cells on an interior face,
// the inner cell on a boundary face), then to alpha^{n+1} with dmu
__device__ __forceinline__ void mu_face_bar(const Geo& g, const AccV& AV, int B, I3 f, double c) {
  f = wr(g, f);
  const int t = face_type(g, B, f);
  if (t == F_INTERIOR) {
    aadd(AV.anp1, IX(g, f), AV.dmu * (0.5 * c));
    aadd(AV.anp1, IX(g, wr(g, sh(f, B, -1))), AV.dmu * (0.5 * c));
  } else if (t == F_BOUNDARY) {
    aadd(AV.anp1, IX(g, f.v[B] == 1 ? f : sh(f, B, -1)), AV.dmu * c);
  }
}

// adjoint of visc_cross<BC, PR, A, B> scaled by cot (VOF: and of its edge
// viscosities)
template <class BC, class PR, bool VOF, int A, int B>
__device__ __forceinline__ void visc_vjp(const Geo& g, const Props& P, const Vel& V, const Acc& U, const AccV& AV, I3 q,
                                         const TV& v0, double cot) {
  const TV vbm = BC::dir(g, V, A, B, q, -1), vbp = BC::dir(g, V, A, B, q, +1);
  const TV wpp = BC::face(g, V, B, sh(q, B, +1)), wmp = BC::face(g, V, B, sh(sh(q, B, +1), A, -1));
  const TV wp0 = BC::face(g, V, B, q), wm0 = BC::face(g, V, B, sh(q, A, -1));
  const double mu_p = PR::mu_edge(g, P, B, wr(g, sh(q, B, +1)), wr(g, sh(sh(q, B, +1), A, -1)));
  const double mu_m = PR::mu_edge(g, P, B, wr(g, q), wr(g, sh(q, A, -1)));
  const double c = cot * g.inv_d[B];  // cotangent of t_p; -c of t_m
  scat(U, wpp, c * mu_p * g.inv_d[A]);
  scat(U, wmp, -c * mu_p * g.inv_d[A]);
  scat(U, vbp, c * mu_p * g.inv_d[B]);
  scat(U, v0, -c * mu_p * g.inv_d[B]);
  scat(U, wp0, -c * mu_m * g.inv_d[A]);
  scat(U, wm0, c * mu_m * g.inv_d[A]);
  scat(U, v0, -c * mu_m * g.inv_d[B]);
  scat(U, vbm, c * mu_m * g.inv_d[B]);
  if constexpr (VOF) {
    // t_p = mu_p X_p with mu_p = 0.5 (muf(f1) + muf(f2)); t_m likewise
    const double xp = (wpp.v - wmp.v) * g.inv_d[A] + (vbp.v - v0.v) * g.inv_d[B];
    const double xm = (wp0.v - wm0.v) * g.inv_d[A] + (v0.v - vbm.v) * g.inv_d[B];
    mu_face_bar(g, AV, B, sh(q, B, +1), 0.5 * c * xp);
    mu_face_bar(g, AV, B, sh(sh(q, B, +1), A, -1), 0.5 * c * xp);
    mu_face_bar(g, AV, B, q, -0.5 * c * xm);
    mu_face_bar(g, AV, B, sh(q, A, -1), -0.5 * c * xm);
  }
}

// (d u*_A(q) / d u^n)^T w, scattered into U (the pressure term is separate).
// VOF: the step after a VoF transport -- mass flux with the accumulated alpha
// flux (its cotangent to AV.Fb) and the property cotangents: beta = 1/rho_f
// and mu of alpha^{n+1}, rho_old of alpha^n
template <class BC, class PR, bool VOF, int A>
__device__ void predict_face_vjp(const Geo& g, const Props& P, const Vel& V, const VofF& M, const double* p,
                                 const Acc& U, const AccV& AV, I3 q, double dt, double w) {
  constexpr int B1 = A == 0 ? 1 : 0;
  constexpr int B2 = A == 2 ? 1 : 2;
  const TV v0 = tmem(V, A, IX(g, q));
  const TV vp1 = BC::face(g, V, A, sh(q, A, +1));
  const TV vm1 = BC::face(g, V, A, sh(q, A, -1));
  const I3 cm = wr(g, sh(q, A, -1));
  const double beta = PR::beta(g, P, A, q);
  const double rho_old_f = PR::rho_old_f(g, P, q, cm);
  // u* = beta (v0 rho_old_f) + dt (beta tmp + g)
  scat(U, v0, beta * rho_old_f * w);
  const double W = dt * beta * w;  // cotangent of tmp
  const double muq = PR::mu(g, P, q), mucm = PR::mu(g, P, cm);
  const double k2 = 2. * g.inv_d2[A] * W;
  scat(U, vp1, k2 * muq);
  scat(U, v0, -k2 * muq - k2 * mucm);
  scat(U, vm1, k2 * mucm);
  visc_vjp<BC, PR, VOF, A, B1>(g, P, V, U, AV, q, v0, W);
  visc_vjp<BC, PR, VOF, A, B2>(g, P, V, U, AV, q, v0, W);
  conv_vjp<BC, VOF, A, 0>(g, V, M, AV, U, q, v0, vp1, vm1, -W);
  conv_vjp<BC, VOF, A, 1>(g, V, M, AV, U, q, v0, vp1, vm1, -W);
  conv_vjp<BC, VOF, A, 2>(g, V, M, AV, U, q, v0, vp1, vm1, -W);
  if constexpr (VOF) {
    const size_t iq = IX(g, q), ic = IX(g, cm);
    // beta = 1/(0.5 (rho_q + rho_cm)): d beta / d rho_c = -0.5 beta^2
    const double tmp = predict_tmp<typename BC::Fwd, PR, MfVof, A>(g, P, V, M, p, q);
    const double bb = w * (v0.v * rho_old_f + dt * tmp);
    const double kr = AV.drho * (-0.5 * beta * beta * bb);
    aadd(AV.anp1, iq, kr);
    aadd(AV.anp1, ic, kr);
    // rho_old_f = 0.5 (rho_old_q + rho_old_cm)
    const double ko = AV.drho * (0.5 * (w * beta * v0.v));
    aadd(AV.an, iq, ko);
    aadd(AV.an, ic, ko);
    // own-axis stresses mu_q (vp1 - v0) and mu_cm (v0 - vm1)
    aadd(AV.anp1, iq, AV.dmu * (k2 * (vp1.v - v0.v)));
    aadd(AV.anp1, ic, AV.dmu * (-k2 * (v0.v - vm1.v)));
  }
}

template <class BC, class PR, bool VOF, int A>
__global__ __launch_bounds__(256) void k_predict_vjp(Geo g, Props P, Vel V, VofF M, const double* __restrict__ p,
                                                     const double* __restrict__ wbar, Acc U, AccV AV, double dt) {
  int lo[3] = {1, 1, 1}, hi[3] = {g.n[0], g.n[1], g.n[2]};
  lo[A] = g.per[A] ? 1 : 2;
  int i = lo[0] + (int)(blockIdx.x * 64 + threadIdx.x);
  int j = lo[1] + (int)(blockIdx.y * 4 + threadIdx.y);
  int k = lo[2] + (int)blockIdx.z;
  if (i > hi[0] || j > hi[1] || k > hi[2]) return;
  I3 q = mk(i, j, k);
  const double w = wbar[IX(g, q)];
  if (w == 0.0) return;
  predict_face_vjp<BC, PR, VOF, A>(g, P, V, M, p, U, AV, q, dt, w);
}

// out_c = pin_c + s * sum over the interior faces f of cell c of +-beta_f ub_f / h
// (s dt D_i(beta ub)); cells natural, beta == nullptr: beta0
__global__ __launch_bounds__(256) void k_adj_div(Geo g, BetaF bf, const double* __restrict__ ub0,
                                                 const double* __restrict__ ub1, const double* __restrict__ ub2,
                                                 const double* __restrict__ pin, double* __restrict__ out, double s) {
  ROW_SETUP
  const double* ub[3] = {ub0, ub1, ub2};
  for (int i = 1 + lane; i <= g.n[0]; i += 64) {
    const int q[3] = {i, j, k};
    double acc = 0.0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      int qp[3] = {i, j, k};
      qp[d] = W(g, d, q[d] + 1);
      const size_t fp = idx(g, qp[0], qp[1], qp[2]), fm = idx(g, i, j, k);
      // a face is interior unless it is a wall face of a wall direction
      const bool inp = g.per[d] || q[d] + 1 <= g.n[d];
      const bool inm = g.per[d] || q[d] >= 2;
      const double bp = bf.b[0] ? bf.b[d][fp] : g.beta0, bm = bf.b[0] ? bf.b[d][fm] : g.beta0;
      double t = 0.0;
      if (inp) t += bp * ub[d][fp];
      if (inm) t -= bm * ub[d][fm];
      acc += t * g.inv_d[d];
    }
    out[rb + i] = pin[rb + i] + s * acc;
  }
}

// ub*_f = ub'_f + (1/dt) (bb_i - bb_(i-1)) / h on interior faces of axis A;
// un_f = ub'_f on wall faces (u^{n+1} keeps u^n there), 0 elsewhere
template <int A>
__global__ __launch_bounds__(256) void k_adj_ustar(Geo g, const double* __restrict__ ubn,
                                                   const double* __restrict__ bb, double* __restrict__ ubs,
                                                   double* __restrict__ un, double inv_dt) {
  const int i = (int)(blockIdx.x * 64 + threadIdx.x);
  const int j = (int)(blockIdx.y * 4 + threadIdx.y);
  const int k = (int)blockIdx.z;
  if (i > g.n[0] + 1 || j > g.n[1] + 1 || k > g.n[2] + 1) return;
  const I3 q = mk(i, j, k);
  const size_t id = IX(g, q);
  int hi[3] = {g.n[0], g.n[1], g.n[2]};
  if (!g.per[A]) hi[A] += 1;
  bool inside = true;
#pragma unroll
  for (int d = 0; d < 3; d++)
    if (q.v[d] < 1 || q.v[d] > hi[d]) inside = false;
  if (!inside) {
    ubs[id] = 0.0;
    un[id] = 0.0;
    return;
  }
  const int fa = q.v[A];
  const bool wall = !g.per[A] && (fa == 1 || fa == g.n[A] + 1);
  if (wall) {
    ubs[id] = 0.0;
    un[id] = ubn[id];
  } else {
    const size_t cm = IX(g, wr(g, sh(q, A, -1)));
    ubs[id] = ubn[id] + inv_dt * ((bb[id] - bb[cm]) * g.inv_d[A]);
    un[id] = 0.0;
  }
}

// canonical row sums of a natural-layout cell field (part0 <- row sums)
__global__ __launch_bounds__(256) void k_rowsum_nat(Geo g, const double* __restrict__ x, double* __restrict__ part0) {
  ROW_SETUP
  double acc = 0.0;
  for (int i = 1 + lane; i <= g.n[0]; i += 64) acc += x[rb + i];
  acc = wave_tree_sum(acc);
  if (lane == 0) part0[row] = acc;
}

template <class BC, class PR, bool VOF>
void predict_vjp_components(const Geo& g, const Props& P, const Vel& V, const VofF& M, const double* p,
                            double* const wbar[3], const Acc& U, const AccV& AV, double dt, hipStream_t st) {
  for (int a = 0; a < 3; a++) {
    int lo = g.per[a] ? 1 : 2;
    int nf[3] = {g.n[0], g.n[1], g.n[2]};
    nf[a] = g.n[a] - lo + 1;
    if (nf[0] <= 0 || nf[1] <= 0 || nf[2] <= 0) continue;
    dim3 cg((unsigned)((nf[0] + 63) / 64), (unsigned)((nf[1] + 3) / 4), (unsigned)nf[2]);
    if (a == 0)
      hipLaunchKernelGGL((k_predict_vjp<BC, PR, VOF, 0>), cg, CELLB_A, 0, st, g, P, V, M, p, wbar[0], U, AV, dt);
    if (a == 1)
      hipLaunchKernelGGL((k_predict_vjp<BC, PR, VOF, 1>), cg, CELLB_A, 0, st, g, P, V, M, p, wbar[1], U, AV, dt);
    if (a == 2)
      hipLaunchKernelGGL((k_predict_vjp<BC, PR, VOF, 2>), cg, CELLB_A, 0, st, g, P, V, M, p, wbar[2], U, AV, dt);
  }
}

}  // namespace

void launch_predict_vjp(const Geo& g, const Props* P, const Vel& V, double* const wbar[3], double* ub[3], double dt,
                        hipStream_t st) {
  Acc U;
  for (int d = 0; d < 3; d++) U.ub[d] = ub[d];
  const VofF M0{};
  const AccV AV0{};
  const bool allper = g.per[0] && g.per[1] && g.per[2];
  if (P) {
    if (allper) predict_vjp_components<PeriodicT, VarP, false>(g, *P, V, M0, nullptr, wbar, U, AV0, dt, st);
    else predict_vjp_components<WallT, VarP, false>(g, *P, V, M0, nullptr, wbar, U, AV0, dt, st);
  } else {
    if (allper) predict_vjp_components<PeriodicT, UniP, false>(g, Props{}, V, M0, nullptr, wbar, U, AV0, dt, st);
    else predict_vjp_components<WallT, UniP, false>(g, Props{}, V, M0, nullptr, wbar, U, AV0, dt, st);
  }
}

void launch_predict_vjp_vof(const Geo& g, const Props& P, const Vel& V, const VofF& M, const double* p,
                            double* const wbar[3], double* ub[3], double* const Fb[3], double* anp1, double* an,
                            double drho, double dmu, double dt, hipStream_t st) {
  Acc U;
  AccV AV;
  for (int d = 0; d < 3; d++) {
    U.ub[d] = ub[d];
    AV.Fb[d] = Fb[d];
  }
  AV.anp1 = anp1;
  AV.an = an;
  AV.drho = drho;
  AV.dmu = dmu;
  const bool allper = g.per[0] && g.per[1] && g.per[2];
  if (allper) predict_vjp_components<PeriodicT, VarP, true>(g, P, V, M, p, wbar, U, AV, dt, st);
  else predict_vjp_components<WallT, VarP, true>(g, P, V, M, p, wbar, U, AV, dt, st);
}

void launch_adj_div(const Geo& g, const BetaF* bf, double* const ub[3], const double* pin, double* out, double s,
                    hipStream_t st) {
  dim3 grid((unsigned)((g.n[1] * g.n[2] + 3) / 4));
  hipLaunchKernelGGL(k_adj_div, grid, dim3(256), 0, st, g, bf ? *bf : BetaF{}, ub[0], ub[1], ub[2], pin, out, s);
}

void launch_adj_ustar(const Geo& g, double* const ubn[3], const double* bb, double* ubs[3], double* un[3],
                      double inv_dt, hipStream_t st) {
  dim3 grid((unsigned)((g.n[0] + 2 + 63) / 64), (unsigned)((g.n[1] + 2 + 3) / 4), (unsigned)(g.n[2] + 2));
  hipLaunchKernelGGL(k_adj_ustar<0>, grid, CELLB_A, 0, st, g, ubn[0], bb, ubs[0], un[0], inv_dt);
  hipLaunchKernelGGL(k_adj_ustar<1>, grid, CELLB_A, 0, st, g, ubn[1], bb, ubs[1], un[1], inv_dt);
  hipLaunchKernelGGL(k_adj_ustar<2>, grid, CELLB_A, 0, st, g, ubn[2], bb, ubs[2], un[2], inv_dt);
}

void launch_rowsum_nat(const Geo& g, const double* x, double* p0, hipStream_t st) {
  dim3 grid((unsigned)((g.n[1] * g.n[2] + 3) / 4));
  hipLaunchKernelGGL(k_rowsum_nat, grid, dim3(256), 0, st, g, x, p0);
}

}  // namespace smac
