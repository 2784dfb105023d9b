// smac_kernels.hip -- gfx950 kernels of the SMAC projection step (FP64 form).
//
//   predictor        k_predict<A>      per face of component A (MUSCL-minmod
//                                      advection + variable-mu stress + old dp/dx);
//                    k_predict_tile    triply periodic one-phase grids (z-march);
//                    k_predict_tw<A>   the two-way (CFD-DEM) form
//   star BC          k_star_bc         wall faces of u*, v*, w*
//   Poisson RHS      k_rhs, k_rhs_eps  b = -(1/dt) div u* (+ the volume-fraction form)
//   PCG vectors      k_init_r, k_add_scalar, k_jacobi, k_copy_dot, k_sum,
//                    k_spmv_dot, k_update, k_dot2, k_xpay, k_dp_flush
//   reductions       k_finalize        canonical 1024-lane tree + scalar ops
//   (multigrid kernels: smac_mg.hip)
//   correction       k_correct_all, k_zero_dp
//
// Row kernels: one wave64 per grid row (j,k); lane l owns i = 1+l, 1+l+64, ...
// Row partial sums follow the canonical order shared with the CPU oracle.
//
// Solver cell vectors (r, z, d, Ad, multigrid e/f/res on every level) use a
// row-split layout: inside row (j,k) the even-i cells come first, then the
// odd-i cells (spos()). A red-black half sweep then streams contiguous half
// rows. Fields seen by users (p, dp, rhs) keep the natural layout.
#include "smac_internal.hpp"
#include "smac_predict.hpp"

namespace smac {

// grid: x = face index along i (block 64), y = j (block 4), z = k
template <class BC, class PR, class MF, int A>
__global__ __launch_bounds__(256) void k_predict(Geo g, Props P, Vel V, VofF M, const double* __restrict__ p,
                                                 double* __restrict__ ustar, double dt) {
  int lo[3] = {1, 1, 1}, hi[3] = {g.n[0], g.n[1], g.n[2]};
  lo[A] = g.per[A] ? 1 : 2;
  int i = lo[0] + (int)(blockIdx.x * 64 + threadIdx.x);
  int j = lo[1] + (int)(blockIdx.y * 4 + threadIdx.y);
  int k = lo[2] + (int)blockIdx.z;
  if (i > hi[0] || j > hi[1] || k > hi[2]) return;
  I3 q = mk(i, j, k);
  ustar[IX(g, q)] = predict_face<BC, PR, MF, A>(g, P, V, M, p, q, dt);
}

// ------------------------------------------------------- two-way (CFD-DEM)
// Mass flux of k_compute_mass_flux_from_alpha_flux_two_way
// (CFD/G_SMACSolver_VoF.cu:2472-2557): mf = rho0 (eps_f u_f) + (rho1-rho0) F h/dt
// on the faces 1..N+1 of axis E and 0 outside them, eps_f the t^{n+1/2}
// volume fraction of the face -- the mean of its two cells, the inner cell on
// faces 1 and N+1 (d_get_void_fraction_half_*face, 1954-1967). F enters when
// a VoF transport preceded (M.F[E] != nullptr).
struct MfEps {
  template <int E>
  static __device__ __forceinline__ double mf(const Geo& g, const VofF& M, I3 r, double fv) {
    const int m = r.v[E];
    if (m < 1 || m > g.n[E] + 1) return 0.0;
    double ef;
    if (m == 1) {
      ef = M.eps_half[IX(g, r)];
    } else if (m == g.n[E] + 1) {
      ef = M.eps_half[IX(g, sh(r, E, -1))];
    } else {
      ef = 0.5 * (M.eps_half[IX(g, sh(r, E, -1))] + M.eps_half[IX(g, r)]);
    }
    const double q = ef * fv;
    if (!M.F[E]) return g.rho0 * q;
    return g.rho0 * q + M.drho * (M.F[E][IX(g, wr(g, r))] * M.hbydt[E]);
  }
};

// Cross viscous term of the two-way predictor: the stresses of visc_cross,
// each weighted by the t^n volume fraction of its edge, the mean of the four
// cells around the edge summed in memory (x-fastest) order
// (CFD/G_SMACSolver.cu:1608-1623, 1826-1844, 2044-2062).
template <class BC, class PR, int A, int B>
__device__ __forceinline__ double visc_cross_tw(const Geo& g, const Props& P, const Vel& V, const EpsF& E, I3 q,
                                                double v0) {
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
  const double* eo = E.eps_old;
  const I3 qa = sh(q, A, -1);
  double ep, em;
  if (B > A) {
    ep = 0.25 * (eo[IX(g, qa)] + eo[IX(g, q)] + eo[IX(g, sh(qa, B, +1))] + eo[IX(g, sh(q, B, +1))]);
    em = 0.25 * (eo[IX(g, sh(qa, B, -1))] + eo[IX(g, sh(q, B, -1))] + eo[IX(g, qa)] + eo[IX(g, q)]);
  } else {
    ep = 0.25 * (eo[IX(g, qa)] + eo[IX(g, sh(qa, B, +1))] + eo[IX(g, q)] + eo[IX(g, sh(q, B, +1))]);
    em = 0.25 * (eo[IX(g, sh(qa, B, -1))] + eo[IX(g, qa)] + eo[IX(g, sh(q, B, -1))] + eo[IX(g, q)]);
  }
  return (ep * t_p - em * t_m) * g.inv_d[B];
}

// One component of k_get_vof_vstar_rhouu_consistent_{x,y,z}_two_way
// (CFD/G_SMACSolver.cu:1538-2194): the one-fluid predictor of predict_face
// with the stresses weighted by the t^n volume fraction (cell values for the
// own-direction stress, edge means for the cross stresses), the mass flux of
// MfEps, the old pressure gradient weighted by the face mean of eps^n, and
//   u* = (beta/eps^{n+1}_f) (eps^n_f rho^n_f u + dt tmp + J/V) + dt g
// with J the particles' momentum impulse on the face. The IBM solid fraction
// factor (1 - 0) of the reference is 1.
template <class BC, class PR, class MF, int A>
__device__ double predict_face_tw(const Geo& g, const Props& P, const Vel& V, const VofF& M, const EpsF& E,
                                  const double* __restrict__ p, I3 q, double dt) {
  constexpr int B1 = A == 0 ? 1 : 0;
  constexpr int B2 = A == 2 ? 1 : 2;
  double v0 = V.c[A][IX(g, q)];
  double vp1 = BC::face(g, V, A, sh(q, A, +1));
  double vm1 = BC::face(g, V, A, sh(q, A, -1));
  const I3 cr = sh(q, A, -1);  // the minus cell at its raw index (volume fractions)
  const I3 cm = wr(g, cr);
  const double* eo = E.eps_old;

  double tau_p = PR::mu(g, P, q) * (vp1 - v0);
  double tau_m = PR::mu(g, P, cm) * (v0 - vm1);
  double tmp = 2. * (eo[IX(g, q)] * tau_p - eo[IX(g, cr)] * tau_m) * g.inv_d2[A];
  tmp += visc_cross_tw<BC, PR, A, B1>(g, P, V, E, q, v0);
  tmp += visc_cross_tw<BC, PR, A, B2>(g, P, V, E, q, v0);
  tmp -= conv_term<BC, MF, A, 0>(g, V, M, q, v0, vp1, vm1);
  tmp -= conv_term<BC, MF, A, 1>(g, V, M, q, v0, vp1, vm1);
  tmp -= conv_term<BC, MF, A, 2>(g, V, M, q, v0, vp1, vm1);

  const double eo_f = 0.5 * (eo[IX(g, cr)] + eo[IX(g, q)]);
  const double en_f = 0.5 * (E.eps[IX(g, cr)] + E.eps[IX(g, q)]);
  tmp -= eo_f * (p[IX(g, q)] - p[IX(g, cm)]) * g.inv_d[A];
  const double beta = PR::beta(g, P, A, q);
  const double rho_old_f = PR::rho_old_f(g, P, q, cm);
  const double jd = E.imp[A] ? E.imp[A][IX(g, q)] * E.inv_vol : 0.0;
  const double inv_eps_rho = beta / en_f;
  return inv_eps_rho * (eo_f * rho_old_f * v0 + dt * tmp + jd) + dt * g.g[A];
}

template <class BC, class PR, int A>
__global__ __launch_bounds__(256) void k_predict_tw(Geo g, Props P, Vel V, VofF M, EpsF E,
                                                    const double* __restrict__ p, double* __restrict__ ustar,
                                                    double dt) {
  int lo[3] = {1, 1, 1}, hi[3] = {g.n[0], g.n[1], g.n[2]};
  lo[A] = g.per[A] ? 1 : 2;
  int i = lo[0] + (int)(blockIdx.x * 64 + threadIdx.x);
  int j = lo[1] + (int)(blockIdx.y * 4 + threadIdx.y);
  int k = lo[2] + (int)blockIdx.z;
  if (i > hi[0] || j > hi[1] || k > hi[2]) return;
  I3 q = mk(i, j, k);
  ustar[IX(g, q)] = predict_face_tw<BC, PR, MfEps, A>(g, P, V, M, E, p, q, dt);
}

// Poisson face coefficients of a two-way grid (k_update_poisson_beta_two_way,
// Coupling/G_CFDDEMCoupling.cu:419-455): on the faces 1..N+1 of axis A with
// the other indices in 1..N, bp = eps_f * beta_f, eps_f the mean of the two
// cells' eps^{n+1}; other faces keep beta_f. beta_f = the face 1/rho_f array,
// or (one phase) beta0 on interior faces and 0 on boundary and ghost faces.
template <int A>
__global__ __launch_bounds__(256) void k_pois_two_way(Geo g, const double* __restrict__ beta,
                                                      const double* __restrict__ eps, double* __restrict__ bp) {
  const int i = (int)(blockIdx.x * 64 + threadIdx.x);
  const int j = (int)(blockIdx.y * 4 + threadIdx.y);
  const int k = (int)blockIdx.z;
  if (i > g.n[0] + 2 || j > g.n[1] + 2) return;
  const I3 q = mk(i, j, k);
  const size_t id = IX(g, q);
  const double b = beta ? beta[id] : (face_type(g, A, q) == F_INTERIOR ? g.beta0 : 0.0);
  bool in = q.v[A] >= 1 && q.v[A] <= g.n[A] + 1;
#pragma unroll
  for (int d = 0; d < 3; d++)
    if (d != A) in = in && q.v[d] >= 1 && q.v[d] <= g.n[d];
  if (in) {
    const double epsf = 0.5 * (eps[IX(g, sh(q, A, -1))] + eps[id]);
    bp[id] = epsf * b;
  } else {
    bp[id] = b;
  }
}

// Periodic one-phase form of predict_face<PeriodicBC, UniP, A> with the whole
// stencil loaded before any branch: the own component at q + s e_d for
// s = -2..2 along each axis, the two cross components on the four faces around
// the edge, p on both sides. The upwind side and the minmod slope are then
// register selects, so no load waits on another load's value. The arithmetic
// is that of predict_face, term for term, so the results are bit-identical.
struct PerIdx {
  int x[5], y[5], z[5];  // wrapped coordinates i-2..i+2, j-2..j+2, k-2..k+2
  int px, py;
  __device__ __forceinline__ size_t at(int di, int dj, int dk) const {
    return ((size_t)z[dk + 2] * (size_t)py + (size_t)y[dj + 2]) * (size_t)px + (size_t)x[di + 2];
  }
};

template <int A>
__device__ __forceinline__ size_t per_at(const PerIdx& X, int d, int s) {
  return d == 0 ? X.at(s, 0, 0) : (d == 1 ? X.at(0, s, 0) : X.at(0, 0, s));
}
template <int A, int B>
__device__ __forceinline__ size_t per_at2(const PerIdx& X, int sb, int sa) {
  // offset sb along B and sa along A
  int o[3] = {0, 0, 0};
  o[B] += sb;
  o[A] += sa;
  return X.at(o[0], o[1], o[2]);
}

// one side of the convective flux along E, MUSCL-minmod on the five values
// c[-2..2] = u_A at q + s e_E (upwind side from the sign of m)
__device__ __forceinline__ double flux_plus(double m, const double c[5]) {
  // + face: up = q (c0) when m > 0, else q + e (c1)
  double dpv, dmv, uw;
  if (m > 0.) {
    dpv = c[3] - c[2];
    dmv = c[2] - c[1];
    uw = c[2];
  } else {
    dpv = c[4] - c[3];
    dmv = c[3] - c[2];
    uw = c[3];
  }
  double sl = minmod(dpv, dmv);
  int dir = sgn(m);
  return m * (uw + dir * 0.5 * sl);
}
__device__ __forceinline__ double flux_minus(double m, const double c[5]) {
  // - face: up = q - e (c-1) when m > 0, else q (c0)
  double dpv, dmv, uw;
  if (m > 0.) {
    dpv = c[2] - c[1];
    dmv = c[1] - c[0];
    uw = c[1];
  } else {
    dpv = c[3] - c[2];
    dmv = c[2] - c[1];
    uw = c[2];
  }
  double sl = minmod(dpv, dmv);
  int dir = sgn(m);
  return m * (uw + dir * 0.5 * sl);
}

// ld(c, di, dj, dk) = component c (0..2 = u, v, w; 3 = p) at (i+di, j+dj, k+dk)
// of the face's cell, periodic images included. The offsets are compile-time
// constants after unrolling, so every loader resolves to a register, an LDS
// or a global read.
template <int A, class LD>
__device__ __forceinline__ double predict_face_gen(const Geo& g, const LD& ld, double dt) {
  constexpr int B1 = A == 0 ? 1 : 0;
  constexpr int B2 = A == 2 ? 1 : 2;
  // offset s along axis d
  auto along = [&](int c, int d, int s) { return ld(c, d == 0 ? s : 0, d == 1 ? s : 0, d == 2 ? s : 0); };
  // offset sb along B and sa along A
  auto along2 = [&](int c, int B, int sb, int sa) {
    int o[3] = {0, 0, 0};
    o[B] += sb;
    o[A] += sa;
    return ld(c, o[0], o[1], o[2]);
  };
  double ca[5], c1[5], c2[5];
#pragma unroll
  for (int s = -2; s <= 2; s++) {
    ca[s + 2] = along(A, A, s);
    c1[s + 2] = s == 0 ? 0.0 : along(A, B1, s);
    c2[s + 2] = s == 0 ? 0.0 : along(A, B2, s);
  }
  const double v0 = ca[2];
  c1[2] = v0;
  c2[2] = v0;
  // cross components on the faces around the edges: (q+eB), (q+eB-eA), q, (q-eA)
  const double b1pp = along2(B1, B1, 1, 0), b1mp = along2(B1, B1, 1, -1);
  const double b1p0 = along2(B1, B1, 0, 0), b1m0 = along2(B1, B1, 0, -1);
  const double b2pp = along2(B2, B2, 1, 0), b2mp = along2(B2, B2, 1, -1);
  const double b2p0 = along2(B2, B2, 0, 0), b2m0 = along2(B2, B2, 0, -1);
  const double pq = ld(3, 0, 0, 0), pm = along(3, A, -1);

  const double rho0 = g.rho0, mu0 = g.mu0;
  const double vp1 = ca[3], vm1 = ca[1];
  double tau_p = mu0 * (vp1 - v0);
  double tau_m = mu0 * (v0 - vm1);
  double tmp = 2. * (tau_p - tau_m) * g.inv_d2[A];
  {  // visc_cross B1
    const double mu_p = 0.5 * (mu0 + mu0);
    double t_p = mu_p * ((b1pp - b1mp) * g.inv_d[A] + (c1[3] - v0) * g.inv_d[B1]);
    const double mu_m = 0.5 * (mu0 + mu0);
    double t_m = mu_m * ((b1p0 - b1m0) * g.inv_d[A] + (v0 - c1[1]) * g.inv_d[B1]);
    tmp += (t_p - t_m) * g.inv_d[B1];
  }
  {  // visc_cross B2
    const double mu_p = 0.5 * (mu0 + mu0);
    double t_p = mu_p * ((b2pp - b2mp) * g.inv_d[A] + (c2[3] - v0) * g.inv_d[B2]);
    const double mu_m = 0.5 * (mu0 + mu0);
    double t_m = mu_m * ((b2p0 - b2m0) * g.inv_d[A] + (v0 - c2[1]) * g.inv_d[B2]);
    tmp += (t_p - t_m) * g.inv_d[B2];
  }
  // convection along x, y, z
#pragma unroll
  for (int E = 0; E < 3; E++) {
    double mp, mm;
    const double* c;
    if (E == A) {
      mp = 0.5 * (rho0 * ca[2] + rho0 * ca[3]);
      mm = 0.5 * (rho0 * ca[1] + rho0 * ca[2]);
      c = ca;
    } else if (E == B1) {
      mp = 0.5 * (rho0 * b1mp + rho0 * b1pp);
      mm = 0.5 * (rho0 * b1p0 + rho0 * b1m0);
      c = c1;
    } else {
      mp = 0.5 * (rho0 * b2mp + rho0 * b2pp);
      mm = 0.5 * (rho0 * b2p0 + rho0 * b2m0);
      c = c2;
    }
    double Mp = flux_plus(mp, c);
    double Mm = flux_minus(mm, c);
    tmp -= (Mp - Mm) * g.inv_d[E];
  }
  const double beta = g.beta0;
  const double rho_old_f = 0.5 * (rho0 + rho0);
  tmp -= (pq - pm) * g.inv_d[A];
  return beta * (v0 * rho_old_f) + dt * (beta * tmp + g.g[A]);
}

// global loads over the wrapped neighbourhood X
template <int A>
__device__ __forceinline__ double predict_face_pu(const Geo& g, const Vel& V, const double* __restrict__ p,
                                                  const PerIdx& X, double dt) {
  auto ld = [&](int c, int di, int dj, int dk) {
    const size_t o = X.at(di, dj, dk);
    return c == 3 ? p[o] : V.c[c][o];
  };
  return predict_face_gen<A>(g, ld, dt);
}

constexpr int PT_X = 64, PT_Y = 4, PT_W = PT_X + 4, PT_PL = (PT_X + 4) * (PT_Y + 4);

// The + side of the face of component A along direction E (periodic, one
// phase): the convective flux through it (MUSCL-minmod, mass flux from the E
// component on the edge) and the viscous stress there (the normal stress when
// E == A, the cross stress otherwise). The - side of a face is the + side of
// the face one cell down along E with the very same operands (only the two
// terms of the mass-flux sum swap places), so face_assemble applied to + sides
// computed here and - sides taken from the neighbours gives predict_face_gen's
// result bit for bit.
template <int A, int E, class LD>
__device__ __forceinline__ void plus_side(const Geo& g, const LD& ld, double& M, double& T) {
  const double rho0 = g.rho0, mu0 = g.mu0;
  auto at = [&](int c, int sE, int sA) {
    int o[3] = {0, 0, 0};
    o[E] += sE;
    o[A] += sA;
    return ld(c, o[0], o[1], o[2]);
  };
  double c[5];
  c[0] = 0.0;  // flux_plus reads c[1..4]
#pragma unroll
  for (int s = -1; s <= 2; s++) c[s + 2] = at(A, s, 0);
  double mp;
  if constexpr (E == A) {
    mp = 0.5 * (rho0 * c[2] + rho0 * c[3]);
    T = mu0 * (c[3] - c[2]);
  } else {
    // the E component on the edge: at q + e_E and q + e_E - e_A
    const double bpp = at(E, 1, 0), bmp = at(E, 1, -1);
    mp = 0.5 * (rho0 * bmp + rho0 * bpp);
    const double mu_p = 0.5 * (mu0 + mu0);
    T = mu_p * ((bpp - bmp) * g.inv_d[A] + (c[3] - c[2]) * g.inv_d[E]);
  }
  M = flux_plus(mp, c);
}

// predict_face_gen's sum from the + and - sides (index E) of the face of
// component A, its own old value v0 and the pressure on both sides.
template <int A>
__device__ __forceinline__ double face_assemble(const Geo& g, const double Mp[3], const double Tp[3],
                                                const double Mm[3], const double Tm[3], double v0, double pq,
                                                double pm, double dt) {
  constexpr int B1 = A == 0 ? 1 : 0;
  constexpr int B2 = A == 2 ? 1 : 2;
  const double rho0 = g.rho0;
  double tmp = 2. * (Tp[A] - Tm[A]) * g.inv_d2[A];
  tmp += (Tp[B1] - Tm[B1]) * g.inv_d[B1];
  tmp += (Tp[B2] - Tm[B2]) * g.inv_d[B2];
#pragma unroll
  for (int E = 0; E < 3; E++) tmp -= (Mp[E] - Mm[E]) * g.inv_d[E];
  const double beta = g.beta0;
  const double rho_old_f = 0.5 * (rho0 + rho0);
  tmp -= (pq - pm) * g.inv_d[A];
  return beta * (v0 * rho_old_f) + dt * (beta * tmp + g.g[A]);
}

template <int A, class LD>
__device__ __forceinline__ void plus_sides(const Geo& g, const LD& ld, double M[3], double T[3]) {
  plus_side<A, 0>(g, ld, M[0], T[0]);
  plus_side<A, 1>(g, ld, M[1], T[1]);
  plus_side<A, 2>(g, ld, M[2], T[2]);
}

// The fused periodic predictor as a z-march over a 64 x 4 column tile: the
// block keeps planes k-1 and k of u and v, k and k+1 of w and k of p (each
// with a two-cell x/y halo) in LDS, and every thread its own column's u, v, w
// at k-2..k+2 and p at k-1..k in registers. Every value a face needs is in
// one of those (the cross components reach one plane up for w and one plane
// down for u and v; all x/y offsets stay in plane k but those of w at k+1
// and u, v at k-1, which are within one cell). A plane step writes the next
// planes (loaded during the previous step) into the ring slots the step no
// longer reads, then issues the loads for the following step.
//
// Each face side is evaluated once (plus_side): a thread computes the + sides
// of its three faces; the - sides come from the cell one plane down
// (registers carried along the march; on a column's first plane they are
// evaluated there), from the lane one cell down in x (a lane shuffle; lane 0
// reads the x halo) and from the row one cell down in y (LDS; row 0 reads the
// y halo). Wave c evaluates component c's halo sides from the LDS planes: along
// y on the row j0 - 1 (all lanes) and along x on the column i0 - 1 (lanes 0..3).
// Plane k is evaluated after the second barrier of step k and assembled after
// the first barrier of step k + 1. Per cell this is about 480 VALU
// instructions against 640 when each face evaluates both of its sides.
//
// Needs nx % 64 == 0, ny % 4 == 0 and nz >= 3 (one periodic wrap reaches k-2
// and k+3); the results are those of predict_face_gen (k_predict_fused) bit
// for bit.
__global__ __launch_bounds__(256, 3) void k_predict_tile(Geo g, Vel V, const double* __restrict__ p,
                                                      double* __restrict__ us0, double* __restrict__ us1,
                                                      double* __restrict__ us2, double dt, int KC) {
  __shared__ double Ush[2 * PT_PL], Vsh[2 * PT_PL], Wsh[2 * PT_PL], Psh[PT_PL];
  // + sides along y of rows j0-1 .. j0+2 (index row - j0 + 1), quantity 2A (M) / 2A+1 (T)
  __shared__ double PY[4 * 6 * PT_X];
  // + sides along x of the column i0 - 1, rows j0 .. j0+3
  __shared__ double PX[PT_Y * 6];
  const int tx = (int)threadIdx.x, ty = (int)threadIdx.y;
  const int t = ty * PT_X + tx;
  const int i0 = 1 + (int)blockIdx.x * PT_X, j0 = 1 + (int)blockIdx.y * PT_Y;
  const int i = i0 + tx, j = j0 + ty;
  const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
  const size_t px = (size_t)g.px, plane = (size_t)g.px * (size_t)g.py;
  const int k0 = 1 + (int)blockIdx.z * KC;
  const int k1 = min(k0 + KC - 1, nz);
  size_t toff[3];
  bool tval[3];
#pragma unroll
  for (int u = 0; u < 3; u++) {
    const int e = t + 256 * u;
    tval[u] = e < PT_PL;
    const int ey = e / PT_W - 2, ex = e % PT_W - 2;
    toff[u] = tval[u] ? (size_t)wrapi(j0 + ey, ny) * px + (size_t)wrapi(i0 + ex, nx) : 0;
  }
  const size_t own_off = (size_t)j * px + (size_t)i;
  const int own = (ty + 2) * PT_W + tx + 2;
  auto kpl = [&](int kk) { return (size_t)wrapi(kk, nz) * plane; };
  auto tile_ld = [&](const double* a, int kk, double out[3]) {
    const size_t kb = kpl(kk);
#pragma unroll
    for (int u = 0; u < 3; u++)
      if (tval[u]) out[u] = a[kb + toff[u]];
  };
  auto tile_st = [&](double* sl, const double in[3]) {
#pragma unroll
    for (int u = 0; u < 3; u++)
      if (tval[u]) sl[t + 256 * u] = in[u];
  };
  const double* ua = V.c[0];
  const double* va = V.c[1];
  const double* wa = V.c[2];
  double uz[5], vz[5], wz[5], pz[2];
  {
    double tu[3], tv[3], tw[3];
    tile_ld(ua, k0 - 1, tu);
    tile_ld(va, k0 - 1, tv);
    tile_ld(wa, k0, tw);
#pragma unroll
    for (int m = 1; m < 5; m++) {
      const size_t o = kpl(k0 - 3 + m) + own_off;
      uz[m] = ua[o];
      vz[m] = va[o];
      wz[m] = wa[o];
    }
    pz[1] = p[kpl(k0 - 1) + own_off];
    tile_st(Ush + ((k0 - 1) & 1) * PT_PL, tu);
    tile_st(Vsh + ((k0 - 1) & 1) * PT_PL, tv);
    tile_st(Wsh + (k0 & 1) * PT_PL, tw);
  }
  double nu[3], nv[3], nw[3], np[3], ou, ov, ow;
  tile_ld(ua, k0, nu);
  tile_ld(va, k0, nv);
  tile_ld(wa, k0 + 1, nw);
  tile_ld(p, k0, np);
  {
    const size_t o = kpl(k0 + 2) + own_off;
    ou = ua[o];
    ov = va[o];
    ow = wa[o];
  }
  // - sides along z of the plane being evaluated (the + sides of the plane below)
  double mzM[3], mzT[3];
  // + sides of the pending plane k - 1, its pressure one cell down in x and y
  double Mu[3], Tu[3], Mv[3], Tv[3], Mw[3], Tw[3], pm0 = 0.0, pm1 = 0.0;
  // plane k is evaluated after the second barrier of step k and assembled
  // after the first barrier of step k + 1, which also orders the exchange
  // through PY / PX (written after the second barrier, overwritten only after
  // the second barrier of the next step)
  for (int k = k0; k <= k1 + 1; k++) {
    __syncthreads();  // plane k - 1's exchange is written; the ring slots below are free
    if (k > k0) {
      double myM[3], myT[3], mxM[3], mxT[3];
      const double* py = PY + ty * 6 * PT_X + tx;
#pragma unroll
      for (int a = 0; a < 3; a++) {
        myM[a] = py[(2 * a) * PT_X];
        myT[a] = py[(2 * a + 1) * PT_X];
      }
      const double xs[6] = {Mu[0], Tu[0], Mv[0], Tv[0], Mw[0], Tw[0]};
      double xm[6];
#pragma unroll
      for (int q = 0; q < 6; q++) {
        const double s = __shfl_up(xs[q], 1, 64);
        xm[q] = tx == 0 ? PX[ty * 6 + q] : s;
      }
#pragma unroll
      for (int a = 0; a < 3; a++) {
        mxM[a] = xm[2 * a];
        mxT[a] = xm[2 * a + 1];
      }
      const size_t id = kpl(k - 1) + own_off;
      const double pq = pz[1];
      {
        const double Mm[3] = {mxM[0], myM[0], mzM[0]}, Tm[3] = {mxT[0], myT[0], mzT[0]};
        us0[id] = face_assemble<0>(g, Mu, Tu, Mm, Tm, uz[2], pq, pm0, dt);
      }
      {
        const double Mm[3] = {mxM[1], myM[1], mzM[1]}, Tm[3] = {mxT[1], myT[1], mzT[1]};
        us1[id] = face_assemble<1>(g, Mv, Tv, Mm, Tm, vz[2], pq, pm1, dt);
      }
      {
        const double Mm[3] = {mxM[2], myM[2], mzM[2]}, Tm[3] = {mxT[2], myT[2], mzT[2]};
        us2[id] = face_assemble<2>(g, Mw, Tw, Mm, Tm, wz[2], pq, pz[0], dt);
      }
      mzM[0] = Mu[2];
      mzT[0] = Tu[2];
      mzM[1] = Mv[2];
      mzT[1] = Tv[2];
      mzM[2] = Mw[2];
      mzT[2] = Tw[2];
    }
    if (k > k1) break;
    double* U0 = Ush + (k & 1) * PT_PL;
    double* V0 = Vsh + (k & 1) * PT_PL;
    double* W1 = Wsh + ((k + 1) & 1) * PT_PL;
    tile_st(U0, nu);
    tile_st(V0, nv);
    tile_st(W1, nw);
    tile_st(Psh, np);
#pragma unroll
    for (int m = 0; m < 4; m++) {
      uz[m] = uz[m + 1];
      vz[m] = vz[m + 1];
      wz[m] = wz[m + 1];
    }
    uz[4] = ou;
    vz[4] = ov;
    wz[4] = ow;
    pz[0] = pz[1];
    __syncthreads();
    pz[1] = Psh[own];
    const double* U1 = Ush + ((k - 1) & 1) * PT_PL;
    const double* V1 = Vsh + ((k - 1) & 1) * PT_PL;
    const double* W0 = Wsh + (k & 1) * PT_PL;
    // LDS planes at LDS index o: plane k of every field, k + 1 of w, k - 1 of u and v
    auto lds = [&](int c, int o, int dk) -> double {
      if (dk == 0) return c == 0 ? U0[o] : (c == 1 ? V0[o] : (c == 2 ? W0[o] : Psh[o]));
      if (dk == 1) return W1[o];
      return c == 0 ? U1[o] : V1[o];
    };
    auto ld = [&](int c, int di, int dj, int dk) -> double {
      if (di == 0 && dj == 0) {
        if (c == 0) return uz[dk + 2];
        if (c == 1) return vz[dk + 2];
        if (c == 2) return wz[dk + 2];
        return pz[dk + 1];
      }
      return lds(c, own + dj * PT_W + di, dk);
    };
    // the next planes load while this one is evaluated
    if (k < k1) {
      tile_ld(ua, k + 1, nu);
      tile_ld(va, k + 1, nv);
      tile_ld(wa, k + 2, nw);
      tile_ld(p, k + 1, np);
      const size_t o = kpl(k + 3) + own_off;
      ou = ua[o];
      ov = va[o];
      ow = wa[o];
    }
    plus_sides<0>(g, ld, Mu, Tu);
    plus_sides<1>(g, ld, Mv, Tv);
    plus_sides<2>(g, ld, Mw, Tw);
    if (k == k0) {
      // first plane of the column: the - sides along z evaluated on the plane below
      auto ldz = [&](int c, int di, int dj, int dk) -> double { return ld(c, di, dj, dk - 1); };
      plus_side<0, 2>(g, ldz, mzM[0], mzT[0]);
      plus_side<1, 2>(g, ldz, mzM[1], mzT[1]);
      plus_side<2, 2>(g, ldz, mzM[2], mzT[2]);
    }
    if (ty < PT_Y - 1) {
      double* py = PY + (ty + 1) * 6 * PT_X + tx;
      py[0 * PT_X] = Mu[1];
      py[1 * PT_X] = Tu[1];
      py[2 * PT_X] = Mv[1];
      py[3 * PT_X] = Tv[1];
      py[4 * PT_X] = Mw[1];
      py[5 * PT_X] = Tw[1];
    }
    // halo sides, component ty on wave ty: along y on row j0 - 1 (all lanes),
    // along x on column i0 - 1, row j0 + tx (lanes 0..3)
    {
      const int yb = PT_W + tx + 2;
      const int xb = (tx + 2) * PT_W + 1;
      auto ldy = [&](int c, int di, int dj, int dk) -> double { return lds(c, yb + dj * PT_W + di, dk); };
      auto ldx = [&](int c, int di, int dj, int dk) -> double { return lds(c, xb + dj * PT_W + di, dk); };
      double hm, ht;
      if (ty == 0) {
        plus_side<0, 1>(g, ldy, hm, ht);
        PY[0 * PT_X + tx] = hm;
        PY[1 * PT_X + tx] = ht;
        if (tx < PT_Y) {
          plus_side<0, 0>(g, ldx, hm, ht);
          PX[tx * 6 + 0] = hm;
          PX[tx * 6 + 1] = ht;
        }
      } else if (ty == 1) {
        plus_side<1, 1>(g, ldy, hm, ht);
        PY[2 * PT_X + tx] = hm;
        PY[3 * PT_X + tx] = ht;
        if (tx < PT_Y) {
          plus_side<1, 0>(g, ldx, hm, ht);
          PX[tx * 6 + 2] = hm;
          PX[tx * 6 + 3] = ht;
        }
      } else if (ty == 2) {
        plus_side<2, 1>(g, ldy, hm, ht);
        PY[4 * PT_X + tx] = hm;
        PY[5 * PT_X + tx] = ht;
        if (tx < PT_Y) {
          plus_side<2, 0>(g, ldx, hm, ht);
          PX[tx * 6 + 4] = hm;
          PX[tx * 6 + 5] = ht;
        }
      }
    }
    // pressure one cell down in x and y, read before the next step's stores
    pm0 = Psh[own - 1];
    pm1 = Psh[own - PT_W];
  }
}

// All three components in one pass (triply periodic, one phase): the thread of
// cell (i,j,k) computes u* on x-face i, v* on y-face j and w* on z-face k (the
// cell's minus faces). The neighbourhoods of the three faces overlap, so their
// loads are shared through L1; per face the arithmetic is exactly
// predict_face<PeriodicBC, UniP, A>. Each thread marches KC planes in z so that
// planes are reused while they are still in L2.
template <class BC>
__global__ __launch_bounds__(256) void k_predict_fused(Geo g, Vel V, const double* __restrict__ p,
                                                       double* __restrict__ us0, double* __restrict__ us1,
                                                       double* __restrict__ us2, double dt, int KC) {
  const int i = 1 + (int)(blockIdx.x * 64 + threadIdx.x);
  const int j = 1 + (int)(blockIdx.y * 4 + threadIdx.y);
  if (i > g.n[0] || j > g.n[1]) return;
  const int k0 = 1 + (int)blockIdx.z * KC;
  const int k1 = min(k0 + KC - 1, g.n[2]);
  const bool xi = g.per[0] || i >= 2;
  const bool yi = g.per[1] || j >= 2;
  PerIdx X;
  X.px = g.px;
  X.py = g.py;
#pragma unroll
  for (int s = -2; s <= 2; s++) {
    X.x[s + 2] = wrapi(i + s, g.n[0]);
    X.y[s + 2] = wrapi(j + s, g.n[1]);
  }
  for (int k = k0; k <= k1; k++) {
    const bool zi = g.per[2] || k >= 2;
#pragma unroll
    for (int s = -2; s <= 2; s++) X.z[s + 2] = wrapi(k + s, g.n[2]);
    const size_t id = X.at(0, 0, 0);
    if (xi) us0[id] = predict_face_pu<0>(g, V, p, X, dt);
    if (yi) us1[id] = predict_face_pu<1>(g, V, p, X, dt);
    if (zi) us2[id] = predict_face_pu<2>(g, V, p, X, dt);
  }
}

// update_vstar_boundary (CFD/G_SMACSolver.cu:1044-1052; G_BoundaryFunctions.cu:3-71,
// getters G_BoundaryFunctions.h:369-486). blockIdx.z = side (0..5); x,y span the plane.
__global__ __launch_bounds__(256) void k_star_bc(Geo g, double* us0, double* us1, double* us2) {
  int side = blockIdx.z;
  int a = side >> 1;
  if (g.per[a]) return;
  int b1 = (a + 1) % 3, b2 = (a + 2) % 3;
  int t1 = 1 + (int)(blockIdx.x * 64 + threadIdx.x);
  int t2 = 1 + (int)(blockIdx.y * 4 + threadIdx.y);
  if (t1 > g.n[b1] || t2 > g.n[b2]) return;
  I3 q;
  q.v[a] = (side & 1) ? g.n[a] + 1 : 1;
  q.v[b1] = t1;
  q.v[b2] = t2;
  double* us = a == 0 ? us0 : (a == 1 ? us1 : us2);
  int bt = btype_of(g, side);
  double val = 0.;
  if (bt == BC_INFLOW) val = (1. - 0.0) * bv_of(g, side, a);
  else if (bt == BC_NOSLIP) val = 0.;
  else if (bt == BC_OUTLET) {
    int internal = (side & 1) ? -1 : 0;
    val = internal < 0 ? us[IX(g, sh(q, a, -1))] : us[IX(g, sh(q, a, +1))];
  }
  us[IX(g, q)] = val;
}

// ------------------------------------------------------------- row kernels

// base_k_make_poisson_rhs with eps = 1, no surface tension
// (CFD/pressure_solver/G_PressureSolverBase.cu:49-118). part0 <- row sums of b.
__global__ __launch_bounds__(256) void k_rhs(Geo g, const double* __restrict__ us0, const double* __restrict__ us1,
                                             const double* __restrict__ us2, double* __restrict__ rhs, double inv_dt,
                                             double* __restrict__ part0) {
  ROW_SETUP
  const size_t yp = idx(g, 0, W(g, 1, j + 1), k), zp = idx(g, 0, j, W(g, 2, k + 1));
  double acc = 0.0;
  for (int i = 1 + lane; i <= g.n[0]; i += 64) {
    int ip = W(g, 0, i + 1);
    double div = (us0[rb + ip] - us0[rb + i]) * g.inv_d[0] + (us1[yp + i] - us1[rb + i]) * g.inv_d[1] +
                 (us2[zp + i] - us2[rb + i]) * g.inv_d[2];
    double b = -inv_dt * div;
    rhs[rb + i] = b;
    acc += b;
  }
  acc = wave_tree_sum(acc);
  if (lane == 0) part0[row] = acc;
}

// base_k_make_poisson_rhs with the volume fraction, no surface tension
// (CFD/pressure_solver/G_PressureSolverBase.cu:49-118):
// b = -(1/dt) sum_d (eps_{d+} u*_{d+} - eps_{d-} u*_{d-})/h_d - (eps - eps_old)/dt^2
// with eps_{d+-} the face means of eps^{n+1} at the raw neighbour cells.
// part0 <- row sums of b.
__global__ __launch_bounds__(256) void k_rhs_eps(Geo g, const double* __restrict__ us0,
                                                 const double* __restrict__ us1, const double* __restrict__ us2,
                                                 const double* __restrict__ eps, const double* __restrict__ eps_old,
                                                 double* __restrict__ rhs, double inv_dt, double* __restrict__ part0) {
  ROW_SETUP
  const size_t yp = idx(g, 0, W(g, 1, j + 1), k), zp = idx(g, 0, j, W(g, 2, k + 1));
  const size_t ey = (size_t)g.px, ez = (size_t)g.px * (size_t)g.py;
  double acc = 0.0;
  for (int i = 1 + lane; i <= g.n[0]; i += 64) {
    const int ip = W(g, 0, i + 1);
    const size_t c = rb + i;
    const double exp_ = 0.5 * (eps[c] + eps[c + 1]), exm = 0.5 * (eps[c - 1] + eps[c]);
    const double eyp = 0.5 * (eps[c] + eps[c + ey]), eym = 0.5 * (eps[c - ey] + eps[c]);
    const double ezp = 0.5 * (eps[c] + eps[c + ez]), ezm = 0.5 * (eps[c - ez] + eps[c]);
    const double div_eps = (exp_ * us0[rb + ip] - exm * us0[c]) * g.inv_d[0] +
                           (eyp * us1[yp + i] - eym * us1[c]) * g.inv_d[1] +
                           (ezp * us2[zp + i] - ezm * us2[c]) * g.inv_d[2];
    const double deps = eps[c] - eps_old[c];
    const double b = -inv_dt * div_eps - deps * inv_dt * inv_dt;
    rhs[c] = b;
    acc += b;
  }
  acc = wave_tree_sum(acc);
  if (lane == 0) part0[row] = acc;
}

// The start of the PCG, r = b - mean b (pure Neumann: mean = S[slot_mean]) on
// the interior cells, r in the split layout (G_PCGSolver.cu:619-640 with
// dp = 0: r = b - A dp = b). SUMS: part0 <- sum v^2, part1 <- sum v of
// v = b - mean (lanes over natural i, the canonical row order). WRITE: r = v,
// or v + a S[slot_shift] (slot_shift >= 0: the second mean removal,
// subtract_cell_mean, G_PressureSolverBase.cu:142-155), f0 = (F) r when f0 is
// given (the fine-level rhs of the mixed V-cycle), dp = 0 when zero_dp (dp is
// otherwise left as the last correction cleared it). The pure-Neumann start
// is two passes, sums then writes (the shift needs the sum of v); a regular
// operator's start is one.
template <typename F, bool SUMS, bool WRITE>
__global__ __launch_bounds__(256) void k_init_r(Geo g, const double* __restrict__ rhs, double* __restrict__ r,
                                                F* __restrict__ f0, double* __restrict__ dp, int zero_dp,
                                                const double* __restrict__ S, int slot_mean, double a,
                                                int slot_shift, double* __restrict__ part0,
                                                double* __restrict__ part1) {
  ROW_SETUP
  const int nx = g.n[0];
  const double mb = slot_mean >= 0 ? S[slot_mean] : 0.0;
  const double sh = slot_shift >= 0 ? S[slot_shift] : 0.0;
  double a0 = 0.0, a1 = 0.0;
  for (int i = 1 + lane; i <= nx; i += 64) {
    double v = rhs[rb + i] - mb;
    if (SUMS) {
      a0 += v * v;
      a1 += v;
    }
    if (WRITE) {
      if (slot_shift >= 0) v = v + a * sh;
      const size_t o = rb + spos(i, nx);
      r[o] = v;
      if (f0) St<F>::st(f0, o, (typename St<F>::C)v);
      if (zero_dp) dp[rb + i] = 0.0;
    }
  }
  if (SUMS) {
    a0 = wave_tree_sum(a0);
    a1 = wave_tree_sum(a1);
    if (lane == 0) {
      part0[row] = a0;
      part1[row] = a1;
    }
  }
}

// x += a * S[slot] on interior cells (subtract_cell_mean, G_PressureSolverBase.cu:142-155)
template <typename Z>
__global__ __launch_bounds__(256) void k_add_scalar(Geo g, Z* __restrict__ x, double a,
                                                    const double* __restrict__ S, int slot) {
  ROW_SETUP
  const double s = S[slot];
  for (int p = 1 + lane; p <= g.n[0]; p += 64)
    St<Z>::st(x, rb + p, (typename St<Z>::C)((double)St<Z>::ld(x, rb + p) + a * s));
}

// part0 <- row sums of x (split layout, canonical natural order)
template <typename Z>
__global__ __launch_bounds__(256) void k_sum(Geo g, const Z* __restrict__ x, double* __restrict__ part0) {
  ROW_SETUP
  double acc = 0.0;
  for (int i = 1 + lane; i <= g.n[0]; i += 64) acc += (double)St<Z>::ld(x, rb + spos(i, g.n[0]));
  acc = wave_tree_sum(acc);
  if (lane == 0) part0[row] = acc;
}

// Jacobi preconditioner z = (1/Adiag) r (G_PCGSolver.cu:270)
template <bool VAR>
__global__ __launch_bounds__(256) void k_jacobi(Geo g, BetaF bf, const double* __restrict__ r,
                                                double* __restrict__ z) {
  ROW_SETUP
  for (int i = 1 + lane; i <= g.n[0]; i += 64) {
    Coef c = VAR ? coef_var(g, bf.b, i, j, k) : coef_uniform(g.n, g.per, g.inv_d2, g.beta0, i, j, k);
    const size_t o = rb + spos(i, g.n[0]);
    z[o] = (1. / c.diag) * r[o];
  }
}

// d = z; part0 <- sum r*z in pair order (lane l: split positions 1+2l+128m,
// then +1; the oracle's RED_DOT_S2, the order of the iteration's r.z).
// slot >= 0: z + a S[slot] first (k_add_scalar's arithmetic) rounded to the
// storage of z, the value a separate pass would store and read back (z itself
// is not written: the next V-cycle overwrites it before anything reads it)
template <typename Z>
__global__ __launch_bounds__(256) void k_copy_dot(Geo g, const Z* __restrict__ z, double* __restrict__ d,
                                                  const double* __restrict__ r, double* __restrict__ part0,
                                                  double a, const double* __restrict__ S, int slot) {
  ROW_SETUP
  const int nx = g.n[0];
  double acc = 0.0;
  const double s = slot >= 0 ? S[slot] : 0.0;
  for (int q0 = 1 + 2 * lane; q0 <= nx; q0 += 128) {
#pragma unroll
    for (int t = 0; t < 2; t++) {
      const int q = q0 + t;
      if (q > nx) continue;
      const size_t o = rb + q;
      double zz = (double)St<Z>::ld(z, o);
      if (slot >= 0) zz = stored_value<Z>((double)(typename St<Z>::C)(zz + a * s));
      d[o] = zz;
      acc += r[o] * zz;
    }
  }
  acc = wave_tree_sum(acc);
  if (lane == 0) part0[row] = acc;
}

// The direction update fused into the next SpMV (DIR): d = z + beta d_old
// (k_update_dir, G_PCGSolver.cu:276-288) is formed on the fly for the cell and
// its neighbours from z and d_old, stored to a second array (a neighbouring
// block may still read d_old), and dp += alpha d_old (the deferred half of
// k_update_p_r) rides along; the lane's pairs of both half rows are four
// consecutive natural cells, two 16-byte dp pairs.
// FIRST (the first SpMV of a solve, in place of k_copy_dot and the SpMV after
// it): d = z, shifted by a S[slot_beta] and rounded as k_copy_dot does when
// slot_beta >= 0; no d_old, no dp; the row's r.d in k_copy_dot's order goes to
// prz.
template <typename Z>
struct DirArgs {
  const Z* z;
  const double* dold;
  double* dp;
  const double* S;
  int slot_alpha, slot_beta;
  const double* r = nullptr;
  double* prz = nullptr;
  double a = 0.0;
};

// Ad = A d, part0 <- sum d*Ad (k_make_pAp, G_PCGSolver.cu:87-124), the dot in
// pair order (lane l: split positions 1+2l+128m, then +1; the oracle's
// RED_DOT_S2). NM = nx/128 in {2, 4}, rows aligned (FAST): the lane holds its
// pairs of both half rows as double2; the x neighbours of a cell are the other
// half's pair in the same lane and, across a pair edge, one rotate shuffle
// (wall neighbours are the zero ghost value). LDSY (ny % 4 == 0): the four
// waves of a block own rows j0..j0+3 of one plane and share them, with rows
// j0-1 and j0+4, through LDS for the y neighbours. NM = 0: any nx, loads per
// neighbour, same order. DIR (NM > 0 only): d is written, not read; its values
// come from da as described above.
template <bool VAR, int NM, bool LDSY, typename Z = double, bool DIR = false, bool FIRST = false>
__global__ __launch_bounds__(256) void k_spmv_dot(Geo g, BetaF bf, double* __restrict__ d,
                                                  double* __restrict__ Ad, double* __restrict__ part0,
                                                  DirArgs<Z> da) {
  static_assert(!DIR || NM > 0, "the fused direction update needs the pair form");
  static_assert(!FIRST || DIR, "the first direction is formed as in DIR");
  constexpr int NMC = NM > 0 ? NM : 1;
  __shared__ double ysm[LDSY ? 6 * 128 * NMC : 1];
  ROW_SETUP
  const int nx = g.n[0], nh = nx >> 1;
  const size_t ypr = idx(g, 0, W(g, 1, j + 1), k), ymr = idx(g, 0, W(g, 1, j - 1), k);
  const size_t zpr = idx(g, 0, j, W(g, 2, k + 1)), zmr = idx(g, 0, j, W(g, 2, k - 1));
  double acc = 0.0;
  if constexpr (NM == 0) {
    for (int q0 = 1 + 2 * lane; q0 <= nx; q0 += 128) {
#pragma unroll
      for (int t = 0; t < 2; t++) {
        const int q = q0 + t;
        if (q > nx) continue;
        const int i = q <= nh ? 2 * q : 2 * (q - nh) - 1;
        Coef c = VAR ? coef_var(g, bf.b, i, j, k) : coef_uniform(g.n, g.per, g.inv_d2, g.beta0, i, j, k);
        double xc = d[rb + q];
        double xp = d[rb + spos(W(g, 0, i + 1), nx)], xm = d[rb + spos(W(g, 0, i - 1), nx)];
        double yp = d[ypr + q], ym = d[ymr + q], zp = d[zpr + q], zm = d[zmr + q];
        double s = c.a[0] * xp + c.a[1] * xm + c.a[2] * yp + c.a[3] * ym + c.a[4] * zp + c.a[5] * zm;
        double v = c.diag * xc - s;
        Ad[rb + q] = v;
        acc += xc * v;
      }
    }
  } else {
    constexpr int NH = NM / 2;  // pair chunks per half row
    constexpr int NX = 128 * NM;
    double xc[NM][2], yp[NM][2], ym[NM][2], zp[NM][2], zm[NM][2];
    using ZC = typename St<Z>::C;
    const double beta = (DIR && !FIRST) ? da.S[da.slot_beta] : 0.0;
    const bool shift = FIRST && da.slot_beta >= 0;
    const double zsum = shift ? da.S[da.slot_beta] : 0.0;
    // (FIRST) the direction value of z, as k_copy_dot forms it
    auto first_d = [&](ZC zc) -> double {
      double v = (double)zc;
      if (shift) v = stored_value<Z>((double)(ZC)(v + da.a * zsum));
      return v;
    };
    // a pair of d at split position a: loaded, or (DIR) z + beta d_old, or
    // (FIRST) the value of z
    auto ld2 = [&](size_t a, double* o2) {
      if constexpr (FIRST) {
        ZC zz[2];
        Vec<Z, 2>::ld(da.z + a, zz);
        o2[0] = first_d(zz[0]);
        o2[1] = first_d(zz[1]);
      } else if constexpr (DIR) {
        ZC zz[2];
        double dd[2];
        Vec<Z, 2>::ld(da.z + a, zz);
        Vec<double, 2>::ld(da.dold + a, dd);
        o2[0] = (double)zz[0] + beta * dd[0];
        o2[1] = (double)zz[1] + beta * dd[1];
      } else {
        Vec<double, 2>::ld(d + a, o2);
      }
    };
    if constexpr (FIRST) {
      // the row itself: d stored and r.d in pair order
      double rz = 0.0;
#pragma unroll
      for (int m = 0; m < NM; m++) {
        const size_t a = rb + 1 + 2 * lane + 128 * m;
        double rr[2];
        ld2(a, xc[m]);
        Vec<double, 2>::ld(da.r + a, rr);
        Vec<double, 2>::st(d + a, xc[m]);
        rz += rr[0] * xc[m][0];
        rz += rr[1] * xc[m][1];
      }
      rz = wave_tree_sum(rz);
      if (lane == 0) da.prz[row] = rz;
    } else if constexpr (DIR) {
      // the row itself: d = z + beta d_old stored, dp += alpha d_old on the
      // natural cells 4l+256m+1..+4 (odd cells from the second half row)
      const double alpha = da.S[da.slot_alpha];
      double dold[NM][2];
#pragma unroll
      for (int m = 0; m < NM; m++) {
        const size_t a = rb + 1 + 2 * lane + 128 * m;
        typename St<Z>::C zz[2];
        Vec<Z, 2>::ld(da.z + a, zz);
        Vec<double, 2>::ld(da.dold + a, dold[m]);
        xc[m][0] = (double)zz[0] + beta * dold[m][0];
        xc[m][1] = (double)zz[1] + beta * dold[m][1];
        Vec<double, 2>::st(d + a, xc[m]);
      }
#pragma unroll
      for (int m = 0; m < NH; m++) {
        double* pa = da.dp + rb + 1 + 4 * lane + 256 * m;
        double v0[2], v1[2];
        Vec<double, 2>::ld(pa, v0);
        Vec<double, 2>::ld(pa + 2, v1);
        v0[0] = v0[0] + alpha * dold[m + NH][0];
        v0[1] = v0[1] + alpha * dold[m][0];
        v1[0] = v1[0] + alpha * dold[m + NH][1];
        v1[1] = v1[1] + alpha * dold[m][1];
        Vec<double, 2>::st(pa, v0);
        Vec<double, 2>::st(pa + 2, v1);
      }
    }
    if constexpr (LDSY) {
      const int w = (int)(threadIdx.x >> 6);
      double hv[NM][2];
      const size_t hr = w == 0 ? ymr : ypr;  // wave 0: row j0-1, wave 3: row j0+4
#pragma unroll
      for (int m = 0; m < NM; m++) {
        const int q = 1 + 2 * lane + 128 * m;
        if constexpr (!DIR) Vec<double, 2>::ld(d + rb + q, xc[m]);
        ld2(zpr + q, zp[m]);
        ld2(zmr + q, zm[m]);
        if (w == 0 || w == 3) ld2(hr + q, hv[m]);
      }
#pragma unroll
      for (int m = 0; m < NM; m++)
#pragma unroll
        for (int t = 0; t < 2; t++) {
          const int o = 2 * lane + 128 * m + t;
          ysm[(w + 1) * NX + o] = xc[m][t];
          if (w == 0) ysm[o] = hv[m][t];
          if (w == 3) ysm[5 * NX + o] = hv[m][t];
        }
      __syncthreads();
#pragma unroll
      for (int m = 0; m < NM; m++)
#pragma unroll
        for (int t = 0; t < 2; t++) {
          const int o = 2 * lane + 128 * m + t;
          ym[m][t] = ysm[w * NX + o];
          yp[m][t] = ysm[(w + 2) * NX + o];
        }
    } else {
#pragma unroll
      for (int m = 0; m < NM; m++) {
        const int q = 1 + 2 * lane + 128 * m;
        if constexpr (!DIR) Vec<double, 2>::ld(d + rb + q, xc[m]);
        ld2(ypr + q, yp[m]);
        ld2(ymr + q, ym[m]);
        ld2(zpr + q, zp[m]);
        ld2(zmr + q, zm[m]);
      }
    }
    const bool perx = g.per[0] != 0;
    const int up = (lane + 1) & 63, dn = (lane + 63) & 63;
    // uniform coefficients: the cells off the x walls share those of i = 2
    const Coef crow = coef_uniform(g.n, g.per, g.inv_d2, g.beta0, 2, j, k);
#pragma unroll
    for (int m = 0; m < NM; m++) {
      // the x neighbour across the pair edge: for an even i (first half) the
      // other half's first value of the next lane (lane 63: lane 0 of the
      // next chunk, wrapping from i = nx to i = 1), for an odd i (second half)
      // the other half's second value of the previous lane
      double nb;
      if (m < NH) {
        const double src = lane == 0 ? xc[(m + 1) % NH + NH][0] : xc[m + NH][0];
        nb = __shfl(src, up);
      } else {
        const double src = lane == 63 ? xc[(m - NH + NH - 1) % NH][1] : xc[m - NH][1];
        nb = __shfl(src, dn);
      }
      double av[2];
#pragma unroll
      for (int t = 0; t < 2; t++) {
        const int q = 1 + 2 * lane + 128 * m + t;
        const int i = m < NH ? 2 * q : 2 * (q - nh) - 1;
        double xp, xm;
        if (m < NH) {  // even i = 2q: i-1 = other half, same slot; i+1 the next slot
          xm = xc[m + NH][t];
          xp = t == 0 ? xc[m + NH][1] : nb;
          if (!perx && i == nx) xp = 0.0;
        } else {       // odd i: i+1 = other half, same slot; i-1 the previous slot
          xp = xc[m - NH][t];
          xm = t == 1 ? xc[m - NH][0] : nb;
          if (!perx && i == 1) xm = 0.0;
        }
        const Coef c = VAR ? coef_var(g, bf.b, i, j, k)
                           : ((!perx && (i == 1 || i == nx)) ? coef_uniform(g.n, g.per, g.inv_d2, g.beta0, i, j, k) : crow);
        double s = c.a[0] * xp + c.a[1] * xm + c.a[2] * yp[m][t] + c.a[3] * ym[m][t] + c.a[4] * zp[m][t] +
                   c.a[5] * zm[m][t];
        double v = c.diag * xc[m][t] - s;
        av[t] = v;
        acc += xc[m][t] * v;
      }
      Vec<double, 2>::st(Ad + rb + 1 + 2 * lane + 128 * m, av);
    }
  }
  acc = wave_tree_sum(acc);
  if (lane == 0) part0[row] = acc;
}


// Row sums over the solver's split vectors in pair order: lane l accumulates the
// split positions q = 1+2l+128m, 1+2l+128m+1 (m = 0, 1, ...; q <= nx) in that
// order (the oracle's RED_*_S2 sums), then the canonical tree over the lanes.
//
// Chunk form (rows of 128 NM cells, NM = 1..4, element 1 of every row on a
// 128-byte line): a wave owns one 128-position chunk c of one row -- lane l the
// split positions 1+2l+128c and +1, one 16-byte access per double array -- and
// a block owns RB whole rows (RB = 4/NM; NM = 3: one row, 192 threads), blocks
// in plane order. A wave walking a whole row with NM accesses per lane per
// array, in the band order of the stencil kernels, streams 20-25% slower. The
// row partials keep the pair order: each chunk wave leaves its lane's two
// values in LDS and the row's chunk-0 wave adds them chunk by chunk.
template <int NM>
struct ChunkGeom {
  static constexpr int RB = NM == 3 ? 1 : 4 / NM;  // rows per block
  static constexpr int THREADS = 64 * NM * RB;
};
#define CHUNK_SETUP(NM)                                                                    \
  constexpr int RB_ = ChunkGeom<NM>::RB;                                                   \
  const int lane = threadIdx.x & 63;                                                       \
  const int wv = (int)(threadIdx.x >> 6);                                                  \
  const int rib = wv / (NM), ch = wv % (NM);                                               \
  const int row = (int)blockIdx.x * RB_ + rib;                                             \
  const bool live = row < g.n[1] * g.n[2];                                                 \
  const int j = live ? row % g.n[1] + 1 : 1, k = live ? row / g.n[1] + 1 : 1;              \
  const size_t rb = ((size_t)k * (size_t)g.py + (size_t)j) * (size_t)g.px;                 \
  const size_t o = rb + 1 + 2 * lane + 128 * ch;                                           \
  (void)o;

// r -= alpha Ad (the residual half of k_update_p_r, G_PCGSolver.cu:66-85);
// part0 <- sum r^2 (pair order), part1 <- max |r|. The dp += alpha d half is
// deferred to the next k_xpay, which reads d anyway, or to k_dp_flush after
// the last iteration (the same dp + alpha d per cell).
// F: multigrid storage; when f0 is given it receives (F) r (the fine-level rhs
// of the next V-cycle in the mixed forms). Row form: any nx.
template <typename F>
__global__ __launch_bounds__(256) void k_update(Geo g, double* __restrict__ r, const double* __restrict__ Ad,
                                                const double* __restrict__ S, int slot_alpha,
                                                double* __restrict__ part0, double* __restrict__ part1,
                                                F* __restrict__ f0) {
  using C = typename St<F>::C;
  ROW_SETUP
  const int nx = g.n[0];
  const double alpha = S[slot_alpha];
  double a0 = 0.0, mx = 0.0;
  for (int q0 = 1 + 2 * lane; q0 <= nx; q0 += 128) {
#pragma unroll
    for (int t = 0; t < 2; t++) {
      const int q = q0 + t;
      if (q <= nx) {
        const size_t o = rb + q;
        double rn = r[o] - alpha * Ad[o];
        r[o] = rn;
        if (f0) St<F>::st(f0, o, (C)rn);
        a0 += rn * rn;
        double ar = fabs(rn);
        if (ar > mx || ar != ar) mx = ar;
      }
    }
  }
  a0 = wave_tree_sum(a0);
  mx = wave_max(mx);
  if (lane == 0) {
    part0[row] = a0;
    part1[row] = mx;
  }
}

// chunk form of k_update (same arithmetic and sums). zs != 0 (uniform
// coefficients): the first half sweep of the V-cycle that follows rides
// along -- it starts from e = 0, so on the red cells ((i+j+k) even)
// e = f/diag, pointwise, with f the value just stored (rounded to F) and the
// division of the level-0 smoother (zs = 1: IEEE division, 2: the FP32
// reciprocal); e0 is level 0's e, L0 its geometry
template <typename F, int NM>
__global__ __launch_bounds__(256) void k_update_ch(Geo g, double* __restrict__ r, const double* __restrict__ Ad,
                                                   const double* __restrict__ S, int slot_alpha,
                                                   double* __restrict__ part0, double* __restrict__ part1,
                                                   F* __restrict__ f0, F* __restrict__ e0, int zs, Lv L0) {
  using C = typename St<F>::C;
  __shared__ double rv[ChunkGeom<NM>::RB][NM][64][2];
  CHUNK_SETUP(NM)
  double rn[2] = {0.0, 0.0};
  if (live) {
    const double alpha = S[slot_alpha];
    double vr[2], vA[2];
    Vec<double, 2>::ld(r + o, vr);
    Vec<double, 2>::ld(Ad + o, vA);
    C fv[2];
#pragma unroll
    for (int t = 0; t < 2; t++) {
      rn[t] = vr[t] - alpha * vA[t];
      fv[t] = (C)rn[t];
    }
    Vec<double, 2>::st(r + o, rn);
    if (f0) Vec<F, 2>::st(f0 + o, fv);
    if (zs) {
      // the lane's pair: split positions q, q+1 of one half, cells of one colour
      const int nh = g.n[0] >> 1;
      const int q = 1 + 2 * lane + 128 * ch;
      const int i0 = q <= nh ? 2 * q : 2 * (q - nh) - 1;
      if (((i0 + j + k) & 1) == 0) {
        C ev[2];
#pragma unroll
        for (int t = 0; t < 2; t++) {
          const C fs = (C)stored_value<F>(rn[t]);
          const C cd = (C)coef_uniform(L0.n, L0.per, L0.inv_d2, L0.beta0, i0 + 2 * t, j, k).diag;
          if constexpr (sizeof(C) == 4) ev[t] = zs == 2 ? fs * __builtin_amdgcn_rcpf((float)cd) : fs / cd;
          else ev[t] = fs / cd;
        }
        Vec<F, 2>::st(e0 + o, ev);
      }
    }
  }
  rv[rib][ch][lane][0] = rn[0];
  rv[rib][ch][lane][1] = rn[1];
  __syncthreads();
  if (live && ch == 0) {
    double a0 = 0.0, mx = 0.0;
#pragma unroll
    for (int m = 0; m < NM; m++)
#pragma unroll
      for (int t = 0; t < 2; t++) {
        const double v = rv[rib][m][lane][t];
        a0 += v * v;
        const double ar = fabs(v);
        if (ar > mx || ar != ar) mx = ar;
      }
    a0 = wave_tree_sum(a0);
    mx = wave_max(mx);
    if (lane == 0) {
      part0[row] = a0;
      part1[row] = mx;
    }
  }
}

// dp += alpha d (dp natural, d split): the deferred half of k_update_p_r
// after the last iteration
__global__ __launch_bounds__(256) void k_dp_flush(Geo g, double* __restrict__ dp, const double* __restrict__ d,
                                                  const double* __restrict__ S, int slot_alpha) {
  ROW_SETUP
  const double alpha = S[slot_alpha];
  for (int i = 1 + lane; i <= g.n[0]; i += 64) dp[rb + i] = dp[rb + i] + alpha * d[rb + spos(i, g.n[0])];
}

// part0 <- sum r*z; part1 <- sum z*Ad (flexible only), both in pair order.
// Row form: any nx.
template <typename Z>
__global__ __launch_bounds__(256) void k_dot2(Geo g, const double* __restrict__ r, const Z* __restrict__ z,
                                              const double* __restrict__ Ad, int flexible,
                                              double* __restrict__ part0, double* __restrict__ part1) {
  ROW_SETUP
  const int nx = g.n[0];
  double a0 = 0.0, a1 = 0.0;
  for (int q0 = 1 + 2 * lane; q0 <= nx; q0 += 128) {
#pragma unroll
    for (int t = 0; t < 2; t++) {
      const int q = q0 + t;
      if (q <= nx) {
        const size_t o = rb + q;
        const double zz = (double)St<Z>::ld(z, o);
        a0 += r[o] * zz;
        if (flexible) a1 += zz * Ad[o];
      }
    }
  }
  a0 = wave_tree_sum(a0);
  if (flexible) a1 = wave_tree_sum(a1);
  if (lane == 0) {
    part0[row] = a0;
    if (flexible) part1[row] = a1;
  }
}

// chunk form of k_dot2 (the same products, summed in the same order)
template <typename Z, int NM>
__global__ __launch_bounds__(256) void k_dot2_ch(Geo g, const double* __restrict__ r, const Z* __restrict__ z,
                                                 const double* __restrict__ Ad, int flexible,
                                                 double* __restrict__ part0, double* __restrict__ part1) {
  using C = typename St<Z>::C;
  __shared__ double pr[ChunkGeom<NM>::RB][NM][64][2];
  __shared__ double pa[ChunkGeom<NM>::RB][NM][64][2];
  CHUNK_SETUP(NM)
  double t0[2] = {0.0, 0.0}, t1[2] = {0.0, 0.0};
  if (live) {
    double vr[2], vA[2];
    C vz[2];
    Vec<double, 2>::ld(r + o, vr);
    Vec<Z, 2>::ld(z + o, vz);
    if (flexible) Vec<double, 2>::ld(Ad + o, vA);
#pragma unroll
    for (int t = 0; t < 2; t++) {
      t0[t] = vr[t] * (double)vz[t];
      if (flexible) t1[t] = (double)vz[t] * vA[t];
    }
  }
  pr[rib][ch][lane][0] = t0[0];
  pr[rib][ch][lane][1] = t0[1];
  pa[rib][ch][lane][0] = t1[0];
  pa[rib][ch][lane][1] = t1[1];
  __syncthreads();
  if (live && ch == 0) {
    double a0 = 0.0, a1 = 0.0;
#pragma unroll
    for (int m = 0; m < NM; m++)
#pragma unroll
      for (int t = 0; t < 2; t++) {
        a0 += pr[rib][m][lane][t];
        if (flexible) a1 += pa[rib][m][lane][t];
      }
    a0 = wave_tree_sum(a0);
    if (flexible) a1 = wave_tree_sum(a1);
    if (lane == 0) {
      part0[row] = a0;
      if (flexible) part1[row] = a1;
    }
  }
}

// dp += alpha d_old (natural dp; the deferred half of k_update_p_r), then
// d = z + beta d_old (k_update_dir, G_PCGSolver.cu:276-288). Lanes run over
// natural i, so the dp stores are contiguous and those of d and the loads of
// z and d are two contiguous half-row runs per instruction (lanes over split
// positions would store dp at stride two, which costs a second write-back of
// every dp line).
template <typename Z, int VPL>
__global__ __launch_bounds__(256) void k_xpay(Geo g, const Z* __restrict__ z, double* __restrict__ d,
                                              double* __restrict__ dp, const double* __restrict__ S, int slot_alpha,
                                              int slot_beta) {
  ROW_SETUP
  const int nx = g.n[0];
  const double alpha = S[slot_alpha];
  const double beta = S[slot_beta];
  constexpr int NV = VPL > 0 ? VPL : 1;
  const int trips = VPL > 0 ? 1 : (nx + 63) / 64;
  for (int t = 0; t < trips; t++) {
    double vz[NV], vd[NV], vdp[NV];
#pragma unroll
    for (int m = 0; m < NV; m++) {
      const int i = 1 + lane + 64 * (VPL > 0 ? m : t);
      if (i <= nx) {
        const size_t o = rb + spos(i, nx);
        vz[m] = (double)St<Z>::ld(z, o);
        vd[m] = d[o];
        vdp[m] = dp[rb + i];
      }
    }
#pragma unroll
    for (int m = 0; m < NV; m++) {
      const int i = 1 + lane + 64 * (VPL > 0 ? m : t);
      if (i <= nx) {
        const size_t o = rb + spos(i, nx);
        dp[rb + i] = vdp[m] + alpha * vd[m];
        d[o] = vz[m] + beta * vd[m];
      }
    }
  }
}

// chunk form of k_xpay: wave c of a row owns the natural cells i = 128c+1+2l
// and +1 (lane l; dp as one 16-byte access), d and z at their split positions
// (the odd cell in the second half, the even one in the first)
template <typename Z, int NM>
__global__ __launch_bounds__(256) void k_xpay_ch(Geo g, const Z* __restrict__ z, double* __restrict__ d,
                                                 double* __restrict__ dp, const double* __restrict__ S,
                                                 int slot_alpha, int slot_beta) {
  CHUNK_SETUP(NM)
  if (!live) return;
  const double alpha = S[slot_alpha];
  const double beta = S[slot_beta];
  const int i0 = 128 * ch + 1 + 2 * lane;             // odd; i0 + 1 even
  const size_t so = rb + (g.n[0] >> 1) + (i0 + 1) / 2;  // spos(i0)
  const size_t se = rb + (i0 + 1) / 2;                  // spos(i0 + 1)
  double vdp[2];
  Vec<double, 2>::ld(dp + rb + i0, vdp);
  const double dO = d[so], dE = d[se];
  const double zO = (double)St<Z>::ld(z, so), zE = (double)St<Z>::ld(z, se);
  const double np[2] = {vdp[0] + alpha * dO, vdp[1] + alpha * dE};
  Vec<double, 2>::st(dp + rb + i0, np);
  d[so] = zO + beta * dO;
  d[se] = zE + beta * dE;
}

// --------------------------------------------------------------- finalize
// Reduce up to four row-partial arrays with the canonical 1024-lane tree
// (kind 0 = sum, 1 = max) into S[out[m]], then apply one scalar op.
// lane t of the canonical tree accumulates rows t, t+1024, ... in order;
// loads are issued in batches (64 in the multi-block first stage, 16 in the
// single-block kernel) so their latencies overlap; the summation order is
// unchanged
// B: loads in flight per lane (the adds stay sequential)
template <int B>
__device__ __forceinline__ double fin_lane(const double* __restrict__ part, int kind, int nrows, int t) {
  double acc = 0.0;
  int r = t;
  for (; r + (B - 1) * 1024 < nrows; r += B * 1024) {
    double v[B];
#pragma unroll
    for (int u = 0; u < B; u++) v[u] = part[r + u * 1024];
#pragma unroll
    for (int u = 0; u < B; u++) {
      if (kind == 0) {
        acc += v[u];
      } else if (v[u] > acc || v[u] != v[u]) {
        acc = v[u];
      }
    }
  }
  for (; r < nrows; r += 1024) {
    double v = part[r];
    if (kind == 0) {
      acc += v;
    } else if (v > acc || v != v) {
      acc = v;
    }
  }
  return acc;
}

// first stage of a large reduction: the 1024 lane sums of array blockIdx.y,
// 64 lanes per block so that many CUs stream the row partials
__global__ __launch_bounds__(64) void k_fin_lanes(FinalizeArgs fa) {
  const int m = blockIdx.y;
  const int t = (int)blockIdx.x * 64 + (int)threadIdx.x;
  fa.stage[m * 1024 + t] = fin_lane<64>(fa.part[m], fa.kind[m], fa.nrows, t);
}

__global__ __launch_bounds__(1024) void k_finalize(FinalizeArgs fa, double* __restrict__ S) {
  __shared__ double lds[1024];
  const int t = threadIdx.x;
  for (int m = 0; m < fa.count; m++) {
    const int kind = fa.kind[m];
    double acc = fa.stage ? fa.stage[m * 1024 + t] : fin_lane<16>(fa.part[m], kind, fa.nrows, t);
    lds[t] = acc;
    __syncthreads();
    for (int s = 512; s >= 64; s >>= 1) {
      if (t < s) {
        if (kind == 0) {
          lds[t] = lds[t] + lds[t + s];
        } else {
          double o = lds[t + s];
          if (o > lds[t] || o != o) lds[t] = o;
        }
      }
      __syncthreads();
    }
    if (t < 64) {
      double v = lds[t];
      v = kind == 0 ? wave_tree_sum(v) : wave_max(v);
      if (t == 0) S[fa.out[m]] = v;
    }
    __syncthreads();
  }
  if (t == 0) {
    switch (fa.op) {
      case OP_MEAN:  // S[dst] = S[a] / c
        S[fa.dst] = S[fa.sa] / fa.c;
        break;
      case OP_DIV:  // S[dst] = S[a] / S[b]
        S[fa.dst] = S[fa.sa] / S[fa.sb];
        break;
      case OP_BETA_PR:  // S[dst] = (-S[alpha] * S[zad]) / S[rz]
        S[fa.dst] = (-S[fa.sa] * S[fa.sb]) / S[fa.sc];
        break;
      default:
        break;
    }
  }
}

// S[dst] = S[src] (e.g. rz <- rz_new)
__global__ void k_scalar_copy(double* S, int dst, int src) { S[dst] = S[src]; }
__global__ void k_scalar_set(double* S, int dst, double v) { S[dst] = v; }

// ------------------------------------------------------------- correction
// The correction of all three components and p += dp in one pass: the
// thread of interior cell (i,j,k) corrects its minus x-, y- and z-faces
// (those that are interior faces) with u = u* + (1/rho_f) dt (0 - (dp_c -
// dp_{c-1})/h) (k_correct_vof_velocity, CFD/G_SMACSolver.cu:1072-1142), then
// adds the cell's dp to p; dp is cleared afterwards by k_zero_dp (the memset
// of CFD/G_SMACSolver.cu:1273), since the neighbouring threads read it.
__global__ __launch_bounds__(256) void k_correct_all(Geo g, BetaF bf, int var, const double* __restrict__ us0,
                                                     const double* __restrict__ us1,
                                                     const double* __restrict__ us2, const double* __restrict__ dp,
                                                     double* __restrict__ u0, double* __restrict__ u1,
                                                     double* __restrict__ u2, double* __restrict__ p, double dt,
                                                     const double* __restrict__ dv, const double* __restrict__ S,
                                                     int slot_alpha) {
  const int i = 1 + (int)(blockIdx.x * 64 + threadIdx.x);
  const int j = 1 + (int)(blockIdx.y * 4 + threadIdx.y);
  const int k = 1 + (int)blockIdx.z;
  if (i > g.n[0] || j > g.n[1]) return;
  const I3 q = mk(i, j, k);
  const size_t id = IX(g, q);
  // slot_alpha >= 0: the solve's last dp += alpha d (d split, per row) is
  // still pending and is added here, cell by cell, with the flush's arithmetic
  const double alpha = slot_alpha >= 0 ? S[slot_alpha] : 0.0;
  auto dpv = [&](const I3& c, size_t ic) {
    double v = dp[ic];
    if (slot_alpha >= 0) v = v + alpha * dv[ic - (size_t)c.v[0] + (size_t)spos(c.v[0], g.n[0])];
    return v;
  };
  const double d0 = dpv(q, id);
  const double* const us[3] = {us0, us1, us2};
  double* const u[3] = {u0, u1, u2};
#pragma unroll
  for (int a = 0; a < 3; a++) {
    if (!g.per[a] && q.v[a] < 2) continue;  // the wall face 1 is not interior
    const I3 cm = wr(g, sh(q, a, -1));
    const double gr = (d0 - dpv(cm, IX(g, cm))) * g.inv_d[a];
    const double ir = var ? bf.b[a][id] : g.beta0;  // 1/rho_f on interior faces
    u[a][id] = us[a][id] + ir * dt * (0.0 - gr);
  }
  p[id] += d0;
}

// dp = 0 on interior cells (the reference's memset of p_delta after the correction)
__global__ __launch_bounds__(256) void k_zero_dp(Geo g, double* __restrict__ dp) {
  ROW_SETUP
  for (int i = 1 + lane; i <= g.n[0]; i += 64) dp[rb + i] = 0.0;
}

// ---------------------------------------------------------------- CFL
// k_calc_cfl (CFD/G_SMACSolver.cu:1365-1407): part0 <- row max of the convective
// CFL, part1 <- row max of the viscous CFL (uniform).
__global__ __launch_bounds__(256) void k_cfl(Geo g, Vel V, double dt, double visc, const double* __restrict__ inv_rho,
                                             const double* __restrict__ mu, double* __restrict__ part0,
                                             double* __restrict__ part1) {
  ROW_SETUP
  const size_t yp = idx(g, 0, W(g, 1, j + 1), k), zp = idx(g, 0, j, W(g, 2, k + 1));
  double m0 = 0.0, m1 = 0.0;
  for (int i = 1 + lane; i <= g.n[0]; i += 64) {
    if (inv_rho) {  // two-phase: (mu/rho) dt sum 1/h^2 per cell
      double cvv = inv_rho[rb + i] * mu[rb + i] * dt * (g.inv_d2[0] + g.inv_d2[1] + g.inv_d2[2]);
      if (cvv > m1 || cvv != cvv) m1 = cvv;
    }
    double a = fabs(V.c[0][rb + W(g, 0, i + 1)]), b = fabs(V.c[0][rb + i]);
    double mxv = a > b ? a : b;
    a = fabs(V.c[1][yp + i]);
    b = fabs(V.c[1][rb + i]);
    double myv = a > b ? a : b;
    a = fabs(V.c[2][zp + i]);
    b = fabs(V.c[2][rb + i]);
    double mzv = a > b ? a : b;
    double cf = dt * (mxv * g.inv_d[0] + myv * g.inv_d[1] + mzv * g.inv_d[2]);
    if (cf > m0 || cf != cf) m0 = cf;
  }
  m0 = wave_max(m0);
  m1 = wave_max(m1);
  if (lane == 0) {
    part0[row] = m0;
    part1[row] = inv_rho ? m1 : visc;
  }
}

// ---------------------------------------------------------------- launchers
static inline dim3 rows_grid(const Geo& g) { return dim3((unsigned)((g.n[1] * g.n[2] + 3) / 4)); }
static const dim3 ROWB(256), CELLB(64, 4, 1);

// ------------------------------------------------- two-phase properties
// update_properties_by_alpha (CFD/G_SMACSolver_VoF.cu:1374-1407) over the
// whole cell box: a ghost cell takes the values of its wrapped (periodic) or
// clamped (wall) interior cell (k_update_cell_boundary_properties_by_alpha,
// 1244-1307), so it is computed from that cell's alpha.
__global__ __launch_bounds__(256) void k_props_cells(Geo g, const double* __restrict__ alpha,
                                                     double* __restrict__ rho, double* __restrict__ rho_old,
                                                     double* __restrict__ inv_rho, double* __restrict__ mu,
                                                     double rho0, double rho1, double mu0, double mu1, int initial) {
  const int i = (int)(blockIdx.x * 64 + threadIdx.x);
  const int j = (int)(blockIdx.y * 4 + threadIdx.y);
  const int k = (int)blockIdx.z;
  if (i > g.n[0] + 1 || j > g.n[1] + 1) return;
  int q[3] = {i, j, k}, sq[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    int v = q[d];
    if (g.per[d]) v = wrapi(v, g.n[d]);
    else v = v < 1 ? 1 : (v > g.n[d] ? g.n[d] : v);
    sq[d] = v;
  }
  const double al = alpha[idx(g, sq[0], sq[1], sq[2])];
  const size_t id = idx(g, i, j, k);
  const double r = (1. - al) * rho0 + al * rho1;
  rho[id] = r;
  inv_rho[id] = 1. / r;
  mu[id] = (1. - al) * mu0 + al * mu1;
  if (initial) rho_old[id] = r;
}

// face properties of axis A over the whole face box: interior faces
// rho_f = (rho_L + rho_R)/2, beta = 1/rho_f, mu_f = (mu_L + mu_R)/2
// (k_update_{x,y,z}_face_properties_by_alpha, CFD/G_SMACSolver_VoF.cu:1409-1535);
// wall faces mu_f = inner mu, beta = 0 (update_boundary_faces, 1798-1941);
// ghost faces beta = 0.
template <int A>
__global__ __launch_bounds__(256) void k_props_faces(Geo g, const double* __restrict__ rho,
                                                     const double* __restrict__ mu, double* __restrict__ beta,
                                                     double* __restrict__ muf) {
  const int i = (int)(blockIdx.x * 64 + threadIdx.x);
  const int j = (int)(blockIdx.y * 4 + threadIdx.y);
  const int k = (int)blockIdx.z;
  if (i > g.n[0] + 2 || j > g.n[1] + 2) return;
  const I3 q = mk(i, j, k);
  const size_t id = IX(g, q);
  const int t = face_type(g, A, q);
  if (t == F_INTERIOR) {
    const size_t cm = IX(g, wr(g, sh(q, A, -1))), cp = IX(g, wr(g, q));
    const double rf = 0.5 * (rho[cp] + rho[cm]);
    beta[id] = 1. / rf;
    muf[id] = 0.5 * (mu[cp] + mu[cm]);
  } else if (t == F_BOUNDARY) {
    const int inner = q.v[A] == 1 ? 0 : -1;
    muf[id] = mu[IX(g, sh(q, A, inner))];
    // k_update_{x,y,z}_face_boundary_properties (CFD/G_SMACSolver_VoF.cu:1798-1941):
    // an outlet face takes 2 beta of the adjacent interior face (p = 0 at the
    // face), formed here with that face's own arithmetic; other boundary faces 0
    double b = 0.;
    if (btype_of(g, face_slot(g, A, q)) == BC_OUTLET) {
      const I3 qi = sh(q, A, inner < 0 ? -1 : 1);
      const double rfi = 0.5 * (rho[IX(g, qi)] + rho[IX(g, sh(qi, A, -1))]);
      b = 2.0 * (1. / rfi);
    }
    beta[id] = b;
  } else {
    beta[id] = 0.;
  }
}

void launch_props(const Geo& g, const double* alpha, double* rho, double* rho_old, double* inv_rho, double* mu,
                  double* beta[3], double* muf[3], double rho0, double rho1, double mu0, double mu1, int initial,
                  hipStream_t st) {
  dim3 cg((unsigned)((g.n[0] + 2 + 63) / 64), (unsigned)((g.n[1] + 2 + 3) / 4), (unsigned)(g.n[2] + 2));
  hipLaunchKernelGGL(k_props_cells, cg, CELLB, 0, st, g, alpha, rho, rho_old, inv_rho, mu, rho0, rho1, mu0, mu1,
                     initial);
  dim3 fg((unsigned)((g.n[0] + 3 + 63) / 64), (unsigned)((g.n[1] + 3 + 3) / 4), (unsigned)(g.n[2] + 3));
  hipLaunchKernelGGL(k_props_faces<0>, fg, CELLB, 0, st, g, rho, mu, beta[0], muf[0]);
  hipLaunchKernelGGL(k_props_faces<1>, fg, CELLB, 0, st, g, rho, mu, beta[1], muf[1]);
  hipLaunchKernelGGL(k_props_faces<2>, fg, CELLB, 0, st, g, rho, mu, beta[2], muf[2]);
}

template <class BC, class PR, class MF>
static void predict_components(const Geo& g, const Props& P, const Vel& V, const VofF& M, const double* p,
                               double* us[3], double dt, hipStream_t st) {
  for (int a = 0; a < 3; a++) {
    int lo = g.per[a] ? 1 : 2;
    int nf[3] = {g.n[0], g.n[1], g.n[2]};
    nf[a] = g.n[a] - lo + 1;
    if (nf[0] <= 0 || nf[1] <= 0 || nf[2] <= 0) continue;
    dim3 cg((unsigned)((nf[0] + 63) / 64), (unsigned)((nf[1] + 3) / 4), (unsigned)nf[2]);
    if (a == 0) hipLaunchKernelGGL((k_predict<BC, PR, MF, 0>), cg, CELLB, 0, st, g, P, V, M, p, us[0], dt);
    if (a == 1) hipLaunchKernelGGL((k_predict<BC, PR, MF, 1>), cg, CELLB, 0, st, g, P, V, M, p, us[1], dt);
    if (a == 2) hipLaunchKernelGGL((k_predict<BC, PR, MF, 2>), cg, CELLB, 0, st, g, P, V, M, p, us[2], dt);
  }
}

void launch_predict(const Geo& g, const Props* P, const Vel& V, const double* p, double* us[3], double dt,
                    hipStream_t st, const VofF* M) {
  const int KC = 16;  // z-planes per thread: the k+-2 halo is re-read once per 16 planes
  const bool allper = g.per[0] && g.per[1] && g.per[2];
  const VofF M0{};
  if (P && M) {
    // two-phase grids right after a VoF transport: mass flux with the alpha flux
    if (allper) predict_components<PeriodicBC, VarP, MfVof>(g, *P, V, *M, p, us, dt, st);
    else predict_components<WallBC, VarP, MfVof>(g, *P, V, *M, p, us, dt, st);
  } else if (P) {
    // two-phase grids: one kernel per component, properties from the arrays
    if (allper) predict_components<PeriodicBC, VarP, MfRho0>(g, *P, V, M0, p, us, dt, st);
    else predict_components<WallBC, VarP, MfRho0>(g, *P, V, M0, p, us, dt, st);
  } else if (allper && g.n[0] % PT_X == 0 && g.n[1] % PT_Y == 0 && g.n[2] >= 3) {
    const int KT = 32;  // z-planes per block column
    dim3 grid((unsigned)(g.n[0] / PT_X), (unsigned)(g.n[1] / PT_Y), (unsigned)((g.n[2] + KT - 1) / KT));
    hipLaunchKernelGGL(k_predict_tile, grid, dim3(PT_X, PT_Y, 1), 0, st, g, V, p, us[0], us[1], us[2], dt, KT);
    return;
  } else if (allper) {
    dim3 grid((unsigned)((g.n[0] + 63) / 64), (unsigned)((g.n[1] + 3) / 4), (unsigned)((g.n[2] + KC - 1) / KC));
    hipLaunchKernelGGL(k_predict_fused<PeriodicBC>, grid, CELLB, 0, st, g, V, p, us[0], us[1], us[2], dt, KC);
    return;
  } else {
    // walls: one kernel per component (the fused wall kernel exceeds the register file)
    predict_components<WallBC, UniP, MfRho0>(g, Props{}, V, M0, p, us, dt, st);
  }
  if (allper) return;
  int m = g.n[0] > g.n[1] ? g.n[0] : g.n[1];
  m = m > g.n[2] ? m : g.n[2];
  dim3 sgrid((unsigned)((m + 63) / 64), (unsigned)((m + 3) / 4), 6);
  hipLaunchKernelGGL(k_star_bc, sgrid, CELLB, 0, st, g, us[0], us[1], us[2]);
}

void launch_predict_two_way(const Geo& g, const Props* P, const Vel& V, const double* p, double* us[3], double dt,
                            hipStream_t st, const VofF* M, const EpsF& E) {
  const bool allper = g.per[0] && g.per[1] && g.per[2];
  VofF MM{};
  if (M) MM = *M;
  const Props PP = P ? *P : Props{};
  for (int a = 0; a < 3; a++) {
    int lo = g.per[a] ? 1 : 2;
    int nf[3] = {g.n[0], g.n[1], g.n[2]};
    nf[a] = g.n[a] - lo + 1;
    if (nf[0] <= 0 || nf[1] <= 0 || nf[2] <= 0) continue;
    dim3 cg((unsigned)((nf[0] + 63) / 64), (unsigned)((nf[1] + 3) / 4), (unsigned)nf[2]);
#define SMAC_TW(BCT, PRT, AX) \
  hipLaunchKernelGGL((k_predict_tw<BCT, PRT, AX>), cg, CELLB, 0, st, g, PP, V, MM, E, p, us[AX], dt)
#define SMAC_TW_A(BCT, PRT)              \
  if (a == 0) SMAC_TW(BCT, PRT, 0);      \
  else if (a == 1) SMAC_TW(BCT, PRT, 1); \
  else SMAC_TW(BCT, PRT, 2);
    if (allper) {
      if (P) { SMAC_TW_A(PeriodicBC, VarP) } else { SMAC_TW_A(PeriodicBC, UniP) }
    } else {
      if (P) { SMAC_TW_A(WallBC, VarP) } else { SMAC_TW_A(WallBC, UniP) }
    }
#undef SMAC_TW_A
#undef SMAC_TW
  }
  if (allper) return;
  int m = g.n[0] > g.n[1] ? g.n[0] : g.n[1];
  m = m > g.n[2] ? m : g.n[2];
  dim3 sgrid((unsigned)((m + 63) / 64), (unsigned)((m + 3) / 4), 6);
  hipLaunchKernelGGL(k_star_bc, sgrid, CELLB, 0, st, g, us[0], us[1], us[2]);
}
void launch_rhs_eps(const Geo& g, double* const us[3], const double* eps, const double* eps_old, double* rhs,
                    double inv_dt, double* part0, hipStream_t st) {
  hipLaunchKernelGGL(k_rhs_eps, rows_grid(g), ROWB, 0, st, g, us[0], us[1], us[2], eps, eps_old, rhs, inv_dt, part0);
}
void launch_pois_two_way(const Geo& g, double* const beta[3], const double* eps, double* bp[3], hipStream_t st) {
  dim3 fg((unsigned)((g.n[0] + 3 + 63) / 64), (unsigned)((g.n[1] + 3 + 3) / 4), (unsigned)(g.n[2] + 3));
  hipLaunchKernelGGL(k_pois_two_way<0>, fg, CELLB, 0, st, g, beta ? beta[0] : nullptr, eps, bp[0]);
  hipLaunchKernelGGL(k_pois_two_way<1>, fg, CELLB, 0, st, g, beta ? beta[1] : nullptr, eps, bp[1]);
  hipLaunchKernelGGL(k_pois_two_way<2>, fg, CELLB, 0, st, g, beta ? beta[2] : nullptr, eps, bp[2]);
}
void launch_rhs(const Geo& g, double* const us[3], double* rhs, double inv_dt, double* part0, hipStream_t st) {
  hipLaunchKernelGGL(k_rhs, rows_grid(g), ROWB, 0, st, g, us[0], us[1], us[2], rhs, inv_dt, part0);
}
template <typename F>
void launch_init_r(const Geo& g, const double* rhs, double* r, F* f0, double* dp, int zero_dp, const double* S,
                   int slot_mean, double a, int slot_shift, double* p0, double* p1, int sums, int write,
                   hipStream_t st) {
  if (sums && write)
    hipLaunchKernelGGL((k_init_r<F, true, true>), rows_grid(g), ROWB, 0, st, g, rhs, r, f0, dp, zero_dp, S,
                       slot_mean, a, slot_shift, p0, p1);
  else if (sums)
    hipLaunchKernelGGL((k_init_r<F, true, false>), rows_grid(g), ROWB, 0, st, g, rhs, r, f0, dp, zero_dp, S,
                       slot_mean, a, slot_shift, p0, p1);
  else
    hipLaunchKernelGGL((k_init_r<F, false, true>), rows_grid(g), ROWB, 0, st, g, rhs, r, f0, dp, zero_dp, S,
                       slot_mean, a, slot_shift, p0, p1);
}
template <typename Z>
void launch_add_scalar(const Geo& g, Z* x, double a, const double* S, int slot, hipStream_t st) {
  hipLaunchKernelGGL(k_add_scalar<Z>, rows_grid(g), ROWB, 0, st, g, x, a, S, slot);
}
template <typename Z>
void launch_sum(const Geo& g, const Z* x, double* p0, hipStream_t st) {
  hipLaunchKernelGGL(k_sum<Z>, rows_grid(g), ROWB, 0, st, g, x, p0);
}
void launch_jacobi(const Geo& g, const BetaF* bf, const double* r, double* z, hipStream_t st) {
  if (bf) hipLaunchKernelGGL(k_jacobi<true>, rows_grid(g), ROWB, 0, st, g, *bf, r, z);
  else hipLaunchKernelGGL(k_jacobi<false>, rows_grid(g), ROWB, 0, st, g, BetaF{}, r, z);
}
template <typename Z>
void launch_copy_dot(const Geo& g, const Z* z, double* d, const double* r, double* p0, hipStream_t st, double a,
                     const double* S, int slot) {
  hipLaunchKernelGGL(k_copy_dot<Z>, rows_grid(g), ROWB, 0, st, g, z, d, r, p0, a, S, slot);
}
// pair-order kernels: the vector form when a row is 128 NM cells (NM = 1..4)
// and the split vectors start every row on a 128-byte line
static inline int pair_nm(const Geo& g, const void* a) {
  const int nx = g.n[0];
  if (nx % 128 != 0 || nx > 512 || g.px % 64 != 0) return 0;
  const uintptr_t e1 = (uintptr_t)a + sizeof(double);  // element 1 of the box
  return e1 % 128 == 0 ? nx / 128 : 0;
}
void launch_spmv_dot(const Geo& g, const BetaF* bf, const double* d, double* Ad, double* p0, hipStream_t st) {
  int nm = pair_nm(g, d);
  if (nm && pair_nm(g, Ad) != nm) nm = 0;
  if (nm != 2 && nm != 4) nm = 0;  // both half rows in whole pair chunks
  const bool ldsy = nm && g.n[1] % 4 == 0;
#define SMAC_SPMV(V, N, L)                                                                               \
  hipLaunchKernelGGL((k_spmv_dot<V, N, L>), rows_grid(g), ROWB, 0, st, g, bf ? *bf : BetaF{}, const_cast<double*>(d), \
                     Ad, p0, DirArgs<double>{})
#define SMAC_SPMV_V(V)                                 \
  if (nm == 2) {                                       \
    if (ldsy) SMAC_SPMV(V, 2, true);                   \
    else SMAC_SPMV(V, 2, false);                       \
  } else if (nm == 4) {                                \
    if (ldsy) SMAC_SPMV(V, 4, true);                   \
    else SMAC_SPMV(V, 4, false);                       \
  } else {                                             \
    SMAC_SPMV(V, 0, false);                            \
  }
  if (bf) {
    SMAC_SPMV_V(true)
  } else {
    SMAC_SPMV_V(false)
  }
#undef SMAC_SPMV_V
#undef SMAC_SPMV
}
// grid and block of the chunk form: RB rows per block
template <int NM>
static inline dim3 chunk_grid(const Geo& g) {
  return dim3((unsigned)((g.n[1] * g.n[2] + ChunkGeom<NM>::RB - 1) / ChunkGeom<NM>::RB));
}
#define SMAC_CHUNK_SWITCH(NMV, LAUNCH) \
  switch (NMV) {                       \
    case 1: LAUNCH(1); break;          \
    case 2: LAUNCH(2); break;          \
    case 3: LAUNCH(3); break;          \
    default: LAUNCH(4); break;         \
  }
template <typename Z>
bool dir_spmv_applies(const Geo& g, const Z* z, const double* dold, const double* dnew, const double* dp,
                      const double* Ad) {
  int nm = pair_nm(g, dold);
  if (nm && (pair_nm(g, dnew) != nm || pair_nm(g, Ad) != nm || pair_nm(g, dp) != nm)) nm = 0;
  if (nm && ((uintptr_t)(z + 1) % 128) != 0) nm = 0;
  return nm == 2 || nm == 4;
}
template <typename Z>
void launch_dir_spmv(const Geo& g, const BetaF* bf, const Z* z, const double* dold, double* dnew, double* dp,
                     double* Ad, double* p0, const double* S, int slot_alpha, int slot_beta, hipStream_t st) {
  const int nm = pair_nm(g, dold);
  const bool ldsy = g.n[1] % 4 == 0;
  const DirArgs<Z> da{z, dold, dp, S, slot_alpha, slot_beta};
#define SMAC_DSP(V, N, L) \
  hipLaunchKernelGGL((k_spmv_dot<V, N, L, Z, true>), rows_grid(g), ROWB, 0, st, g, bf ? *bf : BetaF{}, dnew, Ad, p0, da)
#define SMAC_DSP_V(V)                  \
  if (nm == 2) {                       \
    if (ldsy) SMAC_DSP(V, 2, true);    \
    else SMAC_DSP(V, 2, false);        \
  } else {                             \
    if (ldsy) SMAC_DSP(V, 4, true);    \
    else SMAC_DSP(V, 4, false);        \
  }
  if (bf) {
    SMAC_DSP_V(true)
  } else {
    SMAC_DSP_V(false)
  }
#undef SMAC_DSP_V
#undef SMAC_DSP
}
template <typename Z>
bool first_spmv_applies(const Geo& g, const Z* z, const double* r, const double* d, const double* Ad) {
  int nm = pair_nm(g, d);
  if (nm && (pair_nm(g, r) != nm || pair_nm(g, Ad) != nm)) nm = 0;
  if (nm && ((uintptr_t)(z + 1) % 128) != 0) nm = 0;
  return nm == 2 || nm == 4;
}
template <typename Z>
void launch_first_spmv(const Geo& g, const BetaF* bf, const Z* z, const double* r, double* d, double* Ad,
                       double* p_dad, double* p_rz, double a, const double* S, int slot, hipStream_t st) {
  const int nm = pair_nm(g, d);
  const bool ldsy = g.n[1] % 4 == 0;
  DirArgs<Z> da{z, nullptr, nullptr, S, -1, slot};
  da.r = r;
  da.prz = p_rz;
  da.a = a;
#define SMAC_FSP(V, N, L)                                                                                   \
  hipLaunchKernelGGL((k_spmv_dot<V, N, L, Z, true, true>), rows_grid(g), ROWB, 0, st, g, bf ? *bf : BetaF{}, d, \
                     Ad, p_dad, da)
#define SMAC_FSP_V(V)                  \
  if (nm == 2) {                       \
    if (ldsy) SMAC_FSP(V, 2, true);    \
    else SMAC_FSP(V, 2, false);        \
  } else {                             \
    if (ldsy) SMAC_FSP(V, 4, true);    \
    else SMAC_FSP(V, 4, false);        \
  }
  if (bf) {
    SMAC_FSP_V(true)
  } else {
    SMAC_FSP_V(false)
  }
#undef SMAC_FSP_V
#undef SMAC_FSP
}
template <typename F>
bool launch_update(const Geo& g, double* r, const double* Ad, const double* S, int slot_alpha, double* p0,
                   double* p1, F* f0, hipStream_t st, F* e0, int zs, const Lv* L0) {
  int nm = pair_nm(g, r);
  if (nm && pair_nm(g, Ad) != nm) nm = 0;
  if (nm && f0 && ((uintptr_t)(f0 + 1) % 128) != 0) nm = 0;
  if (nm && e0 && ((uintptr_t)(e0 + 1) % 128) != 0) nm = 0;
  if (!nm) {
    hipLaunchKernelGGL(k_update<F>, rows_grid(g), ROWB, 0, st, g, r, Ad, S, slot_alpha, p0, p1, f0);
    return false;
  }
  if (!e0 || !L0) zs = 0;
  const Lv lz = L0 ? *L0 : Lv{};
#define SMAC_UPD(N)                                                                                         \
  hipLaunchKernelGGL((k_update_ch<F, N>), chunk_grid<N>(g), dim3(ChunkGeom<N>::THREADS), 0, st, g, r, Ad, S, \
                     slot_alpha, p0, p1, f0, e0, zs, lz)
  SMAC_CHUNK_SWITCH(nm, SMAC_UPD)
#undef SMAC_UPD
  return zs != 0;
}
void launch_dp_flush(const Geo& g, double* dp, const double* d, const double* S, int slot_alpha, hipStream_t st) {
  hipLaunchKernelGGL(k_dp_flush, rows_grid(g), ROWB, 0, st, g, dp, d, S, slot_alpha);
}
template <typename Z>
void launch_dot2(const Geo& g, const double* r, const Z* z, const double* Ad, int flexible, double* p0,
                 double* p1, hipStream_t st) {
  int nm = pair_nm(g, r);
  if (nm && flexible && pair_nm(g, Ad) != nm) nm = 0;
  if (nm && ((uintptr_t)(z + 1) % 128) != 0) nm = 0;
  if (!nm) {
    hipLaunchKernelGGL(k_dot2<Z>, rows_grid(g), ROWB, 0, st, g, r, z, Ad, flexible, p0, p1);
    return;
  }
#define SMAC_D2(N)                                                                                         \
  hipLaunchKernelGGL((k_dot2_ch<Z, N>), chunk_grid<N>(g), dim3(ChunkGeom<N>::THREADS), 0, st, g, r, z, Ad, \
                     flexible, p0, p1)
  SMAC_CHUNK_SWITCH(nm, SMAC_D2)
#undef SMAC_D2
}
template <typename Z>
void launch_xpay(const Geo& g, const Z* z, double* d, double* dp, const double* S, int slot_alpha, int slot_beta,
                 hipStream_t st) {
  int nm = pair_nm(g, d);
  if (nm && pair_nm(g, dp) != nm) nm = 0;  // dp: natural rows, 16-byte pairs from element 1
  if (nm) {
#define SMAC_XP(N)                                                                                        \
  hipLaunchKernelGGL((k_xpay_ch<Z, N>), chunk_grid<N>(g), dim3(ChunkGeom<N>::THREADS), 0, st, g, z, d, dp, S, \
                     slot_alpha, slot_beta)
    SMAC_CHUNK_SWITCH(nm, SMAC_XP)
#undef SMAC_XP
    return;
  }
  SMAC_VPL_DISPATCH(g.n[0], hipLaunchKernelGGL((k_xpay<Z, VPL>), rows_grid(g), ROWB, 0, st, g, z, d, dp, S,
                                               slot_alpha, slot_beta))
}
#define SMAC_INST_Z(Z)                                                                                     \
  template void launch_add_scalar<Z>(const Geo&, Z*, double, const double*, int, hipStream_t);            \
  template void launch_sum<Z>(const Geo&, const Z*, double*, hipStream_t);                                \
  template void launch_copy_dot<Z>(const Geo&, const Z*, double*, const double*, double*, hipStream_t,    \
                                   double, const double*, int);                                           \
  template void launch_init_r<Z>(const Geo&, const double*, double*, Z*, double*, int, const double*, int, \
                                 double, int, double*, double*, int, int, hipStream_t);                   \
  template bool launch_update<Z>(const Geo&, double*, const double*, const double*, int, double*, double*, \
                                 Z*, hipStream_t, Z*, int, const Lv*);                                    \
  template void launch_dot2<Z>(const Geo&, const double*, const Z*, const double*, int, double*, double*,  \
                               hipStream_t);                                                              \
  template void launch_xpay<Z>(const Geo&, const Z*, double*, double*, const double*, int, int, hipStream_t); \
  template bool dir_spmv_applies<Z>(const Geo&, const Z*, const double*, const double*, const double*,     \
                                    const double*);                                                       \
  template void launch_dir_spmv<Z>(const Geo&, const BetaF*, const Z*, const double*, double*, double*, double*, \
                                   double*, const double*, int, int, hipStream_t);                         \
  template bool first_spmv_applies<Z>(const Geo&, const Z*, const double*, const double*, const double*);  \
  template void launch_first_spmv<Z>(const Geo&, const BetaF*, const Z*, const double*, double*, double*,   \
                                     double*, double*, double, const double*, int, hipStream_t);
SMAC_INST_Z(double)
SMAC_INST_Z(float)
SMAC_INST_Z(bf16s)
void launch_finalize(const FinalizeArgs& fa, double* S, hipStream_t st) {
  if (fa.stage && fa.count > 0 && fa.nrows >= 16384) {
    // large reductions: lane sums on 16 blocks per array, then the tree
    hipLaunchKernelGGL(k_fin_lanes, dim3(16, (unsigned)fa.count), dim3(64), 0, st, fa);
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, st, fa, S);
  } else {
    FinalizeArgs f1 = fa;
    f1.stage = nullptr;
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, st, f1, S);
  }
}
void launch_scalar_copy(double* S, int dst, int src, hipStream_t st) {
  hipLaunchKernelGGL(k_scalar_copy, dim3(1), dim3(1), 0, st, S, dst, src);
}
void launch_scalar_set(double* S, int dst, double v, hipStream_t st) {
  hipLaunchKernelGGL(k_scalar_set, dim3(1), dim3(1), 0, st, S, dst, v);
}
void launch_correct_all(const Geo& g, const BetaF* bf, double* const us[3], double* dp, double* u[3], double* p,
                        double dt, hipStream_t st, const double* dv, const double* S, int slot_alpha) {
  dim3 grid((unsigned)((g.n[0] + 63) / 64), (unsigned)((g.n[1] + 3) / 4), (unsigned)g.n[2]);
  hipLaunchKernelGGL(k_correct_all, grid, CELLB, 0, st, g, bf ? *bf : BetaF{}, bf ? 1 : 0, us[0], us[1], us[2], dp,
                     u[0], u[1], u[2], p, dt, dv, S, slot_alpha);
  hipLaunchKernelGGL(k_zero_dp, rows_grid(g), ROWB, 0, st, g, dp);
}
void launch_cfl(const Geo& g, const Vel& V, double dt, double visc, const double* inv_rho, const double* mu,
                double* p0, double* p1, hipStream_t st) {
  hipLaunchKernelGGL(k_cfl, rows_grid(g), ROWB, 0, st, g, V, dt, visc, inv_rho, mu, p0, p1);
}

}  // namespace smac
