// smac_kernels.hip -- gfx950 kernels of the SMAC projection step (FP64 form).
//
//   predictor        k_predict<A>      per face of component A (MUSCL-minmod
//                                      advection + variable-mu stress + old dp/dx)
//   star BC          k_star_bc         wall faces of u*, v*, w*
//   Poisson RHS      k_rhs             b = -(1/dt) div u*   (+ row partial sums)
//   PCG vectors      k_init_r, k_add_scalar, k_jacobi, k_copy_dot, k_sum,
//                    k_spmv_dot, k_update, k_dot2, k_xpay
//   reductions       k_finalize        canonical 1024-lane tree + scalar ops
//   multigrid        k_mg_rbgs, k_mg_residual, k_mg_restrict, k_mg_prolong,
//                    k_mg_coarse, k_fill
//   correction       k_correct, k_add_dp
//
// Row kernels: one wave64 per grid row (j,k); lane l owns i = 1+l, 1+l+64, ...
// Row partial sums follow the canonical order shared with the CPU oracle.
//
// Solver cell vectors (r, z, d, Ad, multigrid e/f/res on every level) use a
// row-split layout: inside row (j,k) the even-i cells come first, then the
// odd-i cells (spos()). A red-black half sweep then streams contiguous half
// rows. Fields seen by users (p, dp, rhs) keep the natural layout.
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
template <class BC, int A, int B>
__device__ __forceinline__ double visc_cross(const Geo& g, const Vel& V, I3 q, double v0, double mu0) {
  double vbm = BC::dir(g, V, A, B, q, -1);
  double vbp = BC::dir(g, V, A, B, q, +1);
  double wpp = BC::face(g, V, B, sh(q, B, +1));
  double wmp = BC::face(g, V, B, sh(sh(q, B, +1), A, -1));
  double wp0 = BC::face(g, V, B, q);
  double wm0 = BC::face(g, V, B, sh(q, A, -1));
  double mu_p = 0.5 * (mu0 + mu0);
  double t_p = mu_p * ((wpp - wmp) * g.inv_d[A] + (vbp - v0) * g.inv_d[B]);
  double mu_m = 0.5 * (mu0 + mu0);
  double t_m = mu_m * ((wp0 - wm0) * g.inv_d[A] + (v0 - vbm) * g.inv_d[B]);
  return (t_p - t_m) * g.inv_d[B];
}

// Convective flux difference (M+ - M-)/h_E of component A along axis E,
// MUSCL-minmod with first-order upwind where the stencil leaves the interior.
// Convective flux difference (M+ - M-)/h_E of component A along axis E,
// MUSCL-minmod with first-order upwind where the stencil leaves the interior.
template <class BC, int A, int E>
__device__ __forceinline__ double conv_term(const Geo& g, const Vel& V, I3 q, double v0, double vp1, double vm1,
                                            double rho0) {
  double mp, mm, up_p, up_m;
  if (E == A) {
    mp = 0.5 * (rho0 * BC::face(g, V, A, q) + rho0 * BC::face(g, V, A, sh(q, A, +1)));
    mm = 0.5 * (rho0 * BC::face(g, V, A, sh(q, A, -1)) + rho0 * BC::face(g, V, A, q));
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
    mp = 0.5 * (rho0 * BC::face(g, V, E, sh(sh(q, E, +1), A, -1)) + rho0 * BC::face(g, V, E, sh(q, E, +1)));
    mm = 0.5 * (rho0 * BC::face(g, V, E, q) + rho0 * BC::face(g, V, E, sh(q, A, -1)));
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
// (CFD/G_SMACSolver.cu:139-392, 441-662, 763-1003) with uniform properties:
// mass flux mf = rho0*u_f (CFD/G_SMACSolver_VoF.cu:1708-1790), mu and face mu
// = mu0, beta = beta0, rho_old = rho0. Terms are accumulated in the
// reference's order: own-direction stress, cross stresses in increasing
// axis order, convection along x, y, z, old pressure gradient.
template <class BC, int A>
__device__ double predict_face(const Geo& g, const Vel& V, const double* __restrict__ p, I3 q, double dt) {
  constexpr int B1 = A == 0 ? 1 : 0;
  constexpr int B2 = A == 2 ? 1 : 2;
  const double rho0 = g.rho0, mu0 = g.mu0;
  double v0 = V.c[A][IX(g, q)];
  double vp1 = BC::face(g, V, A, sh(q, A, +1));
  double vm1 = BC::face(g, V, A, sh(q, A, -1));
  I3 cm = wr(g, sh(q, A, -1));

  double tau_p = mu0 * (vp1 - v0);
  double tau_m = mu0 * (v0 - vm1);
  double tmp = 2. * (tau_p - tau_m) * g.inv_d2[A];
  tmp += visc_cross<BC, A, B1>(g, V, q, v0, mu0);
  tmp += visc_cross<BC, A, B2>(g, V, q, v0, mu0);
  tmp -= conv_term<BC, A, 0>(g, V, q, v0, vp1, vm1, rho0);
  tmp -= conv_term<BC, A, 1>(g, V, q, v0, vp1, vm1, rho0);
  tmp -= conv_term<BC, A, 2>(g, V, q, v0, vp1, vm1, rho0);

  double beta = g.beta0;
  double rho_old_f = 0.5 * (rho0 + rho0);
  tmp -= (p[IX(g, q)] - p[IX(g, cm)]) * g.inv_d[A];
  return beta * (v0 * rho_old_f) + dt * (beta * tmp + g.g[A]);
}

// grid: x = face index along i (block 64), y = j (block 4), z = k
template <class BC, int A>
__global__ __launch_bounds__(256) void k_predict(Geo g, Vel V, const double* __restrict__ p,
                                                 double* __restrict__ ustar, double dt) {
  int lo[3] = {1, 1, 1}, hi[3] = {g.n[0], g.n[1], g.n[2]};
  lo[A] = g.per[A] ? 1 : 2;
  int i = lo[0] + (int)(blockIdx.x * 64 + threadIdx.x);
  int j = lo[1] + (int)(blockIdx.y * 4 + threadIdx.y);
  int k = lo[2] + (int)blockIdx.z;
  if (i > hi[0] || j > hi[1] || k > hi[2]) return;
  I3 q = mk(i, j, k);
  ustar[IX(g, q)] = predict_face<BC, A>(g, V, p, q, dt);
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
// Blocks are dispatched round-robin over the 8 XCDs; remap so that each XCD
// works on a contiguous slab of rows and the y/z stencil neighbours of a row
// sit in the same L2.
__device__ __forceinline__ int xcd_block(int b, int nb) {
  const int q = nb >> 3;
  if (b >= (q << 3)) return b;
  return (b & 7) * q + (b >> 3);
}

#define ROW_SETUP_N(NX, NY, NZ, PX, PY)                                        \
  const int lane = threadIdx.x & 63;                                          \
  const int row = xcd_block((int)blockIdx.x, (int)gridDim.x) * 4 + (int)(threadIdx.x >> 6); \
  if (row >= (NY) * (NZ)) return;                                             \
  const int j = row % (NY) + 1;                                               \
  const int k = row / (NY) + 1;                                               \
  const size_t rb = ((size_t)k * (size_t)(PY) + (size_t)j) * (size_t)(PX);    \
  (void)rb;

#define ROW_SETUP ROW_SETUP_N(g.n[0], g.n[1], g.n[2], g.px, g.py)
#define LROW_SETUP ROW_SETUP_N(L.n[0], L.n[1], L.n[2], L.px, L.py)

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

// r = b - mean_b (split), dp = 0 (natural); part0 <- sum r^2, part1 <- sum r
__global__ __launch_bounds__(256) void k_init_r(Geo g, const double* __restrict__ rhs, double* __restrict__ r,
                                                double* __restrict__ dp, const double* __restrict__ S, int slot_mean,
                                                double* __restrict__ part0, double* __restrict__ part1) {
  ROW_SETUP
  const int nx = g.n[0];
  const double mb = slot_mean >= 0 ? S[slot_mean] : 0.0;
  double a0 = 0.0, a1 = 0.0;
  for (int i = 1 + lane; i <= nx; i += 64) {
    double v = rhs[rb + i] - mb;
    r[rb + spos(i, nx)] = v;
    dp[rb + i] = 0.0;
    a0 += v * v;
    a1 += v;
  }
  a0 = wave_tree_sum(a0);
  a1 = wave_tree_sum(a1);
  if (lane == 0) {
    part0[row] = a0;
    part1[row] = a1;
  }
}

// x += a * S[slot] on interior cells (subtract_cell_mean, G_PressureSolverBase.cu:142-155)
__global__ __launch_bounds__(256) void k_add_scalar(Geo g, double* __restrict__ x, double a,
                                                    const double* __restrict__ S, int slot) {
  ROW_SETUP
  const double s = S[slot];
  for (int p = 1 + lane; p <= g.n[0]; p += 64) x[rb + p] += a * s;
}

// part0 <- row sums of x (split layout, canonical natural order)
__global__ __launch_bounds__(256) void k_sum(Geo g, const double* __restrict__ x, double* __restrict__ part0) {
  ROW_SETUP
  double acc = 0.0;
  for (int i = 1 + lane; i <= g.n[0]; i += 64) acc += x[rb + spos(i, g.n[0])];
  acc = wave_tree_sum(acc);
  if (lane == 0) part0[row] = acc;
}

// Jacobi preconditioner z = (1/Adiag) r (G_PCGSolver.cu:270)
__global__ __launch_bounds__(256) void k_jacobi(Geo g, const double* __restrict__ r, double* __restrict__ z) {
  ROW_SETUP
  for (int i = 1 + lane; i <= g.n[0]; i += 64) {
    Coef c = coef_uniform(g.n, g.per, g.inv_d2, g.beta0, i, j, k);
    const size_t o = rb + spos(i, g.n[0]);
    z[o] = (1. / c.diag) * r[o];
  }
}

// d = z; part0 <- sum r*z
__global__ __launch_bounds__(256) void k_copy_dot(Geo g, const double* __restrict__ z, double* __restrict__ d,
                                                  const double* __restrict__ r, double* __restrict__ part0) {
  ROW_SETUP
  double acc = 0.0;
  for (int i = 1 + lane; i <= g.n[0]; i += 64) {
    const size_t o = rb + spos(i, g.n[0]);
    double zz = z[o];
    d[o] = zz;
    acc += r[o] * zz;
  }
  acc = wave_tree_sum(acc);
  if (lane == 0) part0[row] = acc;
}

// Ad = A d, part0 <- sum d*Ad (k_make_pAp, G_PCGSolver.cu:87-124)
__global__ __launch_bounds__(256) void k_spmv_dot(Geo g, const double* __restrict__ d, double* __restrict__ Ad,
                                                  double* __restrict__ part0) {
  ROW_SETUP
  const int nx = g.n[0];
  const size_t ypr = idx(g, 0, W(g, 1, j + 1), k), ymr = idx(g, 0, W(g, 1, j - 1), k);
  const size_t zpr = idx(g, 0, j, W(g, 2, k + 1)), zmr = idx(g, 0, j, W(g, 2, k - 1));
  double acc = 0.0;
  for (int i = 1 + lane; i <= nx; i += 64) {
    Coef c = coef_uniform(g.n, g.per, g.inv_d2, g.beta0, i, j, k);
    const int p = spos(i, nx);
    double xc = d[rb + p];
    double xp = d[rb + spos(W(g, 0, i + 1), nx)], xm = d[rb + spos(W(g, 0, i - 1), nx)];
    double yp = d[ypr + p], ym = d[ymr + p], zp = d[zpr + p], zm = d[zmr + p];
    double s = c.a[0] * xp + c.a[1] * xm + c.a[2] * yp + c.a[3] * ym + c.a[4] * zp + c.a[5] * zm;
    double v = c.diag * xc - s;
    Ad[rb + p] = v;
    acc += xc * v;
  }
  acc = wave_tree_sum(acc);
  if (lane == 0) part0[row] = acc;
}

// dp += alpha d (dp natural); r -= alpha Ad (k_update_p_r, G_PCGSolver.cu:66-85);
// part0 <- sum r^2, part1 <- max |r|
__global__ __launch_bounds__(256) void k_update(Geo g, double* __restrict__ dp, const double* __restrict__ d,
                                                double* __restrict__ r, const double* __restrict__ Ad,
                                                const double* __restrict__ S, int slot_alpha,
                                                double* __restrict__ part0, double* __restrict__ part1) {
  ROW_SETUP
  const double alpha = S[slot_alpha];
  double a0 = 0.0, mx = 0.0;
  for (int i = 1 + lane; i <= g.n[0]; i += 64) {
    const size_t o = rb + spos(i, g.n[0]);
    dp[rb + i] += alpha * d[o];
    double rn = r[o] - alpha * Ad[o];
    r[o] = rn;
    a0 += rn * rn;
    double ar = fabs(rn);
    if (ar > mx || ar != ar) mx = ar;
  }
  a0 = wave_tree_sum(a0);
  mx = wave_max(mx);
  if (lane == 0) {
    part0[row] = a0;
    part1[row] = mx;
  }
}

// part0 <- sum r*z; part1 <- sum z*Ad (flexible only)
__global__ __launch_bounds__(256) void k_dot2(Geo g, const double* __restrict__ r, const double* __restrict__ z,
                                              const double* __restrict__ Ad, int flexible,
                                              double* __restrict__ part0, double* __restrict__ part1) {
  ROW_SETUP
  double a0 = 0.0, a1 = 0.0;
  for (int i = 1 + lane; i <= g.n[0]; i += 64) {
    const size_t o = rb + spos(i, g.n[0]);
    double zz = z[o];
    a0 += r[o] * zz;
    if (flexible) a1 += zz * Ad[o];
  }
  a0 = wave_tree_sum(a0);
  if (flexible) a1 = wave_tree_sum(a1);
  if (lane == 0) {
    part0[row] = a0;
    if (flexible) part1[row] = a1;
  }
}

// d = z + beta d (k_update_dir, G_PCGSolver.cu:276-288)
__global__ __launch_bounds__(256) void k_xpay(Geo g, const double* __restrict__ z, double* __restrict__ d,
                                              const double* __restrict__ S, int slot_beta) {
  ROW_SETUP
  const double beta = S[slot_beta];
  for (int p = 1 + lane; p <= g.n[0]; p += 64) d[rb + p] = z[rb + p] + beta * d[rb + p];
}

// --------------------------------------------------------------- finalize
// Reduce up to four row-partial arrays with the canonical 1024-lane tree
// (kind 0 = sum, 1 = max) into S[out[m]], then apply one scalar op.
__global__ __launch_bounds__(1024) void k_finalize(FinalizeArgs fa, double* __restrict__ S) {
  __shared__ double lds[1024];
  const int t = threadIdx.x;
  for (int m = 0; m < fa.count; m++) {
    const double* part = fa.part[m];
    const int kind = fa.kind[m];
    double acc = 0.0;
    if (kind == 0) {
      for (int r = t; r < fa.nrows; r += 1024) acc += part[r];
    } else {
      for (int r = t; r < fa.nrows; r += 1024) {
        double v = part[r];
        if (v > acc || v != v) acc = v;
      }
    }
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

// ------------------------------------------------------------- correction
// k_correct_vof_velocity (CFD/G_SMACSolver.cu:1072-1142), component A (dp natural)
template <int A>
__global__ __launch_bounds__(256) void k_correct(Geo g, const double* __restrict__ us, const double* __restrict__ dp,
                                                 double* __restrict__ u, double dt) {
  int lo[3] = {1, 1, 1}, hi[3] = {g.n[0], g.n[1], g.n[2]};
  lo[A] = g.per[A] ? 1 : 2;
  int i = lo[0] + (int)(blockIdx.x * 64 + threadIdx.x);
  int j = lo[1] + (int)(blockIdx.y * 4 + threadIdx.y);
  int k = lo[2] + (int)blockIdx.z;
  if (i > hi[0] || j > hi[1] || k > hi[2]) return;
  I3 q = mk(i, j, k);
  size_t id = IX(g, q);
  I3 cm = wr(g, sh(q, A, -1));
  double gr = (dp[id] - dp[IX(g, cm)]) * g.inv_d[A];
  u[id] = us[id] + g.beta0 * dt * (0.0 - gr);
}

// p += dp; dp = 0 on interior cells
__global__ __launch_bounds__(256) void k_add_dp(Geo g, double* __restrict__ p, double* __restrict__ dp) {
  ROW_SETUP
  for (int i = 1 + lane; i <= g.n[0]; i += 64) {
    p[rb + i] += dp[rb + i];
    dp[rb + i] = 0.0;
  }
}

// ---------------------------------------------------------------- CFL
// k_calc_cfl (CFD/G_SMACSolver.cu:1365-1407): part0 <- row max of the convective
// CFL, part1 <- row max of the viscous CFL (uniform).
__global__ __launch_bounds__(256) void k_cfl(Geo g, Vel V, double dt, double visc,
                                             double* __restrict__ part0, double* __restrict__ part1) {
  ROW_SETUP
  const size_t yp = idx(g, 0, W(g, 1, j + 1), k), zp = idx(g, 0, j, W(g, 2, k + 1));
  double m0 = 0.0;
  for (int i = 1 + lane; i <= g.n[0]; i += 64) {
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
  if (lane == 0) {
    part0[row] = m0;
    part1[row] = visc;
  }
}

// ------------------------------------------------------------- multigrid
// Level geometry with uniform coefficients beta0/h_l^2: the coarse face
// coefficient is the mean of the four fine faces covering it, which keeps
// beta0, and walls stay 0. All level arrays use the row-split layout.
__device__ __forceinline__ size_t lrow(const Lv& L, int j, int k) {
  return ((size_t)k * (size_t)L.py + (size_t)j) * (size_t)L.px;
}
__device__ __forceinline__ int LW(const Lv& L, int d, int i) { return L.per[d] ? wrapi(i, L.n[d]) : i; }

// Red-black Gauss-Seidel half sweep of colour col ((i+j+k)&1 == col) on the
// active half row: e = (f + S)/diag. zero_start: e is zero on entry, so the
// half sweep is e = f/diag without reading neighbours.
__global__ __launch_bounds__(256) void k_mg_rbgs(Lv L, double* __restrict__ e, const double* __restrict__ f, int col,
                                                 int zero_start) {
  LROW_SETUP
  const int nx = L.n[0], nE = nx >> 1;
  const int par = (col + j + k) & 1;  // parity of the active i
  const int cnt = par ? nx - nE : nE;
  const int off = par ? nE : 0;
  const size_t ypr = lrow(L, LW(L, 1, j + 1), k), ymr = lrow(L, LW(L, 1, j - 1), k);
  const size_t zpr = lrow(L, j, LW(L, 2, k + 1)), zmr = lrow(L, j, LW(L, 2, k - 1));
  for (int h = 1 + lane; h <= cnt; h += 64) {
    const int i = par ? 2 * h - 1 : 2 * h;
    const int p = off + h;
    Coef c = coef_uniform(L.n, L.per, L.inv_d2, L.beta0, i, j, k);
    if (zero_start) {
      e[rb + p] = f[rb + p] / c.diag;
    } else {
      double xp = e[rb + spos(LW(L, 0, i + 1), nx)], xm = e[rb + spos(LW(L, 0, i - 1), nx)];
      double yp = e[ypr + p], ym = e[ymr + p], zp = e[zpr + p], zm = e[zmr + p];
      double S = c.a[0] * xp + c.a[1] * xm + c.a[2] * yp + c.a[3] * ym + c.a[4] * zp + c.a[5] * zm;
      e[rb + p] = (f[rb + p] + S) / c.diag;
    }
  }
}

// res = f - (diag e - S) on every cell of the level
__global__ __launch_bounds__(256) void k_mg_residual(Lv L, const double* __restrict__ e, const double* __restrict__ f,
                                                     double* __restrict__ res) {
  LROW_SETUP
  const int nx = L.n[0], nE = nx >> 1;
  const size_t ypr = lrow(L, LW(L, 1, j + 1), k), ymr = lrow(L, LW(L, 1, j - 1), k);
  const size_t zpr = lrow(L, j, LW(L, 2, k + 1)), zmr = lrow(L, j, LW(L, 2, k - 1));
  for (int p = 1 + lane; p <= nx; p += 64) {
    const int i = p <= nE ? 2 * p : 2 * (p - nE) - 1;
    Coef c = coef_uniform(L.n, L.per, L.inv_d2, L.beta0, i, j, k);
    double xp = e[rb + spos(LW(L, 0, i + 1), nx)], xm = e[rb + spos(LW(L, 0, i - 1), nx)];
    double yp = e[ypr + p], ym = e[ymr + p], zp = e[zpr + p], zm = e[zmr + p];
    double S = c.a[0] * xp + c.a[1] * xm + c.a[2] * yp + c.a[3] * ym + c.a[4] * zp + c.a[5] * zm;
    res[rb + p] = f[rb + p] - (c.diag * e[rb + p] - S);
  }
}

// 1-D restriction stencil of coarse index I (transpose of cell-centred linear
// prolongation; at a wall the mirrored weight folds onto the boundary cell)
__device__ __forceinline__ int rweights(int per, int nc, int I, int ix[4], double w[4]) {
  int nf = 2 * nc;
  if (per) {
    ix[0] = wrapi(2 * I - 2, nf); w[0] = 0.25;
    ix[1] = 2 * I - 1; w[1] = 0.75;
    ix[2] = 2 * I; w[2] = 0.75;
    ix[3] = wrapi(2 * I + 1, nf); w[3] = 0.25;
    return 4;
  }
  int m = 0;
  if (I > 1) { ix[m] = 2 * I - 2; w[m] = 0.25; m++; }
  ix[m] = 2 * I - 1; w[m] = (I == 1) ? 1.0 : 0.75; m++;
  ix[m] = 2 * I; w[m] = (I == nc) ? 1.0 : 0.75; m++;
  if (I < nc) { ix[m] = 2 * I + 1; w[m] = 0.25; m++; }
  return m;
}

// Restriction, separable: coarse f = (1/8) sum_z wz sum_y wy sum_x wx res
// over the folded 4x4x4 fine neighbourhood. Pass 1 (k_mg_resid_rx) forms the
// residual of one fine row in LDS and restricts it along x into T (coarse x,
// fine y and z, natural order, pitch C.n[0]+3). Pass 2 (k_mg_ryz) restricts
// T along y and z. T lives in the fine level's residual buffer.
__global__ __launch_bounds__(256) void k_mg_resid_rx(Lv F, Lv C, const double* __restrict__ e,
                                                     const double* __restrict__ f, double* __restrict__ T) {
  extern __shared__ double sm[];
  const Lv& L = F;
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int row = xcd_block((int)blockIdx.x, (int)gridDim.x) * 4 + w;
  const bool valid = row < L.n[1] * L.n[2];
  const int nx = L.n[0], nE = nx >> 1;
  double* buf = sm + (size_t)w * (size_t)(nx + 2);
  int j = 1, k = 1;
  if (valid) {
    j = row % L.n[1] + 1;
    k = row / L.n[1] + 1;
    const size_t rb = lrow(L, j, k);
    const size_t ypr = lrow(L, LW(L, 1, j + 1), k), ymr = lrow(L, LW(L, 1, j - 1), k);
    const size_t zpr = lrow(L, j, LW(L, 2, k + 1)), zmr = lrow(L, j, LW(L, 2, k - 1));
    for (int p = 1 + lane; p <= nx; p += 64) {
      const int i = p <= nE ? 2 * p : 2 * (p - nE) - 1;
      Coef c = coef_uniform(L.n, L.per, L.inv_d2, L.beta0, i, j, k);
      double xp = e[rb + spos(LW(L, 0, i + 1), nx)], xm = e[rb + spos(LW(L, 0, i - 1), nx)];
      double yp = e[ypr + p], ym = e[ymr + p], zp = e[zpr + p], zm = e[zmr + p];
      double S = c.a[0] * xp + c.a[1] * xm + c.a[2] * yp + c.a[3] * ym + c.a[4] * zp + c.a[5] * zm;
      buf[i] = f[rb + p] - (c.diag * e[rb + p] - S);
    }
  }
  __syncthreads();
  if (!valid) return;
  const size_t tb = ((size_t)k * (size_t)L.py + (size_t)j) * (size_t)(C.n[0] + 3);
  for (int I = 1 + lane; I <= C.n[0]; I += 64) {
    int ix[4];
    double wx[4];
    int mx = rweights(C.per[0], C.n[0], I, ix, wx);
    double tx = 0.0;
    for (int c = 0; c < mx; c++) tx += wx[c] * buf[ix[c]];
    T[tb + I] = tx;
  }
}

__global__ __launch_bounds__(256) void k_mg_ryz(Lv F, Lv C, const double* __restrict__ T, double* __restrict__ fc) {
  const int I = 1 + (int)(blockIdx.x * 64 + threadIdx.x);
  const int J = 1 + (int)(blockIdx.y * 4 + threadIdx.y);
  const int K = 1 + (int)blockIdx.z;
  if (I > C.n[0] || J > C.n[1] || K > C.n[2]) return;
  int iy[4], iz[4];
  double wy[4], wz[4];
  int my = rweights(C.per[1], C.n[1], J, iy, wy);
  int mz = rweights(C.per[2], C.n[2], K, iz, wz);
  const size_t pxT = (size_t)(C.n[0] + 3);
  double tz = 0.0;
  for (int a = 0; a < mz; a++) {
    double ty = 0.0;
    for (int b = 0; b < my; b++) ty += wy[b] * T[((size_t)iz[a] * (size_t)F.py + (size_t)iy[b]) * pxT + I];
    tz += wz[a] * ty;
  }
  fc[lrow(C, J, K) + spos(I, C.n[0])] = 0.125 * tz;
}

__device__ __forceinline__ void pweights(int per, int nc, int i, int ix[2], double w[2]) {
  int par = (i + 1) >> 1;
  int nb = (i & 1) ? par - 1 : par + 1;
  if (per) nb = wrapi(nb, nc);
  else if (nb < 1 || nb > nc) nb = par;
  ix[0] = par; w[0] = 0.75;
  ix[1] = nb; w[1] = 0.25;
}

// e_f += sum over 8 coarse cells of (wz*wy*wx) e_c (cell-centred trilinear), on
// the fine cells of colour col only: the other colour is overwritten by the
// next half sweep without being read.
__global__ __launch_bounds__(256) void k_mg_prolong(Lv C, Lv F, const double* __restrict__ ec, double* __restrict__ ef,
                                                    int col) {
  const Lv& L = F;
  LROW_SETUP
  const int nx = L.n[0], nE = nx >> 1;
  // col >= 0: the active half row of that colour; col < 0: the whole row
  const int par = (col + j + k) & 1;
  const int cnt = col < 0 ? nx : (par ? nx - nE : nE);
  const int off = (col < 0 || !par) ? 0 : nE;
  int iy[2], iz[2];
  double wy[2], wz[2];
  pweights(C.per[1], C.n[1], j, iy, wy);
  pweights(C.per[2], C.n[2], k, iz, wz);
  size_t cr[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) cr[a][b] = lrow(C, iy[b], iz[a]);
  for (int h = 1 + lane; h <= cnt; h += 64) {
    const int pp = off + h;
    const int i = pp <= nE ? 2 * pp : 2 * (pp - nE) - 1;
    int ix[2];
    double wx[2];
    pweights(C.per[0], C.n[0], i, ix, wx);
    const int c0 = spos(ix[0], C.n[0]), c1 = spos(ix[1], C.n[0]);
    double acc = 0.0;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++) {
        double w0 = wz[a] * wy[b] * wx[0];
        acc += w0 * ec[cr[a][b] + c0];
        double w1 = wz[a] * wy[b] * wx[1];
        acc += w1 * ec[cr[a][b] + c1];
      }
    ef[rb + pp] += acc;
  }
}

// coarsest level: e = Ainv f with the dense pseudo-inverse (rows in natural
// x-fastest order), sequential accumulation per row
__global__ __launch_bounds__(256) void k_mg_coarse(Lv C, const double* __restrict__ ainv, const double* __restrict__ f,
                                                   double* __restrict__ e) {
  const int n = C.n[0] * C.n[1] * C.n[2];
  const int row = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (row >= n) return;
  const double* a = ainv + (size_t)row * (size_t)n;
  double acc = 0.0;
  for (int col = 0; col < n; col++) {
    int ci = col % C.n[0] + 1, cj = (col / C.n[0]) % C.n[1] + 1, ck = col / (C.n[0] * C.n[1]) + 1;
    acc += a[col] * f[lrow(C, cj, ck) + spos(ci, C.n[0])];
  }
  int i = row % C.n[0] + 1, j = (row / C.n[0]) % C.n[1] + 1, k = row / (C.n[0] * C.n[1]) + 1;
  e[lrow(C, j, k) + spos(i, C.n[0])] = acc;
}

// x = v on interior cells of a level (a fill kernel, safe under graph capture)
__global__ __launch_bounds__(256) void k_fill(Lv L, double* __restrict__ x, double v) {
  LROW_SETUP
  for (int p = 1 + lane; p <= L.n[0]; p += 64) x[rb + p] = v;
}

// ---------------------------------------------------------------- launchers
static inline dim3 rows_grid(const Geo& g) { return dim3((unsigned)((g.n[1] * g.n[2] + 3) / 4)); }
static inline dim3 lrows_grid(const Lv& L) { return dim3((unsigned)((L.n[1] * L.n[2] + 3) / 4)); }
static inline dim3 cell_grid(const Lv& L) {
  return dim3((unsigned)((L.n[0] + 63) / 64), (unsigned)((L.n[1] + 3) / 4), (unsigned)L.n[2]);
}
static const dim3 ROWB(256), CELLB(64, 4, 1);

void launch_predict(const Geo& g, const Vel& V, const double* p, double* us[3], double dt, hipStream_t st) {
  for (int a = 0; a < 3; a++) {
    int lo = g.per[a] ? 1 : 2;
    int nf[3] = {g.n[0], g.n[1], g.n[2]};
    nf[a] = g.n[a] - lo + 1;
    dim3 grid((unsigned)((nf[0] + 63) / 64), (unsigned)((nf[1] + 3) / 4), (unsigned)nf[2]);
    if (nf[0] <= 0 || nf[1] <= 0 || nf[2] <= 0) continue;
    const bool allp = g.per[0] && g.per[1] && g.per[2];
    if (allp) {
      if (a == 0) hipLaunchKernelGGL((k_predict<PeriodicBC, 0>), grid, CELLB, 0, st, g, V, p, us[0], dt);
      if (a == 1) hipLaunchKernelGGL((k_predict<PeriodicBC, 1>), grid, CELLB, 0, st, g, V, p, us[1], dt);
      if (a == 2) hipLaunchKernelGGL((k_predict<PeriodicBC, 2>), grid, CELLB, 0, st, g, V, p, us[2], dt);
    } else {
      if (a == 0) hipLaunchKernelGGL((k_predict<WallBC, 0>), grid, CELLB, 0, st, g, V, p, us[0], dt);
      if (a == 1) hipLaunchKernelGGL((k_predict<WallBC, 1>), grid, CELLB, 0, st, g, V, p, us[1], dt);
      if (a == 2) hipLaunchKernelGGL((k_predict<WallBC, 2>), grid, CELLB, 0, st, g, V, p, us[2], dt);
    }
  }
  if (g.per[0] && g.per[1] && g.per[2]) return;
  int m = g.n[0] > g.n[1] ? g.n[0] : g.n[1];
  m = m > g.n[2] ? m : g.n[2];
  dim3 grid((unsigned)((m + 63) / 64), (unsigned)((m + 3) / 4), 6);
  hipLaunchKernelGGL(k_star_bc, grid, CELLB, 0, st, g, us[0], us[1], us[2]);
}

void launch_rhs(const Geo& g, double* const us[3], double* rhs, double inv_dt, double* part0, hipStream_t st) {
  hipLaunchKernelGGL(k_rhs, rows_grid(g), ROWB, 0, st, g, us[0], us[1], us[2], rhs, inv_dt, part0);
}
void launch_init_r(const Geo& g, const double* rhs, double* r, double* dp, const double* S, int slot_mean,
                   double* p0, double* p1, hipStream_t st) {
  hipLaunchKernelGGL(k_init_r, rows_grid(g), ROWB, 0, st, g, rhs, r, dp, S, slot_mean, p0, p1);
}
void launch_add_scalar(const Geo& g, double* x, double a, const double* S, int slot, hipStream_t st) {
  hipLaunchKernelGGL(k_add_scalar, rows_grid(g), ROWB, 0, st, g, x, a, S, slot);
}
void launch_sum(const Geo& g, const double* x, double* p0, hipStream_t st) {
  hipLaunchKernelGGL(k_sum, rows_grid(g), ROWB, 0, st, g, x, p0);
}
void launch_jacobi(const Geo& g, const double* r, double* z, hipStream_t st) {
  hipLaunchKernelGGL(k_jacobi, rows_grid(g), ROWB, 0, st, g, r, z);
}
void launch_copy_dot(const Geo& g, const double* z, double* d, const double* r, double* p0, hipStream_t st) {
  hipLaunchKernelGGL(k_copy_dot, rows_grid(g), ROWB, 0, st, g, z, d, r, p0);
}
void launch_spmv_dot(const Geo& g, const double* d, double* Ad, double* p0, hipStream_t st) {
  hipLaunchKernelGGL(k_spmv_dot, rows_grid(g), ROWB, 0, st, g, d, Ad, p0);
}
void launch_update(const Geo& g, double* dp, const double* d, double* r, const double* Ad, const double* S,
                   int slot_alpha, double* p0, double* p1, hipStream_t st) {
  hipLaunchKernelGGL(k_update, rows_grid(g), ROWB, 0, st, g, dp, d, r, Ad, S, slot_alpha, p0, p1);
}
void launch_dot2(const Geo& g, const double* r, const double* z, const double* Ad, int flexible, double* p0,
                 double* p1, hipStream_t st) {
  hipLaunchKernelGGL(k_dot2, rows_grid(g), ROWB, 0, st, g, r, z, Ad, flexible, p0, p1);
}
void launch_xpay(const Geo& g, const double* z, double* d, const double* S, int slot_beta, hipStream_t st) {
  hipLaunchKernelGGL(k_xpay, rows_grid(g), ROWB, 0, st, g, z, d, S, slot_beta);
}
void launch_finalize(const FinalizeArgs& fa, double* S, hipStream_t st) {
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, st, fa, S);
}
void launch_scalar_copy(double* S, int dst, int src, hipStream_t st) {
  hipLaunchKernelGGL(k_scalar_copy, dim3(1), dim3(1), 0, st, S, dst, src);
}
void launch_correct(const Geo& g, double* const us[3], const double* dp, double* u[3], double dt, hipStream_t st) {
  for (int a = 0; a < 3; a++) {
    int lo = g.per[a] ? 1 : 2;
    int nf[3] = {g.n[0], g.n[1], g.n[2]};
    nf[a] = g.n[a] - lo + 1;
    if (nf[0] <= 0 || nf[1] <= 0 || nf[2] <= 0) continue;
    dim3 grid((unsigned)((nf[0] + 63) / 64), (unsigned)((nf[1] + 3) / 4), (unsigned)nf[2]);
    if (a == 0) hipLaunchKernelGGL(k_correct<0>, grid, CELLB, 0, st, g, us[0], dp, u[0], dt);
    if (a == 1) hipLaunchKernelGGL(k_correct<1>, grid, CELLB, 0, st, g, us[1], dp, u[1], dt);
    if (a == 2) hipLaunchKernelGGL(k_correct<2>, grid, CELLB, 0, st, g, us[2], dp, u[2], dt);
  }
}
void launch_add_dp(const Geo& g, double* p, double* dp, hipStream_t st) {
  hipLaunchKernelGGL(k_add_dp, rows_grid(g), ROWB, 0, st, g, p, dp);
}
void launch_cfl(const Geo& g, const Vel& V, double dt, double visc, double* p0, double* p1, hipStream_t st) {
  hipLaunchKernelGGL(k_cfl, rows_grid(g), ROWB, 0, st, g, V, dt, visc, p0, p1);
}
void launch_mg_rbgs(const Lv& L, double* e, const double* f, int col, int zero_start, hipStream_t st) {
  hipLaunchKernelGGL(k_mg_rbgs, lrows_grid(L), ROWB, 0, st, L, e, f, col, zero_start);
}
void launch_mg_residual(const Lv& L, const double* e, const double* f, double* res, hipStream_t st) {
  hipLaunchKernelGGL(k_mg_residual, lrows_grid(L), ROWB, 0, st, L, e, f, res);
}
void launch_mg_restrict(const Lv& F, const Lv& C, const double* e, const double* f, double* T, double* fc,
                        hipStream_t st) {
  size_t shm = 4 * (size_t)(F.n[0] + 2) * sizeof(double);
  hipLaunchKernelGGL(k_mg_resid_rx, lrows_grid(F), ROWB, shm, st, F, C, e, f, T);
  hipLaunchKernelGGL(k_mg_ryz, cell_grid(C), CELLB, 0, st, F, C, T, fc);
}
void launch_mg_prolong(const Lv& C, const Lv& F, const double* ec, double* ef, int col, hipStream_t st) {
  hipLaunchKernelGGL(k_mg_prolong, lrows_grid(F), ROWB, 0, st, C, F, ec, ef, col);
}
void launch_mg_coarse(const Lv& C, const double* ainv, const double* f, double* e, hipStream_t st) {
  int n = C.n[0] * C.n[1] * C.n[2];
  hipLaunchKernelGGL(k_mg_coarse, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, C, ainv, f, e);
}
void launch_fill(const Lv& L, double* x, double v, hipStream_t st) {
  hipLaunchKernelGGL(k_fill, lrows_grid(L), ROWB, 0, st, L, x, v);
}

}  // namespace smac
