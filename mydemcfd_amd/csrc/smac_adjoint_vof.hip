// smac_adjoint_vof.hip -- adjoint (vector-Jacobian product) of one step of the
// VoF loop: the split THINC/WLIC transport of alpha (smac_transport_alpha)
// followed by the SMAC step whose properties, mass flux and Poisson operator
// follow alpha (smac_step). The velocity part is smac_adjoint.hip's; this file
// holds the alpha-dependent parts:
//
//   properties   rho = (1-a) rho0 + a rho1, mu likewise (cells); beta_f =
//                1/(0.5 (rho_c + rho_cm)) on interior faces; rho_old = rho(a^n)
//   correction   u^{n+1}_f = u*_f + beta_f dt (0 - G dp_f)
//                  -> beta_bar_f = -dt G dp_f ub_f
//   solve        A(beta) dp = b, bb = A^+ (cotangent of dp)
//                  -> beta_bar_f = -G bb_f G dp_f
//   transport    per substep and sweep (axis A): F_f = flux(alpha, u_A f),
//                  alpha' = alpha - (F+ - F-) + c (u+ - u-) dt/h, clipped
//                  after the third sweep (c = [alpha > 0.5] at the substep
//                  start: piecewise constant, no derivative)
//
// The flux derivative follows the branch the forward took (upwind donor,
// first-order or THINC/WLIC, clip): d/d lambda and d/d xi of the THINC
// integral 0.5 (b - a) + gamma/(2 beta) (ln cosh beta(b - xi) - ln cosh beta(a - xi))
// are 0.5 + gamma/2 tanh(beta (b - xi)) and gamma/2 (tanh beta(a - xi) -
// tanh beta(b - xi)); xi0 = ln((e^beta - A)/(A - e^-beta))/(2 beta) with
// A = exp(beta gamma (2 alpha - 1)); the WLIC weight w = 1 - (2 theta/pi)^2,
// theta = acos(x), x = |n_A|/(|n| + EPS), has dw/dx = (8/pi^2) theta/sin theta
// (-> 8/pi^2 as the normal aligns with the axis). Derivatives use the math
// library's functions (the forward's explicit sequences agree with them to a
// few ulp). Scatters into alpha cotangents are FP64 atomics.
#include "smac_internal.hpp"
#include "smac_vof.hpp"

namespace smac {

namespace {

constexpr double WLIC_DW = 8.0 / (VOF_PI * VOF_PI);

__device__ __forceinline__ void aadd2(double* base, size_t i, double v) {
  if (v != 0.0) unsafeAtomicAdd(base + i, v);
}

// beta cotangent of interior face f (axis A, between cells cm and c) from the
// correction and the solve; 0 on other faces
template <int A>
__device__ __forceinline__ double face_beta_bar(const Geo& g, const double* __restrict__ ub,
                                                const double* __restrict__ bb, const double* __restrict__ dp,
                                                I3 c, double dt) {
  if (!g.per[A] && (c.v[A] < 2 || c.v[A] > g.n[A])) return 0.0;
  const size_t ic = IX(g, c), im = IX(g, wr(g, sh(c, A, -1)));
  const double gdp = (dp[ic] - dp[im]) * g.inv_d[A];
  const double gbb = (bb[ic] - bb[im]) * g.inv_d[A];
  return -gdp * (dt * ub[ic] + gbb);
}

// abar_c += sum over the interior faces f of cell c of drho (-0.5 beta_f^2) beta_bar_f
// (beta_f = 1/(0.5 (rho_c + rho_cm))), gathered per cell
__global__ __launch_bounds__(256) void k_adj_beta_cells(Geo g, BetaF bf, const double* __restrict__ ub0,
                                                        const double* __restrict__ ub1,
                                                        const double* __restrict__ ub2,
                                                        const double* __restrict__ bb,
                                                        const double* __restrict__ dp, double dt, double drho,
                                                        double* __restrict__ abar) {
  const int i = 1 + (int)(blockIdx.x * 64 + threadIdx.x);
  const int j = 1 + (int)(blockIdx.y * 4 + threadIdx.y);
  const int k = 1 + (int)blockIdx.z;
  if (i > g.n[0] || j > g.n[1]) return;
  const I3 c = mk(i, j, k);
  double acc = 0.0;
  {
    const I3 cp = wr(g, sh(c, 0, +1));
    const double bm = bf.b[0][IX(g, c)], bp = bf.b[0][IX(g, cp)];
    acc += -0.5 * bm * bm * face_beta_bar<0>(g, ub0, bb, dp, c, dt);
    if (g.per[0] || i + 1 <= g.n[0]) acc += -0.5 * bp * bp * face_beta_bar<0>(g, ub0, bb, dp, cp, dt);
  }
  {
    const I3 cp = wr(g, sh(c, 1, +1));
    const double bm = bf.b[1][IX(g, c)], bp = bf.b[1][IX(g, cp)];
    acc += -0.5 * bm * bm * face_beta_bar<1>(g, ub1, bb, dp, c, dt);
    if (g.per[1] || j + 1 <= g.n[1]) acc += -0.5 * bp * bp * face_beta_bar<1>(g, ub1, bb, dp, cp, dt);
  }
  {
    const I3 cp = wr(g, sh(c, 2, +1));
    const double bm = bf.b[2][IX(g, c)], bp = bf.b[2][IX(g, cp)];
    acc += -0.5 * bm * bm * face_beta_bar<2>(g, ub2, bb, dp, c, dt);
    if (g.per[2] || k + 1 <= g.n[2]) acc += -0.5 * bp * bp * face_beta_bar<2>(g, ub2, bb, dp, cp, dt);
  }
  abar[IX(g, c)] += drho * acc;
}

// Does get_face(A, f) read the stored velocity (so that its cotangent
// reaches u^n)? Interior faces; boundary faces of outlet kind, or of slip
// kind off the x axis (CFD/G_BoundaryFunctions.h:9-97).
__device__ __forceinline__ bool face_tracked(const Geo& g, int A, I3 f) {
  f = wr(g, f);
  const int t = face_type(g, A, f);
  if (t == F_INTERIOR) return true;
  if (t == F_BOUNDARY) {
    const int bt = btype_of(g, face_slot(g, A, f));
    return bt == BC_OUTLET || (bt == BC_SLIP && A != 0);
  }
  return false;
}

// Per interior cell of one sweep (axis A): the cotangent of the cell's value
// before the clip (CLIP: zero where the forward clipped), written to apre and
// to acur (the cotangent of the sweep's input, which the face kernel then
// completes). The pre-clip value is formed as k_vof_transport forms it.
template <int A, int CLIP>
__global__ __launch_bounds__(256) void k_vof_adj_cell(Geo g, Vel V, const double* __restrict__ al,
                                                      const double* __restrict__ F,
                                                      const unsigned char* __restrict__ coeff,
                                                      const double* __restrict__ aout, double* __restrict__ apre,
                                                      double* __restrict__ acur, double dt) {
  const int i = 1 + (int)(blockIdx.x * 64 + threadIdx.x);
  const int j = 1 + (int)(blockIdx.y * 4 + threadIdx.y);
  const int k = 1 + (int)blockIdx.z;
  if (i > g.n[0] || j > g.n[1]) return;
  const I3 q = mk(i, j, k);
  const size_t id = IX(g, q);
  double b = aout[id];
  if (CLIP) {
    double flux = 0.;
    flux += F[IX(g, wr(g, sh(q, A, +1)))] - F[id];
    const double up = get_face(g, V, A, sh(q, A, +1));
    const double um = get_face(g, V, A, q);
    const double cf = (double)coeff[id];
    flux -= cf * (up - um) * dt * g.inv_d[A];
    const double a = al[id] - flux;
    if (a < VOF_EPS || a > 1.) b = 0.0;
  }
  apre[id] = b;
  acur[id] = b;
}

// adjoint of the face flux of k_vof_flux<A> at face q with cotangent Fb:
// alpha cotangents scattered into acur, the velocity's returned
template <int A>
__device__ double vof_flux_vjp(const Geo& g, const Spacing& H, const double* __restrict__ u,
                               const double* __restrict__ al, I3 q, double dt, double Fb,
                               double* __restrict__ acur) {
  const size_t id = IX(g, q);
  const double v = u[id];
  const I3 dn = wr(g, v > 0.0 ? sh(q, A, -1) : q);
  const size_t idn = IX(g, dn);
  const double ad = al[idn];
  const double dtbyd = dt * g.inv_d[A];
  const I3 dp = sh(dn, A, +1), dm = sh(dn, A, -1);
  const bool ip = g.per[A] || dp.v[A] <= g.n[A];
  const bool im = g.per[A] || dm.v[A] >= 1;
  double gam = 0.0;
  bool upwind = !ip || !im;
  if (!upwind) {
    gam = al[IX(g, wr(g, dp))] - al[IX(g, wr(g, dm))];
    upwind = ad < VOF_EPS || ad > 1.0 - VOF_EPS || fabs(gam) < 1e-6;
  }
  if (upwind) {  // F = v ad dt/h
    aadd2(acur, idn, Fb * v * dtbyd);
    return Fb * ad * dtbyd;
  }
  constexpr int T1 = A == 0 ? 1 : 0, T2 = A == 2 ? 1 : 2;
  double nrm[3];
  nrm[A] = -gam * (1. / (2. * H.h[A]));
  nrm[T1] = vof_normal_t<T1>(g, H, al, dn);
  nrm[T2] = vof_normal_t<T2>(g, H, al, dn);
  const double s_sq = nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2];
  const double sn = sqrt(s_sq);
  const double inv_s = 1.0 / (sn + VOF_EPS);
  const double x = fabs(nrm[A]) * inv_s;
  const double theta = vm_acos01(x);
  const double wc = 2. * theta / VOF_PI;
  const double w = 1.0 - wc * wc;
  const double gamma = gam < 0 ? -1. : 1.;
  const double xi0 = thinc_xi0(ad, gamma);
  const double lambda = v * dtbyd;
  const double ft =
      v > 0.0 ? thinc_integral(1.0 - lambda, 1.0, gamma, xi0) : -thinc_integral(0.0, -lambda, gamma, xi0);
  const double fu = lambda * ad;
  // F = w ft + (1 - w) fu
  const double wb = Fb * (ft - fu), ftb = Fb * w, fub = Fb * (1. - w);
  double lb = fub * ad;
  double adb = fub * lambda;
  double xib;
  const double B = THINC_BETA;
  if (v > 0.0) {  // ft = I(1 - lambda, 1)
    const double ta = tanh(B * ((1.0 - lambda) - xi0)), tb = tanh(B * (1.0 - xi0));
    lb += ftb * (0.5 + 0.5 * gamma * ta);
    xib = ftb * (0.5 * gamma * (ta - tb));
  } else {  // ft = -I(0, -lambda)
    const double ta = tanh(B * (0.0 - xi0)), tb = tanh(B * (-lambda - xi0));
    lb += ftb * (0.5 + 0.5 * gamma * tb);
    xib = -ftb * (0.5 * gamma * (ta - tb));
  }
  {  // xi0(alpha): A_ = exp(B gamma (2 alpha - 1))
    const double Ax = exp(B * gamma * (2. * ad - 1.));
    const double dxi = (1. / (2. * B)) * (-1. / (exp(B) - Ax) - 1. / (Ax - exp(-B)));
    adb += xib * dxi * (2. * B * gamma * Ax);
  }
  // WLIC weight through the normal
  const double st = sqrt(fmax(0.0, 1.0 - x * x));
  const double r = st > 1e-12 ? theta / st : 1.0;
  const double xb = wb * WLIC_DW * r;
  if (xb != 0.0 && sn > 0.0) {
    const double an = fabs(nrm[A]);
    const double k2 = an * inv_s * inv_s / sn;
    double nb[3];
    nb[A] = xb * ((nrm[A] < 0 ? -1. : 1.) * inv_s - k2 * nrm[A]);
    nb[T1] = xb * (-k2 * nrm[T1]);
    nb[T2] = xb * (-k2 * nrm[T2]);
    // n_A = -gam/(2 h_A), gam = alpha(dn + e_A) - alpha(dn - e_A)
    const double gb = -nb[A] * (1. / (2. * H.h[A]));
    aadd2(acur, IX(g, wr(g, dp)), gb);
    aadd2(acur, IX(g, wr(g, dm)), -gb);
    // tangential components: the branch of vof_normal_t
#pragma unroll
    for (int t = 0; t < 2; t++) {
      const int T = t == 0 ? T1 : T2;
      const double n = nb[T];
      if (n == 0.0) continue;
      const I3 qp = sh(dn, T, +1), qm = sh(dn, T, -1);
      const bool tp = g.per[T] || (qp.v[T] >= 1 && qp.v[T] <= g.n[T]);
      const bool tm = g.per[T] || (qm.v[T] >= 1 && qm.v[T] <= g.n[T]);
      if (tp && tm) {
        const double c = n * (1. / (2. * H.h[T]));
        aadd2(acur, IX(g, wr(g, qp)), -c);
        aadd2(acur, IX(g, wr(g, qm)), c);
      } else if (!tp && tm) {
        const double c = n * g.inv_d[T];
        aadd2(acur, idn, -c);
        aadd2(acur, IX(g, wr(g, qm)), c);
      } else if (tp && !tm) {
        const double c = n * g.inv_d[T];
        aadd2(acur, IX(g, wr(g, qp)), -c);
        aadd2(acur, idn, c);
      }
    }
  }
  aadd2(acur, idn, adb);
  return lb * dtbyd;
}

// Per A-face of one sweep: the velocity cotangent of the divergence
// correction c (u+ - u-) dt/h of its two cells, and, on the faces whose flux
// the forward computed (2..N, 1..N periodic), the flux adjoint with
// Fb = apre(cell f) - apre(cell f-1) + Fbacc (the accumulated flux's cotangent).
template <int A>
__global__ __launch_bounds__(256) void k_vof_adj_face(Geo g, Spacing H, const double* __restrict__ u,
                                                      const double* __restrict__ al,
                                                      const unsigned char* __restrict__ coeff,
                                                      const double* __restrict__ apre,
                                                      const double* __restrict__ Fbacc, double* __restrict__ acur,
                                                      double* __restrict__ ubar, double dt) {
  int hi[3] = {g.n[0], g.n[1], g.n[2]};
  if (!g.per[A]) hi[A] += 1;
  const int i = 1 + (int)(blockIdx.x * 64 + threadIdx.x);
  const int j = 1 + (int)(blockIdx.y * 4 + threadIdx.y);
  const int k = 1 + (int)blockIdx.z;
  if (i > hi[0] || j > hi[1] || k > hi[2]) return;
  const I3 q = mk(i, j, k);
  const size_t id = IX(g, q);
  const int f = q.v[A];
  const bool hasm = g.per[A] || f >= 2;        // cell f-1 is interior
  const bool hasp = g.per[A] || f <= g.n[A];   // cell f is interior
  const I3 cmq = wr(g, sh(q, A, -1));
  const size_t icm = IX(g, cmq);
  const double dtd = dt * g.inv_d[A];
  double ub = 0.0;
  if (face_tracked(g, A, q)) {
    // cell f-1: alpha' -= -c (u+ ...) -> +c dt/h on its plus face; cell f: -c dt/h
    if (hasm) ub += (double)coeff[icm] * dtd * apre[icm];
    if (hasp) ub -= (double)coeff[id] * dtd * apre[id];
  }
  const bool computed = g.per[A] || (f >= 2 && f <= g.n[A]);
  if (computed) {
    double Fb = apre[id] - apre[icm];
    if (Fbacc) Fb += Fbacc[id];
    if (Fb != 0.0) ub += vof_flux_vjp<A>(g, H, u, al, q, dt, Fb, acur);
  }
  if (ub != 0.0) ubar[id] += ub;
}

template <int A>
void adj_sweep(const Geo& g, const Spacing& H, const Vel& V, const double* al, const double* F,
               const unsigned char* coeff, const double* aout, double* apre, double* acur, const double* Fbacc,
               double* ubar, double dt, bool clip, hipStream_t st) {
  dim3 cg((unsigned)((g.n[0] + 63) / 64), (unsigned)((g.n[1] + 3) / 4), (unsigned)g.n[2]);
  if (clip) hipLaunchKernelGGL((k_vof_adj_cell<A, 1>), cg, dim3(64, 4), 0, st, g, V, al, F, coeff, aout, apre, acur, dt);
  else hipLaunchKernelGGL((k_vof_adj_cell<A, 0>), cg, dim3(64, 4), 0, st, g, V, al, F, coeff, aout, apre, acur, dt);
  int hi[3] = {g.n[0], g.n[1], g.n[2]};
  if (!g.per[A]) hi[A] += 1;
  dim3 fg((unsigned)((hi[0] + 63) / 64), (unsigned)((hi[1] + 3) / 4), (unsigned)hi[2]);
  hipLaunchKernelGGL(k_vof_adj_face<A>, fg, dim3(64, 4), 0, st, g, H, V.c[A], al, coeff, apre, Fbacc, acur, ubar,
                     dt);
}

}  // namespace

void launch_adj_beta_cells(const Geo& g, const BetaF& bf, double* const ub[3], const double* bb, const double* dp,
                           double dt, double drho, double* abar, hipStream_t st) {
  dim3 cg((unsigned)((g.n[0] + 63) / 64), (unsigned)((g.n[1] + 3) / 4), (unsigned)g.n[2]);
  hipLaunchKernelGGL(k_adj_beta_cells, cg, dim3(64, 4), 0, st, g, bf, ub[0], ub[1], ub[2], bb, dp, dt, drho, abar);
}

void launch_vof_adj_sweep(const Geo& g, const double h[3], const Vel& V, int axis, const double* al, const double* F,
                          const unsigned char* coeff, const double* aout, double* apre, double* acur,
                          const double* Fbacc, double* ubar, double dt, bool clip, hipStream_t st) {
  Spacing H;
  for (int d = 0; d < 3; d++) H.h[d] = h[d];
  if (axis == 0) adj_sweep<0>(g, H, V, al, F, coeff, aout, apre, acur, Fbacc, ubar, dt, clip, st);
  else if (axis == 1) adj_sweep<1>(g, H, V, al, F, coeff, aout, apre, acur, Fbacc, ubar, dt, clip, st);
  else adj_sweep<2>(g, H, V, al, F, coeff, aout, apre, acur, Fbacc, ubar, dt, clip, st);
}

}  // namespace smac

namespace smac {

namespace {
__global__ __launch_bounds__(256) void k_add_box(double* __restrict__ x, const double* __restrict__ y, size_t n) {
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t < n) x[t] += y[t];
}
}  // namespace

// x += y over n entries
void launch_add_box(double* x, const double* y, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_add_box, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, x, y, n);
}

}  // namespace smac
