// smac_vof.hip -- VoF transport of alpha by the split THINC/WLIC scheme
// (G_SMACSolver::alpha_flux_thincwlic_split and its kernels,
// CFD/G_SMACSolver_VoF.cu:451-1190,1575-1706).
//
// Per substep: correct_coeff = [alpha > 0.5]; then for each direction of the
// order step % 3 (xyz, yzx, zxy) the face fluxes F (THINC/WLIC on interface
// donors, first-order upwind elsewhere) and the transport
//   alpha_new = alpha - (F+ - F-) + coeff (u+ - u-) dt/h,
// the last direction clipping alpha to {0} u [EPS, 1]. The flux kernel adds F
// to the step's accumulated flux in the same pass (the reference's separate
// k_alpha_flux_accum adds zero on every face it does not compute).
//
// Every kernel is one thread per face or cell over a 64x4 block; interface
// faces are a thin set, so the transcendental branch costs little. The
// transcendentals are the explicit operation sequences below (not the math
// library's), so the transport is bit-identical to the CPU restatement.
#include "smac_internal.hpp"

namespace smac {

#define THINC_BETA 2.0   // BETA, CFD/hardCodedParameters.h:6
#define VOF_EPS 1.E-12   // EPS, CFD/hardCodedParameters.h:7
#define VOF_PI 3.14159265358979323846

// exp, log, cosh and acos of the THINC/WLIC flux as a fixed sequence of IEEE
// double operations -- the same operations in the same order as the CPU
// restatement of the transport, so that both give the same bits. exp: Cody-Waite
// reduction (k = round(x/ln2), ln2 split into a 33-bit head and a tail) and the
// Taylor polynomial of degree 13 on |r| <= ln2/2; log: frexp to m in
// [sqrt(1/2), sqrt(2)) and the atanh series 2 atanh((m-1)/(m+1)) to degree 21;
// acos on [0, 1]: the asin series to degree 53 on [0, 1/2] and
// acos x = 2 asin(sqrt((1-x)/2)) above 1/2. Coefficients are the nearest
// doubles of the exact rationals; accuracy a few ulp.
__device__ __forceinline__ double vm_exp(double x) {
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
__device__ __forceinline__ double vm_log(double x) {
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
__device__ __forceinline__ double vm_cosh(double x) { return 0.5 * (vm_exp(x) + vm_exp(-x)); }
// asin on [0, 1/2]
__device__ __forceinline__ double vm_asin05(double y) {
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
// acos on [0, 1] (NaN above 1, as acos)
__device__ __forceinline__ double vm_acos01(double x) {
  if (x <= 0.5) return (0x1.921fb54442d18p+0 - vm_asin05(x)) + 0x1.1a62633145c07p-54;
  double y = sqrt((1.0 - x) * 0.5);
  return 2.0 * vm_asin05(y);
}

// find_xi0_analytic (CFD/G_SMACSolver_VoF.cu:459-465)
__device__ __forceinline__ double thinc_xi0(double alpha, double gamma) {
  double A = vm_exp(THINC_BETA * gamma * (2. * alpha - 1.));
  return 1. / (2. * THINC_BETA) * vm_log((vm_exp(THINC_BETA) - A) / (A - vm_exp(-THINC_BETA)));
}
// integrate_thinc (CFD/G_SMACSolver_VoF.cu:451-455)
__device__ __forceinline__ double thinc_integral(double a, double b, double gamma, double xi) {
  return 0.5 * (b - a) +
         gamma / (2. * THINC_BETA) * (vm_log(vm_cosh(THINC_BETA * (b - xi))) - vm_log(vm_cosh(THINC_BETA * (a - xi))));
}

// tangential component T of the donor cell's normal (-grad alpha): central,
// one-sided where a neighbour is not interior, else 0
template <int T>
__device__ __forceinline__ double vof_normal_t(const Geo& g, const Spacing& H, const double* __restrict__ al, I3 dn) {
  const I3 qp = sh(dn, T, +1), qm = sh(dn, T, -1);
  const bool ip = g.per[T] || (qp.v[T] >= 1 && qp.v[T] <= g.n[T]);
  const bool im = g.per[T] || (qm.v[T] >= 1 && qm.v[T] <= g.n[T]);
  if (ip && im) return -(al[IX(g, wr(g, qp))] - al[IX(g, wr(g, qm))]) * (1. / (2. * H.h[T]));
  if (!ip && im) return -(al[IX(g, dn)] - al[IX(g, wr(g, qm))]) * g.inv_d[T];
  if (ip && !im) return -(al[IX(g, wr(g, qp))] - al[IX(g, dn)]) * g.inv_d[T];
  return 0.;
}

// k_alpha_flux_thincwlic_{x,y,z} (CFD/G_SMACSolver_VoF.cu:507-666, 668-827,
// 829-984) on the A-faces 2..N (1..N periodic) x interior rows: the donor is
// interior, so the F_BOUNDARY branch has no case
template <int A>
__global__ __launch_bounds__(256) void k_vof_flux(Geo g, Spacing H, const double* __restrict__ u,
                                                  const double* __restrict__ al, double* __restrict__ F,
                                                  double* __restrict__ Facc, double dt) {
  int lo[3] = {1, 1, 1};
  lo[A] = g.per[A] ? 1 : 2;
  const int i = lo[0] + (int)(blockIdx.x * 64 + threadIdx.x);
  const int j = lo[1] + (int)(blockIdx.y * 4 + threadIdx.y);
  const int k = lo[2] + (int)blockIdx.z;
  if (i > g.n[0] || j > g.n[1] || k > g.n[2]) return;
  const I3 q = mk(i, j, k);
  const size_t id = IX(g, q);
  const double v = u[id];
  const I3 dn = wr(g, v > 0.0 ? sh(q, A, -1) : q);
  const double ad = al[IX(g, dn)];
  const double dtbyd = dt * g.inv_d[A];
  const I3 dp = sh(dn, A, +1), dm = sh(dn, A, -1);
  const bool ip = g.per[A] || dp.v[A] <= g.n[A];
  const bool im = g.per[A] || dm.v[A] >= 1;
  double Fv;
  if (!ip || !im) {
    Fv = v * ad * dtbyd;
  } else {
    const double gam = al[IX(g, wr(g, dp))] - al[IX(g, wr(g, dm))];
    if (ad < VOF_EPS || ad > 1.0 - VOF_EPS || fabs(gam) < 1e-6) {
      Fv = v * ad * dtbyd;
    } else {
      constexpr int T1 = A == 0 ? 1 : 0, T2 = A == 2 ? 1 : 2;
      double nrm[3];
      nrm[A] = -gam * (1. / (2. * H.h[A]));
      nrm[T1] = vof_normal_t<T1>(g, H, al, dn);
      nrm[T2] = vof_normal_t<T2>(g, H, al, dn);
      const double s_sq = nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2];
      const double sn = sqrt(s_sq);
      const double inv_s = 1.0 / (sn + VOF_EPS);
      const double theta = vm_acos01(fabs(nrm[A]) * inv_s);
      const double wc = 2. * theta / VOF_PI;
      const double w = 1.0 - wc * wc;
      const double gamma = gam < 0 ? -1. : 1.;
      const double xi0 = thinc_xi0(ad, gamma);
      const double lambda = v * dtbyd;
      const double ft =
          v > 0.0 ? thinc_integral(1.0 - lambda, 1.0, gamma, xi0) : -thinc_integral(0.0, -lambda, gamma, xi0);
      const double fu = lambda * ad;
      Fv = w * ft + (1. - w) * fu;
    }
  }
  F[id] = Fv;
  Facc[id] += Fv;
}

// k_transport_alpha_{x,y,z} (CFD/G_SMACSolver_VoF.cu:1011-1152); CLIP adds
// k_clip_alpha (1173-1190) after the last direction of a substep
template <int A, int CLIP>
__global__ __launch_bounds__(256) void k_vof_transport(Geo g, Vel V, const double* __restrict__ al,
                                                       const double* __restrict__ F,
                                                       const unsigned char* __restrict__ coeff,
                                                       double* __restrict__ an, double dt) {
  const int i = 1 + (int)(blockIdx.x * 64 + threadIdx.x);
  const int j = 1 + (int)(blockIdx.y * 4 + threadIdx.y);
  const int k = 1 + (int)blockIdx.z;
  if (i > g.n[0] || j > g.n[1]) return;
  const I3 q = mk(i, j, k);
  const size_t id = IX(g, q);
  double flux = 0.;
  flux += F[IX(g, wr(g, sh(q, A, +1)))] - F[id];
  const double up = get_face(g, V, A, sh(q, A, +1));
  const double um = get_face(g, V, A, q);
  const double cf = (double)coeff[id];
  flux -= cf * (up - um) * dt * g.inv_d[A];
  double a = al[id] - flux;
  if (CLIP) {
    if (a < VOF_EPS) a = 0.;
    else if (a > 1.) a = 1.;
  }
  an[id] = a;
}

// k_set_correct_coeff (CFD/G_SMACSolver_VoF.cu:986-1009)
__global__ __launch_bounds__(256) void k_vof_coeff(Geo g, const double* __restrict__ al,
                                                   unsigned char* __restrict__ coeff) {
  const int i = 1 + (int)(blockIdx.x * 64 + threadIdx.x);
  const int j = 1 + (int)(blockIdx.y * 4 + threadIdx.y);
  const int k = 1 + (int)blockIdx.z;
  if (i > g.n[0] || j > g.n[1]) return;
  const size_t id = idx(g, i, j, k);
  coeff[id] = al[id] > 0.5 ? 1 : 0;
}

template <int A>
static void vof_direction(const Geo& g, const Spacing& H, const Vel& V, const double* al, double* an, double* F,
                          double* Facc, const unsigned char* coeff, double dt, bool clip, hipStream_t st) {
  int nf[3] = {g.n[0], g.n[1], g.n[2]};
  nf[A] = g.per[A] ? g.n[A] : g.n[A] - 1;
  if (nf[A] > 0) {
    dim3 fg((unsigned)((nf[0] + 63) / 64), (unsigned)((nf[1] + 3) / 4), (unsigned)nf[2]);
    hipLaunchKernelGGL(k_vof_flux<A>, fg, dim3(64, 4), 0, st, g, H, V.c[A], al, F, Facc, dt);
  }
  dim3 cg((unsigned)((g.n[0] + 63) / 64), (unsigned)((g.n[1] + 3) / 4), (unsigned)g.n[2]);
  if (clip) hipLaunchKernelGGL((k_vof_transport<A, 1>), cg, dim3(64, 4), 0, st, g, V, al, F, coeff, an, dt);
  else hipLaunchKernelGGL((k_vof_transport<A, 0>), cg, dim3(64, 4), 0, st, g, V, al, F, coeff, an, dt);
}

void launch_vof_substep(const Geo& g, const double h[3], const Vel& V, double** al, double** an, double* const F[3],
                        double* const Facc[3], unsigned char* coeff, double dt, int step, hipStream_t st) {
  Spacing H;
  for (int d = 0; d < 3; d++) H.h[d] = h[d];
  dim3 cg((unsigned)((g.n[0] + 63) / 64), (unsigned)((g.n[1] + 3) / 4), (unsigned)g.n[2]);
  hipLaunchKernelGGL(k_vof_coeff, cg, dim3(64, 4), 0, st, g, *al, coeff);
  const int first = step % 3;
  for (int s = 0; s < 3; s++) {
    const int a = (first + s) % 3;
    const bool clip = s == 2;
    if (a == 0) vof_direction<0>(g, H, V, *al, *an, F[0], Facc[0], coeff, dt, clip, st);
    else if (a == 1) vof_direction<1>(g, H, V, *al, *an, F[1], Facc[1], coeff, dt, clip, st);
    else vof_direction<2>(g, H, V, *al, *an, F[2], Facc[2], coeff, dt, clip, st);
    double* t = *al;
    *al = *an;
    *an = t;
  }
}

}  // namespace smac
