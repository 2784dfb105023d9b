// smac_vof.hpp -- device helpers of the split THINC/WLIC VoF transport,
// shared by the transport kernels (smac_vof.hip) and their adjoint
// (smac_adjoint_vof.hip).
#pragma once

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

}  // namespace smac
