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
#include "smac_vof.hpp"

namespace smac {

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

namespace smac {

// One substep that keeps what its adjoint needs: sweep s (axis a = (step + s) % 3)
// reads al[s] and writes al[s+1] and its flux into F[a] (al[0], the substep's
// input, is only read). The flux arrays are per axis, as the forward's are:
// the boundary faces of an axis, which no sweep writes, stay zero. The same
// kernels and arithmetic as launch_vof_substep.
void launch_vof_substep_rec(const Geo& g, const double h[3], const Vel& V, double* const al[4], double* const F[3],
                            double* const Facc[3], unsigned char* coeff, double dt, int step, hipStream_t st) {
  Spacing H;
  for (int d = 0; d < 3; d++) H.h[d] = h[d];
  dim3 cg((unsigned)((g.n[0] + 63) / 64), (unsigned)((g.n[1] + 3) / 4), (unsigned)g.n[2]);
  hipLaunchKernelGGL(k_vof_coeff, cg, dim3(64, 4), 0, st, g, al[0], coeff);
  const int first = step % 3;
  for (int s = 0; s < 3; s++) {
    const int a = (first + s) % 3;
    const bool clip = s == 2;
    if (a == 0) vof_direction<0>(g, H, V, al[s], al[s + 1], F[0], Facc[0], coeff, dt, clip, st);
    else if (a == 1) vof_direction<1>(g, H, V, al[s], al[s + 1], F[1], Facc[1], coeff, dt, clip, st);
    else vof_direction<2>(g, H, V, al[s], al[s + 1], F[2], Facc[2], coeff, dt, clip, st);
  }
}

void launch_vof_coeff(const Geo& g, const double* al, unsigned char* coeff, hipStream_t st) {
  dim3 cg((unsigned)((g.n[0] + 63) / 64), (unsigned)((g.n[1] + 3) / 4), (unsigned)g.n[2]);
  hipLaunchKernelGGL(k_vof_coeff, cg, dim3(64, 4), 0, st, g, al, coeff);
}

}  // namespace smac
