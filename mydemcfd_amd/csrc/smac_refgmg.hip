// smac_refgmg.hip -- the reference's geometric multigrid preconditioner
// (G_GMGSolver, CFD/pressure_solver/G_GMGSolver.cu:10-998) on gfx950.
//
// Algorithm (kept as the reference, so the iterates match):
//   * level 0 is already coarse (N/2); a direction stops coarsening when its
//     halved size would drop below 6 (G_GMGSolver.h:20);
//   * coarse face coefficient = mean of the 2x2x2 fine face coefficients,
//     two fine faces along the face's own axis (k_create_level*_coeffs,
//     G_GMGSolver.cu:499-644); level operator a = beta/h^2, diag = sum,
//     inv = 0 when |diag| < 1e-12 (k_create_level_A, 646-685);
//   * weighted Jacobi, omega = 0.6, from zero (692-746);
//   * residual, mean restriction over the children (152-226, 416-493);
//   * constant prolongation-add into the children (285-389);
//   * coarse ghosts stay zero; periodic directions (absent in the reference)
//     wrap face and cell indices.
//
// MI355X form: the finest-grid kernels work on the PCG vectors in their
// row-split layout (wave per row, XCD-aware row order); the residual is fused
// into the restriction (each coarse thread forms its children's residuals in
// the reference's summation order), so no residual array is written. Level
// arrays use a natural x-fastest layout with one ghost layer and pitches
// (n+3); the level count is small and their traffic is a fraction of the fine
// level's.
#include "smac_internal.hpp"

namespace smac {

namespace {

constexpr double RG_OMEGA = 0.6;   // G_GMGSolver.cu:695
constexpr double RG_EPS = 1.E-12;  // CFD/hardCodedParameters.h:7
constexpr int RGB = 256;

__device__ __forceinline__ size_t ridx(const RgLv& L, int i, int j, int k) {
  return ((size_t)k * (size_t)L.py + (size_t)j) * (size_t)L.px + (size_t)i;
}
__device__ __forceinline__ int rwrap(int i, int n) { return i < 1 ? i + n : (i > n ? i - n : i); }

// thread -> cell (i,j,k) of a level, x fastest
__device__ __forceinline__ bool rg_cell(const RgLv& L, int& i, int& j, int& k) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t nxy = (size_t)L.n[0] * (size_t)L.n[1];
  if (t >= nxy * (size_t)L.n[2]) return false;
  k = (int)(t / nxy) + 1;
  const size_t r = t % nxy;
  j = (int)(r / (size_t)L.n[0]) + 1;
  i = (int)(r % (size_t)L.n[0]) + 1;
  return true;
}

// grid face coefficient of a uniform-density grid: beta0 on interior faces,
// 0 on boundary and ghost faces (SMACSolver.cpp:341-420)
__device__ __forceinline__ double grid_beta(const Geo& g, int d, const int f[3]) {
  int fd = f[d];
  if (g.per[d]) return g.beta0;
  if (fd <= 1 || fd >= g.n[d] + 1) return 0.0;
  return g.beta0;
}

// coarse face coefficients of axis d on level C (k_create_level0_coeffs when
// fb == nullptr, else k_create_level_coeffs from level F)
__global__ __launch_bounds__(RGB) void k_rg_beta(Geo g, BetaF gb, RgLv C, RgLv F, const double* __restrict__ fb,
                                                 int d, double* __restrict__ cb) {
  int hi[3] = {C.n[0], C.n[1], C.n[2]};
  hi[d] += 1;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t nxy = (size_t)hi[0] * (size_t)hi[1];
  if (t >= nxy * (size_t)hi[2]) return;
  const int k = (int)(t / nxy) + 1;
  const size_t r = t % nxy;
  const int j = (int)(r / (size_t)hi[0]) + 1;
  const int i = (int)(r % (size_t)hi[0]) + 1;
  const int q[3] = {i, j, k};
  int B[3], R[3];
#pragma unroll
  for (int e = 0; e < 3; e++) {
    B[e] = C.co[e] ? 2 * q[e] - 1 : q[e];
    R[e] = C.co[e] ? 2 : 1;
  }
  double count = 0., tmp = 0.;
  for (int iz = 0; iz < R[2]; iz++)
    for (int iy = 0; iy < R[1]; iy++)
      for (int ix = 0; ix < R[0]; ix++) {
        int f[3] = {B[0] + ix, B[1] + iy, B[2] + iz};
        double v;
        if (fb == nullptr) {
          if (gb.b[0]) {  // the grid's face coefficients (two-phase grids)
            if (g.per[d]) f[d] = rwrap(f[d], g.n[d]);
            v = f[d] > g.n[d] + 1 ? 0.0 : gb.b[d][idx(g, f[0], f[1], f[2])];
          } else {
            v = grid_beta(g, d, f);
          }
        } else {
          if (F.per[d]) f[d] = rwrap(f[d], F.n[d]);
          v = f[d] > F.n[d] + 1 ? 0.0 : fb[ridx(F, f[0], f[1], f[2])];
        }
        tmp += v;
        count += 1.0;
      }
  cb[ridx(C, i, j, k)] = tmp / count;
}

struct RgCoef {
  double a[6];
  double diag, inv;
};
__device__ __forceinline__ RgCoef rg_coef(const RgLv& L, const double* __restrict__ bx,
                                          const double* __restrict__ by, const double* __restrict__ bz, int i,
                                          int j, int k) {
  RgCoef c;
  c.a[0] = bx[ridx(L, i + 1, j, k)] * L.inv_d2[0];
  c.a[1] = bx[ridx(L, i, j, k)] * L.inv_d2[0];
  c.a[2] = by[ridx(L, i, j + 1, k)] * L.inv_d2[1];
  c.a[3] = by[ridx(L, i, j, k)] * L.inv_d2[1];
  c.a[4] = bz[ridx(L, i, j, k + 1)] * L.inv_d2[2];
  c.a[5] = bz[ridx(L, i, j, k)] * L.inv_d2[2];
  c.diag = (c.a[0] + c.a[1] + c.a[2] + c.a[3] + c.a[4] + c.a[5]);
  c.inv = fabs(c.diag) < RG_EPS ? 0. : 1. / c.diag;
  return c;
}
__device__ __forceinline__ double rg_nbsum(const RgLv& L, const double* __restrict__ x, const RgCoef& c, int i,
                                           int j, int k) {
  const int ip = L.per[0] ? rwrap(i + 1, L.n[0]) : i + 1, im = L.per[0] ? rwrap(i - 1, L.n[0]) : i - 1;
  const int jp = L.per[1] ? rwrap(j + 1, L.n[1]) : j + 1, jm = L.per[1] ? rwrap(j - 1, L.n[1]) : j - 1;
  const int kp = L.per[2] ? rwrap(k + 1, L.n[2]) : k + 1, km = L.per[2] ? rwrap(k - 1, L.n[2]) : k - 1;
  return c.a[0] * x[ridx(L, ip, j, k)] + c.a[1] * x[ridx(L, im, j, k)] + c.a[2] * x[ridx(L, i, jp, k)] +
         c.a[3] * x[ridx(L, i, jm, k)] + c.a[4] * x[ridx(L, i, j, kp)] + c.a[5] * x[ridx(L, i, j, km)];
}

// k_jacobi_level_iteration (G_GMGSolver.cu:692-717)
__global__ __launch_bounds__(RGB) void k_rg_jacobi(RgLv L, const double* __restrict__ bx,
                                                   const double* __restrict__ by, const double* __restrict__ bz,
                                                   const double* __restrict__ q, double* __restrict__ qt,
                                                   const double* __restrict__ rhs) {
  int i, j, k;
  if (!rg_cell(L, i, j, k)) return;
  const RgCoef c = rg_coef(L, bx, by, bz, i, j, k);
  const double tmp_p = rg_nbsum(L, q, c, i, j, k);
  const size_t id = ridx(L, i, j, k);
  const double omega = RG_OMEGA;
  qt[id] = (1.0 - omega) * q[id] + omega * (tmp_p + rhs[id]) * c.inv;
}

// k_get_res_levels + k_restrict_level_to_level (G_GMGSolver.cu:152-188,457-493):
// crhs = mean over the children in F of rhs - A q
__global__ __launch_bounds__(RGB) void k_rg_res_restrict(RgLv F, const double* __restrict__ bx,
                                                         const double* __restrict__ by,
                                                         const double* __restrict__ bz,
                                                         const double* __restrict__ q,
                                                         const double* __restrict__ rhs, RgLv C,
                                                         double* __restrict__ crhs) {
  int i, j, k;
  if (!rg_cell(C, i, j, k)) return;
  const int Q[3] = {i, j, k};
  int B[3], R[3];
#pragma unroll
  for (int e = 0; e < 3; e++) {
    B[e] = C.co[e] ? 2 * Q[e] - 1 : Q[e];
    R[e] = C.co[e] ? 2 : 1;
  }
  double count = 0., tmp = 0.;
  for (int iz = 0; iz < R[2]; iz++)
    for (int iy = 0; iy < R[1]; iy++)
      for (int ix = 0; ix < R[0]; ix++) {
        const int fi = B[0] + ix, fj = B[1] + iy, fk = B[2] + iz;
        const RgCoef c = rg_coef(F, bx, by, bz, fi, fj, fk);
        const size_t id = ridx(F, fi, fj, fk);
        const double Ap = q[id] * c.diag - rg_nbsum(F, q, c, fi, fj, fk);
        tmp += rhs[id] - Ap;
        count += 1.0;
      }
  crhs[ridx(C, i, j, k)] = tmp / count;
}

// k_propagate_level_to_level (G_GMGSolver.cu:355-389)
__global__ __launch_bounds__(RGB) void k_rg_prolong(RgLv C, const double* __restrict__ cq, RgLv F,
                                                    double* __restrict__ fq) {
  int i, j, k;
  if (!rg_cell(C, i, j, k)) return;
  const int Q[3] = {i, j, k};
  int B[3], R[3];
#pragma unroll
  for (int e = 0; e < 3; e++) {
    B[e] = C.co[e] ? 2 * Q[e] - 1 : Q[e];
    R[e] = C.co[e] ? 2 : 1;
  }
  const double prop = cq[ridx(C, i, j, k)];
  for (int iz = 0; iz < R[2]; iz++)
    for (int iy = 0; iy < R[1]; iy++)
      for (int ix = 0; ix < R[0]; ix++) fq[ridx(F, B[0] + ix, B[1] + iy, B[2] + iz)] += prop;
}

// finest grid, row-split PCG vectors -----------------------------------------

// k_jacobi_iteration_general (G_GMGSolver.cu:719-746) with the grid operator
// of k_build_vof_poisson_A: out = (1-omega) x + omega (sum a_nb x_nb + r) / Adiag.
// zero_start: x is known to be zero (the first sweep of a V-cycle).
__global__ __launch_bounds__(256) void k_rg_fine_jacobi(Geo g, BetaF gb, const double* __restrict__ x,
                                                        double* __restrict__ out, const double* __restrict__ r,
                                                        int zero_start) {
  ROW_SETUP
  const int nx = g.n[0];
  const size_t ypr = idx(g, 0, W(g, 1, j + 1), k), ymr = idx(g, 0, W(g, 1, j - 1), k);
  const size_t zpr = idx(g, 0, j, W(g, 2, k + 1)), zmr = idx(g, 0, j, W(g, 2, k - 1));
  const double omega = RG_OMEGA;
  for (int i = 1 + lane; i <= nx; i += 64) {
    const Coef c = gb.b[0] ? coef_var(g, gb.b, i, j, k) : coef_uniform(g.n, g.per, g.inv_d2, g.beta0, i, j, k);
    const int p = spos(i, nx);
    double xc = 0.0, xp = 0.0, xm = 0.0, yp = 0.0, ym = 0.0, zp = 0.0, zm = 0.0;
    if (!zero_start) {
      xc = x[rb + p];
      xp = x[rb + spos(W(g, 0, i + 1), nx)];
      xm = x[rb + spos(W(g, 0, i - 1), nx)];
      yp = x[ypr + p];
      ym = x[ymr + p];
      zp = x[zpr + p];
      zm = x[zmr + p];
    }
    const double tmp_p = c.a[0] * xp + c.a[1] * xm + c.a[2] * yp + c.a[3] * ym + c.a[4] * zp + c.a[5] * zm;
    out[rb + p] = (1.0 - omega) * xc + omega * (tmp_p + r[rb + p]) * (1. / c.diag);
  }
}

// k_get_res_general + k_restrict_grid_to_level0 (G_GMGSolver.cu:191-226,416-455)
__global__ __launch_bounds__(RGB) void k_rg_fine_res_restrict(Geo g, BetaF gb, const double* __restrict__ x,
                                                              const double* __restrict__ r, RgLv C,
                                                              double* __restrict__ crhs) {
  int i, j, k;
  if (!rg_cell(C, i, j, k)) return;
  const int Q[3] = {i, j, k};
  int B[3], R[3];
#pragma unroll
  for (int e = 0; e < 3; e++) {
    B[e] = C.co[e] ? 2 * Q[e] - 1 : Q[e];
    R[e] = C.co[e] ? 2 : 1;
  }
  const int nx = g.n[0];
  double count = 0., tmp = 0.;
  for (int iz = 0; iz < R[2]; iz++)
    for (int iy = 0; iy < R[1]; iy++)
      for (int ix = 0; ix < R[0]; ix++) {
        const int fi = B[0] + ix, fj = B[1] + iy, fk = B[2] + iz;
        const Coef c =
            gb.b[0] ? coef_var(g, gb.b, fi, fj, fk) : coef_uniform(g.n, g.per, g.inv_d2, g.beta0, fi, fj, fk);
        const size_t rb = idx(g, 0, fj, fk);
        const int p = spos(fi, nx);
        const double xc = x[rb + p];
        const double xp = x[rb + spos(W(g, 0, fi + 1), nx)], xm = x[rb + spos(W(g, 0, fi - 1), nx)];
        const double yp = x[idx(g, 0, W(g, 1, fj + 1), fk) + p], ym = x[idx(g, 0, W(g, 1, fj - 1), fk) + p];
        const double zp = x[idx(g, 0, fj, W(g, 2, fk + 1)) + p], zm = x[idx(g, 0, fj, W(g, 2, fk - 1)) + p];
        const double s = c.a[0] * xp + c.a[1] * xm + c.a[2] * yp + c.a[3] * ym + c.a[4] * zp + c.a[5] * zm;
        tmp += r[rb + p] - (c.diag * xc - s);
        count += 1.0;
      }
  crhs[ridx(C, i, j, k)] = tmp / count;
}

// k_propagate_level0_to_array (G_GMGSolver.cu:285-316) into the row-split x
__global__ __launch_bounds__(RGB) void k_rg_prolong_fine(RgLv C, const double* __restrict__ cq, Geo g,
                                                         double* __restrict__ x) {
  int i, j, k;
  if (!rg_cell(C, i, j, k)) return;
  const int Q[3] = {i, j, k};
  int B[3], R[3];
#pragma unroll
  for (int e = 0; e < 3; e++) {
    B[e] = C.co[e] ? 2 * Q[e] - 1 : Q[e];
    R[e] = C.co[e] ? 2 : 1;
  }
  const double prop = cq[ridx(C, i, j, k)];
  for (int iz = 0; iz < R[2]; iz++)
    for (int iy = 0; iy < R[1]; iy++)
      for (int ix = 0; ix < R[0]; ix++)
        x[idx(g, 0, B[1] + iy, B[2] + iz) + spos(B[0] + ix, g.n[0])] += prop;
}

__global__ __launch_bounds__(RGB) void k_rg_zero(double* __restrict__ x, size_t n) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) x[t] = 0.0;
}

// r += a*S[slot] on interior cells; part0 <- row sums of r^2, part1 <- row max |r|
// (subtract_cell_mean, G_PressureSolverBase.cu:142-155, with the norms of the
// convergence check taken on the result)
__global__ __launch_bounds__(256) void k_add_scalar_norms(Geo g, double* __restrict__ r, double a,
                                                          const double* __restrict__ S, int slot,
                                                          double* __restrict__ part0,
                                                          double* __restrict__ part1) {
  ROW_SETUP
  const double s = S[slot];
  double a0 = 0.0, mx = 0.0;
  for (int i = 1 + lane; i <= g.n[0]; i += 64) {
    const size_t o = rb + spos(i, g.n[0]);
    const double v = r[o] + a * s;
    r[o] = v;
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

dim3 cells_grid(const RgLv& L) {
  size_t n = (size_t)L.n[0] * L.n[1] * L.n[2];
  return dim3((unsigned)((n + RGB - 1) / RGB));
}
dim3 rows_grid_g(const Geo& g) { return dim3((unsigned)((g.n[1] * g.n[2] + 3) / 4)); }

}  // namespace

void launch_rg_beta(const Geo& g, const BetaF* gb, const RgLv& C, const RgLv& F, const double* fb, int d,
                    double* cb, hipStream_t st) {
  size_t n = (size_t)(C.n[0] + (d == 0)) * (size_t)(C.n[1] + (d == 1)) * (size_t)(C.n[2] + (d == 2));
  hipLaunchKernelGGL(k_rg_beta, dim3((unsigned)((n + RGB - 1) / RGB)), dim3(RGB), 0, st, g, gb ? *gb : BetaF{}, C,
                     F, fb, d, cb);
}
void launch_rg_jacobi(const RgLv& L, double* const b[3], const double* q, double* qt, const double* rhs,
                      hipStream_t st) {
  hipLaunchKernelGGL(k_rg_jacobi, cells_grid(L), dim3(RGB), 0, st, L, b[0], b[1], b[2], q, qt, rhs);
}
void launch_rg_res_restrict(const RgLv& F, double* const b[3], const double* q, const double* rhs, const RgLv& C,
                            double* crhs, hipStream_t st) {
  hipLaunchKernelGGL(k_rg_res_restrict, cells_grid(C), dim3(RGB), 0, st, F, b[0], b[1], b[2], q, rhs, C, crhs);
}
void launch_rg_prolong(const RgLv& C, const double* cq, const RgLv& F, double* fq, hipStream_t st) {
  hipLaunchKernelGGL(k_rg_prolong, cells_grid(C), dim3(RGB), 0, st, C, cq, F, fq);
}
void launch_rg_fine_jacobi(const Geo& g, const BetaF* gb, const double* x, double* out, const double* r,
                           int zero_start, hipStream_t st) {
  hipLaunchKernelGGL(k_rg_fine_jacobi, rows_grid_g(g), dim3(256), 0, st, g, gb ? *gb : BetaF{}, x, out, r,
                     zero_start);
}
void launch_rg_fine_res_restrict(const Geo& g, const BetaF* gb, const double* x, const double* r, const RgLv& C,
                                 double* crhs, hipStream_t st) {
  hipLaunchKernelGGL(k_rg_fine_res_restrict, cells_grid(C), dim3(RGB), 0, st, g, gb ? *gb : BetaF{}, x, r, C,
                     crhs);
}
void launch_rg_prolong_fine(const RgLv& C, const double* cq, const Geo& g, double* x, hipStream_t st) {
  hipLaunchKernelGGL(k_rg_prolong_fine, cells_grid(C), dim3(RGB), 0, st, C, cq, g, x);
}
void launch_rg_zero(double* x, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_rg_zero, dim3((unsigned)((n + RGB - 1) / RGB)), dim3(RGB), 0, st, x, n);
}
void launch_add_scalar_norms(const Geo& g, double* r, double a, const double* S, int slot, double* p0, double* p1,
                             hipStream_t st) {
  hipLaunchKernelGGL(k_add_scalar_norms, rows_grid_g(g), dim3(256), 0, st, g, r, a, S, slot, p0, p1);
}

}  // namespace smac
