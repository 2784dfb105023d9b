// smac_internal.hpp -- kernel launchers and shared structs of the HIP library.
#pragma once

#include <hip/hip_runtime.h>

#include "smac_dev.hpp"

namespace smac {

// multigrid level geometry; b[0..2] == nullptr: uniform coefficient beta0,
// else the level's face coefficients (natural layout, pitches px, py)
struct Lv {
  int n[3];
  int px, py;
  int per[3];
  double inv_d2[3];
  double beta0;
  const double* b[3];
  // the face coefficients in FP32 for the mixed V-cycles (the compute type
  // of their smoothers), or nullptr: half the coefficient bytes per sweep
  const float* bs[3];
};

// level of the reference's multigrid (natural layout, pitches n+3)
struct RgLv {
  int n[3];
  int co[3];   // coarsened from the level above (or the grid) in this direction
  int per[3];
  int px, py;
  double inv_d2[3];
};

// Blocks are dispatched round-robin over the 8 XCDs; remap so that each XCD
// works on a contiguous slab of rows and the y/z stencil neighbours of a row
// sit in the same L2.
__device__ __forceinline__ int xcd_block(int b, int nb) {
  const int q = nb >> 3;
  if (b >= (q << 3)) return b;
  return (b & 7) * q + (b >> 3);
}

// Rows are visited band-major: the y-rows are cut into bands of JB rows and a
// band is swept through all z-planes before the next, so the y- and z-
// neighbour rows of a stencil are read again while they are still in L2
// (a plane-major order needs three whole planes resident, 6 MB per FP64 plane
// triple at 512^2, more than an XCD's 4 MB L2).
__device__ __forceinline__ int band_rows(int ny) {
  return (ny % 32 == 0) ? 32 : ((ny % 16 == 0) ? 16 : ((ny % 8 == 0) ? 8 : ny));
}
// visit index r -> row (j, k), both from 1
__device__ __forceinline__ void band_jk(int r, int ny, int nz, int& j, int& k) {
  const int jb = band_rows(ny);
  const int t = r / jb;
  k = t % nz + 1;
  j = (t / nz) * jb + (r % jb) + 1;
}

// row = (k-1)*NY + (j-1) is the canonical row index (row partial sums are
// stored there whatever the visiting order)
#define ROW_SETUP_N(NX, NY, NZ, PX, PY)                                        \
  const int lane = threadIdx.x & 63;                                          \
  const int vrow_ = xcd_block((int)blockIdx.x, (int)gridDim.x) * 4 + (int)(threadIdx.x >> 6); \
  if (vrow_ >= (NY) * (NZ)) return;                                           \
  int j, k;                                                                   \
  band_jk(vrow_, (NY), (NZ), j, k);                                           \
  const int row = (k - 1) * (NY) + (j - 1);                                   \
  const size_t rb = ((size_t)k * (size_t)(PY) + (size_t)j) * (size_t)(PX);    \
  (void)rb;                                                                   \
  (void)row;

#define ROW_SETUP ROW_SETUP_N(g.n[0], g.n[1], g.n[2], g.px, g.py)

// Cells per lane of a row kernel over rows of n cells (a wave owns a row and
// lane l the cells 1+l+64m): 1, 2, 4 or 8, so that the lane issues all its
// loads before the first use; 0 = runtime loop for longer rows.
inline int row_vpl(int n) { return n <= 64 ? 1 : (n <= 128 ? 2 : (n <= 256 ? 4 : (n <= 512 ? 8 : 0))); }
#define SMAC_VPL_DISPATCH(N, CALL)   \
  switch (row_vpl(N)) {              \
    case 1: {                        \
      constexpr int VPL = 1;         \
      CALL;                          \
    } break;                         \
    case 2: {                        \
      constexpr int VPL = 2;         \
      CALL;                          \
    } break;                         \
    case 4: {                        \
      constexpr int VPL = 4;         \
      CALL;                          \
    } break;                         \
    case 8: {                        \
      constexpr int VPL = 8;         \
      CALL;                          \
    } break;                         \
    default: {                       \
      constexpr int VPL = 0;         \
      CALL;                          \
    } break;                         \
  }
#define LROW_SETUP ROW_SETUP_N(L.n[0], L.n[1], L.n[2], L.px, L.py)


enum { OP_NONE = 0, OP_MEAN = 1, OP_DIV = 2, OP_BETA_PR = 3 };

struct FinalizeArgs {
  const double* part[3];
  int kind[3];  // 0 = sum, 1 = max
  int out[3];   // destination slots in S
  int count;
  int nrows;
  int op, dst, sa, sb, sc;
  double c;
  double* stage;  // 3*1024 doubles: per-lane sums of large reductions (may be null)
};

// device scalar slots
enum {
  S_SUMB = 0, S_MEANB, S_BB, S_SUMR, S_SUMZ, S_RZ, S_RZN, S_PAP, S_ALPHA, S_BETA,
  S_RR, S_LINF, S_ZAD, S_CFLC, S_CFLV, S_COUNT = 32
};

// alpha flux of the VoF transport for the predictor's mass flux:
// mf = rho0 u_f + drho * F h/dt (hbydt = h/dt per axis)
// Mass-flux context of the predictor: the accumulated alpha flux F (with
// drho = rho1 - rho0 and h/dt per axis) and, for the two-way form, the fluid
// volume fraction at t^{n+1/2} on cells (eps_half).
struct VofF {
  const double* F[3];
  double drho;
  double hbydt[3];
  const double* eps_half;
};
// CFD-DEM (two-way) fields of the predictor: fluid volume fraction at t^{n+1}
// (eps) and t^n (eps_old) on cells, the particles' momentum impulse over the
// step on faces (imp, nullptr = none) and 1/(dx dy dz). Volume fractions are
// read at the raw (ghost-inclusive) cell indices the reference reads; the
// caller fills the ghost cells.
struct EpsF {
  const double* eps;
  const double* eps_old;
  const double* imp[3];
  double inv_vol;
};
// Two-way predictor: k_get_vof_vstar_rhouu_consistent_{x,y,z}_two_way with the
// mass flux of k_compute_mass_flux_from_alpha_flux_two_way (M: the alpha flux
// when a VoF transport preceded, else nullptr)
void launch_predict_two_way(const Geo& g, const Props* P, const Vel& V, const double* p, double* us[3], double dt,
                            hipStream_t st, const VofF* M, const EpsF& E);
// Poisson rhs with the volume fraction: b = -(1/dt) D(eps_f u*) - (eps - eps_old)/dt^2
void launch_rhs_eps(const Geo& g, double* const us[3], const double* eps, const double* eps_old, double* rhs,
                    double inv_dt, double* part0, hipStream_t st);
// Poisson face coefficients of a two-way grid: bp = eps_f * beta (beta: the
// face 1/rho_f arrays, or nullptr for one phase: beta0 on interior faces, 0
// elsewhere)
void launch_pois_two_way(const Geo& g, double* const beta[3], const double* eps, double* bp[3], hipStream_t st);
// P == nullptr: one-phase (uniform) properties from Geo. M == nullptr: mass
// flux rho0 u_f; M requires the two-phase properties P
void launch_predict(const Geo& g, const Props* P, const Vel& V, const double* p, double* us[3], double dt,
                    hipStream_t st, const VofF* M = nullptr);
void launch_rhs(const Geo& g, double* const us[3], double* rhs, double inv_dt, double* part0, hipStream_t st);
// r = rhs - S[slot_mean] (+ a S[slot_shift]), f0 = (F) r, dp = 0 when zero_dp
// (write); part0/part1 <- row sums of v^2 and v of v = rhs - S[slot_mean] (sums)
template <typename F>
void launch_init_r(const Geo& g, const double* rhs, double* r, F* f0, double* dp, int zero_dp, const double* S,
                   int slot_mean, double a, int slot_shift, double* p0, double* p1, int sums, int write,
                   hipStream_t st);
template <typename Z>
void launch_add_scalar(const Geo& g, Z* x, double a, const double* S, int slot, hipStream_t st);
template <typename Z>
void launch_sum(const Geo& g, const Z* x, double* p0, hipStream_t st);
// bf == nullptr: one-phase operator (beta0); else face coefficients bf->b
void launch_jacobi(const Geo& g, const BetaF* bf, const double* r, double* z, hipStream_t st);
template <typename Z>
// d = z, part0 <- sum r*z (pair order); slot >= 0: z + a S[slot] first, as stored in Z
void launch_copy_dot(const Geo& g, const Z* z, double* d, const double* r, double* p0, hipStream_t st,
                     double a = 0.0, const double* S = nullptr, int slot = -1);
void launch_spmv_dot(const Geo& g, const BetaF* bf, const double* d, double* Ad, double* p0, hipStream_t st);
// dp += alpha d_old, dnew = z + beta d_old and Ad = A dnew with its pair-order
// row partials in one pass (slots of S hold alpha and beta), for rows of the
// pair form (nx = 256 or 512, every array's rows aligned): dir_spmv_applies
template <typename Z>
bool dir_spmv_applies(const Geo& g, const Z* z, const double* dold, const double* dnew, const double* dp,
                      const double* Ad);
template <typename Z>
void launch_dir_spmv(const Geo& g, const BetaF* bf, const Z* z, const double* dold, double* dnew, double* dp,
                     double* Ad, double* p0, const double* S, int slot_alpha, int slot_beta, hipStream_t st);
// k_copy_dot and the SpMV after it in one pass, the first SpMV of a solve:
// d = z (slot >= 0: z + a S[slot] rounded to the storage of z), Ad = A d,
// p_dad <- the rows' d.Ad and p_rz <- their r.d, both in pair order; the same
// pair form as launch_dir_spmv (every array's rows aligned, z's included):
// first_spmv_applies
template <typename Z>
bool first_spmv_applies(const Geo& g, const Z* z, const double* r, const double* d, const double* Ad);
template <typename Z>
void launch_first_spmv(const Geo& g, const BetaF* bf, const Z* z, const double* r, double* d, double* Ad,
                       double* p_dad, double* p_rz, double a, const double* S, int slot, hipStream_t st);
// r -= alpha Ad with |r|^2, max|r| and f0 = (F) r; e0 != nullptr and zs != 0:
// also the zero-start red half sweep of level 0 (uniform coefficients;
// zs = 1: IEEE division, 2: the FP32 reciprocal). Returns whether it did that
// (the chunk form only)
template <typename F>
bool launch_update(const Geo& g, double* r, const double* Ad, const double* S, int slot_alpha, double* p0,
                   double* p1, F* f0, hipStream_t st, F* e0 = nullptr, int zs = 0, const Lv* L0 = nullptr);
// the division the zero-start half sweep of level L uses (mg_rbgs's dispatch):
// 2 for the FP32 reciprocal, 1 for IEEE division
template <typename T>
int mg_zs_division(const Lv& L, const T* e, const T* f);
void launch_dp_flush(const Geo& g, double* dp, const double* d, const double* S, int slot_alpha, hipStream_t st);
template <typename Z>
void launch_dot2(const Geo& g, const double* r, const Z* z, const double* Ad, int flexible, double* p0,
                 double* p1, hipStream_t st);
template <typename Z>
void launch_xpay(const Geo& g, const Z* z, double* d, double* dp, const double* S, int slot_alpha, int slot_beta,
                 hipStream_t st);
void launch_finalize(const FinalizeArgs& fa, double* S, hipStream_t st);
void launch_scalar_copy(double* S, int dst, int src, hipStream_t st);
void launch_scalar_set(double* S, int dst, double v, hipStream_t st);
void launch_props(const Geo& g, const double* alpha, double* rho, double* rho_old, double* inv_rho, double* mu,
                  double* beta[3], double* muf[3], double rho0, double rho1, double mu0, double mu1, int initial,
                  hipStream_t st);
// the correction of all components, p += dp, then dp = 0 (one pass + a clear)
void launch_correct_all(const Geo& g, const BetaF* bf, double* const us[3], double* dp, double* u[3], double* p,
                        double dt, hipStream_t st, const double* dv = nullptr,
                        const double* S = nullptr, int slot_alpha = -1);
// inv_rho == nullptr: one-phase viscous CFL `visc`; else (mu/rho) dt sum 1/h^2 per cell
void launch_cfl(const Geo& g, const Vel& V, double dt, double visc, const double* inv_rho, const double* mu,
                double* p0, double* p1, hipStream_t st);
// after the step (smac_post.hip): a launch covers planes q[ax] = v[0..nv) of
// one array, the other two axes u, w over [lo, lo+cnt)
struct Plane {
  int ax, nv, v[3];
  int u, ulo, ucnt;
  int w, wlo, wcnt;
};
struct Spacing {
  double h[3];
};
// ghost cells of p; rho == nullptr: uniform density g.rho0
void launch_ghost_pressure(const Geo& g, const double h[3], double* p, const double* rho, hipStream_t st);
void launch_ghost_velocity(const Geo& g, double* const u[3], hipStream_t st);
void launch_outlet_faces(const Geo& g, const double* const us[3], const double* const beta[3], const double* dp,
                         double* const u[3], double dt, hipStream_t st);
void launch_face_gradp(const Geo& g, const double* p, double* const gp[3], hipStream_t st);
// VTI cell data: p[n], alpha[n] (al == nullptr: 0), velocity[3n] as float
void launch_pack_cells(const Geo& g, const double* p, const double* al, const double* const u[3], float* out,
                       hipStream_t st);
// one substep of the split VoF transport (smac_vof.hip); swaps *al and *an
// once per direction, accumulates the face fluxes into Facc
// the same substep recording each sweep's input alpha (al[0..3]) and flux
// (F[axis]) for the adjoint
void launch_vof_substep_rec(const Geo& g, const double h[3], const Vel& V, double* const al[4], double* const F[3],
                            double* const Facc[3], unsigned char* coeff, double dt, int step, hipStream_t st);
void launch_vof_coeff(const Geo& g, const double* al, unsigned char* coeff, hipStream_t st);
// adjoint of the VoF loop's step (smac_adjoint.hip, smac_adjoint_vof.hip)
void launch_predict_vjp_vof(const Geo& g, const Props& P, const Vel& V, const VofF& M, const double* p,
                            double* const wbar[3], double* ub[3], double* const Fb[3], double* anp1, double* an,
                            double drho, double dmu, double dt, hipStream_t st);
void launch_adj_beta_cells(const Geo& g, const BetaF& bf, double* const ub[3], const double* bb, const double* dp,
                           double dt, double drho, double* abar, hipStream_t st);
void launch_vof_adj_sweep(const Geo& g, const double h[3], const Vel& V, int axis, const double* al, const double* F,
                          const unsigned char* coeff, const double* aout, double* apre, double* acur,
                          const double* Fbacc, double* ubar, double dt, bool clip, hipStream_t st);
void launch_add_box(double* x, const double* y, size_t n, hipStream_t st);
void launch_vof_substep(const Geo& g, const double h[3], const Vel& V, double** al, double** an, double* const F[3],
                        double* const Facc[3], unsigned char* coeff, double dt, int step, hipStream_t st);
// multigrid (smac_mg.hip), storage T in {double, float, bf16s}; the restriction
// buffer tb has the compute type of T
template <typename T>
void launch_mg_rbgs(const Lv& L, T* e, const T* f, int col, int zero_start, hipStream_t st);
// the last colour-0 half sweep of a V-cycle with the PCG's dot2 row partials
// (uniform coefficients; the vector form with LDS y rows for FP32 and FP64,
// the two-rows-per-wave form for every storage type)
template <typename T>
bool mg_rbgs_dot_applies(const Lv& L, const T* e, const T* f);
template <typename T>
void launch_mg_rbgs_dot(const Lv& L, T* e, const T* f, const double* r, const double* Ad, int flexible, double* p0,
                        double* p1, hipStream_t st);
// dst = (float) src over n elements (the FP32 copies of the face coefficients)
void launch_d2f(const double* src, float* dst, size_t n, hipStream_t st);
// red != 0: the residual is 0 on black cells (the pre-smoothing ended with a
// black half sweep) and formed on the red cells only
template <typename T>
void launch_mg_restrict(const Lv& F, const Lv& C, const T* e, const T* f, typename St<T>::C* tb, T* fc, int red,
                        hipStream_t st);
template <typename T>
void launch_mg_prolong(const Lv& C, const Lv& F, const T* ec, T* ef, int col, hipStream_t st);
template <typename T>
void launch_mg_coarse(const Lv& C, const double* ainv, const T* f, T* e, hipStream_t st);
template <typename T>
void launch_fill(const Lv& L, T* x, hipStream_t st);
// coarse face coefficients of axis d on level C from level F (k_mg_coarsen_beta)
void launch_mg_coarsen_beta(const Lv& F, const Lv& C, int d, double* cb, hipStream_t st);
// dense (pseudo-)inverse of the coarsest level's operator into out (n x n,
// n = cells of C; work: n x n scratch), on the device
void launch_dense_pinv(const Lv& C, int singular, double* work, double* out, hipStream_t st);

// the reference's multigrid (smac_refgmg.hip); fb == nullptr: coefficients
// of level 0 from the grid
void launch_rg_beta(const Geo& g, const BetaF* gb, const RgLv& C, const RgLv& F, const double* fb, int d,
                    double* cb, hipStream_t st);
void launch_rg_jacobi(const RgLv& L, double* const b[3], const double* q, double* qt, const double* rhs,
                      hipStream_t st);
void launch_rg_res_restrict(const RgLv& F, double* const b[3], const double* q, const double* rhs, const RgLv& C,
                            double* crhs, hipStream_t st);
void launch_rg_prolong(const RgLv& C, const double* cq, const RgLv& F, double* fq, hipStream_t st);
void launch_rg_fine_jacobi(const Geo& g, const BetaF* gb, const double* x, double* out, const double* r,
                           int zero_start, hipStream_t st);
void launch_rg_fine_res_restrict(const Geo& g, const BetaF* gb, const double* x, const double* r, const RgLv& C,
                                 double* crhs, hipStream_t st);
void launch_rg_prolong_fine(const RgLv& C, const double* cq, const Geo& g, double* x, hipStream_t st);
void launch_rg_zero(double* x, size_t n, hipStream_t st);
void launch_add_scalar_norms(const Geo& g, double* r, double a, const double* S, int slot, double* p0, double* p1,
                             hipStream_t st);

// adjoint (smac_adjoint.hip)
void launch_predict_vjp(const Geo& g, const Props* P, const Vel& V, double* const wbar[3], double* ub[3], double dt,
                        hipStream_t st);
void launch_adj_div(const Geo& g, const BetaF* bf, double* const ub[3], const double* pin, double* out, double s,
                    hipStream_t st);
void launch_adj_ustar(const Geo& g, double* const ubn[3], const double* bb, double* ubs[3], double* un[3],
                      double inv_dt, hipStream_t st);
void launch_rowsum_nat(const Geo& g, const double* x, double* p0, hipStream_t st);

}  // namespace smac
