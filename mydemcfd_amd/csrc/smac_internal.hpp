// smac_internal.hpp -- kernel launchers and shared structs of the HIP library.
#pragma once

#include <hip/hip_runtime.h>

#include "smac_dev.hpp"

namespace smac {

// multigrid level geometry (uniform coefficients)
struct Lv {
  int n[3];
  int px, py;
  int per[3];
  double inv_d2[3];
  double beta0;
};

enum { OP_NONE = 0, OP_MEAN = 1, OP_DIV = 2, OP_BETA_PR = 3 };

struct FinalizeArgs {
  const double* part[3];
  int kind[3];  // 0 = sum, 1 = max
  int out[3];   // destination slots in S
  int count;
  int nrows;
  int op, dst, sa, sb, sc;
  double c;
};

// device scalar slots
enum {
  S_SUMB = 0, S_MEANB, S_BB, S_SUMR, S_SUMZ, S_RZ, S_RZN, S_PAP, S_ALPHA, S_BETA,
  S_RR, S_LINF, S_ZAD, S_CFLC, S_CFLV, S_COUNT = 32
};

void launch_predict(const Geo& g, const Vel& V, const double* p, double* us[3], double dt, hipStream_t st);
void launch_rhs(const Geo& g, double* const us[3], double* rhs, double inv_dt, double* part0, hipStream_t st);
void launch_init_r(const Geo& g, const double* rhs, double* r, double* dp, const double* S, int slot_mean,
                   double* p0, double* p1, hipStream_t st);
void launch_add_scalar(const Geo& g, double* x, double a, const double* S, int slot, hipStream_t st);
void launch_sum(const Geo& g, const double* x, double* p0, hipStream_t st);
void launch_jacobi(const Geo& g, const double* r, double* z, hipStream_t st);
void launch_copy_dot(const Geo& g, const double* z, double* d, const double* r, double* p0, hipStream_t st);
void launch_spmv_dot(const Geo& g, const double* d, double* Ad, double* p0, hipStream_t st);
void launch_update(const Geo& g, double* dp, const double* d, double* r, const double* Ad, const double* S,
                   int slot_alpha, double* p0, double* p1, hipStream_t st);
void launch_dot2(const Geo& g, const double* r, const double* z, const double* Ad, int flexible, double* p0,
                 double* p1, hipStream_t st);
void launch_xpay(const Geo& g, const double* z, double* d, const double* S, int slot_beta, hipStream_t st);
void launch_finalize(const FinalizeArgs& fa, double* S, hipStream_t st);
void launch_scalar_copy(double* S, int dst, int src, hipStream_t st);
void launch_correct(const Geo& g, double* const us[3], const double* dp, double* u[3], double dt, hipStream_t st);
void launch_add_dp(const Geo& g, double* p, double* dp, hipStream_t st);
void launch_cfl(const Geo& g, const Vel& V, double dt, double visc, double* p0, double* p1, hipStream_t st);
void launch_mg_rbgs(const Lv& L, double* e, const double* f, int col, int zero_start, hipStream_t st);
void launch_mg_residual(const Lv& L, const double* e, const double* f, double* res, hipStream_t st);
// residual of level F restricted to the rhs fc of level C (T: scratch of F's size)
void launch_mg_restrict(const Lv& F, const Lv& C, const double* e, const double* f, double* T, double* fc,
                        hipStream_t st);
void launch_mg_prolong(const Lv& C, const Lv& F, const double* ec, double* ef, int col, hipStream_t st);
void launch_mg_coarse(const Lv& C, const double* ainv, const double* f, double* e, hipStream_t st);
void launch_fill(const Lv& L, double* x, double v, hipStream_t st);

}  // namespace smac
