// smac_capi.hip -- C-ABI of the MI355X SMAC projection step (include/smac_c_api.h).
//
// The context owns every device array and one HIP stream; all launches go to
// that stream. The pressure solve follows G_PCGSolver::solve_* of the
// reference (CFD/pressure_solver/G_PCGSolver.cu:307-922) with the build's
// preconditioners (Jacobi, RBGS multigrid V-cycle) and stopping rules; device
// scalars (alpha, beta, r.z, ...) stay on the GPU and the host reads them only
// at convergence checks.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <map>
#include <string>
#include <type_traits>
#include <utility>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "../../include/smac_c_api.h"
#include "smac_internal.hpp"

using namespace smac;

namespace {

// One multigrid level: e (correction) and f (rhs) in the storage type of the
// hierarchy's precision, tb (restriction buffer) in its compute type.
struct Level {
  Lv lv;
  void* e = nullptr;
  void* f = nullptr;
  void* tb = nullptr;
  bool own_e = false, own_f = false, own_tb = false;
};

// One level of the reference's multigrid: q ping-pongs with qt (ghosts stay
// zero), rhs from the restriction, face coefficients b[3].
struct RgLevel {
  RgLv lv;
  size_t tot = 0;
  double *q = nullptr, *qt = nullptr, *rhs = nullptr;
  double* b[3] = {nullptr, nullptr, nullptr};
};

struct KStat {
  int64_t n = 0;
  double ms = 0.0;
};

}  // namespace

struct smac_ctx {
  Geo g;
  int n[3];
  int per[3];
  double d[3];
  uint8_t side_bcid[6];
  std::vector<smac_bc_entry> bc;
  smac_fluid_desc fl;
  bool alpha_nonzero = false;
  // an outlet side: its faces carry the Dirichlet coefficient 2 beta of the
  // inner face (p = 0 at the face), so the Poisson operator is regular and
  // the face coefficients live in arrays
  bool has_outlet = false;
  hipStream_t st = nullptr;
  int dev = -1;  // device current at smac_create
  size_t tot = 0;
  double* u[3] = {nullptr, nullptr, nullptr};
  double* us[3] = {nullptr, nullptr, nullptr};
  double *p = nullptr, *dp = nullptr, *rhs = nullptr, *alpha = nullptr;
  double *r = nullptr, *z = nullptr, *dv = nullptr, *Ad = nullptr, *res0 = nullptr;
  // second direction array: the fused direction update + SpMV writes the new
  // direction here while neighbouring blocks still read the old one (dv and
  // dv2 swap after every such pass; dv holds the current direction between calls)
  double* dv2 = nullptr;
  // dp is zero (as the correction leaves it, and at creation): the next solve
  // need not clear it
  bool dp_zero = true;
  // the PCG's dot2 of the next V-cycle's z rides on its last fine half sweep
  // (uniform coefficients, the vector and two-rows-per-wave smoother forms):
  // the operands while dot_req, dot_done once it ran
  bool dot_req = false, dot_done = false;
  const double* dot_r = nullptr;
  const double* dot_Ad = nullptr;
  int dot_flex = 0;
  double *dot_p0 = nullptr, *dot_p1 = nullptr;
  // the residual update has done level 0's zero-start half sweep: the next
  // V-cycle starts with the black half sweep
  bool zs_done = false;
  // smac_step lets the solve leave its last dp += alpha d to the correction
  // (defer_dp), which then adds it cell by cell (dp_deferred)
  bool defer_dp = false, dp_deferred = false;
  double* part[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  int nrows = 0;
  double* S = nullptr;
  double* Sh = nullptr;
  double* fstage = nullptr;  // 3*1024 lane sums of the two-stage reductions
  std::vector<Level> lev;
  double* cinv = nullptr;
  double* cinv_ws = nullptr;  // the factor of the coarsest operator (n x n)
  int cinv_n = 0;
  // the coarsest inverse is rebuilt on a second stream (default priority),
  // next to the predictor; the solve waits for ev_side_out before its first
  // V-cycle while pinv_pending
  hipStream_t st_side = nullptr;
  hipEvent_t ev_side_in = nullptr, ev_side_out = nullptr;
  bool pinv_pending = false;
  int mg_L = 0;
  int mg_prec = -1;
  std::vector<RgLevel> rg;
  int rg_req = 0;
  double rg_beta0 = 0.0;
  uint64_t rg_ver = ~0ull;
  // two-phase properties (allocated on first use; natural layout)
  double *rho = nullptr, *rho_old = nullptr, *inv_rho = nullptr, *mu = nullptr;
  double* beta[3] = {nullptr, nullptr, nullptr};
  double* muf[3] = {nullptr, nullptr, nullptr};
  bool props_ready = false;
  uint64_t props_ver = 1;     // bumped whenever the face coefficients may change
  uint64_t mg_ver = ~0ull;    // props_ver the multigrid hierarchy was built for
  std::vector<double*> mg_b;  // coarse-level face coefficients, 3 per level >= 1, and the FP32 copies of every level
  // the V-cycle below the fine level, captured once as a HIP graph and
  // replayed (its kernels are small and launch-bound); key = nu1, nu2
  hipGraphExec_t cg_exec = nullptr;
  int cg_nu[2] = {-1, -1};
  bool capturing = false;
  bool graphs = true;
  // adjoint work arrays (box layout, allocated on first use): cotangents of
  // u^{n+1}, u*, u^n (3 faces each), of p^{n+1} and of p^n
  double* adj[11] = {nullptr};
  // work arrays of smac_adjoint_vof_step (box layout): the input alpha of
  // every transport sweep (3 S + 1), each substep's flux per axis (3 S), three alpha
  // cotangent buffers, the rho_old part, the accumulated-flux cotangents (3)
  // and the forward dp; vadj_S = the substep count they are sized for
  std::vector<double*> vadj;
  int vadj_S = 0;
  // zero-copy bindings (smac_bind_device): the borrowed array of a field, and
  // the context's own array it stands in for (restored on unbind)
  double* bound[SMAC_FIELD_COUNT] = {};
  double* own[SMAC_FIELD_COUNT] = {};
  double* gradp[3] = {nullptr, nullptr, nullptr};  // f_gradp_{x,y,z}, allocated on first use
  // VoF transport (allocated on first use): alpha_new_, f_F{x,y,z}_,
  // f_F{x,y,z}_accum_, correct_coeff_; vof_armed = the accumulated flux
  // belongs to the coming predictor
  double* alpha_new = nullptr;
  double* vf[3] = {nullptr, nullptr, nullptr};
  double* vfacc[3] = {nullptr, nullptr, nullptr};
  unsigned char* vcoeff = nullptr;
  bool vof_armed = false;
  float* cellout = nullptr;  // device buffer of smac_cell_output (5 floats per cell)
  // CFD-DEM (two-way) fields, allocated on upload: void_fraction_,
  // void_fraction_old_, void_fraction_half_ (cells) and f_coupling_impulse_*
  // (faces); bpois = the two-way Poisson face coefficients eps_f/rho_f, in use
  // (bpois_on) from smac_update_poisson_beta_two_way until the next
  // smac_update_properties (which restores 1/rho_f, as the reference's face
  // property update rewrites f_b)
  double* eps = nullptr;
  double* eps_old = nullptr;
  double* eps_half = nullptr;
  double* imp[3] = {nullptr, nullptr, nullptr};
  double* bpois[3] = {nullptr, nullptr, nullptr};
  bool bpois_on = false;
  double* rgt = nullptr;  // fine ping-pong partner of z in the reference V-cycle
  bool prof = false;
  hipEvent_t ev[8];
  bool ev_ok = false;
  // per-kernel profiling: event pairs recorded around launches, resolved lazily
  std::vector<hipEvent_t> pool;
  size_t pool_used = 0;
  struct Rec {
    int name;
    hipEvent_t a, b;
    int launches;  // launches between the two events
  };
  std::vector<Rec> pending;
  std::map<std::string, int> name_id;
  std::vector<std::string> names;
  std::vector<KStat> totals;
};

static hipEvent_t pool_event(smac_ctx* c) {
  if (c->pool_used == c->pool.size()) {
    hipEvent_t e;
    // timing only (read after the stream is synchronized): a device-scope
    // release, without the system-scope cache writeback that would lengthen
    // the launch it brackets
    if (hipEventCreateWithFlags(&e, hipEventReleaseToDevice) != hipSuccess) return nullptr;
    c->pool.push_back(e);
  }
  return c->pool[c->pool_used++];
}
static void prof_flush(smac_ctx* c) {
  if (c->pending.empty()) return;
  (void)hipStreamSynchronize(c->st);
  for (auto& r : c->pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
      c->totals[r.name].n += r.launches;
      c->totals[r.name].ms += (double)ms;
    }
  }
  c->pending.clear();
  c->pool_used = 0;
}
// returns a record index, or -1 when profiling is off
static int kbeg(smac_ctx* c, const char* name) {
  if (!c->prof || c->capturing) return -1;
  if (c->pool_used + 2 > 65536) prof_flush(c);
  auto it = c->name_id.find(name);
  int id;
  if (it == c->name_id.end()) {
    id = (int)c->names.size();
    c->name_id[name] = id;
    c->names.push_back(name);
    c->totals.push_back(KStat());
  } else {
    id = it->second;
  }
  hipEvent_t a = pool_event(c), b = pool_event(c);
  if (!a || !b) return -1;
  (void)hipEventRecord(a, c->st);
  c->pending.push_back({id, a, b, 1});
  return (int)c->pending.size() - 1;
}
static void kend(smac_ctx* c, int rec, int launches = 1) {
  if (rec < 0) return;
  c->pending[rec].launches = launches;
  (void)hipEventRecord(c->pending[rec].b, c->st);
}
#define KT(c, name, launch)      \
  do {                           \
    int r_ = kbeg((c), (name));  \
    launch;                      \
    kend((c), r_);               \
  } while (0)
// one event pair around `n` consecutive launches of one kernel: the gap
// before a launch that a pair around each launch would add counts once
#define KT_N(c, name, n, launches) \
  do {                             \
    int r_ = kbeg((c), (name));    \
    launches;                      \
    kend((c), r_, (n));            \
  } while (0)

#define HIPOK(x)                         \
  do {                                   \
    hipError_t e_ = (x);                 \
    if (e_ != hipSuccess) {              \
      fprintf(stderr, "smac: %s failed: %s (%s:%d)\n", #x, hipGetErrorString(e_), __FILE__, __LINE__); \
      return SMAC_ERR_HIP;               \
    }                                    \
  } while (0)

// Device arrays are allocated with their element 1 on a 128-byte boundary
// (the pointer handed out is the allocation + 128 - esize bytes): with the
// row pitch a multiple of 64 elements, the first interior cell of every row
// of a box array then starts a cache line. The allocation is zeroed. dfree
// releases by the handed-out pointer.
static std::mutex g_alloc_mu;
static std::unordered_map<void*, void*> g_alloc_raw;
static int balloc(void** p, size_t bytes, size_t esize = 8) {
  void* raw = nullptr;
  const size_t off = 128 - esize;
  *p = nullptr;
  if (hipMalloc(&raw, bytes + 256) != hipSuccess) return SMAC_ERR_OOM;
  // the fill runs on the null stream, which does not order against the
  // context's non-blocking stream: wait for it before any kernel or copy of
  // that stream can touch the buffer (buffers are also allocated lazily, right
  // before their first kernel)
  if (hipMemset(raw, 0, bytes + 256) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
    (void)hipFree(raw);
    return SMAC_ERR_HIP;
  }
  *p = (char*)raw + off;
  std::lock_guard<std::mutex> lk(g_alloc_mu);
  g_alloc_raw[*p] = raw;
  return SMAC_OK;
}
static int dalloc(double** p, size_t n) { return balloc((void**)p, n * sizeof(double), sizeof(double)); }
static void dfree(void* p) {
  if (!p) return;
  void* raw = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_alloc_mu);
    auto it = g_alloc_raw.find(p);
    if (it == g_alloc_raw.end()) return;
    raw = it->second;
    g_alloc_raw.erase(it);
  }
  (void)hipFree(raw);
}
// Scratch of smac_calc_cfl_noalloc: one static device array per GPU holding
// the two row-partial arrays (SMAC_NOALLOC_MAX_ROWS doubles each), the lane
// sums of the two-stage reduction (2 x 1024) and 32 scalar slots. Host code
// never passes the symbol itself to a kernel or a copy (on the host it is a
// stand-in, not a device address): smac_create resolves the array's address
// on each device once, with that device current, and every use goes through
// that address.
namespace {
constexpr size_t SCR_PART1 = (size_t)SMAC_NOALLOC_MAX_ROWS;
constexpr size_t SCR_STAGE = 2 * (size_t)SMAC_NOALLOC_MAX_ROWS;
constexpr size_t SCR_S = SCR_STAGE + 2 * 1024;
constexpr size_t SCR_WORDS = SCR_S + 32;
constexpr int SCR_MAX_DEV = 64;
}  // namespace
static __device__ __align__(256) double scratch[SCR_WORDS];
static std::mutex g_scr_mu;
static double* g_scr_ptr[SCR_MAX_DEV] = {};

// Address of `scratch` on device dev (the current device), resolved once.
static int scratch_resolve(int dev) {
  if (dev < 0 || dev >= SCR_MAX_DEV) return SMAC_ERR_NO_DEVICE;
  std::lock_guard<std::mutex> lk(g_scr_mu);
  if (g_scr_ptr[dev]) return SMAC_OK;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    fprintf(stderr, "smac: hipGetSymbolAddress(scratch) failed on device %d: %s\n", dev, hipGetErrorString(e));
    return SMAC_ERR_HIP;
  }
  g_scr_ptr[dev] = (double*)p;
  return SMAC_OK;
}
static double* scratch_ptr(int dev) {
  if (dev < 0 || dev >= SCR_MAX_DEV) return nullptr;
  std::lock_guard<std::mutex> lk(g_scr_mu);
  return g_scr_ptr[dev];
}

// row pitch of box arrays: a multiple of 64 elements (128 bytes in BF16,
// 256 in FP32, 512 in FP64)
static int box_pitch(int n) { return ((n + 3 + 63) / 64) * 64; }

static int pure_neumann(const smac_ctx* c) {
  for (int d = 0; d < 3; d++) {
    if (c->per[d]) continue;
    for (int s = 0; s < 2; s++) {
      int bt = c->bc[c->side_bcid[2 * d + s]].type;
      if (bt == SMAC_BC_OUTLET || bt == SMAC_BC_INFLOW) return 0;
    }
  }
  return 1;
}

static bool uniform_props(const smac_ctx* c) {
  return c->fl.rho0 == c->fl.rho1 && c->fl.mu0 == c->fl.mu1 && !c->alpha_nonzero && !c->has_outlet;
}

static void set_props(smac_ctx* c) {
  c->g.rho0 = c->fl.rho0;
  c->g.mu0 = c->fl.mu0;
  // face density 0.5*(rho+rho) and beta = 1/rho_f, as k_update_x_face_properties_by_alpha
  double rhof = 0.5 * (c->fl.rho0 + c->fl.rho0);
  c->g.beta0 = 1. / rhof;
  c->g.g[0] = c->fl.gx;
  c->g.g[1] = c->fl.gy;
  c->g.g[2] = c->fl.gz;
}

// two-phase grid (rho0 != rho1, mu0 != mu1 or a nonzero alpha field): the
// properties live in arrays built from alpha
static bool var_props(const smac_ctx* c) { return !uniform_props(c); }

static Props props_of(const smac_ctx* c) {
  Props P;
  P.mu = c->mu;
  P.rho_old = c->rho_old;
  P.inv_rho = c->inv_rho;
  for (int d = 0; d < 3; d++) {
    P.beta[d] = c->beta[d];
    P.muf[d] = c->muf[d];
  }
  return P;
}
static BetaF beta_of(const smac_ctx* c) {
  BetaF b;
  for (int d = 0; d < 3; d++) b.b[d] = c->beta[d];
  return b;
}
// Poisson operator coefficients: the two-way eps_f/rho_f when in use, else
// the face 1/rho_f (variable) or beta0 (uniform)
static bool pois_var(const smac_ctx* c) { return c->bpois_on || var_props(c); }
static BetaF pois_of(const smac_ctx* c) {
  if (!c->bpois_on) return beta_of(c);
  BetaF b;
  for (int d = 0; d < 3; d++) b.b[d] = c->bpois[d];
  return b;
}

// ------------------------------------------------------------------ profiling
static void prof_begin(smac_ctx* c, int slot) {
  if (c->prof && c->ev_ok) (void)hipEventRecord(c->ev[slot], c->st);
}
static double prof_end(smac_ctx* c, int slot) {
  if (!(c->prof && c->ev_ok)) return 0.0;
  (void)hipEventRecord(c->ev[slot + 1], c->st);
  (void)hipEventSynchronize(c->ev[slot + 1]);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, c->ev[slot], c->ev[slot + 1]);
  return (double)ms;
}

// ------------------------------------------------------------------ lifecycle
extern "C" int smac_create(const smac_grid_desc* gd, const smac_fluid_desc* fl, smac_ctx** out) {
  if (!gd || !out) return SMAC_ERR_INVALID_ARG;
  *out = nullptr;
  int n[3] = {gd->nx, gd->ny, gd->nz};
  double sz[3] = {gd->sizex, gd->sizey, gd->sizez};
  for (int d = 0; d < 3; d++) {
    if (n[d] < 2 || !(sz[d] > 0)) return SMAC_ERR_INVALID_GRID;
    if (gd->periodic[d] && n[d] < 4) return SMAC_ERR_INVALID_GRID;
  }
  if (gd->num_bc < 0 || gd->num_bc > 256 || (gd->num_bc > 0 && !gd->bc_table)) return SMAC_ERR_INVALID_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return SMAC_ERR_NO_DEVICE;

  smac_ctx* c = new smac_ctx();
  c->bc.assign(256, smac_bc_entry{SMAC_BC_NOSLIP, 0.0, 0.0, 0.0});
  for (int b = 0; b < gd->num_bc; b++) c->bc[b] = gd->bc_table[b];
  memcpy(c->side_bcid, gd->side_bcid, 6);
  // BC_OUTFLOW has no case in the reference's kernels (and no config
  // spelling, CFD/config/ConfigLoader.cpp:242-267): refused
  for (int s = 0; s < 6; s++) {
    if (gd->periodic[s / 2]) continue;
    int bt = c->bc[c->side_bcid[s]].type;
    if (bt == SMAC_BC_OUTFLOW || bt < SMAC_BC_NOSLIP || bt > SMAC_BC_OUTLET) {
      delete c;
      return SMAC_ERR_UNSUPPORTED;
    }
    if (bt == SMAC_BC_OUTLET) c->has_outlet = true;
  }
  memset(&c->g, 0, sizeof(c->g));
  for (int d = 0; d < 3; d++) {
    c->n[d] = n[d];
    c->per[d] = gd->periodic[d] ? 1 : 0;
    c->d[d] = sz[d] / (double)n[d];
    c->g.n[d] = n[d];
    c->g.per[d] = c->per[d];
    c->g.inv_d[d] = 1. / c->d[d];
    c->g.inv_d2[d] = 1. / (c->d[d] * c->d[d]);
  }
  c->g.px = box_pitch(n[0]);
  c->g.py = n[1] + 3;
  for (int s = 0; s < 6; s++) {
    const smac_bc_entry& e = c->bc[c->side_bcid[s]];
    c->g.btype[s] = e.type;
    c->g.bv[s][0] = e.vx;
    c->g.bv[s][1] = e.vy;
    c->g.bv[s][2] = e.vz;
  }
  c->g.btype[6] = c->bc[0].type;
  c->g.bv[6][0] = c->bc[0].vx;
  c->g.bv[6][1] = c->bc[0].vy;
  c->g.bv[6][2] = c->bc[0].vz;
  if (fl) c->fl = *fl;
  else c->fl = smac_fluid_desc{1.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  set_props(c);

  c->tot = (size_t)c->g.px * (size_t)c->g.py * (size_t)(n[2] + 3);
  c->nrows = n[1] * n[2];
  int rc = SMAC_OK;
  if (hipGetDevice(&c->dev) != hipSuccess) rc = SMAC_ERR_HIP;
  if (rc == SMAC_OK) rc = scratch_resolve(c->dev);
  if (rc == SMAC_OK && hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) rc = SMAC_ERR_HIP;
  if (rc == SMAC_OK && hipStreamCreateWithFlags(&c->st_side, hipStreamNonBlocking) != hipSuccess) rc = SMAC_ERR_HIP;
  if (rc == SMAC_OK && hipEventCreateWithFlags(&c->ev_side_in, hipEventDisableTiming) != hipSuccess) rc = SMAC_ERR_HIP;
  if (rc == SMAC_OK && hipEventCreateWithFlags(&c->ev_side_out, hipEventDisableTiming) != hipSuccess) rc = SMAC_ERR_HIP;
  double** arrs[] = {&c->u[0], &c->u[1], &c->u[2], &c->us[0], &c->us[1], &c->us[2], &c->p, &c->dp, &c->rhs,
                     &c->r, &c->z, &c->dv, &c->Ad, &c->res0};
  for (auto a : arrs)
    if (rc == SMAC_OK) rc = dalloc(a, c->tot);
  for (int m = 0; m < 5 && rc == SMAC_OK; m++) rc = dalloc(&c->part[m], (size_t)c->nrows);
  if (rc == SMAC_OK) rc = dalloc(&c->S, S_COUNT);
  if (rc == SMAC_OK) rc = dalloc(&c->fstage, 3 * 1024);
  if (rc == SMAC_OK && hipHostMalloc((void**)&c->Sh, S_COUNT * sizeof(double)) != hipSuccess) rc = SMAC_ERR_OOM;
  if (rc == SMAC_OK) {
    c->ev_ok = true;
    for (int e = 0; e < 8; e++)
      if (hipEventCreate(&c->ev[e]) != hipSuccess) c->ev_ok = false;
  }
  if (rc != SMAC_OK) {
    smac_destroy(c);
    return rc;
  }
  *out = c;
  return SMAC_OK;
}

static void graph_free(smac_ctx* c) {
  if (c->cg_exec) (void)hipGraphExecDestroy(c->cg_exec);
  c->cg_exec = nullptr;
  c->cg_nu[0] = c->cg_nu[1] = -1;
}

static void mg_free(smac_ctx* c) {
  graph_free(c);
  for (auto b : c->mg_b)
    dfree(b);
  c->mg_b.clear();
  c->mg_ver = ~0ull;
  for (auto& L : c->lev) {
    if (L.own_e) dfree(L.e);
    if (L.own_f) dfree(L.f);
    if (L.own_tb) dfree(L.tb);
  }
  c->lev.clear();
  c->mg_prec = -1;
  // a rebuild of the coarsest inverse may still run on the side stream
  if (c->st_side) (void)hipStreamSynchronize(c->st_side);
  c->pinv_pending = false;
  dfree(c->cinv);
  c->cinv = nullptr;
  dfree(c->cinv_ws);
  c->cinv_ws = nullptr;
  c->cinv_n = 0;
  c->mg_L = 0;
}

static void rg_free(smac_ctx* c) {
  for (auto& L : c->rg) {
    double* arrs[] = {L.q, L.qt, L.rhs, L.b[0], L.b[1], L.b[2]};
    for (auto a : arrs) dfree(a);
  }
  c->rg.clear();
  c->rg_req = 0;
  dfree(c->rgt);
  c->rgt = nullptr;
}

static double** field_slot(smac_ctx* c, int f);

extern "C" void smac_destroy(smac_ctx* c) {
  if (!c) return;
  if (c->st) (void)hipStreamSynchronize(c->st);
  for (int f = 0; f < SMAC_FIELD_COUNT; f++)  // borrowed arrays are the caller's
    if (c->bound[f]) *field_slot(c, f) = c->own[f];
  mg_free(c);
  rg_free(c);
  double* arrs[] = {c->u[0],   c->u[1],    c->u[2],    c->us[0],   c->us[1],  c->us[2],  c->p,      c->dp,
                    c->rhs,    c->alpha,   c->r,       c->z,       c->dv,     c->dv2,    c->Ad,     c->res0,
                    c->S,      c->rho,     c->rho_old, c->inv_rho, c->mu,     c->beta[0], c->beta[1], c->beta[2],
                    c->muf[0], c->muf[1],  c->muf[2], c->fstage, c->gradp[0], c->gradp[1], c->gradp[2],
                    c->alpha_new, c->vf[0], c->vf[1], c->vf[2], c->vfacc[0], c->vfacc[1], c->vfacc[2],
                    c->eps,    c->eps_old, c->eps_half, c->imp[0], c->imp[1], c->imp[2], c->bpois[0],
                    c->bpois[1], c->bpois[2]};
  for (auto a : arrs) dfree(a);
  for (int m = 0; m < 5; m++) dfree(c->part[m]);
  for (int m = 0; m < 11; m++) dfree(c->adj[m]);
  for (auto a : c->vadj) dfree(a);
  dfree(c->vcoeff);
  dfree(c->cellout);
  if (c->Sh) (void)hipHostFree(c->Sh);
  if (c->ev_ok)
    for (int e = 0; e < 8; e++) (void)hipEventDestroy(c->ev[e]);
  if (c->ev_side_in) (void)hipEventDestroy(c->ev_side_in);
  if (c->ev_side_out) (void)hipEventDestroy(c->ev_side_out);
  if (c->st_side) (void)hipStreamDestroy(c->st_side);
  if (c->st) (void)hipStreamDestroy(c->st);
  delete c;
}

extern "C" int smac_set_fluid(smac_ctx* c, const smac_fluid_desc* fl) {
  if (!c || !fl) return SMAC_ERR_INVALID_ARG;
  c->fl = *fl;
  set_props(c);
  c->props_ready = false;
  c->props_ver++;
  return SMAC_OK;
}

// ------------------------------------------------------------------ data movement
static int field_axis(int f) {
  switch (f) {
    case SMAC_FIELD_U: case SMAC_FIELD_USTAR: case SMAC_FIELD_GRADPX: return 0;
    case SMAC_FIELD_V: case SMAC_FIELD_VSTAR: case SMAC_FIELD_GRADPY: return 1;
    case SMAC_FIELD_W: case SMAC_FIELD_WSTAR: case SMAC_FIELD_GRADPZ: return 2;
    case SMAC_FIELD_FLUXX: case SMAC_FIELD_IMPX: return 0;
    case SMAC_FIELD_FLUXY: case SMAC_FIELD_IMPY: return 1;
    case SMAC_FIELD_FLUXZ: case SMAC_FIELD_IMPZ: return 2;
    default: return -1;
  }
}
static void field_shape(const smac_ctx* c, int f, int sh[3]) {
  for (int d = 0; d < 3; d++) sh[d] = c->n[d] + 2;
  int a = field_axis(f);
  if (a >= 0) sh[a]++;
}
static double** field_slot(smac_ctx* c, int f) {
  switch (f) {
    case SMAC_FIELD_U: return &c->u[0];
    case SMAC_FIELD_V: return &c->u[1];
    case SMAC_FIELD_W: return &c->u[2];
    case SMAC_FIELD_USTAR: return &c->us[0];
    case SMAC_FIELD_VSTAR: return &c->us[1];
    case SMAC_FIELD_WSTAR: return &c->us[2];
    case SMAC_FIELD_P: return &c->p;
    case SMAC_FIELD_DP: return &c->dp;
    case SMAC_FIELD_RHS: return &c->rhs;
    case SMAC_FIELD_ALPHA: return &c->alpha;
    case SMAC_FIELD_GRADPX: return &c->gradp[0];
    case SMAC_FIELD_GRADPY: return &c->gradp[1];
    case SMAC_FIELD_GRADPZ: return &c->gradp[2];
    case SMAC_FIELD_FLUXX: return &c->vfacc[0];
    case SMAC_FIELD_FLUXY: return &c->vfacc[1];
    case SMAC_FIELD_FLUXZ: return &c->vfacc[2];
    case SMAC_FIELD_EPS: return &c->eps;
    case SMAC_FIELD_EPS_OLD: return &c->eps_old;
    case SMAC_FIELD_EPS_HALF: return &c->eps_half;
    case SMAC_FIELD_IMPX: return &c->imp[0];
    case SMAC_FIELD_IMPY: return &c->imp[1];
    case SMAC_FIELD_IMPZ: return &c->imp[2];
    default: return nullptr;
  }
}

extern "C" size_t smac_field_count(const smac_ctx* c, int32_t f) {
  if (!c || f < 0 || f >= SMAC_FIELD_COUNT) return 0;
  int sh[3];
  field_shape(c, f, sh);
  return (size_t)sh[0] * (size_t)sh[1] * (size_t)sh[2];
}

// copy between the reference layout (src pitch sh0, plane sh0*sh1) and the
// padded box, one 2-D copy per z-plane
static int copy_field(smac_ctx* c, int f, double* dst_box, const double* src_ref, bool to_box, hipMemcpyKind kind,
                      double* dst_ref, const double* src_box) {
  int sh[3];
  field_shape(c, f, sh);
  size_t row_bytes = (size_t)sh[0] * sizeof(double);
  size_t box_pitch = (size_t)c->g.px * sizeof(double);
  for (int k = 0; k < sh[2]; k++) {
    if (to_box) {
      HIPOK(hipMemcpy2DAsync(dst_box + (size_t)k * c->g.px * c->g.py, box_pitch,
                             src_ref + (size_t)k * sh[0] * sh[1], row_bytes, row_bytes, (size_t)sh[1], kind, c->st));
    } else {
      HIPOK(hipMemcpy2DAsync(dst_ref + (size_t)k * sh[0] * sh[1], row_bytes,
                             src_box + (size_t)k * c->g.px * c->g.py, box_pitch, row_bytes, (size_t)sh[1], kind,
                             c->st));
    }
  }
  HIPOK(hipStreamSynchronize(c->st));
  return SMAC_OK;
}

// fill periodic images (ghost cells, duplicate faces) of a downloaded host array
static void fill_periodic_images(const smac_ctx* c, int f, double* h) {
  int sh[3];
  field_shape(c, f, sh);
  int a = field_axis(f);
  if (!c->per[0] && !c->per[1] && !c->per[2]) return;
  auto src_index = [&](int d, int q) {
    if (!c->per[d]) return q;
    int w = wrapi(q, c->n[d]);
    if (d == a && w > c->n[d]) w -= c->n[d];
    return w;
  };
  for (int k = 0; k < sh[2]; k++)
    for (int j = 0; j < sh[1]; j++) {
      int ks = src_index(2, k), js = src_index(1, j);
      double* row = h + ((size_t)k * sh[1] + j) * sh[0];
      const double* srow = h + ((size_t)ks * sh[1] + js) * sh[0];
      bool inner_row = (ks == k && js == j);
      if (!inner_row) {
        for (int i = 0; i < sh[0]; i++) row[i] = srow[src_index(0, i)];
      } else if (c->per[0]) {
        for (int i = 0; i < sh[0]; i++) {
          int is = src_index(0, i);
          if (is != i) row[i] = srow[is];
        }
      }
    }
}

extern "C" int smac_upload(smac_ctx* c, int32_t f, const double* h, size_t count) {
  if (!c || !h || f < 0 || f >= SMAC_FIELD_COUNT) return SMAC_ERR_INVALID_ARG;
  if (count != smac_field_count(c, f)) return SMAC_ERR_INVALID_ARG;
  if (f == SMAC_FIELD_DP) c->dp_zero = false;
  double** slot = field_slot(c, f);
  if (!*slot) {
    int rc = dalloc(slot, c->tot);
    if (rc) return rc;
  }
  if (f == SMAC_FIELD_ALPHA) {
    bool nz = false;
    for (size_t t = 0; t < count && !nz; t++) nz = h[t] != 0.0;
    c->alpha_nonzero = nz;
    c->props_ready = false;
    c->props_ver++;
  }
  return copy_field(c, f, *slot, h, true, hipMemcpyHostToDevice, nullptr, nullptr);
}

extern "C" int smac_download(smac_ctx* c, int32_t f, double* h, size_t count) {
  if (!c || !h || f < 0 || f >= SMAC_FIELD_COUNT) return SMAC_ERR_INVALID_ARG;
  if (count != smac_field_count(c, f)) return SMAC_ERR_INVALID_ARG;
  double** slot = field_slot(c, f);
  if (!*slot) {
    memset(h, 0, count * sizeof(double));
    return SMAC_OK;
  }
  int rc = copy_field(c, f, nullptr, nullptr, false, hipMemcpyDeviceToHost, h, *slot);
  if (rc) return rc;
  fill_periodic_images(c, f, h);
  return SMAC_OK;
}

extern "C" int smac_upload_device(smac_ctx* c, int32_t f, const double* dsrc, size_t count) {
  if (!c || !dsrc || f < 0 || f >= SMAC_FIELD_COUNT) return SMAC_ERR_INVALID_ARG;
  if (count != smac_field_count(c, f)) return SMAC_ERR_INVALID_ARG;
  if (f == SMAC_FIELD_DP) c->dp_zero = false;
  double** slot = field_slot(c, f);
  if (!*slot) {
    int rc = dalloc(slot, c->tot);
    if (rc) return rc;
  }
  if (f == SMAC_FIELD_ALPHA) {
    // not scanned on the device: the grid takes the property arrays (the
    // same results as the uniform path when alpha is zero)
    c->alpha_nonzero = true;
    c->props_ready = false;
    c->props_ver++;
  }
  return copy_field(c, f, *slot, dsrc, true, hipMemcpyDeviceToDevice, nullptr, nullptr);
}

extern "C" int smac_device_layout(const smac_ctx* c, int64_t* pitch_x, int64_t* pitch_y, int64_t* planes) {
  if (!c || !pitch_x || !pitch_y || !planes) return SMAC_ERR_INVALID_ARG;
  *pitch_x = c->g.px;
  *pitch_y = c->g.py;
  *planes = c->n[2] + 3;
  return SMAC_OK;
}

extern "C" int smac_bind_device(smac_ctx* c, int32_t f, double* dev, size_t count) {
  if (!c || f < 0 || f >= SMAC_FIELD_COUNT) return SMAC_ERR_INVALID_ARG;
  if (f != SMAC_FIELD_U && f != SMAC_FIELD_V && f != SMAC_FIELD_W && f != SMAC_FIELD_P && f != SMAC_FIELD_ALPHA)
    return SMAC_ERR_UNSUPPORTED;
  double** slot = field_slot(c, f);
  if (!dev) {  // unbind: the bound contents move into the context's own array
    if (count != 0) return SMAC_ERR_INVALID_ARG;
    if (!c->bound[f]) return SMAC_OK;
    if (!c->own[f]) {
      int rc = dalloc(&c->own[f], c->tot);
      if (rc) return rc;
    }
    HIPOK(hipMemcpyAsync(c->own[f], c->bound[f], c->tot * sizeof(double), hipMemcpyDeviceToDevice, c->st));
    HIPOK(hipStreamSynchronize(c->st));
    *slot = c->own[f];
    c->own[f] = nullptr;
    c->bound[f] = nullptr;
    return SMAC_OK;
  }
  if (count != c->tot || (reinterpret_cast<uintptr_t>(dev) & 255u) != 0) return SMAC_ERR_INVALID_ARG;
  hipPointerAttribute_t at;
  if (hipPointerGetAttributes(&at, dev) != hipSuccess) {
    (void)hipGetLastError();
    return SMAC_ERR_INVALID_ARG;
  }
  if ((at.type != hipMemoryTypeDevice && at.type != hipMemoryTypeManaged) || at.device != c->dev)
    return SMAC_ERR_INVALID_ARG;
  HIPOK(hipStreamSynchronize(c->st));  // work queued on the old array ends first
  if (!c->bound[f]) c->own[f] = *slot;
  c->bound[f] = dev;
  *slot = dev;
  if (f == SMAC_FIELD_ALPHA) {
    c->alpha_nonzero = true;
    c->props_ready = false;
    c->props_ver++;
  }
  return SMAC_OK;
}

extern "C" int smac_download_device(smac_ctx* c, int32_t f, double* ddst, size_t count) {
  if (!c || !ddst || f < 0 || f >= SMAC_FIELD_COUNT) return SMAC_ERR_INVALID_ARG;
  if (count != smac_field_count(c, f)) return SMAC_ERR_INVALID_ARG;
  double** slot = field_slot(c, f);
  if (!*slot) return SMAC_ERR_INVALID_ARG;
  return copy_field(c, f, nullptr, nullptr, false, hipMemcpyDeviceToDevice, ddst, *slot);
}

// ------------------------------------------------------------------ the step
// update_properties_by_alpha + update_boundary_faces
// (CFD/G_SMACSolver_VoF.cu:1544-1566,1944-1950). initial = 0 keeps the
// previous density as rho_old (the rho/rho_old swap of k_swap_rho).
extern "C" int smac_update_properties(smac_ctx* c, int32_t initial) {
  if (!c) return SMAC_ERR_INVALID_ARG;
  set_props(c);
  if (c->bpois_on) {
    c->bpois_on = false;
    c->props_ver++;
  }
  if (!var_props(c)) {
    c->props_ready = true;
    return SMAC_OK;
  }
  int rc = SMAC_OK;
  double** arrs[] = {&c->rho,    &c->rho_old, &c->inv_rho, &c->mu,     &c->beta[0],
                     &c->beta[1], &c->beta[2], &c->muf[0],  &c->muf[1], &c->muf[2]};
  for (auto a : arrs)
    if (!rc && !*a) rc = dalloc(a, c->tot);
  if (!rc && !c->alpha) rc = dalloc(&c->alpha, c->tot);
  if (rc) return rc;
  const bool init = initial || !c->props_ready;
  if (!init) std::swap(c->rho, c->rho_old);
  KT(c, "props", launch_props(c->g, c->alpha, c->rho, c->rho_old, c->inv_rho, c->mu, c->beta, c->muf, c->fl.rho0,
                              c->fl.rho1, c->fl.mu0, c->fl.mu1, init ? 1 : 0, c->st));
  HIPOK(hipGetLastError());
  c->props_ready = true;
  return SMAC_OK;
}

static int ensure_props(smac_ctx* c) { return c->props_ready ? SMAC_OK : smac_update_properties(c, 1); }

extern "C" int smac_compute_mass_flux(smac_ctx* c, double dt) {
  (void)dt;
  if (!c) return SMAC_ERR_INVALID_ARG;
  // without a VoF flux mf = rho0*u_f, formed inside the predictor
  return SMAC_OK;
}

extern "C" int smac_predictor(smac_ctx* c, double dt) {
  if (!c || !(dt > 0)) return SMAC_ERR_INVALID_ARG;
  int rc = ensure_props(c);
  if (rc) return rc;
  Vel V;
  for (int a = 0; a < 3; a++) V.c[a] = c->u[a];
  const Props P = props_of(c);
  // right after a VoF transport of a two-phase grid the mass flux carries the
  // accumulated alpha flux, mf = rho0 u + (rho1 - rho0) F h/dt
  const double drho = c->fl.rho1 - c->fl.rho0;
  VofF M{};
  const bool vof = c->vof_armed && var_props(c) && drho != 0.0;
  if (vof) {
    const double inv_dt = 1. / dt;
    for (int a = 0; a < 3; a++) {
      M.F[a] = c->vfacc[a];
      M.hbydt[a] = c->d[a] * inv_dt;
    }
    M.drho = drho;
  }
  c->vof_armed = false;
  KT(c, "predict", launch_predict(c->g, var_props(c) ? &P : nullptr, V, c->p, c->us, dt, c->st, vof ? &M : nullptr));
  HIPOK(hipGetLastError());
  return SMAC_OK;
}

// --- two-way (CFD-DEM) variants, Coupling/main.cu:531-540,661-670
// G_SMACSolver::compute_mass_flux_from_alpha_flux_two_way +
// get_vof_vstar_rhouu_consistent_two_way + update_vstar_boundary
// (CFD/G_SMACSolver_VoF.cu:2472-2557; CFD/G_SMACSolver.cu:1538-2203,1044-1052)
extern "C" int smac_predictor_two_way(smac_ctx* c, double dt) {
  if (!c || !(dt > 0)) return SMAC_ERR_INVALID_ARG;
  if (!c->eps || !c->eps_old || !c->eps_half) return SMAC_ERR_INVALID_ARG;
  int rc = ensure_props(c);
  if (rc) return rc;
  Vel V;
  for (int a = 0; a < 3; a++) V.c[a] = c->u[a];
  const Props P = props_of(c);
  const double drho = c->fl.rho1 - c->fl.rho0;
  const double inv_dt = 1. / dt;
  VofF M{};
  const bool vof = c->vof_armed && var_props(c) && drho != 0.0;
  for (int a = 0; a < 3; a++) {
    M.F[a] = vof ? c->vfacc[a] : nullptr;
    M.hbydt[a] = c->d[a] * inv_dt;
  }
  M.drho = drho;
  M.eps_half = c->eps_half;
  c->vof_armed = false;
  EpsF E{};
  E.eps = c->eps;
  E.eps_old = c->eps_old;
  for (int a = 0; a < 3; a++) E.imp[a] = c->imp[a];
  E.inv_vol = c->g.inv_d[0] * c->g.inv_d[1] * c->g.inv_d[2];
  KT(c, "predict_two_way",
     launch_predict_two_way(c->g, var_props(c) ? &P : nullptr, V, c->p, c->us, dt, c->st, &M, E));
  HIPOK(hipGetLastError());
  return SMAC_OK;
}

// G_CFDDEMCoupling::update_poisson_beta_two_way (Coupling/G_CFDDEMCoupling.cu:419-459)
extern "C" int smac_update_poisson_beta_two_way(smac_ctx* c) {
  if (!c) return SMAC_ERR_INVALID_ARG;
  if (!c->eps) return SMAC_ERR_INVALID_ARG;
  int rc = ensure_props(c);
  if (rc) return rc;
  for (int a = 0; a < 3 && !rc; a++)
    if (!c->bpois[a]) rc = dalloc(&c->bpois[a], c->tot);
  if (rc) return rc;
  KT(c, "pois_two_way", launch_pois_two_way(c->g, var_props(c) ? c->beta : nullptr, c->eps, c->bpois, c->st));
  HIPOK(hipGetLastError());
  c->bpois_on = true;
  c->props_ver++;  // the multigrid rebuilds its coarse coefficients
  return SMAC_OK;
}

// --- multigrid setup
static int mg_auto_levels(const smac_ctx* c) {
  int n[3] = {c->n[0], c->n[1], c->n[2]};
  int L = 1;
  while (L < 12) {
    long cells = (long)n[0] * n[1] * n[2];
    if (cells <= 512) break;
    if ((n[0] & 1) || (n[1] & 1) || (n[2] & 1)) break;
    if (n[0] < 4 || n[1] < 4 || n[2] < 4) break;
    n[0] /= 2;
    n[1] /= 2;
    n[2] /= 2;
    L++;
  }
  return L;
}
static int mg_check_levels(const smac_ctx* c, int L) {
  if (L < 1 || L > 12) return 0;
  int n[3] = {c->n[0], c->n[1], c->n[2]};
  for (int l = 1; l < L; l++)
    for (int d = 0; d < 3; d++) {
      if (n[d] & 1) return 0;
      n[d] /= 2;
      if (n[d] < 2) return 0;
    }
  if ((long)n[0] * n[1] * n[2] > 1024) return 0;
  return 1;
}

static size_t prec_bytes(int prec) { return prec == SMAC_PREC_FP64 ? 8 : (prec == SMAC_PREC_MIXED_FP32 ? 4 : 2); }
static size_t prec_cbytes(int prec) { return prec == SMAC_PREC_FP64 ? 8 : 4; }

// Coarse face coefficients (two-phase grids) and the dense coarsest inverse.
static int mg_coeffs(smac_ctx* c) {
  const int L = (int)c->lev.size();
  const bool var = c->lev[0].lv.b[0] != nullptr;
  for (int l = 1; l < L && var; l++)
    for (int d = 0; d < 3; d++)
      launch_mg_coarsen_beta(c->lev[l - 1].lv, c->lev[l].lv, d, (double*)c->lev[l].lv.b[d], c->st);
  // the mixed forms' FP32 copies of every level's coefficients
  for (int l = 0; l < L && var; l++)
    for (int d = 0; d < 3; d++)
      if (c->lev[l].lv.bs[d]) {
        const Lv& V = c->lev[l].lv;
        launch_d2f(V.b[d], (float*)V.bs[d], (size_t)V.px * V.py * (size_t)(V.n[2] + 3), c->st);
      }
  HIPOK(hipGetLastError());
  const Lv& C = c->lev[L - 1].lv;
  const int n = C.n[0] * C.n[1] * C.n[2];
  if (!c->cinv) {
    int rc = dalloc(&c->cinv, (size_t)n * n);
    if (rc) return rc;
  }
  if (!c->cinv_ws) {
    int rc = dalloc(&c->cinv_ws, (size_t)n * n);
    if (rc) return rc;
  }
  // an outlet makes the operator regular: a plain inverse, else the
  // pseudo-inverse of the singular pure-Neumann operator; on the device
  // (row-major: the wave of row r reads the row's entries at consecutive
  // addresses), on the side stream after everything enqueued so far on the
  // context stream (the coarse coefficients above, the last solve's reads of
  // the old inverse); the pressure solve waits for ev_side_out before its
  // first V-cycle
  HIPOK(hipEventRecord(c->ev_side_in, c->st));
  HIPOK(hipStreamWaitEvent(c->st_side, c->ev_side_in, 0));
  launch_dense_pinv(C, c->has_outlet ? 0 : 1, c->cinv_ws, c->cinv, c->st_side);
  HIPOK(hipGetLastError());
  HIPOK(hipEventRecord(c->ev_side_out, c->st_side));
  c->pinv_pending = true;
  c->cinv_n = n;
  c->mg_ver = c->props_ver;
  return SMAC_OK;
}

static int mg_setup(smac_ctx* c, int L, int prec) {
  const bool var = pois_var(c);
  const BetaF pb = pois_of(c);
  if (c->mg_L == L && c->mg_prec == prec && !c->lev.empty() && c->lev[0].lv.beta0 == c->g.beta0 &&
      (c->lev[0].lv.b[0] != nullptr) == var && (!var || c->lev[0].lv.b[0] == pb.b[0])) {
    if (!var || c->mg_ver == c->props_ver) return SMAC_OK;
    return mg_coeffs(c);  // same hierarchy, new coefficients
  }
  mg_free(c);
  c->lev.resize(L);
  const size_t sb = prec_bytes(prec), cb = prec_cbytes(prec);
  for (int l = 0; l < L; l++) {
    Level& V = c->lev[l];
    memset(&V.lv, 0, sizeof(V.lv));
    for (int d = 0; d < 3; d++) {
      V.lv.n[d] = c->n[d] >> l;
      V.lv.per[d] = c->per[d];
      double h = c->d[d] * (double)(1 << l);
      V.lv.inv_d2[d] = 1.0 / (h * h);
    }
    // level 0 shares arrays with the fine grid (its pitch); coarser levels
    // take the same 64-element pitch rule
    V.lv.px = l == 0 ? c->g.px : box_pitch(V.lv.n[0]);
    V.lv.py = V.lv.n[1] + 3;
    V.lv.beta0 = c->g.beta0;
    size_t t = (size_t)V.lv.px * V.lv.py * (size_t)(V.lv.n[2] + 3);
    int rc = SMAC_OK;
    if (l == 0 && prec == SMAC_PREC_FP64) {
      // the FP64 fine level works on the PCG vectors themselves
      V.e = c->z;
      V.f = c->r;
      V.tb = c->res0;
    } else {
      rc = balloc(&V.e, t * sb, sb);
      V.own_e = rc == SMAC_OK;
      if (!rc) {
        rc = balloc(&V.f, t * sb, sb);
        V.own_f = rc == SMAC_OK;
      }
      if (!rc) {
        rc = balloc(&V.tb, t * cb, cb);
        V.own_tb = rc == SMAC_OK;
      }
    }
    if (var) {
      for (int d = 0; d < 3 && !rc; d++) {
        if (l == 0) {
          V.lv.b[d] = pb.b[d];
        } else {
          double* b = nullptr;
          rc = dalloc(&b, t);
          if (!rc) {
            c->mg_b.push_back(b);
            V.lv.b[d] = b;
          }
        }
        if (!rc && prec != SMAC_PREC_FP64) {
          void* bs = nullptr;
          rc = balloc(&bs, t * sizeof(float), sizeof(float));
          if (!rc) {
            c->mg_b.push_back((double*)bs);
            V.lv.bs[d] = (const float*)bs;
          }
        }
      }
    }
    if (rc) return rc;
  }
  c->mg_L = L;
  c->mg_prec = prec;
  return mg_coeffs(c);
}

template <typename T>
static void vcycle(smac_ctx* c, int l, int nu1, int nu2);

// Levels 1.. of the V-cycle as one graph launch. The graph is captured on the
// context stream the first time (and again when the sweep counts change); it
// holds only kernel nodes (the fills are kernels), and its arguments (level
// buffers, coefficients, coarse inverse) stay fixed until mg_free.
template <typename T>
static bool vcycle_graph(smac_ctx* c, int nu1, int nu2) {
  if (!c->graphs) return false;
  if (!c->cg_exec || c->cg_nu[0] != nu1 || c->cg_nu[1] != nu2) {
    graph_free(c);
    hipGraph_t gr = nullptr;
    if (hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      c->graphs = false;
      return false;
    }
    c->capturing = true;
    vcycle<T>(c, 1, nu1, nu2);
    c->capturing = false;
    hipError_t e1 = hipStreamEndCapture(c->st, &gr);
    hipError_t e2 = e1 == hipSuccess ? hipGraphInstantiate(&c->cg_exec, gr, nullptr, nullptr, 0) : e1;
    if (gr) (void)hipGraphDestroy(gr);
    if (e2 != hipSuccess) {
      c->cg_exec = nullptr;
      c->graphs = false;  // the launches above were captured, not run: fall back for good
      (void)hipGetLastError();
      return false;
    }
    c->cg_nu[0] = nu1;
    c->cg_nu[1] = nu2;
  }
  KT(c, "mg_coarse_levels", (void)hipGraphLaunch(c->cg_exec, c->st));
  return true;
}

template <typename T>
static void vcycle(smac_ctx* c, int l, int nu1, int nu2) {
  using Cc = typename St<T>::C;
  if (l == 1 && !c->capturing && (int)c->lev.size() > 2 && vcycle_graph<T>(c, nu1, nu2)) return;
  Level& V = c->lev[l];
  T* e = (T*)V.e;
  T* f = (T*)V.f;
  const bool fine = l == 0;
  if (l == (int)c->lev.size() - 1) {
    KT(c, "mg_coarse_solve", launch_mg_coarse<T>(V.lv, c->cinv, f, e, c->st));
    return;
  }
  // the first red half sweep starts from e = 0 and writes every red cell; the
  // black half sweep that follows reads only red cells, so no fill is needed
  if (nu1 == 0) KT(c, fine ? "mg_fill_l0" : "mg_coarse_levels", launch_fill<T>(V.lv, e, c->st));
  if (fine) {
    // the first red half sweep starts from zero (reads f, writes e: one word
    // less per cell than the others), timed under its own name, unless the
    // residual update has done it; the other pre-smoothing half sweeps run
    // back to back and are timed together
    if (nu1 > 0) {
      if (!c->zs_done) KT(c, "mg_rbgs_l0_zs", launch_mg_rbgs<T>(V.lv, e, f, 0, 1, c->st));
      c->zs_done = false;
      KT_N(c, "mg_rbgs_l0", 2 * nu1 - 1, {
        launch_mg_rbgs<T>(V.lv, e, f, 1, 0, c->st);
        for (int s = 1; s < nu1; s++) {
          launch_mg_rbgs<T>(V.lv, e, f, 0, 0, c->st);
          launch_mg_rbgs<T>(V.lv, e, f, 1, 0, c->st);
        }
      });
    }
  } else {
    for (int s = 0; s < nu1; s++) {
      KT(c, "mg_coarse_levels", launch_mg_rbgs<T>(V.lv, e, f, 0, s == 0, c->st));
      KT(c, "mg_coarse_levels", launch_mg_rbgs<T>(V.lv, e, f, 1, 0, c->st));
    }
  }
  Level& C = c->lev[l + 1];
  KT(c, fine ? "mg_restrict_l0" : "mg_coarse_levels",
     launch_mg_restrict<T>(V.lv, C.lv, e, f, (Cc*)V.tb, (T*)C.f, nu1 > 0, c->st));
  vcycle<T>(c, l + 1, nu1, nu2);
  // trilinear prolongation: with post-smoothing the black cells are recomputed
  // before they are read, so only the red cells take the correction
  KT(c, fine ? "mg_prolong_l0" : "mg_coarse_levels",
     launch_mg_prolong<T>(C.lv, V.lv, (const T*)C.e, e, nu2 > 0 ? 0 : -1, c->st));
  if (fine) {
    // the last half sweep also forms the PCG dots of the new z when asked
    const bool dot = c->dot_req && nu2 > 0 && mg_rbgs_dot_applies<T>(V.lv, e, f);
    if (nu2 > 0)
      KT_N(c, "mg_rbgs_l0", dot ? 2 * nu2 - 1 : 2 * nu2, {
        for (int s = 0; s < nu2; s++) {
          launch_mg_rbgs<T>(V.lv, e, f, 1, 0, c->st);
          if (!dot || s + 1 < nu2) launch_mg_rbgs<T>(V.lv, e, f, 0, 0, c->st);
        }
      });
    if (dot) {
      KT(c, "mg_rbgs_dot",
         launch_mg_rbgs_dot<T>(V.lv, e, f, c->dot_r, c->dot_Ad, c->dot_flex, c->dot_p0, c->dot_p1, c->st));
      c->dot_done = true;
    }
  } else {
    for (int s = 0; s < nu2; s++) {
      KT(c, "mg_coarse_levels", launch_mg_rbgs<T>(V.lv, e, f, 1, 0, c->st));
      KT(c, "mg_coarse_levels", launch_mg_rbgs<T>(V.lv, e, f, 0, 0, c->st));
    }
  }
}

// G_GMGSolver::initialize + create_coeffs (CFD/pressure_solver/G_GMGSolver.cu:10-147,913-922).
// Returns the level count or a negative status: an odd size in a coarsened
// direction (the reference aborts) or a single level (the reference then
// prolongs through levels_[-1], G_GMGSolver.cu:971).
static int rg_setup(smac_ctx* c, int req) {
  if (req < 1 || req > 16) return SMAC_ERR_INVALID_ARG;
  const uint64_t ver = pois_var(c) ? c->props_ver : 0;
  if (!c->rg.empty() && c->rg_req == req && c->rg_beta0 == c->g.beta0 && c->rg_ver == ver) return (int)c->rg.size();
  rg_free(c);
  std::vector<RgLv> lv;
  int prev[3] = {c->n[0], c->n[1], c->n[2]};
  double pinv[3] = {c->g.inv_d2[0], c->g.inv_d2[1], c->g.inv_d2[2]};
  for (int l = 0; l < req; l++) {
    RgLv V;
    memset(&V, 0, sizeof(V));
    int any = 0;
    for (int d = 0; d < 3; d++) {
      V.co[d] = !(prev[d] / 2 < 6);
      any |= V.co[d];
    }
    if (l > 0 && !any) break;
    for (int d = 0; d < 3; d++)
      if (V.co[d] && (prev[d] & 1)) return SMAC_ERR_INVALID_GRID;
    for (int d = 0; d < 3; d++) {
      V.n[d] = V.co[d] ? prev[d] / 2 : prev[d];
      V.inv_d2[d] = V.co[d] ? pinv[d] * 0.25 : pinv[d];
      V.per[d] = c->per[d];
      prev[d] = V.n[d];
      pinv[d] = V.inv_d2[d];
    }
    V.px = V.n[0] + 3;
    V.py = V.n[1] + 3;
    lv.push_back(V);
  }
  if (lv.size() < 2) return SMAC_ERR_INVALID_GRID;
  c->rg.resize(lv.size());
  int rc = dalloc(&c->rgt, c->tot);
  for (size_t l = 0; l < lv.size() && !rc; l++) {
    RgLevel& V = c->rg[l];
    V.lv = lv[l];
    V.tot = (size_t)V.lv.px * V.lv.py * (size_t)(V.lv.n[2] + 3);
    double** arrs[] = {&V.q, &V.qt, &V.rhs, &V.b[0], &V.b[1], &V.b[2]};
    for (auto a : arrs)
      if (!rc) rc = dalloc(a, V.tot);
  }
  if (rc) {
    rg_free(c);
    return rc;
  }
  // face coefficients (create_coeffs); rebuilt when the grid's coefficients change
  const BetaF gb = pois_of(c);
  for (size_t l = 0; l < c->rg.size(); l++)
    for (int d = 0; d < 3; d++)
      launch_rg_beta(c->g, ver ? &gb : nullptr, c->rg[l].lv, l ? c->rg[l - 1].lv : c->rg[l].lv,
                     l ? c->rg[l - 1].b[d] : nullptr, d, c->rg[l].b[d], c->st);
  HIPOK(hipGetLastError());
  c->rg_req = req;
  c->rg_beta0 = c->g.beta0;
  c->rg_ver = ver;
  return (int)c->rg.size();
}

// v_cycle_as_preconditioner (CFD/pressure_solver/G_GMGSolver.cu:925-998): z = M^{-1} r
static void rg_vcycle(smac_ctx* c, const smac_solver_opts& o) {
  const Geo& g = c->g;
  hipStream_t s = c->st;
  const int nf = o.ref_sweeps, nc = o.ref_coarse_sweeps, L = (int)c->rg.size();
  const BetaF gb = pois_of(c);
  const BetaF* gbp = pois_var(c) ? &gb : nullptr;
  double* x = c->z;
  double* t = c->rgt;
  for (auto& V : c->rg) KT(c, "rg_levels", launch_rg_zero(V.q, V.tot, s));
  if (nf == 0) KT(c, "rg_fine", launch_rg_zero(x, c->tot, s));
  for (int it = 0; it < nf; it++) {
    KT(c, "rg_fine_jacobi", launch_rg_fine_jacobi(g, gbp, x, t, c->r, it == 0, s));
    std::swap(x, t);
  }
  KT(c, "rg_fine_restrict", launch_rg_fine_res_restrict(g, gbp, x, c->r, c->rg[0].lv, c->rg[0].rhs, s));
  for (int l = 0; l < L - 1; l++) {
    RgLevel& V = c->rg[l];
    for (int it = 0; it < nf; it++) {
      KT(c, "rg_levels", launch_rg_jacobi(V.lv, V.b, V.q, V.qt, V.rhs, s));
      std::swap(V.q, V.qt);
    }
    KT(c, "rg_levels", launch_rg_res_restrict(V.lv, V.b, V.q, V.rhs, c->rg[l + 1].lv, c->rg[l + 1].rhs, s));
  }
  {
    RgLevel& V = c->rg[L - 1];
    for (int it = 0; it < nc; it++) {
      KT(c, "rg_levels", launch_rg_jacobi(V.lv, V.b, V.q, V.qt, V.rhs, s));
      std::swap(V.q, V.qt);
    }
    KT(c, "rg_levels", launch_rg_prolong(V.lv, V.q, c->rg[L - 2].lv, c->rg[L - 2].q, s));
  }
  for (int l = L - 2; l >= 0; l--) {
    RgLevel& V = c->rg[l];
    for (int it = 0; it < nf; it++) {
      KT(c, "rg_levels", launch_rg_jacobi(V.lv, V.b, V.q, V.qt, V.rhs, s));
      std::swap(V.q, V.qt);
    }
    if (l != 0) KT(c, "rg_levels", launch_rg_prolong(V.lv, V.q, c->rg[l - 1].lv, c->rg[l - 1].q, s));
    else KT(c, "rg_fine_prolong", launch_rg_prolong_fine(V.lv, V.q, g, x, s));
  }
  for (int it = 0; it < nf; it++) {
    KT(c, "rg_fine_jacobi", launch_rg_fine_jacobi(g, gbp, x, t, c->r, 0, s));
    std::swap(x, t);
  }
  // 2*nf sweeps: the result is back in z
}

static FinalizeArgs fin(const smac_ctx* c) {
  FinalizeArgs fa;
  memset(&fa, 0, sizeof(fa));
  fa.nrows = c->nrows;
  fa.stage = c->fstage;
  return fa;
}
static void fin_add(FinalizeArgs& fa, const double* part, int kind, int out) {
  fa.part[fa.count] = part;
  fa.kind[fa.count] = kind;
  fa.out[fa.count] = out;
  fa.count++;
}
static int read_scalars(smac_ctx* c) {
  HIPOK(hipMemcpyAsync(c->Sh, c->S, S_COUNT * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HIPOK(hipStreamSynchronize(c->st));
  return SMAC_OK;
}

extern "C" void smac_default_opts(smac_solver_opts* o) {
  if (!o) return;
  o->precond = SMAC_PRECOND_MG;
  o->precision = SMAC_PREC_FP64;
  o->stop_rule = SMAC_STOP_REL_L2;
  o->tol = 1e-8;
  o->check_every = 1;
  o->max_iter = 2000;
  o->mg_levels = 0;
  o->nu1 = 2;
  o->nu2 = 2;
  o->flexible = 0;
  o->ref_sweeps = 4;
  o->ref_coarse_sweeps = 30;
  o->stationary = 0;
}

// The PCG of G_PCGSolver::solve_* (CFD/pressure_solver/G_PCGSolver.cu:307-922)
// with storage T for the preconditioner vectors: z and the multigrid levels in
// T, r, d, Ad and dp in FP64. For T = double z is c->z and the fine level works
// on r directly; otherwise the fine-level rhs f0 = (T) r is refreshed by the
// update kernel.
// Algorithmic HBM bytes per interior cell of one pressure solve: each array
// read or written once per pass (the canonical model of the step; halo re-reads
// and the implementation's extra passes are not counted). W = smoother width.
static double solve_bytes_model(const smac_ctx* c, const smac_solver_opts& o, int L, int iters, bool rhs_given) {
  const bool var = pois_var(c);
  const double W = o.precision == SMAC_PREC_MIXED_FP32 ? 4.0 : o.precision == SMAC_PREC_MIXED_BF16 ? 2.0 : 8.0;
  double P;  // one application of the preconditioner
  if (o.precond == SMAC_PRECOND_MG) {
    P = 17.43 * W + (var ? (15.0 * 8.0 / 7.0) * W : 0.0);
  } else if (o.precond == SMAC_PRECOND_JACOBI) {
    P = 16.0 + (var ? 24.0 : 0.0);
  } else {  // the reference V-cycle: 8 sweeps of 3 words + 4.25 words of transfers per level, 30 coarsest sweeps
    double words = 0.0, f = 1.0;
    for (int l = 0; l < L - 1; l++, f /= 8.0) words += f * ((8.0 * 3.0 + 4.25) + (var ? 8.0 * 3.0 + 3.0 : 0.0));
    words += f * 30.0 * (var ? 6.0 : 3.0);
    P = words * 8.0;
  }
  const double rhs = rhs_given ? 8.0 : 32.0 + ((c->eps && c->eps_old) ? 16.0 : 0.0);
  const double it = 104.0 + P + (o.flexible ? 16.0 : 0.0) + (var ? 24.0 : 0.0);
  return rhs + (32.0 + P) + iters * it;
}

template <typename T>
static int solve_t(smac_ctx* c, double dt, const smac_solver_opts& o, smac_solve_stats* st, int L, bool rhs_given) {
  const Geo& g = c->g;
  double Ncell = (double)c->n[0] * (double)c->n[1] * (double)c->n[2];
  int pure = pure_neumann(c);
  double inv_dt = 1. / dt;
  const bool mg = o.precond == SMAC_PRECOND_MG;
  const bool refg = o.precond == SMAC_PRECOND_REF_GMG;  // FP64 only (checked by the caller)
  T* z = mg ? (T*)c->lev[0].e : (T*)c->z;
  T* f0 = (mg && (void*)c->lev[0].f != (void*)c->r) ? (T*)c->lev[0].f : nullptr;
  const bool var = pois_var(c);
  const BetaF bf = pois_of(c);
  const BetaF* bfp = var ? &bf : nullptr;
  auto precond = [&]() {
    if (refg) {
      rg_vcycle(c, o);
      return;
    }
    if (!mg) {
      KT(c, "jacobi", launch_jacobi(c->g, bfp, c->r, c->z, c->st));
      return;
    }
    vcycle<T>(c, 0, o.nu1, o.nu2);
  };
  hipStream_t s = c->st;
  double** P = c->part;

  if (rhs_given) launch_rowsum_nat(g, c->rhs, P[0], s);  // b supplied in c->rhs (adjoint solve)
  else if (c->eps && c->eps_old)  // with the fluid volume fraction (CFD-DEM grids)
    KT(c, "rhs", launch_rhs_eps(g, c->us, c->eps, c->eps_old, c->rhs, inv_dt, P[0], s));
  else KT(c, "rhs", launch_rhs(g, c->us, c->rhs, inv_dt, P[0], s));
  if (pure) {
    FinalizeArgs fa = fin(c);
    fin_add(fa, P[0], 0, S_SUMB);
    fa.op = OP_MEAN;
    fa.dst = S_MEANB;
    fa.sa = S_SUMB;
    fa.c = Ncell;
    launch_finalize(fa, c->S, s);
  }
  // r = b - mean b; |r|^2 and sum r; then (pure Neumann) r -= mean r, and
  // f0 = (T) r. dp is zeroed only when the last correction has not cleared it.
  double size_inv = 1. / Ncell;
  const int zero_dp = c->dp_zero ? 0 : 1;
  c->dp_zero = false;
  launch_init_r<T>(g, c->rhs, c->r, f0, c->dp, zero_dp, c->S, pure ? S_MEANB : -1, 0.0, -1, P[0], P[1], 1,
                   pure ? 0 : 1, s);
  {
    FinalizeArgs fa = fin(c);
    fin_add(fa, P[0], 0, S_BB);
    fin_add(fa, P[1], 0, S_SUMR);
    launch_finalize(fa, c->S, s);
  }
  if (pure)
    launch_init_r<T>(g, c->rhs, c->r, f0, c->dp, zero_dp, c->S, S_MEANB, -1. * size_inv, S_SUMR, P[0], P[1], 0, 1,
                     s);
  precond();
  if (pure) {
    launch_sum<T>(g, z, P[0], s);
    FinalizeArgs fa = fin(c);
    fin_add(fa, P[0], 0, S_SUMZ);
    launch_finalize(fa, c->S, s);
  }
  // z -= mean(z) (pure Neumann) in the same pass as d = z and r.z; rows of the
  // pair form also take the first Ad = A d there (its d.Ad partials in P[4])
  const bool first_fused = first_spmv_applies<T>(g, z, c->r, c->dv, c->Ad);
  if (first_fused)
    KT(c, "first_spmv", launch_first_spmv<T>(g, bfp, z, c->r, c->dv, c->Ad, P[4], P[0], -1. * size_inv, c->S,
                                             pure ? S_SUMZ : -1, s));
  else
    launch_copy_dot<T>(g, z, c->dv, c->r, P[0], s, -1. * size_inv, c->S, pure ? S_SUMZ : -1);
  {
    FinalizeArgs fa = fin(c);
    fin_add(fa, P[0], 0, S_RZ);
    launch_finalize(fa, c->S, s);
  }
  HIPOK(hipGetLastError());
  int rc = read_scalars(c);
  if (rc) return rc;
  double bb = c->Sh[S_BB];
  double bnorm = sqrt(bb);
  int iter = 0, conv = 0;
  double rr = bb, linf = 0.0;
  if (bb == 0.0) conv = 1;
  // a NaN/Inf right-hand side (a NaN or overflow in u*) stops the solve at once
  bool nan = !std::isfinite(bb);
  bool have_update = false, dp_pending = false;
  // stationary iteration: alpha = 1, beta = 0 (d = z) throughout
  const bool stat = o.stationary != 0;
  if (stat) {
    launch_scalar_set(c->S, S_ALPHA, 1.0, s);
    launch_scalar_set(c->S, S_BETA, 0.0, s);
  }
  // the direction update of an iteration rides on the next iteration's SpMV
  // where the rows have the pair form; pAp partials then come in P[4]
  if (!c->dv2 && !nan && !conv) {
    rc = dalloc(&c->dv2, c->tot);
    if (rc) return rc;
  }
  const bool fuse = c->dv2 && dir_spmv_applies<T>(g, z, c->dv, c->dv2, c->dp, c->Ad);
  // the zero-start red half sweep of level 0 rides on the residual update
  // (uniform coefficients: pointwise f/diag)
  const int zs_mode = (mg && !var && o.nu1 >= 1)
                          ? mg_zs_division<T>(c->lev[0].lv, (const T*)c->lev[0].e, (const T*)c->lev[0].f)
                          : 0;
  c->zs_done = false;
  bool ad_ready = first_fused;
  for (iter = 0; iter < o.max_iter && !conv && !nan; ++iter) {
    if (!ad_ready) KT(c, "spmv_dot", launch_spmv_dot(g, bfp, c->dv, c->Ad, P[0], s));
    double* pap = ad_ready ? P[4] : P[0];
    ad_ready = false;
    if (!stat) {
      FinalizeArgs fa = fin(c);
      fin_add(fa, pap, 0, S_PAP);
      fa.op = OP_DIV;
      fa.dst = S_ALPHA;
      fa.sa = S_RZ;
      fa.sb = S_PAP;
      launch_finalize(fa, c->S, s);
    }
    if (zs_mode)
      KT(c, "update_zs",
         c->zs_done = launch_update<T>(g, c->r, c->Ad, c->S, S_ALPHA, P[0], P[1], f0, s, z, zs_mode, &c->lev[0].lv));
    else
      KT(c, "update", launch_update<T>(g, c->r, c->Ad, c->S, S_ALPHA, P[0], P[1], f0, s));
    have_update = true;
    dp_pending = true;  // dp += alpha d waits for the next xpay (or the flush)
    if (refg && pure) {
      // solve_pureneumann_gmgpcg removes the mean of r after every update
      // (G_PCGSolver.cu:866); the check norms are taken after it
      launch_sum<double>(g, c->r, P[2], s);
      FinalizeArgs fa = fin(c);
      fin_add(fa, P[2], 0, S_SUMR);
      launch_finalize(fa, c->S, s);
      KT(c, "add_scalar_norms", launch_add_scalar_norms(g, c->r, -1. * size_inv, c->S, S_SUMR, P[0], P[1], s));
    }
    c->dot_req = !stat;
    c->dot_done = false;
    c->dot_r = c->r;
    c->dot_Ad = c->Ad;
    c->dot_flex = o.flexible;
    c->dot_p0 = P[2];
    c->dot_p1 = P[3];
    precond();
    c->dot_req = false;
    if (!stat && !c->dot_done) KT(c, "dot2", launch_dot2<T>(g, c->r, z, c->Ad, o.flexible, P[2], P[3], s));
    bool check = (iter % o.check_every) == 0;
    {
      if (!stat) {
        FinalizeArgs fa = fin(c);
        fin_add(fa, P[2], 0, S_RZN);
        if (o.flexible) fin_add(fa, P[3], 0, S_ZAD);
        launch_finalize(fa, c->S, s);
      }
      FinalizeArgs fb = fin(c);
      if (check) {
        fin_add(fb, P[0], 0, S_RR);
        fin_add(fb, P[1], 1, S_LINF);
      }
      if (stat) {
        // no scalar operation: alpha and beta stay 1 and 0
      } else if (o.flexible) {
        fb.op = OP_BETA_PR;
        fb.dst = S_BETA;
        fb.sa = S_ALPHA;
        fb.sb = S_ZAD;
        fb.sc = S_RZ;
      } else {
        fb.op = OP_DIV;
        fb.dst = S_BETA;
        fb.sa = S_RZN;
        fb.sb = S_RZ;
      }
      if (fb.count > 0 || fb.op != OP_NONE) launch_finalize(fb, c->S, s);
    }
    if (check) {
      HIPOK(hipGetLastError());
      rc = read_scalars(c);
      if (rc) return rc;
      if (!std::isfinite(c->Sh[S_RR]) || !std::isfinite(c->Sh[S_LINF])) {
        iter++;
        nan = true;
        break;
      }
      if (o.stop_rule == SMAC_STOP_REF_LINF_DT) {
        linf = c->Sh[S_LINF] * dt;
        if (linf < o.tol) {
          iter++;
          conv = 1;
          break;
        }
      } else {
        rr = c->Sh[S_RR];
        if (sqrt(rr) < o.tol * bnorm) {
          iter++;
          conv = 1;
          break;
        }
      }
    }
    if (fuse && iter + 1 < o.max_iter) {
      KT(c, "dir_spmv", launch_dir_spmv<T>(g, bfp, z, c->dv, c->dv2, c->dp, c->Ad, P[4], c->S, S_ALPHA, S_BETA, s));
      std::swap(c->dv, c->dv2);
      ad_ready = true;
    } else {
      KT(c, "xpay", launch_xpay<T>(g, z, c->dv, c->dp, c->S, S_ALPHA, S_BETA, s));
    }
    dp_pending = false;
    if (!stat) launch_scalar_copy(c->S, S_RZ, S_RZN, s);
  }
  // the last update's dp += alpha d (S_ALPHA still holds its alpha)
  if (dp_pending) {
    if (c->defer_dp) c->dp_deferred = true;
    else KT(c, "dp_flush", launch_dp_flush(g, c->dp, c->dv, c->S, S_ALPHA, s));
  }
  HIPOK(hipGetLastError());
  if (st) {
    memset(st, 0, sizeof(*st));
    st->iters = iter;
    st->converged = conv;
    st->pure_neumann = pure;
    st->mg_levels = (mg || refg) ? L : 0;
    st->bnorm = bnorm;
    if (have_update) {
      FinalizeArgs fa = fin(c);
      fin_add(fa, P[0], 0, S_RR);
      fin_add(fa, P[1], 1, S_LINF);
      launch_finalize(fa, c->S, s);
      rc = read_scalars(c);
      if (rc) return rc;
      rr = c->Sh[S_RR];
      st->linf_res_dt = c->Sh[S_LINF] * dt;
    } else {
      rr = bb;
      st->linf_res_dt = 0.0;
    }
    st->rel_res = bnorm > 0 ? sqrt(rr) / bnorm : 0.0;
    st->nan_detected = nan || !(rr == rr);
    st->bytes_model = solve_bytes_model(c, o, L, iter, rhs_given) * Ncell;
  }
  if (nan || !(bb == bb) || !(rr == rr)) return SMAC_ERR_NAN;
  return SMAC_OK;
}

static int pressure_solve(smac_ctx* c, double dt, const smac_solver_opts* opts, smac_solve_stats* st,
                          bool rhs_given);

extern "C" int smac_pressure_solve(smac_ctx* c, double dt, const smac_solver_opts* opts, smac_solve_stats* st) {
  return pressure_solve(c, dt, opts, st, false);
}

static int pressure_solve(smac_ctx* c, double dt, const smac_solver_opts* opts, smac_solve_stats* st,
                          bool rhs_given) {
  if (!c || !(dt > 0)) return SMAC_ERR_INVALID_ARG;
  smac_solver_opts o;
  if (opts) o = *opts;
  else smac_default_opts(&o);
  if (o.check_every < 1) o.check_every = 1;
  if (o.precond != SMAC_PRECOND_JACOBI && o.precond != SMAC_PRECOND_MG && o.precond != SMAC_PRECOND_REF_GMG)
    return SMAC_ERR_INVALID_ARG;
  if (o.precision < SMAC_PREC_FP64 || o.precision > SMAC_PREC_MIXED_BF16) return SMAC_ERR_INVALID_ARG;
  if (o.precond != SMAC_PRECOND_MG && o.precision != SMAC_PREC_FP64) return SMAC_ERR_UNSUPPORTED;
  if (o.precond == SMAC_PRECOND_REF_GMG && (o.ref_sweeps < 0 || o.ref_coarse_sweeps < 0))
    return SMAC_ERR_INVALID_ARG;
  if (c->n[0] > 4000) return SMAC_ERR_INVALID_GRID;  // restriction row buffer in LDS
  {
    int rc = ensure_props(c);
    if (rc) return rc;
  }
  int L = 1;
  if (o.precond == SMAC_PRECOND_MG) {
    L = o.mg_levels > 0 ? o.mg_levels : mg_auto_levels(c);
    if (!mg_check_levels(c, L)) return SMAC_ERR_INVALID_GRID;
    int rc = mg_setup(c, L, o.precision);
    if (rc) return rc;
    if (c->pinv_pending) {  // the coarsest inverse of the current coefficients
      HIPOK(hipStreamWaitEvent(c->st, c->ev_side_out, 0));
      c->pinv_pending = false;
    }
  } else if (o.precond == SMAC_PRECOND_REF_GMG) {
    L = rg_setup(c, o.mg_levels > 0 ? o.mg_levels : 4);
    if (L < 0) return L;
  }
  switch (o.precision) {
    case SMAC_PREC_MIXED_FP32: return solve_t<float>(c, dt, o, st, L, rhs_given);
    case SMAC_PREC_MIXED_BF16: return solve_t<bf16s>(c, dt, o, st, L, rhs_given);
    default: return solve_t<double>(c, dt, o, st, L, rhs_given);
  }
}

extern "C" int smac_correct(smac_ctx* c, double dt) {
  if (!c || !(dt > 0)) return SMAC_ERR_INVALID_ARG;
  int rc = ensure_props(c);
  if (rc) return rc;
  const BetaF bf = beta_of(c);
  KT(c, "correct", launch_correct_all(c->g, var_props(c) ? &bf : nullptr, c->us, c->dp, c->u, c->p, dt, c->st,
                                      c->dv, c->S, c->dp_deferred ? S_ALPHA : -1));
  c->dp_deferred = false;
  c->dp_zero = true;
  HIPOK(hipGetLastError());
  return SMAC_OK;
}

extern "C" int smac_step(smac_ctx* c, double dt, const smac_solver_opts* o, smac_solve_stats* st) {
  if (!c) return SMAC_ERR_INVALID_ARG;
  smac_solve_stats local;
  smac_solve_stats* sp = st ? st : &local;
  prof_begin(c, 0);
  // properties: built once, then renewed every step of a two-phase grid
  // (the order of the reference loop, CFD/main.cu:270-317)
  int rc = SMAC_OK;
  if (!c->props_ready) rc = smac_update_properties(c, 1);
  else if (var_props(c) && (c->fl.rho0 != c->fl.rho1 || c->fl.mu0 != c->fl.mu1)) rc = smac_update_properties(c, 0);
  if (rc) return rc;
  // the multigrid coefficients of the new properties before the predictor,
  // so that the coarsest inverse builds on the side stream meanwhile
  if (o && o->precond == SMAC_PRECOND_MG && o->precision >= SMAC_PREC_FP64 && o->precision <= SMAC_PREC_MIXED_BF16 &&
      pois_var(c) && c->n[0] <= 4000) {
    const int L = o->mg_levels > 0 ? o->mg_levels : mg_auto_levels(c);
    if (mg_check_levels(c, L)) {
      rc = mg_setup(c, L, o->precision);
      if (rc) return rc;
    }
  }
  rc = smac_predictor(c, dt);
  if (rc) return rc;
  double t_pred = prof_end(c, 0);
  prof_begin(c, 2);
  c->defer_dp = true;
  rc = smac_pressure_solve(c, dt, o, sp);
  c->defer_dp = false;
  if (rc) {
    if (c->dp_deferred) {  // leave dp complete on an error return
      launch_dp_flush(c->g, c->dp, c->dv, c->S, S_ALPHA, c->st);
      c->dp_deferred = false;
    }
    return rc;
  }
  double t_solve = prof_end(c, 2);
  prof_begin(c, 4);
  rc = smac_correct(c, dt);
  if (c->dp_deferred) {  // the correction returned early: leave dp complete
    launch_dp_flush(c->g, c->dp, c->dv, c->S, S_ALPHA, c->st);
    c->dp_deferred = false;
  }
  double t_corr = prof_end(c, 4);
  sp->ms_predictor = t_pred;
  sp->ms_solve = t_solve;
  sp->ms_correct = t_corr;
  // predictor 56 and correction 72 bytes per cell (+24 each with property arrays)
  sp->bytes_model += (128.0 + (var_props(c) ? 48.0 : 0.0)) * (double)c->n[0] * (double)c->n[1] * (double)c->n[2];
  return rc;
}

// ------------------------------------------------------------------ VoF transport
static int vof_alloc(smac_ctx* c) {
  if (c->alpha_new) return SMAC_OK;
  int rc = c->alpha ? SMAC_OK : dalloc(&c->alpha, c->tot);
  if (!rc) rc = dalloc(&c->alpha_new, c->tot);
  for (int a = 0; a < 3 && !rc; a++) rc = dalloc(&c->vf[a], c->tot);
  for (int a = 0; a < 3 && !rc; a++) rc = dalloc(&c->vfacc[a], c->tot);
  if (!rc) rc = balloc((void**)&c->vcoeff, c->tot, 1);
  return rc;
}

extern "C" int smac_clear_alpha_flux_accum(smac_ctx* c) {
  if (!c) return SMAC_ERR_INVALID_ARG;
  int rc = vof_alloc(c);
  if (rc) return rc;
  for (int a = 0; a < 3; a++) HIPOK(hipMemsetAsync(c->vfacc[a], 0, c->tot * sizeof(double), c->st));
  c->vof_armed = false;
  return SMAC_OK;
}

extern "C" int smac_alpha_flux_split(smac_ctx* c, double sub_dt, int32_t step) {
  if (!c || !(sub_dt > 0) || step < 0) return SMAC_ERR_INVALID_ARG;
  int rc = vof_alloc(c);
  if (rc) return rc;
  Vel V;
  for (int a = 0; a < 3; a++) V.c[a] = c->u[a];
  KT(c, "vof_substep", launch_vof_substep(c->g, c->d, V, &c->alpha, &c->alpha_new, c->vf, c->vfacc, c->vcoeff,
                                          sub_dt, step, c->st));
  HIPOK(hipGetLastError());
  if (c->bound[SMAC_FIELD_ALPHA] && c->alpha != c->bound[SMAC_FIELD_ALPHA]) {
    // the sweeps ping-pong alpha and its scratch array: a bound alpha takes the result back
    HIPOK(hipMemcpyAsync(c->alpha_new, c->alpha, c->tot * sizeof(double), hipMemcpyDeviceToDevice, c->st));
    std::swap(c->alpha, c->alpha_new);
  }
  c->vof_armed = true;
  c->props_ver++;  // the face coefficients follow the new alpha at the next property update
  return SMAC_OK;
}

extern "C" int smac_transport_alpha(smac_ctx* c, double dt, int32_t substeps, int32_t step) {
  if (!c || !(dt > 0) || substeps < 1 || step < 0) return SMAC_ERR_INVALID_ARG;
  int rc = smac_clear_alpha_flux_accum(c);
  const double sub_dt = dt / (double)substeps;
  for (int s = 0; s < substeps && !rc; s++) rc = smac_alpha_flux_split(c, sub_dt, step);
  return rc;
}

// ------------------------------------------------------------------ after the step
extern "C" int smac_update_boundary_ghost(smac_ctx* c, double dt) {
  if (!c) return SMAC_ERR_INVALID_ARG;
  int rc = ensure_props(c);
  if (rc) return rc;
  KT(c, "ghost_pressure", launch_ghost_pressure(c->g, c->d, c->p, var_props(c) ? c->rho : nullptr, c->st));
  // the outlet faces first: the velocity ghosts read them as their inner rows
  if (c->has_outlet) KT(c, "outlet_faces", launch_outlet_faces(c->g, c->us, c->beta, c->dp, c->u, dt, c->st));
  KT(c, "ghost_velocity", launch_ghost_velocity(c->g, c->u, c->st));
  HIPOK(hipGetLastError());
  return SMAC_OK;
}

extern "C" int smac_make_face_gradp(smac_ctx* c) {
  if (!c) return SMAC_ERR_INVALID_ARG;
  for (int a = 0; a < 3; a++)
    if (!c->gradp[a]) {
      int rc = dalloc(&c->gradp[a], c->tot);
      if (rc) return rc;
    }
  KT(c, "face_gradp", launch_face_gradp(c->g, c->p, c->gradp, c->st));
  HIPOK(hipGetLastError());
  return SMAC_OK;
}

extern "C" int smac_cell_output(smac_ctx* c, float* h, size_t count) {
  if (!c || !h) return SMAC_ERR_INVALID_ARG;
  const size_t n = (size_t)c->n[0] * (size_t)c->n[1] * (size_t)c->n[2];
  if (count != 5 * n) return SMAC_ERR_INVALID_ARG;
  if (!c->cellout) {
    int rc = balloc((void**)&c->cellout, 5 * n * sizeof(float), sizeof(float));
    if (rc) return rc;
  }
  const double* u[3] = {c->u[0], c->u[1], c->u[2]};
  KT(c, "pack_cells", launch_pack_cells(c->g, c->p, c->alpha, u, c->cellout, c->st));
  HIPOK(hipGetLastError());
  HIPOK(hipMemcpyAsync(h, c->cellout, 5 * n * sizeof(float), hipMemcpyDeviceToHost, c->st));
  HIPOK(hipStreamSynchronize(c->st));
  return SMAC_OK;
}

// Adjoint of one smac_step (see smac_adjoint.hip for the derivation).
extern "C" int smac_adjoint_step(smac_ctx* c, double dt, const smac_solver_opts* opts, const double* const u_n[3],
                                 double* const ubar[3], double* pbar, smac_solve_stats* st) {
  if (!c || !(dt > 0) || !u_n || !ubar || !pbar) return SMAC_ERR_INVALID_ARG;
  c->dp_zero = false;  // the adjoint's solves and copies leave dp as they please
  for (int a = 0; a < 3; a++)
    if (!u_n[a] || !ubar[a]) return SMAC_ERR_INVALID_ARG;
  // the adjoint holds alpha fixed (mass flux rho0 u): a pending VoF flux is
  // outside it
  if (c->vof_armed) return SMAC_ERR_UNSUPPORTED;
  // the adjoint differentiates the one-fluid step: no volume fraction terms
  if (c->eps || c->bpois_on) return SMAC_ERR_UNSUPPORTED;
  // outlet sides (the star-BC copy and the Dirichlet faces) are outside it
  if (c->has_outlet) return SMAC_ERR_UNSUPPORTED;
  int rc = ensure_props(c);
  if (rc) return rc;
  for (int m = 0; m < 11 && !rc; m++)
    if (!c->adj[m]) rc = dalloc(&c->adj[m], c->tot);
  if (rc) return rc;
  double* ubn[3] = {c->adj[0], c->adj[1], c->adj[2]};
  double* ubs[3] = {c->adj[3], c->adj[4], c->adj[5]};
  double* un[3] = {c->adj[6], c->adj[7], c->adj[8]};
  double* pb = c->adj[9];
  double* pbo = c->adj[10];
  const Geo& g = c->g;
  hipStream_t s = c->st;
  // the state of the step and the cotangents of its output, in the box layout
  for (int a = 0; a < 3; a++) {
    rc = copy_field(c, SMAC_FIELD_U + a, c->u[a], u_n[a], true, hipMemcpyDeviceToDevice, nullptr, nullptr);
    if (!rc) rc = copy_field(c, SMAC_FIELD_U + a, ubn[a], ubar[a], true, hipMemcpyDeviceToDevice, nullptr, nullptr);
    if (rc) return rc;
  }
  rc = copy_field(c, SMAC_FIELD_P, pb, pbar, true, hipMemcpyDeviceToDevice, nullptr, nullptr);
  if (rc) return rc;
  const bool var = var_props(c);
  const BetaF bf = beta_of(c);
  const BetaF* bfp = var ? &bf : nullptr;
  // cotangent of dp: pb' + dt D_i(beta ub'), the right-hand side of the adjoint solve
  KT(c, "adj_div", launch_adj_div(g, bfp, ubn, pb, c->rhs, dt, s));
  launch_rg_zero(c->dp, c->tot, s);
  rc = pressure_solve(c, dt, opts, st, true);
  if (rc) return rc;
  // ub* = ub' + (1/dt) G bb on interior faces; wall faces pass ub' to u^n
  KT(c, "adj_ustar", launch_adj_ustar(g, ubn, c->dp, ubs, un, 1. / dt, s));
  // pb = pb' + dt D_i(beta ub*) (the old-pressure gradient of the predictor)
  KT(c, "adj_div", launch_adj_div(g, bfp, ubs, pb, pbo, dt, s));
  // ub += (dF/du)^T ub*
  Vel V;
  for (int a = 0; a < 3; a++) V.c[a] = c->u[a];
  const Props P = props_of(c);
  KT(c, "predict_vjp", launch_predict_vjp(g, var ? &P : nullptr, V, ubs, un, dt, s));
  launch_rg_zero(c->dp, c->tot, s);
  HIPOK(hipGetLastError());
  for (int a = 0; a < 3; a++) {
    rc = copy_field(c, SMAC_FIELD_U + a, nullptr, nullptr, false, hipMemcpyDeviceToDevice, ubar[a], un[a]);
    if (rc) return rc;
  }
  return copy_field(c, SMAC_FIELD_P, nullptr, nullptr, false, hipMemcpyDeviceToDevice, pbar, pbo);
}

// Adjoint of one step of the VoF loop (smac_transport_alpha(dt, S, step) then
// smac_step(dt, opts)) from the state (alpha^n, u^n, p^n): the step is taken
// again from that state with the transport's sweeps recorded, then the
// cotangents run back through the correction, the solve (a second solve), the
// predictor, the properties and the transport (smac_adjoint_vof.hip).
extern "C" int smac_adjoint_vof_step(smac_ctx* c, double dt, int32_t S, int32_t step, const smac_solver_opts* opts,
                                     const double* alpha_n, const double* const u_n[3], const double* p_n,
                                     double* alpha_bar, double* const ubar[3], double* pbar, smac_solve_stats* st) {
  if (!c || !(dt > 0) || S < 1 || S > 64 || step < 0 || !alpha_n || !u_n || !p_n || !alpha_bar || !ubar || !pbar)
    return SMAC_ERR_INVALID_ARG;
  c->dp_zero = false;  // the adjoint's solves and copies leave dp as they please
  for (int a = 0; a < 3; a++)
    if (!u_n[a] || !ubar[a]) return SMAC_ERR_INVALID_ARG;
  if (c->eps || c->bpois_on || c->has_outlet) return SMAC_ERR_UNSUPPORTED;
  int rc = vof_alloc(c);
  for (int m = 0; m < 11 && !rc; m++)
    if (!c->adj[m]) rc = dalloc(&c->adj[m], c->tot);
  if (rc) return rc;
  const int nrec = 3 * S;
  if (c->vadj_S != S) {
    for (auto a : c->vadj) dfree(a);
    c->vadj.assign((size_t)(2 * nrec + 9), nullptr);
    c->vadj_S = 0;
    for (auto& a : c->vadj)
      if (!rc) rc = dalloc(&a, c->tot);
    if (rc) return rc;
    c->vadj_S = S;
  }
  double** ral = c->vadj.data();          // alpha before sweep r (r = 0..3S)
  double** rF = c->vadj.data() + nrec + 1;  // flux of axis a in substep t: rF[3 t + a]
  double* X = c->vadj[2 * nrec + 1];
  double* Y = c->vadj[2 * nrec + 2];
  double* Z = c->vadj[2 * nrec + 3];
  double* anx = c->vadj[2 * nrec + 4];
  double* Fb[3] = {c->vadj[2 * nrec + 5], c->vadj[2 * nrec + 6], c->vadj[2 * nrec + 7]};
  double* dpf = c->vadj[2 * nrec + 8];
  double* ubn[3] = {c->adj[0], c->adj[1], c->adj[2]};
  double* ubs[3] = {c->adj[3], c->adj[4], c->adj[5]};
  double* un[3] = {c->adj[6], c->adj[7], c->adj[8]};
  double* pb = c->adj[9];
  double* pbo = c->adj[10];
  const Geo& g = c->g;
  hipStream_t s = c->st;
  const size_t bytes = c->tot * sizeof(double);
  // ---- the step again, from (alpha^n, u^n, p^n)
  rc = copy_field(c, SMAC_FIELD_ALPHA, ral[0], alpha_n, true, hipMemcpyDeviceToDevice, nullptr, nullptr);
  for (int a = 0; a < 3 && !rc; a++)
    rc = copy_field(c, SMAC_FIELD_U + a, c->u[a], u_n[a], true, hipMemcpyDeviceToDevice, nullptr, nullptr);
  if (!rc) rc = copy_field(c, SMAC_FIELD_P, c->p, p_n, true, hipMemcpyDeviceToDevice, nullptr, nullptr);
  if (rc) return rc;
  HIPOK(hipMemcpyAsync(c->alpha, ral[0], bytes, hipMemcpyDeviceToDevice, s));
  c->alpha_nonzero = true;  // the property arrays of a two-phase grid
  c->props_ver++;           // new face coefficients: the multigrid rebuilds its coarse ones
  rc = smac_update_properties(c, 1);  // rho_old of the step = rho(alpha^n)
  if (rc) return rc;
  for (int a = 0; a < 3; a++) HIPOK(hipMemsetAsync(c->vfacc[a], 0, bytes, s));
  Vel V;
  for (int a = 0; a < 3; a++) V.c[a] = c->u[a];
  const double sdt = dt / (double)S;
  for (int t = 0; t < S; t++)
    KT(c, "vof_rec", launch_vof_substep_rec(g, c->d, V, ral + 3 * t, rF + 3 * t, c->vfacc, c->vcoeff, sdt, step, s));
  HIPOK(hipMemcpyAsync(c->alpha, ral[nrec], bytes, hipMemcpyDeviceToDevice, s));
  c->vof_armed = true;
  c->props_ver++;
  rc = smac_update_properties(c, 0);
  if (!rc) rc = smac_predictor(c, dt);
  if (!rc) rc = pressure_solve(c, dt, opts, st, false);
  if (rc) return rc;
  HIPOK(hipMemcpyAsync(dpf, c->dp, bytes, hipMemcpyDeviceToDevice, s));
  // ---- cotangents of the step's output, box layout
  for (int a = 0; a < 3 && !rc; a++)
    rc = copy_field(c, SMAC_FIELD_U + a, ubn[a], ubar[a], true, hipMemcpyDeviceToDevice, nullptr, nullptr);
  if (!rc) rc = copy_field(c, SMAC_FIELD_P, pb, pbar, true, hipMemcpyDeviceToDevice, nullptr, nullptr);
  if (!rc) rc = copy_field(c, SMAC_FIELD_ALPHA, X, alpha_bar, true, hipMemcpyDeviceToDevice, nullptr, nullptr);
  if (rc) return rc;
  launch_rg_zero(anx, c->tot, s);
  for (int a = 0; a < 3; a++) launch_rg_zero(Fb[a], c->tot, s);
  const BetaF bf = beta_of(c);
  const double drho = c->fl.rho1 - c->fl.rho0, dmu = c->fl.mu1 - c->fl.mu0;
  // correction and solve: dp cotangent pb' + dt D_i(beta ub'), second solve
  KT(c, "adj_div", launch_adj_div(g, &bf, ubn, pb, c->rhs, dt, s));
  launch_rg_zero(c->dp, c->tot, s);
  rc = pressure_solve(c, dt, opts, st, true);
  if (rc) return rc;
  // beta cotangents of the correction and the solve -> alpha^{n+1}
  KT(c, "adj_beta", launch_adj_beta_cells(g, bf, ubn, c->dp, dpf, dt, drho, X, s));
  KT(c, "adj_ustar", launch_adj_ustar(g, ubn, c->dp, ubs, un, 1. / dt, s));
  KT(c, "adj_div", launch_adj_div(g, &bf, ubs, pb, pbo, dt, s));
  // predictor: velocity, accumulated alpha flux and property cotangents
  {
    const Props P = props_of(c);
    VofF M{};
    for (int a = 0; a < 3; a++) {
      M.F[a] = c->vfacc[a];
      M.hbydt[a] = c->d[a] * (1. / dt);
    }
    M.drho = drho;
    KT(c, "predict_vjp", launch_predict_vjp_vof(g, P, V, M, c->p, ubs, un, Fb, X, anx, drho, dmu, dt, s));
  }
  // transport, sweeps in reverse
  const int first = step % 3;
  for (int t = S - 1; t >= 0; t--) {
    launch_vof_coeff(g, ral[3 * t], c->vcoeff, s);
    for (int k = 2; k >= 0; k--) {
      const int a = (first + k) % 3;
      const int r = 3 * t + k;
      launch_rg_zero(Z, c->tot, s);  // the ghost entries of the result stay zero
      KT(c, "vof_adj", launch_vof_adj_sweep(g, c->d, V, a, ral[r], rF[3 * t + a], c->vcoeff, X, Y, Z, Fb[a], un[a],
                                            sdt, k == 2, s));
      std::swap(X, Z);
    }
  }
  launch_add_box(X, anx, c->tot, s);
  launch_rg_zero(c->dp, c->tot, s);
  HIPOK(hipGetLastError());
  for (int a = 0; a < 3 && !rc; a++)
    rc = copy_field(c, SMAC_FIELD_U + a, nullptr, nullptr, false, hipMemcpyDeviceToDevice, ubar[a], un[a]);
  if (!rc) rc = copy_field(c, SMAC_FIELD_P, nullptr, nullptr, false, hipMemcpyDeviceToDevice, pbar, pbo);
  if (!rc) rc = copy_field(c, SMAC_FIELD_ALPHA, nullptr, nullptr, false, hipMemcpyDeviceToDevice, alpha_bar, X);
  return rc;
}

extern "C" int smac_calc_cfl(smac_ctx* c, double dt, double* cc, double* cv) {
  if (!c) return SMAC_ERR_INVALID_ARG;
  int rc = ensure_props(c);
  if (rc) return rc;
  Vel V;
  for (int a = 0; a < 3; a++) V.c[a] = c->u[a];
  double visc = (1. / c->fl.rho0) * c->fl.mu0 * dt * (c->g.inv_d2[0] + c->g.inv_d2[1] + c->g.inv_d2[2]);
  const bool var = var_props(c);
  launch_cfl(c->g, V, dt, visc, var ? c->inv_rho : nullptr, var ? c->mu : nullptr, c->part[0], c->part[1], c->st);
  FinalizeArgs fa = fin(c);
  fin_add(fa, c->part[0], 1, S_CFLC);
  fin_add(fa, c->part[1], 1, S_CFLV);
  launch_finalize(fa, c->S, c->st);
  HIPOK(hipGetLastError());
  rc = read_scalars(c);
  if (rc) return rc;
  if (cc) *cc = c->Sh[S_CFLC];
  if (cv) *cv = c->Sh[S_CFLV];
  return SMAC_OK;
}

// smac_calc_cfl with its row partials, lane sums and result slots in the
// device's static scratch array instead of the context's buffers; no
// allocation (the result comes back through the context's pinned host slots,
// allocated by smac_create).
extern "C" int smac_calc_cfl_noalloc(smac_ctx* c, double dt, double* cc, double* cv) {
  if (!c) return SMAC_ERR_INVALID_ARG;
  if (c->nrows > SMAC_NOALLOC_MAX_ROWS) return SMAC_ERR_INVALID_GRID;
  const bool var = var_props(c);
  if (!c->props_ready) {
    if (var) return SMAC_ERR_UNSUPPORTED;  // building them would allocate
    int rc = smac_update_properties(c, 1);  // one-phase: sets constants only
    if (rc) return rc;
  }
  static_assert(S_CFLV == S_CFLC + 1, "the two CFL slots are copied as one pair");
  double* scr = scratch_ptr(c->dev);
  if (!scr) return SMAC_ERR_HIP;
  Vel V;
  for (int a = 0; a < 3; a++) V.c[a] = c->u[a];
  double visc = (1. / c->fl.rho0) * c->fl.mu0 * dt * (c->g.inv_d2[0] + c->g.inv_d2[1] + c->g.inv_d2[2]);
  launch_cfl(c->g, V, dt, visc, var ? c->inv_rho : nullptr, var ? c->mu : nullptr, scr, scr + SCR_PART1, c->st);
  FinalizeArgs fa;
  memset(&fa, 0, sizeof(fa));
  fa.nrows = c->nrows;
  fa.stage = scr + SCR_STAGE;
  fin_add(fa, scr, 1, 0);
  fin_add(fa, scr + SCR_PART1, 1, 1);
  launch_finalize(fa, scr + SCR_S, c->st);
  HIPOK(hipGetLastError());
  HIPOK(hipMemcpyAsync(c->Sh + S_CFLC, scr + SCR_S, 2 * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HIPOK(hipStreamSynchronize(c->st));
  if (cc) *cc = c->Sh[S_CFLC];
  if (cv) *cv = c->Sh[S_CFLV];
  return SMAC_OK;
}

extern "C" int smac_set_profiling(smac_ctx* c, int32_t on) {
  if (!c) return SMAC_ERR_INVALID_ARG;
  c->prof = on != 0;
  return SMAC_OK;
}
extern "C" int smac_kernel_time(smac_ctx* c, const char* name, int64_t* n, double* ms) {
  if (!c || !name) return SMAC_ERR_INVALID_ARG;
  prof_flush(c);
  auto it = c->name_id.find(name);
  if (it == c->name_id.end()) {
    if (n) *n = 0;
    if (ms) *ms = 0.0;
    return SMAC_ERR_INVALID_ARG;
  }
  if (n) *n = c->totals[it->second].n;
  if (ms) *ms = c->totals[it->second].ms;
  return SMAC_OK;
}
extern "C" int smac_kernel_names(smac_ctx* c, char* buf, size_t len) {
  if (!c || !buf || len == 0) return SMAC_ERR_INVALID_ARG;
  std::string all;
  for (size_t i = 0; i < c->names.size(); i++) {
    if (i) all += ",";
    all += c->names[i];
  }
  if (all.size() + 1 > len) return SMAC_ERR_INVALID_ARG;
  memcpy(buf, all.c_str(), all.size() + 1);
  return SMAC_OK;
}
extern "C" int smac_reset_kernel_times(smac_ctx* c) {
  if (!c) return SMAC_ERR_INVALID_ARG;
  prof_flush(c);
  for (auto& t : c->totals) t = KStat();
  return SMAC_OK;
}
extern "C" int smac_sync(smac_ctx* c) {
  if (!c) return SMAC_ERR_INVALID_ARG;
  HIPOK(hipStreamSynchronize(c->st));
  return SMAC_OK;
}
extern "C" void* smac_stream(smac_ctx* c) { return c ? (void*)c->st : nullptr; }
extern "C" const char* smac_backend_name(void) { return "hip-gfx950"; }
extern "C" const char* smac_strerror(int s) {
  switch (s) {
    case SMAC_OK: return "ok";
    case SMAC_ERR_INVALID_ARG: return "invalid argument";
    case SMAC_ERR_INVALID_GRID: return "grid sizes not supported by the multigrid hierarchy";
    case SMAC_ERR_HIP: return "HIP runtime error";
    case SMAC_ERR_NO_DEVICE: return "no device";
    case SMAC_ERR_OOM: return "out of device memory";
    case SMAC_ERR_NAN: return "NaN in solve";
    case SMAC_ERR_UNSUPPORTED: return "unsupported option";
    default: return "unknown error";
  }
}
