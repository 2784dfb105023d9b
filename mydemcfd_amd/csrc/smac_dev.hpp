// smac_dev.hpp -- device-side geometry, indexing and boundary getters of the
// MI355X SMAC step. Every array (cells and the three face families) uses one
// padded box (nx+3)(ny+3)(nz+3), x fastest; reference index (i,j,k) keeps its
// meaning (cell i in 1..nx interior, x-face i between cells i-1 and i).
//
// Arithmetic contract: the library is compiled with -ffp-contract=off and every
// expression below is evaluated in the order written, so FP64 results are
// bit-identical to the CPU restatement of the reference kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace smac {

enum { F_INTERIOR = 0, F_BOUNDARY = 1, F_GHOST = 2 };  // CFD/Enums.h:3-8
enum { BC_NOSLIP = 0, BC_SLIP = 1, BC_INFLOW = 2, BC_OUTFLOW = 3, BC_OUTLET = 4 };  // CFD/Enums.h:10-16

// Geometry + boundary table + uniform fluid properties, passed by value.
struct Geo {
  int n[3];        // interior cells
  int px, py;      // pitches (nx+3, ny+3)
  int per[3];      // periodic flags
  double inv_d[3], inv_d2[3];
  // bc slots: 0..5 = the six sides (xmin,xmax,ymin,ymax,zmin,zmax), 6 = bc id 0
  int btype[7];
  double bv[7][3];
  // uniform properties (single phase): rho0, mu0, beta0 = 1/(0.5*(rho0+rho0))
  double rho0, mu0, beta0, g[3];
};

struct I3 {
  int v[3];
};

__host__ __device__ __forceinline__ I3 mk(int i, int j, int k) {
  I3 r;
  r.v[0] = i;
  r.v[1] = j;
  r.v[2] = k;
  return r;
}
__host__ __device__ __forceinline__ I3 sh(I3 a, int d, int s) {
  a.v[d] += s;
  return a;
}
__host__ __device__ __forceinline__ int wrapi(int i, int n) { return i < 1 ? i + n : (i > n ? i - n : i); }
__host__ __device__ __forceinline__ int W(const Geo& g, int d, int i) { return g.per[d] ? wrapi(i, g.n[d]) : i; }
__host__ __device__ __forceinline__ I3 wr(const Geo& g, I3 a) {
  a.v[0] = W(g, 0, a.v[0]);
  a.v[1] = W(g, 1, a.v[1]);
  a.v[2] = W(g, 2, a.v[2]);
  return a;
}
__host__ __device__ __forceinline__ size_t idx(const Geo& g, int i, int j, int k) {
  return ((size_t)k * (size_t)g.py + (size_t)j) * (size_t)g.px + (size_t)i;
}
__host__ __device__ __forceinline__ size_t IX(const Geo& g, I3 a) { return idx(g, a.v[0], a.v[1], a.v[2]); }

// Row-split position of natural x-index i (0..nx+1) inside its row: even i
// first (positions 1..nx/2), then odd i. Ghost i = 0 maps to slot 0.
__host__ __device__ __forceinline__ int spos(int i, int nx) {
  return (i & 1) ? (nx >> 1) + ((i + 1) >> 1) : (i >> 1);
}

// Analytic face type of a box without internal walls
// (SMACSolver::set_face_type, CFD/SMACSolver.cpp:104-202); q already wrapped.
__device__ __forceinline__ int face_type(const Geo& g, int a, I3 q) {
#pragma unroll
  for (int d = 0; d < 3; d++) {
    if (d == a) continue;
    if (q.v[d] < 1 || q.v[d] > g.n[d]) return F_GHOST;
  }
  int i = q.v[a], n = g.n[a];
  if (g.per[a]) return (i >= 1 && i <= n) ? F_INTERIOR : F_GHOST;
  if (i == 1 || i == n + 1) return F_BOUNDARY;
  if (i >= 2 && i <= n) return F_INTERIOR;
  return F_GHOST;
}
// bc slot of an a-face (6 = bc id 0 off the walls)
__device__ __forceinline__ int face_slot(const Geo& g, int a, I3 q) {
  if (g.per[a]) return 6;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    if (d == a) continue;
    if (q.v[d] < 1 || q.v[d] > g.n[d]) return 6;
  }
  if (q.v[a] == 1) return 2 * a;
  if (q.v[a] == g.n[a] + 1) return 2 * a + 1;
  return 6;
}

// bc table lookups by slot with a select chain (a dynamic index into the
// by-value kernel argument would force a private scratch copy)
__device__ __forceinline__ int btype_of(const Geo& g, int s) {
  int t = g.btype[6];
  t = s == 0 ? g.btype[0] : t;
  t = s == 1 ? g.btype[1] : t;
  t = s == 2 ? g.btype[2] : t;
  t = s == 3 ? g.btype[3] : t;
  t = s == 4 ? g.btype[4] : t;
  t = s == 5 ? g.btype[5] : t;
  return t;
}
__device__ __forceinline__ double bv_of(const Geo& g, int s, int a) {
  double v = g.bv[6][a];
  v = s == 0 ? g.bv[0][a] : v;
  v = s == 1 ? g.bv[1][a] : v;
  v = s == 2 ? g.bv[2][a] : v;
  v = s == 3 ? g.bv[3][a] : v;
  v = s == 4 ? g.bv[4][a] : v;
  v = s == 5 ? g.bv[5][a] : v;
  return v;
}

// Velocity arrays of the three components.
struct Vel {
  const double* c[3];
};

// d_get_v{x,y,z}_{x,y,z}face (CFD/G_BoundaryFunctions.h:9-97)
__device__ __forceinline__ double get_face(const Geo& g, const Vel& V, int a, I3 q) {
  q = wr(g, q);
  int t = face_type(g, a, q);
  if (t == F_INTERIOR) return V.c[a][IX(g, q)];
  if (t == F_BOUNDARY) {
    int s = face_slot(g, a, q);
    int bt = btype_of(g, s);
    if (bt == BC_INFLOW) return (1. - 0.0) * bv_of(g, s, a);
    if (bt == BC_NOSLIP) return 0.;
    if (bt == BC_OUTLET) return V.c[a][IX(g, q)];
    if (bt == BC_SLIP && a != 0) return V.c[a][IX(g, q)];
  }
  return 0.;
}

// d_get_v{a}_{b}dir (CFD/G_BoundaryFunctions.h:100-367)
__device__ __forceinline__ double get_dir(const Geo& g, const Vel& V, int a, int b, I3 q, int s) {
  q = wr(g, q);
  double vin = V.c[a][IX(g, q)];
  I3 q2 = wr(g, sh(q, b, s));
  if (face_type(g, a, q2) == F_INTERIOR) return V.c[a][IX(g, q2)];
  I3 qf = q;
  qf.v[b] = s > 0 ? q.v[b] + 1 : q.v[b];
  double vw = 0.0;
  int count = 0;
  if (face_type(g, b, qf) == F_BOUNDARY) {
    vw += bv_of(g, face_slot(g, b, qf), a);
    count++;
  }
  I3 qf2 = wr(g, sh(qf, a, -1));
  if (face_type(g, b, qf2) == F_BOUNDARY) {
    vw += bv_of(g, face_slot(g, b, qf2), a);
    count++;
  }
  if (count > 0) vw /= (double)count;
  int bt = btype_of(g, face_slot(g, b, qf));
  if (bt == BC_SLIP || bt == BC_OUTLET) return vin;
  return 2.0 * vw - vin;
}

// d_minmod (CFD/G_SMACSolver.cu:67-88)
__device__ __forceinline__ double minmod(double dp, double dm) {
  double prod = dp * dm;
  double s = 0.;
  if (prod > 0.) s = fabs(dp) < fabs(dm) ? dp : dm;
  return s;
}
// sgn (CFD/G_SMACSolver.h:130-139)
__device__ __forceinline__ int sgn(double a) { return a < 0 ? -1 : 1; }

// Uniform-property Poisson coefficients: a = beta0/h^2 on interior faces, the
// diagonal sums all six (wall faces carry beta = 0), off-diagonals vanish at
// walls (k_build_vof_poisson_A, CFD/pressure_solver/G_PCGSolver.cu:126-212).
struct Coef {
  double a[6];  // xp, xm, yp, ym, zp, zm
  double diag;
};
__device__ __forceinline__ bool wall_face(const int per, int i, int n) { return !per && (i == 1 || i == n + 1); }

__device__ __forceinline__ Coef coef_uniform(const int n[3], const int per[3], const double inv_d2[3],
                                             double beta0, int i, int j, int k) {
  Coef c;
  int q[3] = {i, j, k};
#pragma unroll
  for (int d = 0; d < 3; d++) {
    double bp = wall_face(per[d], q[d] + 1, n[d]) ? 0.0 : beta0;
    double bm = wall_face(per[d], q[d], n[d]) ? 0.0 : beta0;
    c.a[2 * d] = bp * inv_d2[d];
    c.a[2 * d + 1] = bm * inv_d2[d];
  }
  c.diag = c.a[0] + c.a[1] + c.a[2] + c.a[3] + c.a[4] + c.a[5];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    if (wall_face(per[d], q[d] + 1, n[d])) c.a[2 * d] = 0.;
    if (wall_face(per[d], q[d], n[d])) c.a[2 * d + 1] = 0.;
  }
  return c;
}

// Variable-coefficient Poisson operator of a two-phase grid: the same rule on
// the face coefficients beta = 1/rho_f (natural layout; 0 on wall and ghost
// faces) of k_build_vof_poisson_A.
__device__ __forceinline__ Coef coef_var(const Geo& g, const double* const b[3], int i, int j, int k) {
  Coef c;
  int q[3] = {i, j, k};
#pragma unroll
  for (int d = 0; d < 3; d++) {
    int qp[3] = {i, j, k};
    qp[d] = W(g, d, q[d] + 1);
    c.a[2 * d] = b[d][idx(g, qp[0], qp[1], qp[2])] * g.inv_d2[d];
    c.a[2 * d + 1] = b[d][idx(g, i, j, k)] * g.inv_d2[d];
  }
  c.diag = c.a[0] + c.a[1] + c.a[2] + c.a[3] + c.a[4] + c.a[5];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    if (wall_face(g.per[d], q[d] + 1, g.n[d])) c.a[2 * d] = 0.;
    if (wall_face(g.per[d], q[d], g.n[d])) c.a[2 * d + 1] = 0.;
  }
  return c;
}

struct BetaF {
  const double* b[3];
};

// Cell and face properties of a two-phase grid (natural layout):
// mu, rho_old, 1/rho on cells (ghosts filled), beta = 1/rho_f and mu_f on faces
// (k_update_*_properties_by_alpha, CFD/G_SMACSolver_VoF.cu:1244-1535).
struct Props {
  const double* mu;
  const double* rho_old;
  const double* inv_rho;
  const double* beta[3];
  const double* muf[3];
};

// Property policies of the predictor and the correction. UniP: one phase,
// constants from Geo (mu_f = 0.5(mu0+mu0), rho_f = 0.5(rho0+rho0), as the
// face formulas give them); VarP: the arrays of Props.
struct UniP {
  static __device__ __forceinline__ double mu(const Geo& g, const Props&, I3) { return g.mu0; }
  static __device__ __forceinline__ double mu_edge(const Geo& g, const Props&, int, I3, I3) {
    return 0.5 * (g.mu0 + g.mu0);
  }
  static __device__ __forceinline__ double beta(const Geo& g, const Props&, int, I3) { return g.beta0; }
  static __device__ __forceinline__ double rho_old_f(const Geo& g, const Props&, I3, I3) {
    return 0.5 * (g.rho0 + g.rho0);
  }
};
struct VarP {
  static __device__ __forceinline__ double mu(const Geo& g, const Props& P, I3 c) { return P.mu[IX(g, c)]; }
  static __device__ __forceinline__ double mu_edge(const Geo& g, const Props& P, int b, I3 f1, I3 f2) {
    return 0.5 * (P.muf[b][IX(g, f1)] + P.muf[b][IX(g, f2)]);
  }
  static __device__ __forceinline__ double beta(const Geo& g, const Props& P, int a, I3 q) {
    return P.beta[a][IX(g, q)];
  }
  static __device__ __forceinline__ double rho_old_f(const Geo& g, const Props& P, I3 q, I3 cm) {
    return 0.5 * (P.rho_old[IX(g, q)] + P.rho_old[IX(g, cm)]);
  }
};

// Storage traits of the multigrid vectors: storage type T, compute type C.
// bf16 is kept as raw 16-bit words (upper half of an IEEE float), rounded to
// nearest-even on store.
struct bf16s {
  uint16_t u;
};
template <typename T>
struct St;
template <>
struct St<double> {
  using C = double;
  static __device__ __forceinline__ double ld(const double* p, size_t i) { return p[i]; }
  static __device__ __forceinline__ void st(double* p, size_t i, double v) { p[i] = v; }
};
template <>
struct St<float> {
  using C = float;
  static __device__ __forceinline__ float ld(const float* p, size_t i) { return p[i]; }
  static __device__ __forceinline__ void st(float* p, size_t i, float v) { p[i] = v; }
};
// float -> bf16 bits, round to nearest even (inf stays inf, NaN stays NaN):
// the language conversion, which gfx950 does in one instruction
// (v_cvt_pk_bf16_f32) instead of an integer rounding sequence
__device__ __forceinline__ unsigned bf16_bits(float v) {
  return (unsigned)__builtin_bit_cast(unsigned short, (__bf16)v);
}
template <>
struct St<bf16s> {
  using C = float;
  static __device__ __forceinline__ float ld(const bf16s* p, size_t i) {
    return __uint_as_float(((unsigned)p[i].u) << 16);
  }
  static __device__ __forceinline__ void st(bf16s* p, size_t i, float v) { p[i].u = (uint16_t)bf16_bits(v); }
};

// the value v has once stored as Z (from its compute type) and read back
template <typename Z>
__device__ __forceinline__ double stored_value(double v);
template <>
__device__ __forceinline__ double stored_value<double>(double v) { return v; }
template <>
__device__ __forceinline__ double stored_value<float>(double v) { return (double)(float)v; }
template <>
__device__ __forceinline__ double stored_value<bf16s>(double v) {
  return (double)__uint_as_float(bf16_bits((float)v) << 16);
}

// N consecutive storage elements in one aligned access (N * sizeof(T) <= 16,
// the address a multiple of N * sizeof(T)), converted to / from the compute type
template <typename T, int N>
struct Vec;
template <int N>
struct Vec<double, N> {
  static __device__ __forceinline__ void ld(const double* p, double* o) {
    if constexpr (N == 2) {
      const double2 v = *reinterpret_cast<const double2*>(p);
      o[0] = v.x;
      o[1] = v.y;
    } else {
      o[0] = p[0];
    }
  }
  static __device__ __forceinline__ void st(double* p, const double* v) {
    if constexpr (N == 2) {
      *reinterpret_cast<double2*>(p) = make_double2(v[0], v[1]);
    } else {
      p[0] = v[0];
    }
  }
};
template <int N>
struct Vec<float, N> {
  static __device__ __forceinline__ void ld(const float* p, float* o) {
    if constexpr (N == 4) {
      const float4 v = *reinterpret_cast<const float4*>(p);
      o[0] = v.x;
      o[1] = v.y;
      o[2] = v.z;
      o[3] = v.w;
    } else if constexpr (N == 2) {
      const float2 v = *reinterpret_cast<const float2*>(p);
      o[0] = v.x;
      o[1] = v.y;
    } else {
      o[0] = p[0];
    }
  }
  static __device__ __forceinline__ void st(float* p, const float* v) {
    if constexpr (N == 4) {
      *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
    } else if constexpr (N == 2) {
      *reinterpret_cast<float2*>(p) = make_float2(v[0], v[1]);
    } else {
      p[0] = v[0];
    }
  }
};
template <int N>
struct Vec<bf16s, N> {
  static __device__ __forceinline__ void ld(const bf16s* p, float* o) {
    if constexpr (N == 8) {
      const uint4 v = *reinterpret_cast<const uint4*>(p);
      const unsigned w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int q = 0; q < 4; q++) {
        o[2 * q] = __uint_as_float(w[q] << 16);
        o[2 * q + 1] = __uint_as_float(w[q] & 0xffff0000u);
      }
    } else if constexpr (N == 4) {
      const uint2 v = *reinterpret_cast<const uint2*>(p);
      o[0] = __uint_as_float(v.x << 16);
      o[1] = __uint_as_float(v.x & 0xffff0000u);
      o[2] = __uint_as_float(v.y << 16);
      o[3] = __uint_as_float(v.y & 0xffff0000u);
    } else if constexpr (N == 2) {
      const unsigned v = *reinterpret_cast<const unsigned*>(p);
      o[0] = __uint_as_float(v << 16);
      o[1] = __uint_as_float(v & 0xffff0000u);
    } else {
      o[0] = St<bf16s>::ld(p, 0);
    }
  }
  static __device__ __forceinline__ void st(bf16s* p, const float* v) {
    if constexpr (N == 8) {
      uint4 w;
      w.x = bf16_bits(v[0]) | (bf16_bits(v[1]) << 16);
      w.y = bf16_bits(v[2]) | (bf16_bits(v[3]) << 16);
      w.z = bf16_bits(v[4]) | (bf16_bits(v[5]) << 16);
      w.w = bf16_bits(v[6]) | (bf16_bits(v[7]) << 16);
      *reinterpret_cast<uint4*>(p) = w;
    } else if constexpr (N == 4) {
      uint2 w;
      w.x = bf16_bits(v[0]) | (bf16_bits(v[1]) << 16);
      w.y = bf16_bits(v[2]) | (bf16_bits(v[3]) << 16);
      *reinterpret_cast<uint2*>(p) = w;
    } else if constexpr (N == 2) {
      *reinterpret_cast<unsigned*>(p) = bf16_bits(v[0]) | (bf16_bits(v[1]) << 16);
    } else {
      St<bf16s>::st(p, 0, v[0]);
    }
  }
};

// Canonical wave64 halving tree: lane l += lane l+s, s = 32..1; lane 0 holds
// the result. Identical association to the CPU oracle's tree().
__device__ __forceinline__ double wave_tree_sum(double v) {
#pragma unroll
  for (int s = 32; s >= 1; s >>= 1) v = v + __shfl_down(v, s, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int s = 32; s >= 1; s >>= 1) {
    double o = __shfl_down(v, s, 64);
    if (o > v || o != o) v = o;
  }
  return v;
}

}  // namespace smac
