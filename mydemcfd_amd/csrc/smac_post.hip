// smac_post.hip -- the driver loop's work after the correction
// (CFD/main.cu:316-317): ghost pressure, tangential velocity ghosts
// (G_SMACSolver::update_boundary_ghost, CFD/G_SMACSolver.cu:1345-1362) and the
// face pressure gradient (G_SMACSolver::make_face_gradp, 1277-1343).
//
// The ghosts are a shell one cell wide around the box, so the kernels run over
// planes of the shell only, not over the whole box as the reference's nine
// full-grid launches do. Every value is computed from interior entries that no
// thread of these launches writes, so the launches carry no ordering between
// them and the edge rule (the mean of the two adjacent wall ghosts) evaluates
// those two ghosts itself instead of reading them back.
#include "smac_internal.hpp"

namespace smac {

// d_get_velocity_ghost_value (CFD/G_BoundaryFunctions.cu:257-275) with the IBM
// solid fraction lambda = 0: bval is the face one row inside the wall, inval
// the face two rows inside
__device__ __forceinline__ double ghost_value(int bt, double bv, double bval, double inval) {
  if (bt == BC_INFLOW) return 2.0 * (1.0 - 0.0) * bv - inval;
  if (bt == BC_NOSLIP) return -inval;
  if (bt == BC_OUTLET) return 2.0 * bval - inval;
  if (bt == BC_SLIP) return inval;
  return bval;
}

// ghost of component a across the wall of tangential axis t at q (q.v[t] is 0
// or n+1), bc id of the a-face one row inside: the bodies of
// k_update_v{x,y,z}_ghost (CFD/G_BoundaryFunctions.cu:277-336,365-424,454-513)
__device__ __forceinline__ double wall_ghost(const Geo& g, const double* __restrict__ v, int a, int t, I3 q) {
  const bool lower = q.v[t] == 0;
  I3 qb = q, qi = q;
  qb.v[t] = lower ? 1 : g.n[t];
  qi.v[t] = lower ? 2 : g.n[t] - 1;
  const int s = face_slot(g, a, qb);
  return ghost_value(btype_of(g, s), bv_of(g, s, a), v[IX(g, qb)], v[IX(g, qi)]);
}

// component a at position q of its face array after the update: periodic
// directions resolve to their image, one wall gives wall_ghost, two walls the
// mean of the two adjacent wall ghosts, the lower-numbered axis first
// (k_update_v{x,y,z}_ghost_corner, CFD/G_BoundaryFunctions.cu:338-363,426-452,515-540)
__device__ __forceinline__ double ghost_vel(const Geo& g, const double* __restrict__ v, int a, I3 q) {
#pragma unroll
  for (int d = 0; d < 3; d++) {
    if (!g.per[d]) continue;
    int x = wrapi(q.v[d], g.n[d]);
    if (d == a && x > g.n[d]) x -= g.n[d];
    q.v[d] = x;
  }
  const int t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2;
  const bool g1 = q.v[t1] == 0 || q.v[t1] == g.n[t1] + 1;
  const bool g2 = q.v[t2] == 0 || q.v[t2] == g.n[t2] + 1;
  if (!g1 && !g2) return v[IX(g, q)];
  if (g1 && !g2) return wall_ghost(g, v, a, t1, q);
  if (!g1 && g2) return wall_ghost(g, v, a, t2, q);
  I3 q1 = q, q2 = q;
  q1.v[t2] = q.v[t2] == 0 ? 1 : g.n[t2];
  q2.v[t1] = q.v[t1] == 0 ? 1 : g.n[t1];
  return 0.5 * (wall_ghost(g, v, a, t1, q1) + wall_ghost(g, v, a, t2, q2));
}

// ghost cell q of the pressure: k_update_cell_ghost_pressure
// (CFD/G_SMACSolver.cu:1227-1267), p(q) = p(s) + rho(s)*(r.g), s the nearest
// interior cell (periodic directions: the image, r = 0)
__device__ __forceinline__ double ghost_p(const Geo& g, const double* __restrict__ p, const double* __restrict__ rho,
                                          const Spacing& H, I3 q) {
  I3 s = q;
  double r[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    int x = q.v[d];
    if (g.per[d]) x = wrapi(x, g.n[d]);
    else x = x < 1 ? 1 : (x > g.n[d] ? g.n[d] : x);
    s.v[d] = x;
    r[d] = g.per[d] ? 0.0 : (double)(q.v[d] - x) * H.h[d];
  }
  const size_t is = IX(g, s);
  const double rs = rho ? rho[is] : g.rho0;
  return p[is] + rs * (r[0] * g.g[0] + r[1] * g.g[1] + r[2] * g.g[2]);
}

// One launch covers the planes q[ax] = Q.v[0..nv) (blockIdx.z), with the
// other two axes u, w over [lo, lo+cnt). KIND 0: pressure ghosts; 1: velocity
// ghosts of component A.
template <int KIND, int A>
__global__ __launch_bounds__(256) void k_ghost_planes(Geo g, Plane Q, Spacing H, double* __restrict__ f,
                                                      const double* __restrict__ rho) {
  const int iu = (int)(blockIdx.x * 64 + threadIdx.x);
  const int iw = (int)(blockIdx.y * 4 + threadIdx.y);
  if (iu >= Q.ucnt || iw >= Q.wcnt) return;
  const int z = (int)blockIdx.z;
  const int pv = z == 0 ? Q.v[0] : (z == 1 ? Q.v[1] : Q.v[2]);
  I3 q;
#pragma unroll
  for (int d = 0; d < 3; d++) q.v[d] = d == Q.ax ? pv : (d == Q.u ? Q.ulo + iu : Q.wlo + iw);
  double val;
  if (KIND == 0) val = ghost_p(g, f, rho, H, q);
  else val = ghost_vel(g, f, A, q);
  f[IX(g, q)] = val;
}

// k_update_v{x,y,z}_outlet (CFD/G_BoundaryFunctions.cu:143-255) on the faces
// of one outlet side (side = 2*axis + high): u = u* + s*beta_f*dt*dp(c)/h with
// c the inner cell, s = +1 on the high side and -1 on the low side, beta_f the
// face's Dirichlet coefficient. (In the reference loop the correction has
// cleared dp before this runs, so the face keeps u*.) Faces with the other
// two indices in 1..N, as the reference's F_BOUNDARY test selects.
__global__ __launch_bounds__(256) void k_outlet_face(Geo g, int side, const double* __restrict__ us,
                                                     const double* __restrict__ beta, const double* __restrict__ dp,
                                                     double* __restrict__ u, double dt) {
  const int a = side >> 1;
  const int b1 = (a + 1) % 3, b2 = (a + 2) % 3;
  const int t1 = 1 + (int)(blockIdx.x * 64 + threadIdx.x);
  const int t2 = 1 + (int)(blockIdx.y * 4 + threadIdx.y);
  if (t1 > g.n[b1] || t2 > g.n[b2]) return;
  I3 q;
  q.v[a] = (side & 1) ? g.n[a] + 1 : 1;
  q.v[b1] = t1;
  q.v[b2] = t2;
  const int shift = (side & 1) ? -1 : 0;  // f_{x,y,z}internal_id_
  const double s = (side & 1) ? 1.0 : -1.0;
  const size_t id = IX(g, q);
  u[id] = us[id] + s * beta[id] * dt * dp[IX(g, sh(q, a, shift))] * g.inv_d[a];
}

void launch_outlet_faces(const Geo& g, const double* const us[3], const double* const beta[3], const double* dp,
                         double* const u[3], double dt, hipStream_t st) {
  for (int side = 0; side < 6; side++) {
    const int a = side >> 1;
    if (g.per[a] || g.btype[side] != BC_OUTLET) continue;
    const int b1 = (a + 1) % 3, b2 = (a + 2) % 3;
    dim3 grid((unsigned)((g.n[b1] + 63) / 64), (unsigned)((g.n[b2] + 3) / 4), 1);
    hipLaunchKernelGGL(k_outlet_face, grid, dim3(64, 4), 0, st, g, side, us[a], beta[a], dp, u[a], dt);
  }
}

// f_gradp_{x,y,z} = (p(c) - p(c-1))/h on faces with all indices in 1..N+1
__global__ __launch_bounds__(256) void k_face_gradp(Geo g, const double* __restrict__ p, double* __restrict__ gx,
                                                    double* __restrict__ gy, double* __restrict__ gz) {
  const int i = 1 + (int)(blockIdx.x * 64 + threadIdx.x);
  const int j = 1 + (int)(blockIdx.y * 4 + threadIdx.y);
  const int k = 1 + (int)blockIdx.z;
  if (i > g.n[0] + 1 || j > g.n[1] + 1) return;
  const size_t id = idx(g, i, j, k);
  const double pc = p[id];
  const double pxm = p[id - 1];
  const double pym = p[id - (size_t)g.px];
  const double pzm = p[id - (size_t)g.px * (size_t)g.py];
  gx[id] = (pc - pxm) * g.inv_d[0];
  gy[id] = (pc - pym) * g.inv_d[1];
  gz[id] = (pc - pzm) * g.inv_d[2];
}

// cell output of the VTI writer: p, alpha, face-mean velocity as float, x
// fastest over the interior (FileInOut::output_vti_binary_cellData,
// CFD/FileInOut.cpp:204-236)
__global__ __launch_bounds__(256) void k_pack_cells(Geo g, const double* __restrict__ p,
                                                    const double* __restrict__ al, const double* __restrict__ u,
                                                    const double* __restrict__ v, const double* __restrict__ w,
                                                    float* __restrict__ out) {
  const int i = 1 + (int)(blockIdx.x * 64 + threadIdx.x);
  const int j = 1 + (int)(blockIdx.y * 4 + threadIdx.y);
  const int k = 1 + (int)blockIdx.z;
  if (i > g.n[0] || j > g.n[1]) return;
  const size_t n = (size_t)g.n[0] * (size_t)g.n[1] * (size_t)g.n[2];
  const size_t m = ((size_t)(k - 1) * (size_t)g.n[1] + (size_t)(j - 1)) * (size_t)g.n[0] + (size_t)(i - 1);
  const size_t c = idx(g, i, j, k);
  out[m] = (float)p[c];
  out[n + m] = al ? (float)al[c] : 0.0f;
  out[2 * n + 3 * m + 0] = (float)(0.5 * (u[c] + u[c + 1]));
  out[2 * n + 3 * m + 1] = (float)(0.5 * (v[c] + v[c + (size_t)g.px]));
  out[2 * n + 3 * m + 2] = (float)(0.5 * (w[c] + w[c + (size_t)g.px * (size_t)g.py]));
}

void launch_pack_cells(const Geo& g, const double* p, const double* al, const double* const u[3], float* out,
                       hipStream_t st) {
  dim3 grid((unsigned)((g.n[0] + 63) / 64), (unsigned)((g.n[1] + 3) / 4), (unsigned)g.n[2]);
  hipLaunchKernelGGL(k_pack_cells, grid, dim3(64, 4), 0, st, g, p, al, u[0], u[1], u[2], out);
}

template <int KIND, int A>
static void launch_plane(const Geo& g, const Plane& Q, const Spacing& H, double* f, const double* rho,
                         hipStream_t st) {
  if (Q.nv == 0 || Q.ucnt <= 0 || Q.wcnt <= 0) return;
  dim3 grid((unsigned)((Q.ucnt + 63) / 64), (unsigned)((Q.wcnt + 3) / 4), (unsigned)Q.nv);
  hipLaunchKernelGGL((k_ghost_planes<KIND, A>), grid, dim3(64, 4), 0, st, g, Q, H, f, rho);
}

static Plane mkplane(int ax, int nv, const int* v, int u, int ulo, int ucnt, int w, int wlo, int wcnt) {
  Plane Q;
  Q.ax = ax;
  Q.nv = nv;
  for (int t = 0; t < 3; t++) Q.v[t] = t < nv ? v[t] : 0;
  Q.u = u;
  Q.ulo = ulo;
  Q.ucnt = ucnt;
  Q.w = w;
  Q.wlo = wlo;
  Q.wcnt = wcnt;
  return Q;
}

void launch_ghost_pressure(const Geo& g, const double h[3], double* p, const double* rho, hipStream_t st) {
  Spacing H;
  for (int d = 0; d < 3; d++) H.h[d] = h[d];
  const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
  const int vz[2] = {0, nz + 1}, vy[2] = {0, ny + 1}, vx[2] = {0, nx + 1};
  // z planes whole; y planes for interior z; x planes for interior y and z
  launch_plane<0, 0>(g, mkplane(2, 2, vz, 0, 0, nx + 2, 1, 0, ny + 2), H, p, rho, st);
  launch_plane<0, 0>(g, mkplane(1, 2, vy, 0, 0, nx + 2, 2, 1, nz), H, p, rho, st);
  launch_plane<0, 0>(g, mkplane(0, 2, vx, 1, 1, ny, 2, 1, nz), H, p, rho, st);
}

void launch_ghost_velocity(const Geo& g, double* const u[3], hipStream_t st) {
  Spacing H;
  for (int d = 0; d < 3; d++) H.h[d] = 0.0;
  for (int a = 0; a < 3; a++) {
    Plane P[3];
    int npl = 2;
    const int t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2;
    const int na = g.n[a], n1 = g.n[t1], n2 = g.n[t2];
    // own-axis range: 1..n+1 at walls (faces 0 and n+2 untouched), all of
    // 0..n+2 when periodic
    const int alo = g.per[a] ? 0 : 1, acnt = g.per[a] ? na + 3 : na + 1;
    const int v1[2] = {0, n1 + 1}, v2[2] = {0, n2 + 1}, va[3] = {0, na + 1, na + 2};
    // (u, w) of each plane are the remaining axes in increasing order, with
    // the lower-numbered one along threadIdx.x
    auto pair = [&](int ax, int& uu, int& ww) {
      uu = ax == 0 ? 1 : 0;
      ww = ax == 2 ? 1 : 2;
    };
    auto lo_cnt = [&](int d, int& lo, int& cnt, bool t_all) {
      if (d == a) {
        lo = alo;
        cnt = acnt;
      } else if (t_all) {
        lo = 0;
        cnt = g.n[d] + 2;
      } else {
        lo = 1;
        cnt = g.n[d];
      }
    };
    int uu, ww, ulo, ucnt, wlo, wcnt;
    // planes across t1: t2 over 0..n2+1, own axis over its range
    pair(t1, uu, ww);
    lo_cnt(uu, ulo, ucnt, true);
    lo_cnt(ww, wlo, wcnt, true);
    P[0] = mkplane(t1, 2, v1, uu, ulo, ucnt, ww, wlo, wcnt);
    // planes across t2: t1 over 1..n1
    pair(t2, uu, ww);
    lo_cnt(uu, ulo, ucnt, false);
    lo_cnt(ww, wlo, wcnt, false);
    P[1] = mkplane(t2, 2, v2, uu, ulo, ucnt, ww, wlo, wcnt);
    // periodic own axis: its image planes over interior t1, t2
    if (g.per[a]) {
      pair(a, uu, ww);
      P[2] = mkplane(a, 3, va, uu, 1, g.n[uu], ww, 1, g.n[ww]);
      npl = 3;
    }
    for (int t = 0; t < npl; t++) {
      if (a == 0) launch_plane<1, 0>(g, P[t], H, u[0], nullptr, st);
      else if (a == 1) launch_plane<1, 1>(g, P[t], H, u[1], nullptr, st);
      else launch_plane<1, 2>(g, P[t], H, u[2], nullptr, st);
    }
  }
}

void launch_face_gradp(const Geo& g, const double* p, double* const gp[3], hipStream_t st) {
  dim3 grid((unsigned)((g.n[0] + 1 + 63) / 64), (unsigned)((g.n[1] + 1 + 3) / 4), (unsigned)(g.n[2] + 1));
  hipLaunchKernelGGL(k_face_gradp, grid, dim3(64, 4), 0, st, g, p, gp[0], gp[1], gp[2]);
}

}  // namespace smac
