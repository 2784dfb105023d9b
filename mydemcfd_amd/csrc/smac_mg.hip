// smac_mg.hip -- geometric multigrid V-cycle kernels (RBGS smoother,
// separable restriction, trilinear prolongation, dense coarsest solve).
//
// Templated on the storage type T of the level vectors: double (the FP64
// form, bit-identical to the CPU restatement), float (MIXED_FP32) and bf16s
// (MIXED_BF16, 16-bit storage with FP32 arithmetic). The coarsest solve always
// accumulates in FP64. All level arrays use the row-split layout (even i, then
// odd i, per row) so that a half sweep streams contiguous half rows.
//
// Coefficients are uniform per level: beta0/h_l^2 on interior faces, zero across
// walls (the coarse face coefficient is the mean of the four fine faces covering
// it, which keeps beta0 and keeps walls at zero).
#include "smac_internal.hpp"

namespace smac {

__device__ __forceinline__ size_t lrow(const Lv& L, int j, int k) {
  return ((size_t)k * (size_t)L.py + (size_t)j) * (size_t)L.px;
}
__device__ __forceinline__ int LW(const Lv& L, int d, int i) { return L.per[d] ? wrapi(i, L.n[d]) : i; }

template <typename C>
struct CoefC {
  C a[6];
  C diag;
};
// VB: the level carries face coefficients L.b (variable beta); else the
// uniform beta0 (a separate instantiation, so the uniform kernels do not hold
// the registers of the variable path)
template <typename C, bool VB>
__device__ __forceinline__ CoefC<C> level_coef(const Lv& L, int i, int j, int k) {
  Coef c;
  if (VB) {
    // variable coefficients: a = beta/h^2 per face, diag = sum of all six,
    // off-diagonals zero across walls
    const int q[3] = {i, j, k};
#pragma unroll
    for (int d = 0; d < 3; d++) {
      int qp[3] = {i, j, k};
      qp[d] = L.per[d] ? wrapi(q[d] + 1, L.n[d]) : q[d] + 1;
      const size_t op = lrow(L, qp[1], qp[2]) + qp[0], om = lrow(L, j, k) + i;
      // the mixed V-cycles (compute type float) read the FP32 copies, which
      // the multigrid set-up allocates on every level and axis of a
      // variable-coefficient hierarchy in those forms; chosen at compile time,
      // so the six loads of a cell have no branch between them
      if constexpr (sizeof(C) == 4) {
        c.a[2 * d] = (double)L.bs[d][op] * L.inv_d2[d];
        c.a[2 * d + 1] = (double)L.bs[d][om] * L.inv_d2[d];
      } else {
        c.a[2 * d] = L.b[d][op] * L.inv_d2[d];
        c.a[2 * d + 1] = L.b[d][om] * L.inv_d2[d];
      }
    }
    c.diag = c.a[0] + c.a[1] + c.a[2] + c.a[3] + c.a[4] + c.a[5];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      if (wall_face(L.per[d], q[d] + 1, L.n[d])) c.a[2 * d] = 0.;
      if (wall_face(L.per[d], q[d], L.n[d])) c.a[2 * d + 1] = 0.;
    }
  } else {
    c = coef_uniform(L.n, L.per, L.inv_d2, L.beta0, i, j, k);
  }
  CoefC<C> r;
#pragma unroll
  for (int m = 0; m < 6; m++) r.a[m] = (C)c.a[m];
  r.diag = (C)c.diag;
  return r;
}

// Red-black Gauss-Seidel half sweep of colour col ((i+j+k)&1 == col) on the
// active half row: e = (f + S)/diag with S = sum of a_nb e_nb in the order
// xp, xm, yp, ym, zp, zm. zero_start: e is zero on entry, so e = f/diag.
// VPL > 0: a lane owns VPL cells h = 1+lane+64m, all loaded before any is
// computed (one memory round per wave instead of VPL); VPL = 0: any row length.
// FAST (half row = 64 VPL): the lane loads the other half of its row once;
// an active cell's two x neighbours are that half's value in the same lane
// and, one lane over, a rotate shuffle (wall neighbours are the zero ghost
// value). LDSY (FAST and ny % 4 == 0): the four waves of a block own rows
// j0..j0+3 of one plane; the other half a wave loads is exactly the y
// neighbour half of rows j0+w+-1, so it goes through LDS together with the
// matching halves of rows j0-1 and j0+4.
template <typename T, int VPL, bool VB, bool FAST, bool LDSY>
__global__ __launch_bounds__(256) void k_mg_rbgs(Lv L, T* __restrict__ e, const T* __restrict__ f, int col,
                                                 int zero_start) {
  using C = typename St<T>::C;
  constexpr int NV = VPL > 0 ? VPL : 1;
  __shared__ C ysm[LDSY ? 6 * 64 * NV : 1];
  LROW_SETUP
  const int nx = L.n[0], nE = nx >> 1;
  const int par = (col + j + k) & 1;  // parity of the active i
  const int cnt = par ? nx - nE : nE;
  const int off = par ? nE : 0;
  const size_t ypr = lrow(L, LW(L, 1, j + 1), k), ymr = lrow(L, LW(L, 1, j - 1), k);
  const size_t zpr = lrow(L, j, LW(L, 2, k + 1)), zmr = lrow(L, j, LW(L, 2, k - 1));
  if constexpr (FAST) {
    // nx = 128 VPL: both halves hold 64 VPL cells
    const int oo = par ? 0 : nE;  // the other half
    C fv[NV], ov[NV], yp[NV], ym[NV], zp[NV], zm[NV];
    if (zero_start) {
#pragma unroll
      for (int m = 0; m < NV; m++) fv[m] = St<T>::ld(f, rb + off + 1 + lane + 64 * m);
#pragma unroll
      for (int m = 0; m < NV; m++) {
        const int h = 1 + lane + 64 * m;
        const int i = par ? 2 * h - 1 : 2 * h;
        CoefC<C> c = level_coef<C, VB>(L, i, j, k);
        St<T>::st(e, rb + off + h, fv[m] / c.diag);
      }
      return;
    }
    if constexpr (LDSY) {
      const int w = (int)(threadIdx.x >> 6);
      C hv[NV];
      const size_t hr = w == 0 ? ymr : ypr;  // wave 0: row j0-1, wave 3: row j0+4
#pragma unroll
      for (int m = 0; m < NV; m++) {
        const int h = 1 + lane + 64 * m;
        fv[m] = St<T>::ld(f, rb + off + h);
        ov[m] = St<T>::ld(e, rb + oo + h);
        zp[m] = St<T>::ld(e, zpr + off + h);
        zm[m] = St<T>::ld(e, zmr + off + h);
        if (w == 0 || w == 3) hv[m] = St<T>::ld(e, hr + off + h);
      }
#pragma unroll
      for (int m = 0; m < NV; m++) {
        const int o = lane + 64 * m;
        ysm[(w + 1) * 64 * NV + o] = ov[m];
        if (w == 0) ysm[o] = hv[m];
        if (w == 3) ysm[5 * 64 * NV + o] = hv[m];
      }
      __syncthreads();
#pragma unroll
      for (int m = 0; m < NV; m++) {
        const int o = lane + 64 * m;
        ym[m] = ysm[w * 64 * NV + o];
        yp[m] = ysm[(w + 2) * 64 * NV + o];
      }
    } else {
#pragma unroll
      for (int m = 0; m < NV; m++) {
        const int h = 1 + lane + 64 * m;
        fv[m] = St<T>::ld(f, rb + off + h);
        ov[m] = St<T>::ld(e, rb + oo + h);
        yp[m] = St<T>::ld(e, ypr + off + h);
        ym[m] = St<T>::ld(e, ymr + off + h);
        zp[m] = St<T>::ld(e, zpr + off + h);
        zm[m] = St<T>::ld(e, zmr + off + h);
      }
    }
    const bool perx = L.per[0] != 0;
    const int up = (lane + 1) & 63, dn = (lane + 63) & 63;
#pragma unroll
    for (int m = 0; m < NV; m++) {
      const int h = 1 + lane + 64 * m;
      const int i = par ? 2 * h - 1 : 2 * h;
      C xp, xm;
      if (!par) {
        // even i = 2h: i-1 is the other half's slot h (this lane), i+1 slot
        // h+1 (next lane; lane 63 takes lane 0 of the next slot, i = nx wraps to 1)
        const C src = lane == 0 ? ov[m + 1 < NV ? m + 1 : 0] : ov[m];
        xm = ov[m];
        xp = __shfl(src, up);
        if (!perx && i == nx) xp = (C)0;
      } else {
        // odd i = 2h-1: i+1 is slot h (this lane), i-1 slot h-1 (previous
        // lane; lane 0 takes lane 63 of the previous slot, i = 1 wraps to nx)
        const C src = lane == 63 ? ov[m >= 1 ? m - 1 : NV - 1] : ov[m];
        xp = ov[m];
        xm = __shfl(src, dn);
        if (!perx && i == 1) xm = (C)0;
      }
      CoefC<C> c = level_coef<C, VB>(L, i, j, k);
      C S = c.a[0] * xp + c.a[1] * xm + c.a[2] * yp[m] + c.a[3] * ym[m] + c.a[4] * zp[m] + c.a[5] * zm[m];
      St<T>::st(e, rb + off + h, (fv[m] + S) / c.diag);
    }
    return;
  }
  if (VPL == 0) {
    for (int h = 1 + lane; h <= cnt; h += 64) {
      const int i = par ? 2 * h - 1 : 2 * h;
      const int p = off + h;
      CoefC<C> c = level_coef<C, VB>(L, i, j, k);
      if (zero_start) {
        St<T>::st(e, rb + p, St<T>::ld(f, rb + p) / c.diag);
      } else {
        C xp = St<T>::ld(e, rb + spos(LW(L, 0, i + 1), nx)), xm = St<T>::ld(e, rb + spos(LW(L, 0, i - 1), nx));
        C yp = St<T>::ld(e, ypr + p), ym = St<T>::ld(e, ymr + p);
        C zp = St<T>::ld(e, zpr + p), zm = St<T>::ld(e, zmr + p);
        C S = c.a[0] * xp + c.a[1] * xm + c.a[2] * yp + c.a[3] * ym + c.a[4] * zp + c.a[5] * zm;
        St<T>::st(e, rb + p, (St<T>::ld(f, rb + p) + S) / c.diag);
      }
    }
    return;
  }
  C fv[NV], xp[NV], xm[NV], yp[NV], ym[NV], zp[NV], zm[NV];
#pragma unroll
  for (int m = 0; m < NV; m++) {
    const int h = 1 + lane + 64 * m;
    if (h <= cnt) {
      const int i = par ? 2 * h - 1 : 2 * h;
      const int p = off + h;
      fv[m] = St<T>::ld(f, rb + p);
      if (!zero_start) {
        xp[m] = St<T>::ld(e, rb + spos(LW(L, 0, i + 1), nx));
        xm[m] = St<T>::ld(e, rb + spos(LW(L, 0, i - 1), nx));
        yp[m] = St<T>::ld(e, ypr + p);
        ym[m] = St<T>::ld(e, ymr + p);
        zp[m] = St<T>::ld(e, zpr + p);
        zm[m] = St<T>::ld(e, zmr + p);
      }
    }
  }
#pragma unroll
  for (int m = 0; m < NV; m++) {
    const int h = 1 + lane + 64 * m;
    if (h <= cnt) {
      const int i = par ? 2 * h - 1 : 2 * h;
      const int p = off + h;
      CoefC<C> c = level_coef<C, VB>(L, i, j, k);
      if (zero_start) {
        St<T>::st(e, rb + p, fv[m] / c.diag);
      } else {
        C S = c.a[0] * xp[m] + c.a[1] * xm[m] + c.a[2] * yp[m] + c.a[3] * ym[m] + c.a[4] * zp[m] + c.a[5] * zm[m];
        St<T>::st(e, rb + p, (fv[m] + S) / c.diag);
      }
    }
  }
}


// The smoother's FAST form with each lane owning CPL consecutive cells of a
// half row (h = 1 + CPL*lane + 64*CPL*m + t): every array access is one
// aligned vector of CPL elements (rows start on 128-byte lines and the row
// pitch is a multiple of 64 elements, checked by the launcher), and only the
// x neighbour across a chunk boundary comes from another lane (one rotate
// shuffle per chunk). The y neighbours go through LDS as in k_mg_rbgs, and the
// arithmetic per cell is the same.
template <typename T, int VPL, bool VB, int CPL, bool LDSY>
__global__ __launch_bounds__(256) void k_mg_rbgs_vec(Lv L, T* __restrict__ e, const T* __restrict__ f, int col,
                                                     int zero_start) {
  using C = typename St<T>::C;
  constexpr int NM = VPL / CPL;  // chunks per lane
  constexpr int NH = 64 * VPL;   // cells per half row
  __shared__ C ysm[LDSY ? 6 * NH : 1];
  LROW_SETUP
  const int nx = L.n[0], nE = nx >> 1;
  const int par = (col + j + k) & 1;  // parity of the active i
  const int off = par ? nE : 0;
  const int oo = par ? 0 : nE;  // the other half
  const size_t ypr = lrow(L, LW(L, 1, j + 1), k), ymr = lrow(L, LW(L, 1, j - 1), k);
  const size_t zpr = lrow(L, j, LW(L, 2, k + 1)), zmr = lrow(L, j, LW(L, 2, k - 1));
  const bool perx = L.per[0] != 0;
  // uniform coefficients: the cells off the x walls share those of i = 2
  const CoefC<C> crow = level_coef<C, false>(L, 2, j, k);
  auto coef = [&](int i) -> CoefC<C> {
    if (VB || (!perx && (i == 1 || i == nx))) return level_coef<C, VB>(L, i, j, k);
    return crow;
  };
  C fv[NM][CPL];
#pragma unroll
  for (int m = 0; m < NM; m++) Vec<T, CPL>::ld(f + rb + off + 1 + CPL * lane + 64 * CPL * m, fv[m]);
  if (zero_start) {
#pragma unroll
    for (int m = 0; m < NM; m++) {
      C ev[CPL];
#pragma unroll
      for (int t = 0; t < CPL; t++) {
        const int h = 1 + CPL * lane + 64 * CPL * m + t;
        const int i = par ? 2 * h - 1 : 2 * h;
        CoefC<C> c = coef(i);
        ev[t] = fv[m][t] / c.diag;
      }
      Vec<T, CPL>::st(e + rb + off + 1 + CPL * lane + 64 * CPL * m, ev);
    }
    return;
  }
  C ov[NM][CPL], yp[NM][CPL], ym[NM][CPL], zp[NM][CPL], zm[NM][CPL];
  // variable coefficients of the mixed forms: read with the values, ahead of
  // the barrier (FP64 keeps them after it: 14 more registers per cell)
  constexpr bool HC = VB && sizeof(C) == 4;
  CoefC<C> cv[HC ? NM : 1][HC ? CPL : 1];
  if constexpr (HC) {
#pragma unroll
    for (int m = 0; m < NM; m++)
#pragma unroll
      for (int t = 0; t < CPL; t++) {
        const int h = 1 + CPL * lane + 64 * CPL * m + t;
        cv[m][t] = coef(par ? 2 * h - 1 : 2 * h);
      }
  }
  if constexpr (LDSY) {
    const int w = (int)(threadIdx.x >> 6);
    C hv[NM][CPL];
    const size_t hr = w == 0 ? ymr : ypr;  // wave 0: row j0-1, wave 3: row j0+4
#pragma unroll
    for (int m = 0; m < NM; m++) {
      const int p = 1 + CPL * lane + 64 * CPL * m;
      Vec<T, CPL>::ld(e + rb + oo + p, ov[m]);
      Vec<T, CPL>::ld(e + zpr + off + p, zp[m]);
      Vec<T, CPL>::ld(e + zmr + off + p, zm[m]);
      if (w == 0 || w == 3) Vec<T, CPL>::ld(e + hr + off + p, hv[m]);
    }
#pragma unroll
    for (int m = 0; m < NM; m++)
#pragma unroll
      for (int t = 0; t < CPL; t++) {
        const int o = CPL * lane + 64 * CPL * m + t;
        ysm[(w + 1) * NH + o] = ov[m][t];
        if (w == 0) ysm[o] = hv[m][t];
        if (w == 3) ysm[5 * NH + o] = hv[m][t];
      }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < NM; m++)
#pragma unroll
      for (int t = 0; t < CPL; t++) {
        const int o = CPL * lane + 64 * CPL * m + t;
        ym[m][t] = ysm[w * NH + o];
        yp[m][t] = ysm[(w + 2) * NH + o];
      }
  } else {
#pragma unroll
    for (int m = 0; m < NM; m++) {
      const int p = 1 + CPL * lane + 64 * CPL * m;
      Vec<T, CPL>::ld(e + rb + oo + p, ov[m]);
      Vec<T, CPL>::ld(e + ypr + off + p, yp[m]);
      Vec<T, CPL>::ld(e + ymr + off + p, ym[m]);
      Vec<T, CPL>::ld(e + zpr + off + p, zp[m]);
      Vec<T, CPL>::ld(e + zmr + off + p, zm[m]);
    }
  }
  const int up = (lane + 1) & 63, dn = (lane + 63) & 63;
#pragma unroll
  for (int m = 0; m < NM; m++) {
    // the x neighbour across the chunk boundary: for even i the other half's
    // first cell of the next lane (lane 63: lane 0 of the next chunk; the last
    // chunk wraps to the first), for odd i its last cell of the previous lane
    C nb;
    if (!par) {
      const C src = lane == 0 ? ov[(m + 1) % NM][0] : ov[m][0];
      nb = __shfl(src, up);
    } else {
      const C src = lane == 63 ? ov[(m + NM - 1) % NM][CPL - 1] : ov[m][CPL - 1];
      nb = __shfl(src, dn);
    }
    C ev[CPL];
#pragma unroll
    for (int t = 0; t < CPL; t++) {
      const int h = 1 + CPL * lane + 64 * CPL * m + t;
      const int i = par ? 2 * h - 1 : 2 * h;
      C xp, xm;
      if (!par) {
        xm = ov[m][t];
        xp = t + 1 < CPL ? ov[m][t + 1 < CPL ? t + 1 : 0] : nb;
        if (!perx && i == nx) xp = (C)0;
      } else {
        xp = ov[m][t];
        xm = t > 0 ? ov[m][t > 0 ? t - 1 : 0] : nb;
        if (!perx && i == 1) xm = (C)0;
      }
      const CoefC<C> c = HC ? cv[HC ? m : 0][HC ? t : 0] : coef(i);
      C S = c.a[0] * xp + c.a[1] * xm + c.a[2] * yp[m][t] + c.a[3] * ym[m][t] + c.a[4] * zp[m][t] + c.a[5] * zm[m][t];
      ev[t] = (fv[m][t] + S) / c.diag;
    }
    Vec<T, CPL>::st(e + rb + off + 1 + CPL * lane + 64 * CPL * m, ev);
  }
}

// The last fine-level half sweep of a V-cycle (colour 0) with the PCG's two
// dot products of the new z in the same pass: k_mg_rbgs_vec<T, VPL, false,
// 16 / sizeof(T), true> (T = float or double) followed by k_dot2_ch's row
// partials. After this sweep the
// row's z is final in both halves, the colour-0 half computed here and the
// other half the values the sweep loaded as x neighbours; the row's z goes
// to LDS, and each lane then reads r, Ad and z at the split positions
// 1 + 2l + 128m (+1) of k_dot2_ch and forms r z and z Ad, summed in that
// kernel's order (m, then the pair; wave tree over the lanes), so the
// partials are bit-identical and z is not read again from memory.
template <typename T, int VPL>
__global__ __launch_bounds__(256) void k_mg_rbgs_dot(Lv L, T* __restrict__ e, const T* __restrict__ f,
                                                     const double* __restrict__ r, const double* __restrict__ Ad,
                                                     int flexible, double* __restrict__ part0,
                                                     double* __restrict__ part1) {
  using C = T;
  constexpr int CPL = (int)(16 / sizeof(T));
  constexpr int NM = VPL / CPL;  // smoother chunks per lane
  constexpr int NH = 64 * VPL;   // cells per half row
  constexpr int NX = 2 * NH;
  constexpr int DM = NX / 128;  // dot chunks (k_dot2_ch: 128 positions each)
  __shared__ C ysm[6 * NH];
  __shared__ C zs[4][NX];
  LROW_SETUP
  const int w = (int)(threadIdx.x >> 6);
  const int nx = L.n[0], nE = nx >> 1;
  const int par = (j + k) & 1;  // parity of the active i (colour 0)
  const int off = par ? nE : 0;
  const int oo = par ? 0 : nE;
  const size_t ypr = lrow(L, LW(L, 1, j + 1), k), ymr = lrow(L, LW(L, 1, j - 1), k);
  const size_t zpr = lrow(L, j, LW(L, 2, k + 1)), zmr = lrow(L, j, LW(L, 2, k - 1));
  const bool perx = L.per[0] != 0;
  const CoefC<C> crow = level_coef<C, false>(L, 2, j, k);
  auto coef = [&](int i) -> CoefC<C> {
    if (!perx && (i == 1 || i == nx)) return level_coef<C, false>(L, i, j, k);
    return crow;
  };
  // the dot operands at the lane's dot positions, loaded first
  double vr[DM][2], vA[DM][2];
#pragma unroll
  for (int m = 0; m < DM; m++) {
    const size_t o = rb + 1 + 2 * lane + 128 * m;
    Vec<double, 2>::ld(r + o, vr[m]);
    if (flexible) Vec<double, 2>::ld(Ad + o, vA[m]);
  }
  C fv[NM][CPL], ov[NM][CPL], yp[NM][CPL], ym[NM][CPL], zp[NM][CPL], zm[NM][CPL], hv[NM][CPL];
  const size_t hr = w == 0 ? ymr : ypr;  // wave 0: row j0-1, wave 3: row j0+4
#pragma unroll
  for (int m = 0; m < NM; m++) {
    const int p = 1 + CPL * lane + 64 * CPL * m;
    Vec<T, CPL>::ld(f + rb + off + p, fv[m]);
    Vec<T, CPL>::ld(e + rb + oo + p, ov[m]);
    Vec<T, CPL>::ld(e + zpr + off + p, zp[m]);
    Vec<T, CPL>::ld(e + zmr + off + p, zm[m]);
    if (w == 0 || w == 3) Vec<T, CPL>::ld(e + hr + off + p, hv[m]);
  }
#pragma unroll
  for (int m = 0; m < NM; m++)
#pragma unroll
    for (int t = 0; t < CPL; t++) {
      const int o = CPL * lane + 64 * CPL * m + t;
      ysm[(w + 1) * NH + o] = ov[m][t];
      if (w == 0) ysm[o] = hv[m][t];
      if (w == 3) ysm[5 * NH + o] = hv[m][t];
      zs[w][oo + o] = ov[m][t];  // the other half's z (split position oo + 1 + o, stored at oo + o)
    }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < NM; m++)
#pragma unroll
    for (int t = 0; t < CPL; t++) {
      const int o = CPL * lane + 64 * CPL * m + t;
      ym[m][t] = ysm[w * NH + o];
      yp[m][t] = ysm[(w + 2) * NH + o];
    }
  const int up = (lane + 1) & 63, dn = (lane + 63) & 63;
#pragma unroll
  for (int m = 0; m < NM; m++) {
    C nb;
    if (!par) {
      const C src = lane == 0 ? ov[(m + 1) % NM][0] : ov[m][0];
      nb = __shfl(src, up);
    } else {
      const C src = lane == 63 ? ov[(m + NM - 1) % NM][CPL - 1] : ov[m][CPL - 1];
      nb = __shfl(src, dn);
    }
    C ev[CPL];
#pragma unroll
    for (int t = 0; t < CPL; t++) {
      const int h = 1 + CPL * lane + 64 * CPL * m + t;
      const int i = par ? 2 * h - 1 : 2 * h;
      C xp, xm;
      if (!par) {
        xm = ov[m][t];
        xp = t + 1 < CPL ? ov[m][t + 1 < CPL ? t + 1 : 0] : nb;
        if (!perx && i == nx) xp = (C)0;
      } else {
        xp = ov[m][t];
        xm = t > 0 ? ov[m][t > 0 ? t - 1 : 0] : nb;
        if (!perx && i == 1) xm = (C)0;
      }
      const CoefC<C> c = coef(i);
      C S = c.a[0] * xp + c.a[1] * xm + c.a[2] * yp[m][t] + c.a[3] * ym[m][t] + c.a[4] * zp[m][t] + c.a[5] * zm[m][t];
      ev[t] = (fv[m][t] + S) / c.diag;
      zs[w][off + h - 1] = ev[t];
    }
    Vec<T, CPL>::st(e + rb + off + 1 + CPL * lane + 64 * CPL * m, ev);
  }
  __syncthreads();
  double a0 = 0.0, a1 = 0.0;
#pragma unroll
  for (int m = 0; m < DM; m++)
#pragma unroll
    for (int t = 0; t < 2; t++) {
      const double zv = (double)zs[w][2 * lane + 128 * m + t];
      a0 += vr[m][t] * zv;
      if (flexible) a1 += zv * vA[m][t];
    }
  a0 = wave_tree_sum(a0);
  if (flexible) a1 = wave_tree_sum(a1);
  if (lane == 0) {
    part0[row] = a0;
    if (flexible) part1[row] = a1;
  }
}

// a / b in the smoother: IEEE division in FP64 (the oracle's arithmetic),
// reciprocal times a in FP32 (within 1.5 ulp; the mixed forms store the
// result in FP32 or BF16 and their V-cycle is a preconditioner only)
__device__ __forceinline__ double sm_div(double a, double b) { return a / b; }
__device__ __forceinline__ float sm_div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// Half sweep with RPW = 2 or 4 rows per wave for rows of 128 CPL / RPW cells,
// CPL = the elements of one 16-byte vector (RPW 2: BF16 512-cell rows, FP32
// 256, FP64 128; RPW 4: half those): each group of LPR = 64 / RPW lanes owns a
// row and lane l of a group the cells h = 1 + CPL*l .. CPL*(l+1) of the half
// row, so every access is one 16-byte vector even where a full wave per row
// would move 8 bytes or less per lane. The four waves of a block own rows
// j0..j0+4 RPW-1 of one plane (ny % (4 RPW) == 0) and share them, with the
// rows just below and above, through LDS for the y neighbours; the x
// neighbour across a lane edge is one rotate shuffle within the row's lanes.
// The arithmetic per cell is that of k_mg_rbgs.
template <typename T, bool VB, int CPL, int RPW>
__global__ __launch_bounds__(256) void k_mg_rbgs_v2(Lv L, T* __restrict__ e, const T* __restrict__ f, int col,
                                                    int zero_start) {
  using C = typename St<T>::C;
  constexpr int LPR = 64 / RPW;     // lanes per row
  constexpr int RB = 4 * RPW;       // rows per block
  constexpr int NH = LPR * CPL;     // cells per half row
  __shared__ C ysm[(RB + 2) * NH];
  const int lane = threadIdx.x & 63;
  const int w = (int)(threadIdx.x >> 6);
  const int sub = lane / LPR, l = lane % LPR;
  const int r8 = RPW * w + sub;  // row of the block, 0..RB-1
  const int vrow = xcd_block((int)blockIdx.x, (int)gridDim.x) * RB + r8;
  int j, k;
  band_jk(vrow, L.n[1], L.n[2], j, k);
  const size_t rb = lrow(L, j, k);
  const int nx = L.n[0], nE = nx >> 1;
  const int par = (col + j + k) & 1;  // parity of the active i
  const int off = par ? nE : 0;
  const int oo = par ? 0 : nE;  // the other half
  const size_t ypr = lrow(L, LW(L, 1, j + 1), k), ymr = lrow(L, LW(L, 1, j - 1), k);
  const size_t zpr = lrow(L, j, LW(L, 2, k + 1)), zmr = lrow(L, j, LW(L, 2, k - 1));
  const int p = 1 + CPL * l;
  const bool perx = L.per[0] != 0;
  // uniform coefficients: every cell of the row off the x walls has those of
  // cell i = 2 (computed once per row, the same values cell by cell)
  const CoefC<C> crow = level_coef<C, false>(L, 2, j, k);
  auto coef = [&](int i) -> CoefC<C> {
    if (VB || (!perx && (i == 1 || i == nx))) return level_coef<C, VB>(L, i, j, k);
    return crow;
  };
  C fv[CPL];
  Vec<T, CPL>::ld(f + rb + off + p, fv);
  if (zero_start) {
    C ev[CPL];
#pragma unroll
    for (int t = 0; t < CPL; t++) {
      const int h = p + t;
      const int i = par ? 2 * h - 1 : 2 * h;
      CoefC<C> c = coef(i);
      ev[t] = sm_div(fv[t], c.diag);
    }
    Vec<T, CPL>::st(e + rb + off + p, ev);
    return;
  }
  C ov[CPL], yp[CPL], ym[CPL], zp[CPL], zm[CPL], hv[CPL];
  const bool lo = r8 == 0, hi = r8 == RB - 1;  // rows that also load the halo rows j0-1 and j0+RB
  Vec<T, CPL>::ld(e + rb + oo + p, ov);
  Vec<T, CPL>::ld(e + zpr + off + p, zp);
  Vec<T, CPL>::ld(e + zmr + off + p, zm);
  if (lo || hi) Vec<T, CPL>::ld(e + (lo ? ymr : ypr) + off + p, hv);
  // variable coefficients: read with the values, ahead of the barrier
  CoefC<C> cv[VB ? CPL : 1];
  if constexpr (VB) {
#pragma unroll
    for (int t = 0; t < CPL; t++) {
      const int h = p + t;
      cv[t] = coef(par ? 2 * h - 1 : 2 * h);
    }
  }
#pragma unroll
  for (int t = 0; t < CPL; t++) {
    const int o = CPL * l + t;
    ysm[(r8 + 1) * NH + o] = ov[t];
    if (lo) ysm[o] = hv[t];
    if (hi) ysm[(RB + 1) * NH + o] = hv[t];
  }
  __syncthreads();
#pragma unroll
  for (int t = 0; t < CPL; t++) {
    const int o = CPL * l + t;
    ym[t] = ysm[r8 * NH + o];
    yp[t] = ysm[(r8 + 2) * NH + o];
  }
  // the x neighbour across the lane edge: for even i the other half's first
  // cell of the next lane (the row's last lane: its first, i = nx wraps to
  // i = 1), for odd i its last cell of the previous lane
  C nb;
  if (!par) nb = __shfl(ov[0], sub * LPR + (l + 1) % LPR);
  else nb = __shfl(ov[CPL - 1], sub * LPR + (l + LPR - 1) % LPR);
  C ev[CPL];
#pragma unroll
  for (int t = 0; t < CPL; t++) {
    const int h = p + t;
    const int i = par ? 2 * h - 1 : 2 * h;
    C xp, xm;
    if (!par) {
      xm = ov[t];
      xp = t + 1 < CPL ? ov[t + 1 < CPL ? t + 1 : 0] : nb;
      if (!perx && i == nx) xp = (C)0;
    } else {
      xp = ov[t];
      xm = t > 0 ? ov[t > 0 ? t - 1 : 0] : nb;
      if (!perx && i == 1) xm = (C)0;
    }
    const CoefC<C> c = VB ? cv[VB ? t : 0] : coef(i);
    C S = c.a[0] * xp + c.a[1] * xm + c.a[2] * yp[t] + c.a[3] * ym[t] + c.a[4] * zp[t] + c.a[5] * zm[t];
    ev[t] = sm_div(fv[t] + S, c.diag);
  }
  Vec<T, CPL>::st(e + rb + off + p, ev);
}

// k_mg_rbgs_v2<T, false, CPL, 2> as the V-cycle's last fine half sweep
// (colour 0) with dot2's row partials, as k_mg_rbgs_dot does for the vector
// form: the eight rows' z (the stored values) go to LDS, then each wave forms
// the partials of its two rows with all 64 lanes at k_dot2_ch's positions.
template <typename T, int CPL>
__global__ __launch_bounds__(256) void k_mg_rbgs_v2_dot(Lv L, T* __restrict__ e, const T* __restrict__ f,
                                                        const double* __restrict__ r, const double* __restrict__ Ad,
                                                        int flexible, double* __restrict__ part0,
                                                        double* __restrict__ part1) {
  using C = typename St<T>::C;
  constexpr int RPW = 2;
  constexpr int LPR = 64 / RPW;  // lanes per row
  constexpr int RB = 4 * RPW;    // rows per block
  constexpr int NH = LPR * CPL;  // cells per half row
  constexpr int NX = 2 * NH;
  constexpr int DM = NX / 128;   // dot chunks
  __shared__ C ysm[(RB + 2) * NH];
  __shared__ C zs[RB][NX];
  const int lane = threadIdx.x & 63;
  const int w = (int)(threadIdx.x >> 6);
  const int sub = lane / LPR, l = lane % LPR;
  const int r8 = RPW * w + sub;
  const int vb = xcd_block((int)blockIdx.x, (int)gridDim.x) * RB;
  int j, k;
  band_jk(vb + r8, L.n[1], L.n[2], j, k);
  const size_t rb = lrow(L, j, k);
  const int nx = L.n[0], nE = nx >> 1;
  const int par = (j + k) & 1;
  const int off = par ? nE : 0;
  const int oo = par ? 0 : nE;
  const size_t ypr = lrow(L, LW(L, 1, j + 1), k), ymr = lrow(L, LW(L, 1, j - 1), k);
  const size_t zpr = lrow(L, j, LW(L, 2, k + 1)), zmr = lrow(L, j, LW(L, 2, k - 1));
  const int p = 1 + CPL * l;
  const bool perx = L.per[0] != 0;
  const CoefC<C> crow = level_coef<C, false>(L, 2, j, k);
  auto coef = [&](int i) -> CoefC<C> {
    if (!perx && (i == 1 || i == nx)) return level_coef<C, false>(L, i, j, k);
    return crow;
  };
  C fv[CPL], ov[CPL], yp[CPL], ym[CPL], zp[CPL], zm[CPL], hv[CPL];
  const bool lo = r8 == 0, hi = r8 == RB - 1;
  Vec<T, CPL>::ld(f + rb + off + p, fv);
  Vec<T, CPL>::ld(e + rb + oo + p, ov);
  Vec<T, CPL>::ld(e + zpr + off + p, zp);
  Vec<T, CPL>::ld(e + zmr + off + p, zm);
  if (lo || hi) Vec<T, CPL>::ld(e + (lo ? ymr : ypr) + off + p, hv);
#pragma unroll
  for (int t = 0; t < CPL; t++) {
    const int o = CPL * l + t;
    ysm[(r8 + 1) * NH + o] = ov[t];
    if (lo) ysm[o] = hv[t];
    if (hi) ysm[(RB + 1) * NH + o] = hv[t];
    zs[r8][oo + o] = ov[t];  // the other half's z (split position oo + 1 + o)
  }
  __syncthreads();
#pragma unroll
  for (int t = 0; t < CPL; t++) {
    const int o = CPL * l + t;
    ym[t] = ysm[r8 * NH + o];
    yp[t] = ysm[(r8 + 2) * NH + o];
  }
  C nb;
  if (!par) nb = __shfl(ov[0], sub * LPR + (l + 1) % LPR);
  else nb = __shfl(ov[CPL - 1], sub * LPR + (l + LPR - 1) % LPR);
  C ev[CPL];
#pragma unroll
  for (int t = 0; t < CPL; t++) {
    const int h = p + t;
    const int i = par ? 2 * h - 1 : 2 * h;
    C xp, xm;
    if (!par) {
      xm = ov[t];
      xp = t + 1 < CPL ? ov[t + 1 < CPL ? t + 1 : 0] : nb;
      if (!perx && i == nx) xp = (C)0;
    } else {
      xp = ov[t];
      xm = t > 0 ? ov[t > 0 ? t - 1 : 0] : nb;
      if (!perx && i == 1) xm = (C)0;
    }
    const CoefC<C> c = coef(i);
    C S = c.a[0] * xp + c.a[1] * xm + c.a[2] * yp[t] + c.a[3] * ym[t] + c.a[4] * zp[t] + c.a[5] * zm[t];
    ev[t] = sm_div(fv[t] + S, c.diag);
    zs[r8][off + CPL * l + t] = (C)stored_value<T>((double)ev[t]);
  }
  Vec<T, CPL>::st(e + rb + off + p, ev);
  // the dot operands of the wave's two rows (issued before the barrier)
  double vr[RPW][DM][2], vA[RPW][DM][2];
  int drow[RPW];
#pragma unroll
  for (int q = 0; q < RPW; q++) {
    int jq, kq;
    band_jk(vb + RPW * w + q, L.n[1], L.n[2], jq, kq);
    drow[q] = (kq - 1) * L.n[1] + (jq - 1);
    const size_t rq = lrow(L, jq, kq);
#pragma unroll
    for (int m = 0; m < DM; m++) {
      const size_t o = rq + 1 + 2 * lane + 128 * m;
      Vec<double, 2>::ld(r + o, vr[q][m]);
      if (flexible) Vec<double, 2>::ld(Ad + o, vA[q][m]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < RPW; q++) {
    double a0 = 0.0, a1 = 0.0;
#pragma unroll
    for (int m = 0; m < DM; m++)
#pragma unroll
      for (int t = 0; t < 2; t++) {
        const double zv = (double)zs[RPW * w + q][2 * lane + 128 * m + t];
        a0 += vr[q][m][t] * zv;
        if (flexible) a1 += zv * vA[q][m][t];
      }
    a0 = wave_tree_sum(a0);
    if (flexible) a1 = wave_tree_sum(a1);
    if (lane == 0) {
      part0[drow[q]] = a0;
      if (flexible) part1[drow[q]] = a1;
    }
  }
}

// 1-D restriction stencil of coarse index I (transpose of cell-centred linear
// prolongation; at a wall the mirrored weight folds onto the boundary cell)
__device__ __forceinline__ int rweights(int per, int nc, int I, int ix[4], double w[4]) {
  int nf = 2 * nc;
  if (per) {
    ix[0] = wrapi(2 * I - 2, nf); w[0] = 0.25;
    ix[1] = 2 * I - 1; w[1] = 0.75;
    ix[2] = 2 * I; w[2] = 0.75;
    ix[3] = wrapi(2 * I + 1, nf); w[3] = 0.25;
    return 4;
  }
  int m = 0;
  if (I > 1) { ix[m] = 2 * I - 2; w[m] = 0.25; m++; }
  ix[m] = 2 * I - 1; w[m] = (I == 1) ? 1.0 : 0.75; m++;
  ix[m] = 2 * I; w[m] = (I == nc) ? 1.0 : 0.75; m++;
  if (I < nc) { ix[m] = 2 * I + 1; w[m] = 0.25; m++; }
  return m;
}

// rweights padded to four taps (weight 0 on the last valid index, last),
// written with static indices so that the taps stay in registers
__device__ __forceinline__ void rtaps4(int per, int nc, int I, int ix[4], double w[4]) {
  const int nf = 2 * nc;
  if (per) {
    ix[0] = wrapi(2 * I - 2, nf); w[0] = 0.25;
    ix[1] = 2 * I - 1; w[1] = 0.75;
    ix[2] = 2 * I; w[2] = 0.75;
    ix[3] = wrapi(2 * I + 1, nf); w[3] = 0.25;
    return;
  }
  const bool lo = I == 1, hi = I == nc;
  if (!lo && !hi) {
    ix[0] = 2 * I - 2; w[0] = 0.25;
    ix[1] = 2 * I - 1; w[1] = 0.75;
    ix[2] = 2 * I; w[2] = 0.75;
    ix[3] = 2 * I + 1; w[3] = 0.25;
  } else if (lo && !hi) {
    ix[0] = 1; w[0] = 1.0;
    ix[1] = 2; w[1] = 0.75;
    ix[2] = 3; w[2] = 0.25;
    ix[3] = 3; w[3] = 0.0;
  } else if (!lo && hi) {
    ix[0] = 2 * I - 2; w[0] = 0.25;
    ix[1] = 2 * I - 1; w[1] = 0.75;
    ix[2] = 2 * I; w[2] = 1.0;
    ix[3] = 2 * I; w[3] = 0.0;
  } else {
    ix[0] = 1; w[0] = 1.0;
    ix[1] = 2; w[1] = 1.0;
    ix[2] = 2; w[2] = 0.0;
    ix[3] = 2; w[3] = 0.0;
  }
}

// Restriction, separable: coarse f = (1/8) sum_z wz sum_y wy sum_x wx res over
// the folded 4x4x4 fine neighbourhood. Pass 1 forms the residual
// res = f - (diag e - S) of one fine row in LDS and restricts it along x into
// tb (coarse x, fine y and z, natural order, pitch C.n[0]+3). Pass 2 restricts
// tb along y and z.
// RED: the pre-smoothing ended with a black half sweep, which leaves the black
// equations satisfied, so the residual is taken as 0 on black cells and formed
// on the red half row only (half the loads; n even on every restricted level).
// FAST (RED, half row = 64 VPL, ny % 4 == 0): the smoother's access pattern --
// the lane loads the other half of its row once (x neighbours by a rotate
// shuffle), and the four rows of a block share those halves through LDS as
// each other's y neighbours (with rows j0-1 and j0+4); same arithmetic.
template <typename T, int VPL, bool VB, bool RED, bool FAST, int CPL = 1>
__global__ __launch_bounds__(256) void k_mg_resid_rx(Lv F, Lv Cl, const T* __restrict__ e, const T* __restrict__ f,
                                                     typename St<T>::C* __restrict__ tb) {
  using C = typename St<T>::C;
  constexpr int NVF = VPL > 0 ? VPL : 1;
  __shared__ C ysm[FAST ? 6 * 64 * NVF : 1];
  extern __shared__ unsigned char sm_raw[];
  C* sm = reinterpret_cast<C*>(sm_raw);
  const Lv& L = F;
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int row = xcd_block((int)blockIdx.x, (int)gridDim.x) * 4 + w;
  const bool valid = row < L.n[1] * L.n[2];
  const int nx = L.n[0], nE = nx >> 1;
  C* buf = sm + (size_t)w * (size_t)(nx + 2);
  int j = 1, k = 1;
  if (valid) {
    band_jk(row, L.n[1], L.n[2], j, k);
    const size_t rb = lrow(L, j, k);
    const size_t ypr = lrow(L, LW(L, 1, j + 1), k), ymr = lrow(L, LW(L, 1, j - 1), k);
    const size_t zpr = lrow(L, j, LW(L, 2, k + 1)), zmr = lrow(L, j, LW(L, 2, k - 1));
    constexpr int NV = VPL > 0 ? VPL : 1;
    // RED: active positions p = off + h, h = 1..nE, of colour-0 cells
    const int par = (j + k) & 1;
    const int off = par ? nE : 0;
    const int cnt = RED ? nE : nx;
    if constexpr (FAST) {
      // CPL > 1: the lane owns CPL consecutive cells per chunk, h = 1 +
      // CPL*lane + 64*CPL*m + t, read and written as aligned vectors (the
      // launcher checks the row alignment); CPL = 1 is one cell per chunk
      constexpr int NM = NV / CPL;
      const int oo = par ? 0 : nE;
      C ec[NM][CPL], ef[NM][CPL], ov[NM][CPL], yp[NM][CPL], ym[NM][CPL], zp[NM][CPL], zm[NM][CPL], hv[NM][CPL];
      const size_t hr = w == 0 ? ymr : ypr;  // wave 0: row j0-1, wave 3: row j0+4
#pragma unroll
      for (int m = 0; m < NM; m++) {
        const int hb = 1 + CPL * lane + 64 * CPL * m;
        Vec<T, CPL>::ld(f + rb + off + hb, ef[m]);
        Vec<T, CPL>::ld(e + rb + off + hb, ec[m]);
        Vec<T, CPL>::ld(e + rb + oo + hb, ov[m]);
        Vec<T, CPL>::ld(e + zpr + off + hb, zp[m]);
        Vec<T, CPL>::ld(e + zmr + off + hb, zm[m]);
        if (w == 0 || w == 3) Vec<T, CPL>::ld(e + hr + off + hb, hv[m]);
      }
      // variable coefficients of the mixed forms: read with the values, ahead
      // of the barrier
      constexpr bool HC = VB && sizeof(C) == 4;
      CoefC<C> cv[HC ? NM : 1][HC ? CPL : 1];
      if constexpr (HC) {
#pragma unroll
        for (int m = 0; m < NM; m++)
#pragma unroll
          for (int t = 0; t < CPL; t++) {
            const int h = 1 + CPL * lane + 64 * CPL * m + t;
            cv[m][t] = level_coef<C, VB>(L, par ? 2 * h - 1 : 2 * h, j, k);
          }
      }
#pragma unroll
      for (int m = 0; m < NM; m++)
#pragma unroll
        for (int t = 0; t < CPL; t++) {
          const int o = CPL * lane + 64 * CPL * m + t;
          ysm[(w + 1) * 64 * NV + o] = ov[m][t];
          if (w == 0) ysm[o] = hv[m][t];
          if (w == 3) ysm[5 * 64 * NV + o] = hv[m][t];
        }
      __syncthreads();
#pragma unroll
      for (int m = 0; m < NM; m++)
#pragma unroll
        for (int t = 0; t < CPL; t++) {
          const int o = CPL * lane + 64 * CPL * m + t;
          ym[m][t] = ysm[w * 64 * NV + o];
          yp[m][t] = ysm[(w + 2) * 64 * NV + o];
        }
      const bool perx = L.per[0] != 0;
      const int up = (lane + 1) & 63, dn = (lane + 63) & 63;
      // uniform coefficients: the cells off the x walls share those of i = 2
      const CoefC<C> crow = level_coef<C, false>(L, 2, j, k);
#pragma unroll
      for (int m = 0; m < NM; m++) {
        // the x neighbour across the chunk boundary (as in k_mg_rbgs_vec)
        C nb;
        if (!par) {
          const C src = lane == 0 ? ov[(m + 1) % NM][0] : ov[m][0];
          nb = __shfl(src, up);
        } else {
          const C src = lane == 63 ? ov[(m + NM - 1) % NM][CPL - 1] : ov[m][CPL - 1];
          nb = __shfl(src, dn);
        }
#pragma unroll
        for (int t = 0; t < CPL; t++) {
          const int h = 1 + CPL * lane + 64 * CPL * m + t;
          const int i = par ? 2 * h - 1 : 2 * h;
          C xp, xm;
          if (!par) {  // even i = 2h: i-1 in this cell's slot, i+1 the next slot
            xm = ov[m][t];
            xp = t + 1 < CPL ? ov[m][t + 1 < CPL ? t + 1 : 0] : nb;
            if (!perx && i == nx) xp = (C)0;
          } else {     // odd i = 2h-1: i+1 in this cell's slot, i-1 the previous slot
            xp = ov[m][t];
            xm = t > 0 ? ov[m][t > 0 ? t - 1 : 0] : nb;
            if (!perx && i == 1) xm = (C)0;
          }
          const CoefC<C> c = HC ? cv[HC ? m : 0][HC ? t : 0]
                                : ((VB || (!perx && (i == 1 || i == nx))) ? level_coef<C, VB>(L, i, j, k) : crow);
          C S = c.a[0] * xp + c.a[1] * xm + c.a[2] * yp[m][t] + c.a[3] * ym[m][t] + c.a[4] * zp[m][t] +
                c.a[5] * zm[m][t];
          buf[i] = ef[m][t] - (c.diag * ec[m][t] - S);
          buf[par ? i + 1 : i - 1] = (C)0;  // the black neighbour of the same pair
        }
      }
    } else {
    const int trips = VPL > 0 ? 1 : (cnt + 63) / 64;
    for (int t = 0; t < trips; t++) {
      // VPL > 0: the lane's cells h = 1+lane+64m, all loads before the arithmetic
      C ec[NV], ef[NV], xp[NV], xm[NV], yp[NV], ym[NV], zp[NV], zm[NV];
#pragma unroll
      for (int m = 0; m < NV; m++) {
        const int h = 1 + lane + 64 * (VPL > 0 ? m : t);
        const int p = RED ? off + h : h;
        if (h <= cnt) {
          const int i = RED ? (par ? 2 * h - 1 : 2 * h) : (p <= nE ? 2 * p : 2 * (p - nE) - 1);
          xp[m] = St<T>::ld(e, rb + spos(LW(L, 0, i + 1), nx));
          xm[m] = St<T>::ld(e, rb + spos(LW(L, 0, i - 1), nx));
          yp[m] = St<T>::ld(e, ypr + p);
          ym[m] = St<T>::ld(e, ymr + p);
          zp[m] = St<T>::ld(e, zpr + p);
          zm[m] = St<T>::ld(e, zmr + p);
          ef[m] = St<T>::ld(f, rb + p);
          ec[m] = St<T>::ld(e, rb + p);
        }
      }
#pragma unroll
      for (int m = 0; m < NV; m++) {
        const int h = 1 + lane + 64 * (VPL > 0 ? m : t);
        const int p = RED ? off + h : h;
        if (h <= cnt) {
          const int i = RED ? (par ? 2 * h - 1 : 2 * h) : (p <= nE ? 2 * p : 2 * (p - nE) - 1);
          CoefC<C> c = level_coef<C, VB>(L, i, j, k);
          C S = c.a[0] * xp[m] + c.a[1] * xm[m] + c.a[2] * yp[m] + c.a[3] * ym[m] + c.a[4] * zp[m] + c.a[5] * zm[m];
          buf[i] = ef[m] - (c.diag * ec[m] - S);
          if (RED) buf[par ? i + 1 : i - 1] = (C)0;  // the black neighbour of the same pair
        }
      }
    }
    }
  }
  __syncthreads();
  if (!valid) return;
  const size_t tr = ((size_t)k * (size_t)L.py + (size_t)j) * (size_t)(Cl.n[0] + 3);
  if (Cl.per[0]) {
    // periodic: the four taps of rweights with its constant weights, in its
    // order (only the end taps wrap)
    const int nf = L.n[0];
    const C q = (C)0.25, h = (C)0.75;
    for (int I = 1 + lane; I <= Cl.n[0]; I += 64) {
      const int i0 = I == 1 ? nf : 2 * I - 2, i3 = 2 * I + 1 > nf ? 1 : 2 * I + 1;
      C tx = 0;
      tx += q * buf[i0];
      tx += h * buf[2 * I - 1];
      tx += h * buf[2 * I];
      tx += q * buf[i3];
      tb[tr + I] = tx;
    }
  } else {
    for (int I = 1 + lane; I <= Cl.n[0]; I += 64) {
      int ix[4];
      double wx[4];
      int mx = rweights(Cl.per[0], Cl.n[0], I, ix, wx);
      C tx = 0;
      for (int c = 0; c < mx; c++) tx += (C)wx[c] * buf[ix[c]];
      tb[tr + I] = tx;
    }
  }
}


template <typename T>
__global__ __launch_bounds__(256) void k_mg_ryz(Lv F, Lv Cl, const typename St<T>::C* __restrict__ tb,
                                                T* __restrict__ fc, int KC) {
  using C = typename St<T>::C;
  // a thread owns the coarse column (I, J) over the coarse planes K0..K0+KC-1;
  // blocks in x-fastest order, remapped so that each XCD takes a contiguous
  // range of chunks (neighbouring coarse rows and chunks share fine rows of tb)
  const int nbx = (int)gridDim.x, nby = (int)gridDim.y;
  const int lin = xcd_block((int)(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)),
                            (int)(gridDim.x * gridDim.y * gridDim.z));
  const int bx = lin % nbx, by = (lin / nbx) % nby, bz = lin / (nbx * nby);
  const int I = 1 + bx * 64 + (int)threadIdx.x;
  const int J = 1 + by * 4 + (int)threadIdx.y;
  const int K0 = 1 + bz * KC;
  if (I > Cl.n[0] || J > Cl.n[1]) return;
  // four taps (padded with weight 0 on a valid index: the padded terms come
  // last and add exact zeros, so the sums are those of the 3- or 4-term stencil)
  int iy[4];
  double wy[4];
  rtaps4(Cl.per[1], Cl.n[1], J, iy, wy);
  const size_t pxT = (size_t)(Cl.n[0] + 3);
  // the y-restricted value of fine plane z, ty(z) = sum_b wy[b] tb(I, iy[b], z),
  // computed once per plane: the last four planes are kept
  // (static register indices only: the newest plane enters at the end)
  int cz[4] = {-1, -1, -1, -1};
  C cv[4] = {(C)0, (C)0, (C)0, (C)0};
  auto ty_of = [&](int z) -> C {
    C hit = (C)0;
    bool found = false;
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (!found && cz[u] == z) {
        hit = cv[u];
        found = true;
      }
    if (found) return hit;
    C v[4];
#pragma unroll
    for (int b = 0; b < 4; b++) v[b] = tb[((size_t)z * (size_t)F.py + (size_t)iy[b]) * pxT + I];
    C ty = 0;
#pragma unroll
    for (int b = 0; b < 4; b++) ty += (C)wy[b] * v[b];
#pragma unroll
    for (int u = 0; u < 3; u++) {
      cz[u] = cz[u + 1];
      cv[u] = cv[u + 1];
    }
    cz[3] = z;
    cv[3] = ty;
    return ty;
  };
  const int K1 = min(K0 + KC, Cl.n[2] + 1);
  for (int K = K0; K < K1; K++) {
    int iz[4];
    double wz[4];
    rtaps4(Cl.per[2], Cl.n[2], K, iz, wz);
    C tz = 0;
#pragma unroll
    for (int a = 0; a < 4; a++) tz += (C)wz[a] * ty_of(iz[a]);
    St<T>::st(fc, lrow(Cl, J, K) + spos(I, Cl.n[0]), (C)0.125 * tz);
  }
}

__device__ __forceinline__ void pweights(int per, int nc, int i, int ix[2], double w[2]) {
  int par = (i + 1) >> 1;
  int nb = (i & 1) ? par - 1 : par + 1;
  if (per) nb = wrapi(nb, nc);
  else if (nb < 1 || nb > nc) nb = par;
  ix[0] = par; w[0] = 0.75;
  ix[1] = nb; w[1] = 0.25;
}

// e_f += sum over 8 coarse cells of (wz*wy*wx) e_c (cell-centred trilinear).
// col >= 0: only the fine cells of that colour (the other colour is
// recomputed by the next half sweep without being read); col < 0: all cells.
template <typename T, int VPL>
__global__ __launch_bounds__(256) void k_mg_prolong(Lv Cl, Lv F, const T* __restrict__ ec, T* __restrict__ ef,
                                                    int col) {
  using C = typename St<T>::C;
  const Lv& L = F;
  LROW_SETUP
  const int nx = L.n[0], nE = nx >> 1;
  const int par = (col + j + k) & 1;
  const int cnt = col < 0 ? nx : (par ? nx - nE : nE);
  const int off = (col < 0 || !par) ? 0 : nE;
  int iy[2], iz[2];
  double wy[2], wz[2];
  pweights(Cl.per[1], Cl.n[1], j, iy, wy);
  pweights(Cl.per[2], Cl.n[2], k, iz, wz);
  size_t cr[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) cr[a][b] = lrow(Cl, iy[b], iz[a]);
  constexpr int NV = VPL > 0 ? VPL : 1;
  const int trips = VPL > 0 ? 1 : (cnt + 63) / 64;
  for (int t = 0; t < trips; t++) {
    // VPL > 0: all coarse and fine loads of the lane's cells first
    C cv[NV][2][2][2], fv[NV];
    double wxs[NV][2];
#pragma unroll
    for (int m = 0; m < NV; m++) {
      const int h = 1 + lane + 64 * (VPL > 0 ? m : t);
      if (h <= cnt) {
        const int pp = off + h;
        const int i = pp <= nE ? 2 * pp : 2 * (pp - nE) - 1;
        int ix[2];
        pweights(Cl.per[0], Cl.n[0], i, ix, wxs[m]);
        const int c0 = spos(ix[0], Cl.n[0]), c1 = spos(ix[1], Cl.n[0]);
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
          for (int b = 0; b < 2; b++) {
            cv[m][a][b][0] = St<T>::ld(ec, cr[a][b] + c0);
            cv[m][a][b][1] = St<T>::ld(ec, cr[a][b] + c1);
          }
        fv[m] = St<T>::ld(ef, rb + pp);
      }
    }
#pragma unroll
    for (int m = 0; m < NV; m++) {
      const int h = 1 + lane + 64 * (VPL > 0 ? m : t);
      if (h <= cnt) {
        const int pp = off + h;
        C acc = 0;
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
          for (int b = 0; b < 2; b++) {
            C w0 = (C)(wz[a] * wy[b] * wxs[m][0]);
            acc += w0 * cv[m][a][b][0];
            C w1 = (C)(wz[a] * wy[b] * wxs[m][1]);
            acc += w1 * cv[m][a][b][1];
          }
        St<T>::st(ef, rb + pp, fv[m] + acc);
      }
    }
  }
}

// The same one-colour prolongation for a block of four fine rows j0..j0+3
// (j0 = 4b+1) of one plane, whose coarse cells lie in the coarse rows
// 2b..2b+3 (periodic images or wall ghosts at the ends) of the two coarse
// planes of k: those eight coarse rows are staged once in LDS in natural x
// order (coalesced loads of the split rows), and each fine cell reads its
// eight coarse values there. Needs col >= 0, ny % 4 == 0 and a half row of
// 64 VPL cells; same arithmetic, in the same order, as k_mg_prolong.
template <typename T, int VPL, int CPL = 1>
__global__ __launch_bounds__(256) void k_mg_prolong_lds(Lv Cl, Lv F, const T* __restrict__ ec,
                                                        T* __restrict__ ef, int col) {
  using C = typename St<T>::C;
  constexpr int NC = 64 * VPL;  // coarse row length
  __shared__ C cs[8 * NC];
  const Lv& L = F;
  LROW_SETUP
  const int nE = L.n[0] >> 1;
  const int w = (int)(threadIdx.x >> 6);
  const int par = (col + j + k) & 1;
  const int off = par ? nE : 0;
  int iy[2], iz[2];
  double wy[2], wz[2];
  pweights(Cl.per[1], Cl.n[1], j, iy, wy);
  pweights(Cl.per[2], Cl.n[2], k, iz, wz);
  const int j0 = j - w;
  const int jb = (j0 - 1) >> 1;  // coarse row of LDS slot 0: 2b
  // LDS slot of each coarse y index of this row: J - 2b, or, for the
  // neighbour row, the unwrapped neighbour - 2b (its wrapped image is what
  // slot 0 or 3 holds); a wall-clamped neighbour is J itself
  int sy[2];
  {
    const int J = (j + 1) >> 1;
    const int nb = (j & 1) ? J - 1 : J + 1;
    sy[0] = J - jb;
    sy[1] = (Cl.per[1] || (nb >= 1 && nb <= Cl.n[1])) ? nb - jb : sy[0];
  }
  // stage coarse rows (2b + s, iz[a]) -> slot a*4 + s, natural order; all
  // loads (and the fine cells') are issued before the first LDS store
  const int t = (int)threadIdx.x;
  constexpr int PER = (NC + 255) / 256;
  C stg[8][PER], fv[VPL];
#pragma unroll
  for (int r = 0; r < 8; r++) {
    const size_t crow = lrow(Cl, LW(Cl, 1, jb + (r & 3)), iz[r >> 2]);
#pragma unroll
    for (int u = 0; u < PER; u++) {
      const int q = 1 + t + 256 * u;
      if (q <= NC) stg[r][u] = St<T>::ld(ec, crow + q);
    }
  }
  // the lane's fine cells: h = 1 + CPL*lane + 64*CPL*mm + t, CPL consecutive
  // cells read and written as one aligned vector (CPL > 1: rows checked aligned)
  constexpr int NM = VPL / CPL;
#pragma unroll
  for (int mm = 0; mm < NM; mm++) Vec<T, CPL>::ld(ef + rb + off + 1 + CPL * lane + 64 * CPL * mm, fv + mm * CPL);
#pragma unroll
  for (int r = 0; r < 8; r++)
#pragma unroll
    for (int u = 0; u < PER; u++) {
      const int q = 1 + t + 256 * u;
      if (q <= NC) cs[r * NC + (q <= NC / 2 ? 2 * q : 2 * (q - NC / 2) - 1) - 1] = stg[r][u];
    }
  __syncthreads();
  // the eight weights wz*wy*wx are the same for every cell of the row (wx is
  // 3/4 on the cell's own coarse column and 1/4 on the neighbour, walls only
  // clamp the index): products formed once, in the per-cell order
  C wq[2][2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) {
      wq[a][b][0] = (C)(wz[a] * wy[b] * 0.75);
      wq[a][b][1] = (C)(wz[a] * wy[b] * 0.25);
    }
#pragma unroll
  for (int mm = 0; mm < NM; mm++) {
    C out[CPL];
#pragma unroll
    for (int t = 0; t < CPL; t++) {
      const int h = 1 + CPL * lane + 64 * CPL * mm + t;
      const int pp = off + h;
      const int i = pp <= nE ? 2 * pp : 2 * (pp - nE) - 1;
      int ix[2];
      double wx[2];
      pweights(Cl.per[0], Cl.n[0], i, ix, wx);
      C acc = 0;
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
          const C* rowp = cs + (a * 4 + sy[b]) * NC;
          acc += wq[a][b][0] * rowp[ix[0] - 1];
          acc += wq[a][b][1] * rowp[ix[1] - 1];
        }
      out[t] = fv[mm * CPL + t] + acc;
    }
    Vec<T, CPL>::st(ef + rb + off + 1 + CPL * lane + 64 * CPL * mm, out);
  }
}

// coarsest level: e = Ainv f with the dense pseudo-inverse (rows and columns
// in natural x-fastest order, row-major). One wave per row: lane l
// accumulates the columns l, l+64, ... in order (FP64), and the 64 lane sums
// are combined by the canonical halving tree -- the oracle's order.
template <typename T>
__global__ __launch_bounds__(256) void k_mg_coarse(Lv Cl, const double* __restrict__ ainv, const T* __restrict__ f,
                                                   T* __restrict__ e) {
  const int n = Cl.n[0] * Cl.n[1] * Cl.n[2];
  const int lane = threadIdx.x & 63;
  const int row = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6);
  if (row >= n) return;
  const double* a = ainv + (size_t)row * (size_t)n;
  const int nxy = Cl.n[0] * Cl.n[1];
  double acc = 0.0;
  constexpr int U = 8;  // columns per lane whose loads are issued before their (sequential) adds
  int col = lane;
  for (; col + 64 * (U - 1) < n; col += 64 * U) {
    double av[U], fv[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int cc = col + 64 * u;
      const int ci = cc % Cl.n[0] + 1, cj = (cc / Cl.n[0]) % Cl.n[1] + 1, ck = cc / nxy + 1;
      av[u] = a[cc];
      fv[u] = (double)St<T>::ld(f, lrow(Cl, cj, ck) + spos(ci, Cl.n[0]));
    }
#pragma unroll
    for (int u = 0; u < U; u++) acc += av[u] * fv[u];
  }
  for (; col < n; col += 64) {
    const int ci = col % Cl.n[0] + 1, cj = (col / Cl.n[0]) % Cl.n[1] + 1, ck = col / nxy + 1;
    acc += a[col] * (double)St<T>::ld(f, lrow(Cl, cj, ck) + spos(ci, Cl.n[0]));
  }
  acc = wave_tree_sum(acc);
  if (lane == 0) {
    const int i = row % Cl.n[0] + 1, j = (row / Cl.n[0]) % Cl.n[1] + 1, k = row / nxy + 1;
    St<T>::st(e, lrow(Cl, j, k) + spos(i, Cl.n[0]), (typename St<T>::C)acc);
  }
}

// x = 0 on interior cells of a level (a fill kernel, safe under graph capture)
template <typename T>
__global__ __launch_bounds__(256) void k_fill(Lv L, T* __restrict__ x) {
  LROW_SETUP
  for (int p = 1 + lane; p <= L.n[0]; p += 64) St<T>::st(x, rb + p, (typename St<T>::C)0);
}

// ---------------------------------------------------------------- launchers
static inline dim3 lrows_grid(const Lv& L) { return dim3((unsigned)((L.n[1] * L.n[2] + 3) / 4)); }
static inline dim3 cell_grid(const Lv& L) {
  return dim3((unsigned)((L.n[0] + 63) / 64), (unsigned)((L.n[1] + 3) / 4), (unsigned)L.n[2]);
}
static const dim3 ROWB(256), CELLB(64, 4, 1);


// vector accesses need every row's element 1 on a 128-byte line
template <typename T>
static bool rows_aligned(const Lv& L, const T* a) {
  return L.px % 64 == 0 && ((uintptr_t)(a + 1) % 128) == 0;
}
// cells per lane of the vector forms: 16 bytes, at most the lane's share
template <typename T>
static constexpr int vec_cpl(int vpl) {
  return (int)(16 / sizeof(T)) < vpl ? (int)(16 / sizeof(T)) : vpl;
}

template <typename T>
int mg_zs_division(const Lv& L, const T* e, const T* f) {
  constexpr int C16 = (int)(16 / sizeof(T));
  if (sizeof(T) == 8) return 1;
  const bool al = rows_aligned(L, e) && rows_aligned(L, f);
  // the two- and four-rows-per-wave forms divide with sm_div (the reciprocal)
  if (al && L.n[0] == 64 * C16 && L.n[1] % 8 == 0) return 2;
  if (al && L.n[0] == 32 * C16 && L.n[1] % 16 == 0) return 2;
  return 1;
}
template <typename T>
void launch_mg_rbgs(const Lv& L, T* e, const T* f, int col, int zero_start, hipStream_t st) {
  const int half = L.n[0] - (L.n[0] >> 1);  // longest active half row
  const int vpl = row_vpl(half);
  const bool fast = vpl >= 1 && L.n[0] == 128 * vpl;
  const bool ldsy = fast && L.n[1] % 4 == 0;
  constexpr int C16 = (int)(16 / sizeof(T));  // elements of a 16-byte vector
  const bool al = rows_aligned(L, e) && rows_aligned(L, f);
  if (al && L.n[0] == 64 * C16 && L.n[1] % 8 == 0) {
    const dim3 grid((unsigned)(L.n[1] * L.n[2] / 8));
    if (L.b[0]) hipLaunchKernelGGL((k_mg_rbgs_v2<T, true, C16, 2>), grid, ROWB, 0, st, L, e, f, col, zero_start);
    else hipLaunchKernelGGL((k_mg_rbgs_v2<T, false, C16, 2>), grid, ROWB, 0, st, L, e, f, col, zero_start);
    return;
  }
  if (al && L.n[0] == 32 * C16 && L.n[1] % 16 == 0) {
    const dim3 grid((unsigned)(L.n[1] * L.n[2] / 16));
    if (L.b[0]) hipLaunchKernelGGL((k_mg_rbgs_v2<T, true, C16, 4>), grid, ROWB, 0, st, L, e, f, col, zero_start);
    else hipLaunchKernelGGL((k_mg_rbgs_v2<T, false, C16, 4>), grid, ROWB, 0, st, L, e, f, col, zero_start);
    return;
  }
  if (fast && vpl >= 2 && rows_aligned(L, e) && rows_aligned(L, f)) {
#define SMAC_RV(V, P, Y) \
  hipLaunchKernelGGL((k_mg_rbgs_vec<T, P, V, vec_cpl<T>(P), Y>), lrows_grid(L), ROWB, 0, st, L, e, f, col, zero_start)
#define SMAC_RV_V(V)                                                              \
  switch (vpl) {                                                                  \
    case 2: if (ldsy) SMAC_RV(V, 2, true); else SMAC_RV(V, 2, false); break;      \
    case 4: if (ldsy) SMAC_RV(V, 4, true); else SMAC_RV(V, 4, false); break;      \
    default: if (ldsy) SMAC_RV(V, 8, true); else SMAC_RV(V, 8, false); break;     \
  }
    if (L.b[0]) {
      SMAC_RV_V(true)
    } else {
      SMAC_RV_V(false)
    }
#undef SMAC_RV_V
#undef SMAC_RV
    return;
  }
#define SMAC_RBGS(V, P, F, Y) \
  hipLaunchKernelGGL((k_mg_rbgs<T, P, V, F, Y>), lrows_grid(L), ROWB, 0, st, L, e, f, col, zero_start)
#define SMAC_RBGS_V(V)                                                \
  if (fast) {                                                         \
    switch (vpl) {                                                    \
      case 1: if (ldsy) SMAC_RBGS(V, 1, true, true); else SMAC_RBGS(V, 1, true, false); break; \
      case 2: if (ldsy) SMAC_RBGS(V, 2, true, true); else SMAC_RBGS(V, 2, true, false); break; \
      case 4: if (ldsy) SMAC_RBGS(V, 4, true, true); else SMAC_RBGS(V, 4, true, false); break; \
      default: if (ldsy) SMAC_RBGS(V, 8, true, true); else SMAC_RBGS(V, 8, true, false); break; \
    }                                                                 \
  } else {                                                            \
    SMAC_VPL_DISPATCH(half, SMAC_RBGS(V, VPL, false, false))          \
  }
  if (L.b[0]) {
    SMAC_RBGS_V(true)
  } else {
    SMAC_RBGS_V(false)
  }
#undef SMAC_RBGS_V
#undef SMAC_RBGS
}

// which fused form takes the level's colour-0 half sweep (the form
// launch_mg_rbgs would pick, uniform coefficients): 2 = two rows per wave,
// 1 = the vector form (FP32, FP64), 0 = none
template <typename T>
static int rbgs_dot_form(const Lv& L, const T* e, const T* f) {
  constexpr int C16 = (int)(16 / sizeof(T));
  if (L.b[0] || !rows_aligned(L, e) || !rows_aligned(L, f)) return 0;
  if (L.n[0] == 64 * C16 && L.n[1] % 8 == 0) return 2;
  if (L.n[0] == 32 * C16 && L.n[1] % 16 == 0) return 0;
  const int vpl = row_vpl(L.n[0] - (L.n[0] >> 1));
  if (sizeof(T) >= 4 && (vpl == 4 || vpl == 8) && L.n[0] == 128 * vpl && L.n[1] % 4 == 0) return 1;
  return 0;
}
template <typename T>
bool mg_rbgs_dot_applies(const Lv& L, const T* e, const T* f) {
  return rbgs_dot_form(L, e, f) != 0;
}
template <typename T>
void launch_mg_rbgs_dot(const Lv& L, T* e, const T* f, const double* r, const double* Ad, int flexible, double* p0,
                        double* p1, hipStream_t st) {
  constexpr int C16 = (int)(16 / sizeof(T));
  if (rbgs_dot_form(L, e, f) == 2) {
    hipLaunchKernelGGL((k_mg_rbgs_v2_dot<T, C16>), dim3((unsigned)(L.n[1] * L.n[2] / 8)), ROWB, 0, st, L, e, f, r,
                       Ad, flexible, p0, p1);
    return;
  }
  if constexpr (sizeof(T) >= 4) {
    const int vpl = row_vpl(L.n[0] - (L.n[0] >> 1));
    if (vpl == 4)
      hipLaunchKernelGGL((k_mg_rbgs_dot<T, 4>), lrows_grid(L), ROWB, 0, st, L, e, f, r, Ad, flexible, p0, p1);
    else
      hipLaunchKernelGGL((k_mg_rbgs_dot<T, 8>), lrows_grid(L), ROWB, 0, st, L, e, f, r, Ad, flexible, p0, p1);
  }
}
template bool mg_rbgs_dot_applies<float>(const Lv&, const float*, const float*);
template bool mg_rbgs_dot_applies<double>(const Lv&, const double*, const double*);
template bool mg_rbgs_dot_applies<bf16s>(const Lv&, const bf16s*, const bf16s*);
template void launch_mg_rbgs_dot<float>(const Lv&, float*, const float*, const double*, const double*, int, double*,
                                        double*, hipStream_t);
template void launch_mg_rbgs_dot<double>(const Lv&, double*, const double*, const double*, const double*, int,
                                         double*, double*, hipStream_t);
template void launch_mg_rbgs_dot<bf16s>(const Lv&, bf16s*, const bf16s*, const double*, const double*, int, double*,
                                        double*, hipStream_t);

template <typename T>
void launch_mg_restrict(const Lv& F, const Lv& C, const T* e, const T* f, typename St<T>::C* tb, T* fc, int red,
                        hipStream_t st) {
  size_t shm = 4 * (size_t)(F.n[0] + 2) * sizeof(typename St<T>::C);
  const int nx = F.n[0];
  const int vpl = row_vpl(nx / 2);
  const bool fast = red && vpl >= 1 && nx == 128 * vpl && F.n[1] % 4 == 0;
  const bool vec = fast && rows_aligned(F, e) && rows_aligned(F, f);
#define SMAC_RX(V, P, R, FA) \
  hipLaunchKernelGGL((k_mg_resid_rx<T, P, V, R, FA>), lrows_grid(F), ROWB, shm, st, F, C, e, f, tb)
#define SMAC_RXV(V, P) \
  hipLaunchKernelGGL((k_mg_resid_rx<T, P, V, true, true, vec_cpl<T>(P)>), lrows_grid(F), ROWB, shm, st, F, C, e, f, tb)
#define SMAC_RX_V(V)                                                         \
  if (fast) {                                                                \
    switch (vpl) {                                                           \
      case 1: SMAC_RX(V, 1, true, true); break;                              \
      case 2: if (vec) SMAC_RXV(V, 2); else SMAC_RX(V, 2, true, true); break; \
      case 4: if (vec) SMAC_RXV(V, 4); else SMAC_RX(V, 4, true, true); break; \
      default: if (vec) SMAC_RXV(V, 8); else SMAC_RX(V, 8, true, true); break; \
    }                                                                        \
  } else if (red) {                                                          \
    SMAC_VPL_DISPATCH(nx / 2, SMAC_RX(V, VPL, true, false))                  \
  } else {                                                                   \
    SMAC_VPL_DISPATCH(nx, SMAC_RX(V, VPL, false, false))                     \
  }
  if (F.b[0]) {
    SMAC_RX_V(true)
  } else {
    SMAC_RX_V(false)
  }
#undef SMAC_RX_V
#undef SMAC_RXV
#undef SMAC_RX
  // coarse planes per thread: 8 on large levels (half the loads), 1 on small
  // ones, whose launch needs every thread it can get
  const int KC = (size_t)C.n[0] * C.n[1] * C.n[2] >= ((size_t)1 << 21) ? 8 : 1;
  const dim3 grid((unsigned)((C.n[0] + 63) / 64), (unsigned)((C.n[1] + 3) / 4), (unsigned)((C.n[2] + KC - 1) / KC));
  hipLaunchKernelGGL(k_mg_ryz<T>, grid, CELLB, 0, st, F, C, tb, fc, KC);
}
template <typename T>
void launch_mg_prolong(const Lv& C, const Lv& F, const T* ec, T* ef, int col, hipStream_t st) {
  const int cntmax = col < 0 ? F.n[0] : F.n[0] - (F.n[0] >> 1);
  const int vpl = row_vpl(cntmax);
  if (col >= 0 && vpl >= 1 && F.n[0] == 128 * vpl && F.n[1] % 4 == 0) {
    const bool vec = rows_aligned(F, ef);
#define SMAC_PROLONG_LDS(P) \
  hipLaunchKernelGGL((k_mg_prolong_lds<T, P>), lrows_grid(F), ROWB, 0, st, C, F, ec, ef, col)
#define SMAC_PROLONG_LDSV(P) \
  hipLaunchKernelGGL((k_mg_prolong_lds<T, P, vec_cpl<T>(P)>), lrows_grid(F), ROWB, 0, st, C, F, ec, ef, col)
    switch (vpl) {
      case 1: SMAC_PROLONG_LDS(1); break;
      case 2: if (vec) SMAC_PROLONG_LDSV(2); else SMAC_PROLONG_LDS(2); break;
      case 4: if (vec) SMAC_PROLONG_LDSV(4); else SMAC_PROLONG_LDS(4); break;
      default: if (vec) SMAC_PROLONG_LDSV(8); else SMAC_PROLONG_LDS(8); break;
    }
#undef SMAC_PROLONG_LDSV
#undef SMAC_PROLONG_LDS
    return;
  }
  SMAC_VPL_DISPATCH(cntmax, hipLaunchKernelGGL((k_mg_prolong<T, VPL>), lrows_grid(F), ROWB, 0, st, C, F, ec, ef, col))
}
template <typename T>
void launch_mg_coarse(const Lv& C, const double* ainv, const T* f, T* e, hipStream_t st) {
  int n = C.n[0] * C.n[1] * C.n[2];
  hipLaunchKernelGGL(k_mg_coarse<T>, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, st, C, ainv, f, e);
}
// coarse face coefficient of axis d (mg_coarsen_beta of the oracle): faces
// off the level's face range are 0
__global__ __launch_bounds__(256) void k_mg_coarsen_beta(Lv F, Lv C, int d, double* __restrict__ cb) {
  const int I = (int)(blockIdx.x * 64 + threadIdx.x);
  const int J = (int)(blockIdx.y * 4 + threadIdx.y);
  const int K = (int)blockIdx.z;
  if (I >= C.px || J >= C.py) return;
  const int q[3] = {I, J, K};
  const size_t id = lrow(C, J, K) + I;
  bool ok = true;
#pragma unroll
  for (int e = 0; e < 3; e++) {
    const int hi = C.n[e] + (e == d ? 1 : 0);
    if (q[e] < 1 || q[e] > hi) ok = false;
  }
  if (!ok) {
    cb[id] = 0.;
    return;
  }
  const int t1 = (d + 1) % 3, t2 = (d + 2) % 3;
  // layer sums S(m) of the 2x2 fine faces 2q-1+m along d (m = -1, 0, 1;
  // periodic: wrapped); a face on a wall or outlet side takes its own layer,
  // an interior face the weights 1/4, 1/2, 1/4 along the normal (the coupling
  // of the two coarse cells, the fine faces inside them included: with a
  // density jump inside a coarse cell, the coplanar layer alone leaves the
  // V-cycle four times weaker at 1000:1)
  const bool bnd = !C.per[d] && (q[d] == 1 || q[d] == C.n[d] + 1);
  double S3[3] = {0., 0., 0.};
#pragma unroll
  for (int m = -1; m <= 1; m++) {
    if (bnd && m != 0) continue;
    int fq[3];
    fq[d] = 2 * q[d] - 1 + m;
    if (m != 0 && C.per[d]) fq[d] = wrapi(fq[d], F.n[d]);
    double v[2][2];
#pragma unroll
    for (int s2 = 0; s2 < 2; s2++)
#pragma unroll
      for (int s1 = 0; s1 < 2; s1++) {
        fq[t1] = 2 * q[t1] - 1 + s1;
        fq[t2] = 2 * q[t2] - 1 + s2;
        v[s2][s1] = F.b[d][lrow(F, fq[1], fq[2]) + fq[0]];
      }
    S3[m + 1] = (v[0][0] + v[0][1]) + (v[1][0] + v[1][1]);
  }
  cb[id] = bnd ? S3[1] * 0.25 : ((0.25 * S3[0] + 0.5 * S3[1]) + 0.25 * S3[2]) * 0.25;
}

__global__ __launch_bounds__(256) void k_d2f(const double* __restrict__ src, float* __restrict__ dst, size_t n) {
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t < n) dst[t] = (float)src[t];
}
void launch_d2f(const double* src, float* dst, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_d2f, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, src, dst, n);
}

void launch_mg_coarsen_beta(const Lv& F, const Lv& C, int d, double* cb, hipStream_t st) {
  dim3 g((unsigned)((C.px + 63) / 64), (unsigned)((C.py + 3) / 4), (unsigned)(C.n[2] + 3));
  hipLaunchKernelGGL(k_mg_coarsen_beta, g, dim3(64, 4), 0, st, F, C, d, cb);
}

// Dense (pseudo-)inverse of the coarsest operator in one workgroup: the
// operator from the level's coefficients (the uniform rule, or the face
// coefficient arrays), shifted by c J (c = A00 / n) when singular, its
// Cholesky factor column by column, the inverse column by column by forward
// and back substitution, minus J / (c n^2). Every sum runs in the order of
// the sequential algorithm (dot products in increasing k), with one thread per
// row (factor) or per column (substitution), so the result is that of the
// plain host loops bit for bit. A and out are n x n, row-major.
__device__ void coarsest_coef(const Lv& L, int i, int j, int k, double a6[6], double* diag) {
  const int q[3] = {i, j, k};
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const bool wp = !L.per[d] && (q[d] + 1 == 1 || q[d] + 1 == L.n[d] + 1);
    const bool wm = !L.per[d] && (q[d] == 1 || q[d] == L.n[d] + 1);
    double bp, bm;
    if (L.b[0]) {
      int qp[3] = {i, j, k};
      qp[d] = L.per[d] ? wrapi(q[d] + 1, L.n[d]) : q[d] + 1;
      bp = L.b[d][((size_t)qp[2] * L.py + qp[1]) * L.px + qp[0]];
      bm = L.b[d][((size_t)k * L.py + j) * L.px + i];
    } else {
      bp = wp ? 0.0 : L.beta0;
      bm = wm ? 0.0 : L.beta0;
    }
    a6[2 * d] = bp * L.inv_d2[d];
    a6[2 * d + 1] = bm * L.inv_d2[d];
  }
  *diag = a6[0] + a6[1] + a6[2] + a6[3] + a6[4] + a6[5];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const bool wp = !L.per[d] && (q[d] + 1 == 1 || q[d] + 1 == L.n[d] + 1);
    const bool wm = !L.per[d] && (q[d] == 1 || q[d] == L.n[d] + 1);
    if (wp) a6[2 * d] = 0.;
    if (wm) a6[2 * d + 1] = 0.;
  }
}

constexpr int PINV_LN = 128, PINV_LB = 64;
__device__ __forceinline__ int tri(int i, int k) { return i * (i + 1) / 2 + k; }
// acc - a(k0) b(k0) - a(k0+1) b(k0+1) - ... - a(k1-1) b(k1-1), subtracted in
// that order; the operands are read 16 at a time ahead of the subtractions
template <class FA, class FB>
__device__ __forceinline__ double sub_dot(double acc, int k0, int k1, const FA& a, const FB& b) {
  int k = k0;
  for (; k + 16 <= k1; k += 16) {
    double x[16], y[16];
#pragma unroll
    for (int u = 0; u < 16; u++) {
      x[u] = a(k + u);
      y[u] = b(k + u);
    }
#pragma unroll
    for (int u = 0; u < 16; u++) acc -= x[u] * y[u];
  }
  for (; k < k1; k++) acc -= a(k) * b(k);
  return acc;
}
constexpr int PINV_T = 1024;
__global__ __launch_bounds__(PINV_T) void k_dense_pinv(Lv L, int singular, double* __restrict__ A,
                                                    double* __restrict__ out) {
  const int n = L.n[0] * L.n[1] * L.n[2];
  const int t = (int)threadIdx.x;
  const size_t nn = (size_t)n * n;
  __shared__ double djj_s, cshift_s;
  for (size_t e = t; e < nn; e += PINV_T) A[e] = 0.0;
  __syncthreads();
  for (int row = t; row < n; row += PINV_T) {
    const int i = row % L.n[0] + 1, j = (row / L.n[0]) % L.n[1] + 1, k = row / (L.n[0] * L.n[1]) + 1;
    double a6[6], diag;
    coarsest_coef(L, i, j, k, a6, &diag);
    A[(size_t)row * n + row] += diag;
    const int nb[6][3] = {{i + 1, j, k}, {i - 1, j, k}, {i, j + 1, k}, {i, j - 1, k}, {i, j, k + 1}, {i, j, k - 1}};
    for (int m = 0; m < 6; m++) {
      if (a6[m] == 0.) continue;
      int q[3] = {nb[m][0], nb[m][1], nb[m][2]};
      for (int d = 0; d < 3; d++)
        if (L.per[d]) q[d] = wrapi(q[d], L.n[d]);
      const int col = ((q[2] - 1) * L.n[1] + (q[1] - 1)) * L.n[0] + (q[0] - 1);
      A[(size_t)row * n + col] -= a6[m];
    }
  }
  __syncthreads();
  if (t == 0) cshift_s = singular ? A[0] / (double)n : 0.0;
  __syncthreads();
  const double cshift = cshift_s;
  if (singular) {
    for (size_t e = t; e < nn; e += PINV_T) A[e] += cshift;
    __syncthreads();
  }
  for (int jc = 0; jc < n; jc++) {
    if (t == 0) {
      const double* aj = A + (size_t)jc * n;
      auto lj = [&](int k) { return aj[k]; };
      const double s = sub_dot(aj[jc], 0, jc, lj, lj);
      const double d = sqrt(s);
      A[(size_t)jc * n + jc] = d;
      djj_s = d;
    }
    __syncthreads();
    const double djj = djj_s;
    const double* aj = A + (size_t)jc * n;
    for (int i = jc + 1 + t; i < n; i += PINV_T) {
      double* ai = A + (size_t)i * n;
      const double tt = sub_dot(ai[jc], 0, jc, [&](int k) { return ai[k]; }, [&](int k) { return aj[k]; });
      ai[jc] = tt / djj;
    }
    __syncthreads();
  }
  for (int col = t; col < n; col += PINV_T) {
    // y = the column of the inverse, kept in out's column col
    auto yk = [&](int k) { return out[(size_t)k * n + col]; };
    for (int i = 0; i < n; i++) {
      const double* ai = A + (size_t)i * n;
      const double tt = sub_dot((i == col) ? 1.0 : 0.0, 0, i, [&](int k) { return ai[k]; }, yk);
      out[(size_t)i * n + col] = tt / ai[i];
    }
    for (int i = n - 1; i >= 0; i--) {
      const double tt = sub_dot(out[(size_t)i * n + col], i + 1, n, [&](int k) { return A[(size_t)k * n + i]; }, yk);
      out[(size_t)i * n + col] = tt / A[(size_t)i * n + i];
    }
  }
  if (singular) {
    __syncthreads();
    const double corr = 1.0 / (cshift * (double)n * (double)n);
    for (size_t e = t; e < nn; e += PINV_T) out[e] -= corr;
  }
}
// The same for n <= 128 (a coarsest level of up to 128 cells, rebuilt with
// every change of the coefficients) with the operations on LDS: the packed
// lower triangle (all the factorization reads; the assembled operator is
// symmetric entry for entry) and the solution columns 64 at a time (k-major,
// so the column threads read consecutive words). Same arithmetic, same order.
__global__ __launch_bounds__(PINV_T) void k_dense_pinv_lds(Lv L, int singular, double* __restrict__ out) {
  __shared__ double Ls[PINV_LN * (PINV_LN + 1) / 2];
  __shared__ double Ys[PINV_LN * PINV_LB];
  __shared__ double djj_s, cshift_s;
  const int n = L.n[0] * L.n[1] * L.n[2];
  const int t = (int)threadIdx.x;
  const int ntri = n * (n + 1) / 2;
  for (int e = t; e < ntri; e += PINV_T) Ls[e] = 0.0;
  __syncthreads();
  for (int row = t; row < n; row += PINV_T) {
    const int i = row % L.n[0] + 1, j = (row / L.n[0]) % L.n[1] + 1, k = row / (L.n[0] * L.n[1]) + 1;
    double a6[6], diag;
    coarsest_coef(L, i, j, k, a6, &diag);
    Ls[tri(row, row)] += diag;
    const int nb[6][3] = {{i + 1, j, k}, {i - 1, j, k}, {i, j + 1, k}, {i, j - 1, k}, {i, j, k + 1}, {i, j, k - 1}};
    for (int m = 0; m < 6; m++) {
      if (a6[m] == 0.) continue;
      int q[3] = {nb[m][0], nb[m][1], nb[m][2]};
      for (int d = 0; d < 3; d++)
        if (L.per[d]) q[d] = wrapi(q[d], L.n[d]);
      const int col = ((q[2] - 1) * L.n[1] + (q[1] - 1)) * L.n[0] + (q[0] - 1);
      if (col <= row) Ls[tri(row, col)] -= a6[m];
    }
  }
  __syncthreads();
  if (t == 0) cshift_s = singular ? Ls[0] / (double)n : 0.0;
  __syncthreads();
  const double cshift = cshift_s;
  if (singular) {
    for (int e = t; e < ntri; e += PINV_T) Ls[e] += cshift;
    __syncthreads();
  }
  for (int jc = 0; jc < n; jc++) {
    if (t == 0) {
      const int rj = tri(jc, 0);
      auto lj = [&](int k) { return Ls[rj + k]; };
      const double s = sub_dot(Ls[tri(jc, jc)], 0, jc, lj, lj);
      const double d = sqrt(s);
      Ls[tri(jc, jc)] = d;
      djj_s = d;
    }
    __syncthreads();
    const double djj = djj_s;
    const int rj = tri(jc, 0);
    for (int i = jc + 1 + t; i < n; i += PINV_T) {
      const int ri = tri(i, 0);
      const double tt = sub_dot(Ls[ri + jc], 0, jc, [&](int k) { return Ls[ri + k]; }, [&](int k) { return Ls[rj + k]; });
      Ls[ri + jc] = tt / djj;
    }
    __syncthreads();
  }
  for (int c0 = 0; c0 < n; c0 += PINV_LB) {
    const int cl = t, col = c0 + t;
    if (t < PINV_LB && col < n) {
      auto yk = [&](int k) { return Ys[k * PINV_LB + cl]; };
      for (int i = 0; i < n; i++) {
        const int ri = tri(i, 0);
        const double tt = sub_dot((i == col) ? 1.0 : 0.0, 0, i, [&](int k) { return Ls[ri + k]; }, yk);
        Ys[i * PINV_LB + cl] = tt / Ls[ri + i];
      }
      for (int i = n - 1; i >= 0; i--) {
        const double tt = sub_dot(Ys[i * PINV_LB + cl], i + 1, n, [&](int k) { return Ls[tri(k, i)]; }, yk);
        Ys[i * PINV_LB + cl] = tt / Ls[tri(i, i)];
      }
      for (int i = 0; i < n; i++) out[(size_t)i * n + col] = Ys[i * PINV_LB + cl];
    }
  }
  if (singular) {
    __syncthreads();
    const double corr = 1.0 / (cshift * (double)n * (double)n);
    for (size_t e = t; e < (size_t)n * n; e += PINV_T) out[e] -= corr;
  }
}
void launch_dense_pinv(const Lv& C, int singular, double* work, double* out, hipStream_t st) {
  const int n = C.n[0] * C.n[1] * C.n[2];
  if (n <= PINV_LN)
    hipLaunchKernelGGL(k_dense_pinv_lds, dim3(1), dim3(PINV_T), 0, st, C, singular, out);
  else
    hipLaunchKernelGGL(k_dense_pinv, dim3(1), dim3(PINV_T), 0, st, C, singular, work, out);
}

template <typename T>
void launch_fill(const Lv& L, T* x, hipStream_t st) {
  hipLaunchKernelGGL(k_fill<T>, lrows_grid(L), ROWB, 0, st, L, x);
}

#define SMAC_INST_MG(T)                                                                                 \
  template void launch_mg_rbgs<T>(const Lv&, T*, const T*, int, int, hipStream_t);                     \
  template int mg_zs_division<T>(const Lv&, const T*, const T*);                                         \
  template void launch_mg_restrict<T>(const Lv&, const Lv&, const T*, const T*, typename St<T>::C*, T*, int, \
                                      hipStream_t);                                                    \
  template void launch_mg_prolong<T>(const Lv&, const Lv&, const T*, T*, int, hipStream_t);            \
  template void launch_mg_coarse<T>(const Lv&, const double*, const T*, T*, hipStream_t);              \
  template void launch_fill<T>(const Lv&, T*, hipStream_t);
SMAC_INST_MG(double)
SMAC_INST_MG(float)
SMAC_INST_MG(bf16s)

}  // namespace smac
