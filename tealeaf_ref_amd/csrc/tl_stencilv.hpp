// Vector streaming walker for the 5-point operator, generic in the storage
// type T and the number of consecutive columns V a lane owns (FP64: V = 2,
// FP32: V = 4; both give 16-byte loads).  A wave owns 64*V columns and a
// block of 4 waves 256*V columns.  Each value is loaded once: row neighbours
// come from a three-row register window, column neighbours from the lane's
// own vector or the adjacent lane (wave shuffle); only the two cells just
// outside the wave are loaded separately.  Boundaries are MIRROR (see
// tl_internal.hpp).  Arithmetic is done in T: FP64 for the plain solver,
// FP32 for the inner iterations of the mixed-precision solver; reductions
// are always accumulated in FP64 by the callers.
#pragma once

#include "tl_stencil.hpp"

namespace tl {
namespace dev {

template <typename T, int V>
struct Vec {
  typedef T type __attribute__((ext_vector_type(V)));
};

template <bool NT, typename T, int V>
__device__ inline typename Vec<T, V>::type ldv(const T *p) {
  typedef typename Vec<T, V>::type VT;
  if (NT) return __builtin_nontemporal_load((const VT *)p);
  return *(const VT *)p;
}

template <bool NT, typename T>
__device__ inline void st_one(T *p, T v) {
  if (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// Store the elements of the V-vector at column c that lie inside [0, nx).
template <bool NT, typename T, int V>
__device__ inline void stv(T *row, int c, int nx, typename Vec<T, V>::type v) {
  typedef typename Vec<T, V>::type VT;
  if (c + V <= nx) {
    if (NT) __builtin_nontemporal_store(v, (VT *)(row + c));
    else *(VT *)(row + c) = v;
  } else {
#pragma unroll
    for (int e = 0; e < V; ++e)
      if (c + e < nx) st_one<NT>(row + c + e, v[e]);
  }
}

template <typename T>
__device__ inline T apply_t(T c, T l, T r, T d, T u, T kxl, T kxr, T kyd, T kyu, T rx, T ry) {
  const T di = T(1) + ry * (kyu + kyd) + rx * (kxr + kxl);
  return di * c - ry * (kyu * u + kyd * d) - rx * (kxr * r + kxl * l);
}

// The same operator in difference form,
//   (A x)_c = x_c + rx (Kl (x_c - x_l) + Kr (x_c - x_r)) + ry (Kd (x_c - x_d) + Ku (x_c - x_u)),
// used by the mixed-precision solver: the differences of neighbouring values
// are formed first, so rounding is relative to the fluxes instead of to
// Di x_c (no loss proportional to the condition number for smooth x).
template <typename T>
__device__ inline T apply_diff(T c, T l, T r, T d, T u, T kxl, T kxr, T kyd, T kyu, T rx, T ry) {
  return c + rx * (kxl * (c - l) + kxr * (c - r)) + ry * (kyd * (c - d) + kyu * (c - u));
}

template <int V>
struct TileV {
  int x0;  // first column of this wave
  int c;   // first column of this lane
  int y0, y1;
};

template <int V>
__device__ inline TileV<V> tilev_of(const Geo &g, long t) {
  TileV<V> tt;
  const int tx = (int)(t % g.tiles_x), ty = (int)(t / g.tiles_x);
  const int wave = threadIdx.x / WAVE, lane = threadIdx.x & (WAVE - 1);
  tt.x0 = tx * (BLOCK * V) + wave * (WAVE * V);
  tt.c = tt.x0 + V * lane;
  tt.y0 = ty * g.rows_per_tile;
  tt.y1 = min(tt.y0 + g.rows_per_tile, g.ny);
  return tt;
}

// ldv_op(col, row) -> V-vector of operand values; ld1_op(col, row) -> one
// value (only for the two cells outside the wave).  DIFF selects the
// difference form of the operator.  body(y, v, w) gets the
// operand vector and A applied to it.  Lanes wholly outside [0, nx) run the
// loop too (their values feed the shuffles) but must not store.
template <bool NT, typename T, int V, bool DIFF = false, typename LDV, typename LD1,
          typename BODY>
__device__ inline void walkv(const Geo &g, const TileV<V> &tt, const T *__restrict__ Kx,
                             const T *__restrict__ Ky, T rx, T ry, LDV ldv_op, LD1 ld1_op,
                             BODY body) {
  typedef typename Vec<T, V>::type VT;
  const long P = g.pitch;
  const int lane = threadIdx.x & (WAVE - 1);
  const int c = tt.c, nx = g.nx;
  // lanes past the mesh clamp their column so addresses stay in the row
  const int cl = c <= nx ? c : (nx / V) * V;
  const int xr = tt.x0 + WAVE * V;  // first column right of this wave
  const bool need_left = lane == 0 && tt.x0 > 0;
  const bool need_right = lane == WAVE - 1 && xr < nx;
  const bool need_kxr = lane == WAVE - 1 && xr <= nx;
  VT vd = ldv_op(cl, tt.y0 > 0 ? tt.y0 - 1 : tt.y0);
  VT vc = ldv_op(cl, tt.y0);
  VT kyd = ldv<NT, T, V>(Ky + (long)tt.y0 * P + cl);
  for (int y = tt.y0; y < tt.y1; ++y) {
    const bool last_row = y + 1 >= g.ny;
    const VT vu = last_row ? vc : ldv_op(cl, y + 1);
    const VT kyu = ldv<NT, T, V>(Ky + (long)(y + 1) * P + cl);
    const VT kx = ldv<NT, T, V>(Kx + (long)y * P + cl);
    T left = __shfl_up(vc[V - 1], 1, WAVE);
    T right = __shfl_down(vc[0], 1, WAVE);
    T kxr = __shfl_down(kx[0], 1, WAVE);
    if (need_left) left = ld1_op(tt.x0 - 1, y);
    if (need_right) right = ld1_op(xr, y);
    if (need_kxr) kxr = Kx[(long)y * P + xr];
    if (c == 0) left = vc[0];  // mirror at the left edge
    VT w;
#pragma unroll
    for (int e = 0; e < V; ++e) {
      const int col = c + e;
      const T l = (e == 0) ? left : vc[e - 1];
      T r;
      if (e + 1 < V) r = (col + 1 < nx) ? vc[e + 1] : vc[e];
      else r = (col + 1 < nx) ? right : vc[e];
      const T kl = kx[e];
      const T kr = (e + 1 < V) ? kx[e + 1] : kxr;
      w[e] = DIFF ? apply_diff<T>(vc[e], l, r, vd[e], vu[e], kl, kr, kyd[e], kyu[e], rx, ry)
                  : apply_t<T>(vc[e], l, r, vd[e], vu[e], kl, kr, kyd[e], kyu[e], rx, ry);
    }
    body(y, vc, w);
    vd = vc;
    vc = vu;
    kyd = kyu;
  }
}

}  // namespace dev
}  // namespace tl
