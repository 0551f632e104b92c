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

#include <type_traits>

#include "tl_stencil.hpp"

namespace tl {
namespace dev {

template <typename T, int V>
struct Vec {
  typedef T type __attribute__((ext_vector_type(V)));
};

#ifndef TL_NT_LOADS
#define TL_NT_LOADS 1
#endif

template <bool NT, typename T, int V>
__device__ inline typename Vec<T, V>::type ldv(const T *p) {
  typedef typename Vec<T, V>::type VT;
  if (NT && TL_NT_LOADS) return __builtin_nontemporal_load((const VT *)p);
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

// Per-row data a walk body needs besides the operand (none).
struct NoExtra {};

// The walk's value per cell: A applied to the operand (the default), or a
// cell rule cell(extra, e, c, l, r, d, u, kl, kr, kd, ku) of the extra data of
// the row, the column within the lane's vector, the operand at the cell and
// its four neighbours, and the four face coefficients.
struct ApplyA {};

// ldv_op(col, row) -> V-vector of operand values; ld1_op(col, row) -> one
// value (only for the two cells outside the wave); extra(row) -> the other
// per-row loads of the body.  body(y, v, w, e) gets the operand vector, A
// applied to it and extra(y).  DIFF selects the difference form of the
// operator.  With PF = 1 every load of row y+1's step (operand row y+2,
// faces, edge cells, extra) is issued before row y is computed, so one row
// of loads is always in flight behind the arithmetic; with PF = d > 1 those
// of the d steps ahead (d rows in flight, in more registers).  The body
// itself must not load.
// dir = -1 walks the rows from the top of the tile down (a pass that follows
// one which ended at the top then reads the freshest rows first).
// Lanes wholly outside [0, nx) run the loop too (their values feed the
// shuffles) but must not store.
// With LDSX the cells just outside a wave that belong to another wave of the
// same block are exchanged through LDS (one barrier per row) instead of being
// loaded again; only the block's outer edges are loaded.  All waves of the
// block must walk the same rows.
// Outside the mesh the operand is mirrored (EDGE = 0), zero (EDGE = 1) or
// read from the halo cells as they are in memory (EDGE = 2).
template <bool NT, typename T, int V, bool DIFF, int PF, bool LDSX = false, int EDGE = 0,
          typename LDV, typename LD1, typename EXTRA, typename BODY, typename CELL = ApplyA>
__device__ inline void walkv(const Geo &g, const TileV<V> &tt, const T *__restrict__ Kx,
                             const T *__restrict__ Ky, T rx, T ry, LDV ldv_op, LD1 ld1_op,
                             EXTRA extra, BODY body, int dir = 1, CELL cell = CELL()) {
  typedef typename Vec<T, V>::type VT;
  typedef decltype(extra(0)) E;
  constexpr int NW = BLOCK / WAVE;
  __shared__ T xch[2][NW][3];  // per row parity and wave: first value, last value, first Kx
  const long P = g.pitch;
  const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
  const int c = tt.c, nx = g.nx, ny = g.ny;
  const int count = tt.y1 - tt.y0;
  if (count <= 0) return;
  const bool fwd = dir > 0;
  const int ys = fwd ? tt.y0 : tt.y1 - 1;  // first row walked
  // lanes past the mesh clamp their column so addresses stay in the row
  const int cl = c <= nx ? c : (nx / V) * V;
  const int xr = tt.x0 + WAVE * V;  // first column right of this wave
  const bool edge_l = !LDSX || wave == 0, edge_r = !LDSX || wave == NW - 1;
  const bool need_left = lane == 0 && (tt.x0 > 0 || EDGE == 2) && edge_l;
  const bool need_right = lane == WAVE - 1 && (xr < nx || (EDGE == 2 && xr == nx)) && edge_r;
  const bool need_kxr = lane == WAVE - 1 && xr <= nx && edge_r;
  // Ky face between row y and the next row in walking direction
  auto face_ahead = [&](int y) -> long { return (long)(fwd ? y + 1 : y) * P + cl; };
  // state of row y: operand rows behind (y - dir), y and ahead (y + dir),
  // mirrored at the mesh edge; the Ky faces on both sides; Kx(y); the cells
  // left and right of the wave; Kx right of the wave; the extra data
  VT vc = ldv_op(cl, ys);
  const int yb = ys - dir, ya = ys + dir;
  VT vb = (EDGE == 2 || (yb >= 0 && yb < ny)) ? ldv_op(cl, yb) : (EDGE == 1 ? VT(0) : vc);
  VT va = (EDGE == 2 || (ya >= 0 && ya < ny)) ? ldv_op(cl, ya) : (EDGE == 1 ? VT(0) : vc);
  VT kb = ldv<NT, T, V>(Ky + (long)(fwd ? ys : ys + 1) * P + cl);
  VT ka = ldv<NT, T, V>(Ky + face_ahead(ys));
  VT kx = ldv<NT, T, V>(Kx + (long)ys * P + cl);
  T el = need_left ? ld1_op(tt.x0 - 1, ys) : T(0);
  T er = need_right ? ld1_op(xr, ys) : T(0);
  T ek = need_kxr ? Kx[(long)ys * P + xr] : T(0);
  E ex = extra(ys);
  // the loads that iteration j takes over when it becomes current: the
  // operand row beyond y_j (`in` false where it lies outside the mesh and
  // is resolved on taking over, from rows already loaded), the Ky face ahead
  // of y_j, Kx, the edge cells and the extra data of y_j
  struct Stage {
    VT v, ka, kx;
    T el, er, ek;
    E ex;
    bool in;
  };
  auto load_stage = [&](int j, Stage &o) {
    const int yj = ys + j * dir, yo = yj + dir;
    o.in = EDGE == 2 || (yo >= 0 && yo < ny);
    o.v = o.in ? ldv_op(cl, yo) : VT(0);
    o.ka = ldv<NT, T, V>(Ky + face_ahead(yj));
    o.kx = ldv<NT, T, V>(Kx + (long)yj * P + cl);
    o.el = need_left ? ld1_op(tt.x0 - 1, yj) : T(0);
    o.er = need_right ? ld1_op(xr, yj) : T(0);
    o.ek = need_kxr ? Kx[(long)yj * P + xr] : T(0);
    o.ex = extra(yj);
  };
  // q[i]: iteration k + 1 + i, the last one loaded during iteration k
  constexpr int D = PF > 0 ? PF : 1;
  Stage q[D];
#pragma unroll
  for (int i = 0; i + 1 < D; ++i)
    if (1 + i < count) load_stage(1 + i, q[i]);
  for (int k = 0; k < count; ++k) {
    const int y = ys + k * dir;
    const bool more = k + 1 < count;
    if (PF > 0 && k + PF < count) load_stage(k + PF, q[D - 1]);
    T left = from_lower_lane(vc[V - 1]);
    T right = from_upper_lane(vc[0]);
    T kxr = from_upper_lane(kx[0]);
    if (LDSX) {
      const int par = k & 1;
      if (lane == 0) {
        xch[par][wave][0] = vc[0];
        xch[par][wave][2] = kx[0];
      }
      if (lane == WAVE - 1) xch[par][wave][1] = vc[V - 1];
      __syncthreads();
      if (lane == 0 && wave > 0) left = xch[par][wave - 1][1];
      if (lane == WAVE - 1 && wave < NW - 1) {
        right = xch[par][wave + 1][0];
        kxr = xch[par][wave + 1][2];
      }
    }
    if (need_left) left = el;
    if (need_right) right = er;
    if (need_kxr) kxr = ek;
    if (c == 0 && EDGE != 2) left = EDGE == 1 ? T(0) : vc[0];  // the left edge
    const VT &vd = fwd ? vb : va, &vu = fwd ? va : vb;
    const VT &kyd = fwd ? kb : ka, &kyu = fwd ? ka : kb;
    VT w;
#pragma unroll
    for (int e = 0; e < V; ++e) {
      const int col = c + e;
      const T l = (e == 0) ? left : vc[e - 1];
      T r;
      if (e + 1 < V) r = (col + 1 < nx || EDGE == 2) ? vc[e + 1] : (EDGE == 1 ? T(0) : vc[e]);
      else r = (col + 1 < nx || EDGE == 2) ? right : (EDGE == 1 ? T(0) : vc[e]);
      const T kl = kx[e];
      const T kr = (e + 1 < V) ? kx[e + 1] : kxr;
      if constexpr (std::is_same<CELL, ApplyA>::value)
        w[e] = DIFF ? apply_diff<T>(vc[e], l, r, vd[e], vu[e], kl, kr, kyd[e], kyu[e], rx, ry)
                    : apply_t<T>(vc[e], l, r, vd[e], vu[e], kl, kr, kyd[e], kyu[e], rx, ry);
      else
        w[e] = cell(ex, e, vc[e], l, r, vd[e], vu[e], kl, kr, kyd[e], kyu[e]);
    }
    body(y, vc, w, ex);
    if (!more) break;
    if (PF == 0) load_stage(k + 1, q[0]);
    // iteration k + 1 takes over: outside the mesh the operand row is the
    // mirror of the row at the edge (EDGE = 0) or zero (EDGE = 1)
    const VT nv = q[0].in ? q[0].v : (EDGE == 1 ? VT(0) : va);
    vb = vc;
    vc = va;
    va = nv;
    kb = ka;
    ka = q[0].ka;
    kx = q[0].kx;
    el = q[0].el;
    er = q[0].er;
    ek = q[0].ek;
    ex = q[0].ex;
#pragma unroll
    for (int i = 0; i + 1 < D; ++i) q[i] = q[i + 1];
  }
}

}  // namespace dev
}  // namespace tl
