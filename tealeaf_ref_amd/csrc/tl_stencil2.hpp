// Two-columns-per-lane streaming walker for the 5-point operator.
//
// A wave owns 128 consecutive columns (lane l: columns x0+2l and x0+2l+1,
// one 16-byte load per field per row); a block of 4 waves owns 512 columns
// and `rows_per_tile` rows.  Every interior value is loaded once: the row
// neighbours come from the register window, the column neighbours from the
// adjacent lane (wave shuffle), and only the two cells just outside the
// wave's 128 columns are loaded separately.  Large meshes use nontemporal
// loads and stores (streamed through the caches): measured on MI355X they
// lift a 4-in/2-out FP64 stream from ~5.2 to ~5.85 TB/s.
#pragma once

#include "tl_stencil.hpp"

namespace tl {
namespace dev {

typedef double d2 __attribute__((ext_vector_type(2)));

constexpr int COLS_PER_WAVE = 2 * WAVE;
constexpr int COLS_PER_BLOCK = 2 * BLOCK;

template <bool NT>
__device__ inline d2 ld2(const double *p) {
  if (NT) return __builtin_nontemporal_load((const d2 *)p);
  return *(const d2 *)p;
}
template <bool NT>
__device__ inline double ld1(const double *p) {
  if (NT) return __builtin_nontemporal_load(p);
  return *p;
}
template <bool NT>
__device__ inline void st2(double *p, d2 v) {
  if (NT) __builtin_nontemporal_store(v, (d2 *)p);
  else *(d2 *)p = v;
}
template <bool NT>
__device__ inline void st1(double *p, double v) {
  if (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// Store the valid elements of a pair at column c (c even).
template <bool NT>
__device__ inline void st_pair(double *row, int c, int nx, d2 v) {
  if (c + 1 < nx) st2<NT>(row + c, v);
  else if (c < nx) st1<NT>(row + c, v.x);
}

struct Tile2 {
  int x0;      // first column of this wave
  int c;       // first column of this lane's pair
  int y0, y1;
};

__device__ inline Tile2 tile2_of(const Geo &g, long t) {
  Tile2 tt;
  const int tx = (int)(t % g.tiles_x), ty = (int)(t / g.tiles_x);
  const int wave = threadIdx.x / WAVE, lane = threadIdx.x & (WAVE - 1);
  tt.x0 = tx * COLS_PER_BLOCK + wave * COLS_PER_WAVE;
  tt.c = tt.x0 + 2 * lane;
  tt.y0 = ty * g.rows_per_tile;
  tt.y1 = min(tt.y0 + g.rows_per_tile, g.ny);
  return tt;
}

// Walk rows [y0, y1) of this lane's column pair under MIRROR boundaries.
// `ldp(col, row)` returns the operand pair (columns col, col+1) of a row and
// `ld1p(col, row)` one operand value; body(y, c, v, w) receives the pair of
// operand values v and the pair of A v.  Lanes whose pair lies wholly
// outside [0, nx) still run the loop (their values feed the shuffles) but
// must not store.
template <bool NT, typename LD2, typename LD1, typename BODY>
__device__ inline void walk2(const Geo &g, const Tile2 &tt, const double *__restrict__ Kx,
                             const double *__restrict__ Ky, double rx, double ry, LD2 ldp,
                             LD1 ld1p, BODY body) {
  const long P = g.pitch;
  const int lane = threadIdx.x & (WAVE - 1);
  const int c = tt.c, nx = g.nx;
  // pairs beyond the last needed column load nothing real; clamp the column
  // so their addresses stay inside the row
  const int cl = c <= nx ? c : (nx & ~1);
  const int xr = tt.x0 + COLS_PER_WAVE;  // first column right of this wave
  const bool need_left = lane == 0 && tt.x0 > 0;
  const bool need_right = lane == WAVE - 1 && xr < nx;
  const bool need_kxr = lane == WAVE - 1 && xr <= nx;
  d2 vd = ldp(cl, tt.y0 > 0 ? tt.y0 - 1 : tt.y0);
  d2 vc = ldp(cl, tt.y0);
  d2 kyd = ld2<NT>(Ky + (long)tt.y0 * P + cl);
  for (int y = tt.y0; y < tt.y1; ++y) {
    const bool last_row = y + 1 >= g.ny;
    const d2 vu = last_row ? vc : ldp(cl, y + 1);
    const d2 kyu = ld2<NT>(Ky + (long)(y + 1) * P + cl);
    const d2 kx = ld2<NT>(Kx + (long)y * P + cl);
    // column neighbours of the pair from the adjacent lanes
    double left = from_lower_lane(vc.y);
    double right = from_upper_lane(vc.x);
    double kxr = from_upper_lane(kx.x);
    if (need_left) left = ld1p(tt.x0 - 1, y);
    if (need_right) right = ld1p(xr, y);
    if (need_kxr) kxr = ld1<NT>(Kx + (long)y * P + xr);
    if (c == 0) left = vc.x;                        // mirror at the left edge
    const double r0 = (c + 1 < nx) ? vc.y : vc.x;   // right neighbour of column c
    const double r1 = (c + 2 < nx) ? right : vc.y;  // right neighbour of column c+1
    d2 w;
    w.x = apply_a(vc.x, left, r0, vd.x, vu.x, kx.x, kx.y, kyd.x, kyu.x, rx, ry);
    w.y = apply_a(vc.y, vc.x, r1, vd.y, vu.y, kx.y, kxr, kyd.y, kyu.y, rx, ry);
    body(y, vc, w);
    vd = vc;
    vc = vu;
    kyd = kyu;
  }
}

}  // namespace dev
}  // namespace tl

