// Device-side pieces of the matrix-free 5-point operator
//   (A x)_{j,k} = Di x_{j,k} - ry (Ky_{j,k+1} x_{j,k+1} + Ky_{j,k} x_{j,k-1})
//                            - rx (Kx_{j+1,k} x_{j+1,k} + Kx_{j,k} x_{j-1,k})
//   Di = 1 + ry (Ky_{j,k+1} + Ky_{j,k}) + rx (Kx_{j+1,k} + Kx_{j,k})
// (kernels/tea_leaf_cg_kernel.f90:145-147, Di as kernels/tea_leaf_cheby_kernel.f90:155-157).
// Di is recomputed from the face coefficients instead of being streamed.
#pragma once

#include "tl_internal.hpp"

namespace tl {
namespace dev {

// Geometry passed by value to every streaming kernel.
struct Geo {
  int nx, ny;
  long pitch;
  int tiles_x, rows_per_tile;
  long tiles;
  int blocks;
};

__device__ inline double diag(double kxl, double kxr, double kyd, double kyu, double rx, double ry) {
  return 1.0 + ry * (kyu + kyd) + rx * (kxr + kxl);
}

__device__ inline double apply_a(double c, double l, double r, double d, double u, double kxl,
                                 double kxr, double kyd, double kyu, double rx, double ry) {
  const double di = diag(kxl, kxr, kyd, kyu, rx, ry);
  return di * c - ry * (kyu * u + kyd * d) - rx * (kxr * r + kxl * l);
}

template <typename T>
__device__ inline double ld(const T *__restrict__ a, long i) { return (double)a[i]; }

// Column walk of one thread over rows [y0, y1) of column x: the caller keeps
// a three-row window; these give the index of the row/column neighbours
// under a boundary mode.
template <int MODE>
__device__ inline int xl_of(int x) { return MODE == MIRROR ? (x > 0 ? x - 1 : 0) : x - 1; }
template <int MODE>
__device__ inline int xr_of(int x, int nx) { return MODE == MIRROR ? (x < nx - 1 ? x + 1 : nx - 1) : x + 1; }
template <int MODE>
__device__ inline int yd_of(int y) { return MODE == MIRROR ? (y > 0 ? y - 1 : 0) : y - 1; }
template <int MODE>
__device__ inline int yu_of(int y, int ny) { return MODE == MIRROR ? (y < ny - 1 ? y + 1 : ny - 1) : y + 1; }

// Value of field a at (x, y) under the boundary mode (only ZERO differs from
// a plain load: cells outside the interior read as 0).
template <int MODE, typename T>
__device__ inline double ldm(const T *__restrict__ a, int x, int y, int nx, int ny, long pitch) {
  if (MODE == ZERO && (x < 0 || x >= nx || y < 0 || y >= ny)) return 0.0;
  return (double)a[(long)y * pitch + x];
}

// Walk rows [y0, y1) of column x applying the operator to the field that
// `ldx(xx, yy)` loads, keeping a three-row window in registers.  For every
// row calls body(y, c, l, r, d, u, kxl, kxr, kyd, kyu) with the centre value,
// the four neighbours (under MODE) and the four face coefficients.
template <int MODE, typename KT, typename LD, typename BODY>
__device__ inline void walk(const Geo &g, int x, int y0, int y1, LD ldx,
                            const KT *__restrict__ Kx, const KT *__restrict__ Ky, BODY body) {
  const long P = g.pitch;
  const int xl = xl_of<MODE>(x), xr = xr_of<MODE>(x, g.nx);
  auto get = [&](int xx, int yy) -> double {
    if (MODE == ZERO && (xx < 0 || xx >= g.nx || yy < 0 || yy >= g.ny)) return 0.0;
    return ldx(xx, yy);
  };
  double dn = get(x, yd_of<MODE>(y0));
  double cc = get(x, y0);
  double kyd = (double)Ky[(long)y0 * P + x];
  for (int y = y0; y < y1; ++y) {
    const double up = get(x, yu_of<MODE>(y, g.ny));
    const double kyu = (double)Ky[(long)(y + 1) * P + x];
    const double l = get(xl, y), r = get(xr, y);
    const double kxl = (double)Kx[(long)y * P + x], kxr = (double)Kx[(long)y * P + x + 1];
    body(y, cc, l, r, dn, up, kxl, kxr, kyd, kyu);
    dn = cc;
    cc = up;
    kyd = kyu;
  }
}

// Decompose a tile index.
struct Tile {
  int x, y0, y1;
};
__device__ inline Tile tile_of(const Geo &g, long t) {
  Tile tt;
  const int tx = (int)(t % g.tiles_x), ty = (int)(t / g.tiles_x);
  tt.x = tx * BLOCK + (int)threadIdx.x;
  tt.y0 = ty * g.rows_per_tile;
  tt.y1 = min(tt.y0 + g.rows_per_tile, g.ny);
  return tt;
}

}  // namespace dev
}  // namespace tl
