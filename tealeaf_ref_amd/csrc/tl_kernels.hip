// Per-routine entry points of libtealeaf_hip.so: one C function per
// reference driver wrapper (tea_leaf_common.f90, tea_leaf_cg.f90,
// tea_leaf_cheby.f90, tea_leaf_ppcg.f90, tea_leaf_jacobi.f90, update_halo.f90,
// field_summary.f90).  Scalars come back synchronously, as the reference
// wrappers return them.
#include <cmath>
#include <cstdlib>

#include "tl_internal.hpp"
#include "tl_stencil.hpp"

namespace tl {
dev::Geo geo(int cols_per_tile = BLOCK);

dev::Geo geo(int cols_per_tile) {
  State &s = S();
  // the paired (512-column) kernels run best with ~1024 tall tiles: fewer
  // concurrent row streams keep HBM pages open (measured on MI355X; at
  // 8192^2, 1024 tiles of 128 rows made a CG step 5 % and a PPCG step 4 %
  // faster than 2048 of 64 rows, and 256 rows gained nothing more)
  static int target = -1, maxrows = -1, remap = -1;
  if (target < 0) {
    const char *e = getenv("TL_TILES");
    target = e ? atoi(e) : 1024;
    e = getenv("TL_MAXROWS");
    maxrows = e ? atoi(e) : 128;
    e = getenv("TL_XCD_REMAP");
    remap = e ? atoi(e) : 1;
  }
  const Grid g = cols_per_tile > BLOCK ? grid_for(s.nx, s.ny, cols_per_tile, target, maxrows)
                                       : grid_for(s.nx, s.ny, cols_per_tile);
  dev::Geo G;
  G.nx = s.nx;
  G.ny = s.ny;
  G.pitch = s.pitch;
  G.tiles_x = g.tiles_x;
  G.rows_per_tile = g.rows_per_tile;
  G.tiles = g.tiles;
  G.blocks = g.blocks;
  // a negative count tells the kernels to walk tiles in plain block order
  if (cols_per_tile > BLOCK && !remap) G.blocks = -g.blocks;
  return G;
}

// Tiles of cols_per_tile columns, as many as blocks_per_cu blocks on every
// CU (all resident at once for a kernel that fits blocks_per_cu blocks per
// CU): rows per tile rounded up, so the count does not pass that number.
dev::Geo geo_resident(int cols_per_tile, int blocks_per_cu) {
  State &s = S();
  const long want = (long)blocks_per_cu * (s.cus > 0 ? s.cus : 256);
  Grid g;
  g.tiles_x = (s.nx + cols_per_tile - 1) / cols_per_tile;
  long r = ((long)g.tiles_x * s.ny + want - 1) / want;
  if (r < 1) r = 1;
  if (r > 128) r = 128;
  g.rows_per_tile = (int)r;
  g.tiles_y = (s.ny + g.rows_per_tile - 1) / g.rows_per_tile;
  g.tiles = (long)g.tiles_x * g.tiles_y;
  long b = g.tiles < MAX_GRID ? g.tiles : MAX_GRID;
  if (b >= NUM_XCD) b -= b % NUM_XCD;
  dev::Geo G;
  G.nx = s.nx;
  G.ny = s.ny;
  G.pitch = s.pitch;
  G.tiles_x = g.tiles_x;
  G.rows_per_tile = g.rows_per_tile;
  G.tiles = g.tiles;
  G.blocks = (int)b;
  return G;
}

}  // namespace tl

using namespace tl;
using dev::Geo;

#define TILE_LOOP(G)                                     \
  dev::TileIter it__((G).tiles, (G).blocks);             \
  for (long t__ = it__.t; t__ < it__.t_end; ++t__) {     \
    const dev::Tile tt = dev::tile_of((G), t__);         \
    if (tt.x < (G).nx) {                                 \
      const int x = tt.x;
#define TILE_LOOP_END \
  }                   \
  }

#define IDX(y, x) ((long)(y) * G.pitch + (x))

// ---------------------------------------------------------------------------
// pointwise / reduction kernels
// ---------------------------------------------------------------------------
// kernels/tea_leaf_common_kernel.f90:67-72
__global__ __launch_bounds__(BLOCK) void k_init_u(Geo G, const double *__restrict__ e1,
                                                   const double *__restrict__ rho,
                                                   double *__restrict__ u, double *__restrict__ u0) {
  TILE_LOOP(G)
  for (int y = tt.y0; y < tt.y1; ++y) {
    const long i = IDX(y, x);
    const double v = e1[i] * rho[i];
    u[i] = v;
    u0[i] = v;
  }
  TILE_LOOP_END
}

// kernels/tea_leaf_common_kernel.f90:75-141 over the haloed range
__global__ __launch_bounds__(BLOCK) void k_init_k(int nx, int ny, int h, long pitch,
                                                   const double *__restrict__ rho,
                                                   double *__restrict__ Kx, double *__restrict__ Ky,
                                                   int recip, int reflective, int4 zb) {
  const int w = nx + 2 * h - 1, hh = ny + 2 * h - 1;
  const long n = (long)w * hh;
  for (long i = blockIdx.x * (long)BLOCK + threadIdx.x; i < n; i += (long)gridDim.x * BLOCK) {
    const int x = 1 - h + (int)(i % w), y = 1 - h + (int)(i / w);
    const long c = (long)y * pitch + x;
    const double rc = rho[c], rl = rho[c - 1], rd = rho[c - pitch];
    const double wc = recip ? 1.0 / rc : rc, wl = recip ? 1.0 / rl : rl, wd = recip ? 1.0 / rd : rd;
    double kx = (wl + wc) / (2.0 * wl * wc);
    double ky = (wd + wc) / (2.0 * wd * wc);
    if (!reflective) {
      if ((zb.x && x <= 0) || (zb.y && x >= nx)) kx = 0.0;
      if ((zb.z && y <= 0) || (zb.w && y >= ny)) ky = 0.0;
    }
    Kx[c] = kx;
    Ky[c] = ky;
  }
}

// kernels/tea_leaf_common_kernel.f90:144-152 and tea_diag_init :294-305
__global__ __launch_bounds__(BLOCK) void k_init_di(int nx, int ny, int h, long pitch,
                                                    const double *__restrict__ Kx,
                                                    const double *__restrict__ Ky,
                                                    double *__restrict__ Di, double *__restrict__ Mi,
                                                    double rx, double ry, int diag_prec) {
  const int w = nx + 2 * h - 2, hh = ny + 2 * h - 2;
  const long n = (long)w * hh;
  for (long i = blockIdx.x * (long)BLOCK + threadIdx.x; i < n; i += (long)gridDim.x * BLOCK) {
    const int x = 1 - h + (int)(i % w), y = 1 - h + (int)(i / w);
    const long c = (long)y * pitch + x;
    const double di = (1.0 + ry * (Ky[c + pitch] + Ky[c]) + rx * (Kx[c + 1] + Kx[c]));
    Di[c] = di;
    if (diag_prec) Mi[c] = (di != 0.0) ? 1.0 / di : 0.0;
  }
}

// tea_block_init, kernels/tea_leaf_common_kernel.f90:337-376 (4-row blocks)
__global__ __launch_bounds__(BLOCK) void k_block_init(int nx, int ny, long pitch,
                                                       const double *__restrict__ Ky,
                                                       const double *__restrict__ Di,
                                                       double *__restrict__ cp, double *__restrict__ bfp,
                                                       double ry) {
  const int nb = (ny + 3) / 4;
  const long n = (long)nx * nb;
  for (long i = blockIdx.x * (long)BLOCK + threadIdx.x; i < n; i += (long)gridDim.x * BLOCK) {
    const int x = (int)(i % nx), b = (int)(i / nx);
    const int bottom = 4 * b, top = min(bottom + 3, ny - 1);
    long c = (long)bottom * pitch + x;
    cp[c] = (-Ky[c + pitch] * ry) / Di[c];
    for (int y = bottom + 1; y <= top; ++y) {
      c = (long)y * pitch + x;
      bfp[c] = 1.0 / (Di[c] - (-Ky[c] * ry) * cp[c - pitch]);
      cp[c] = (-Ky[c + pitch] * ry) * bfp[c];
    }
  }
}

// tea_block_solve, kernels/tea_leaf_common_kernel.f90:378-468: z = M^-1 r
__global__ __launch_bounds__(BLOCK) void k_block_solve(int nx, int ny, long pitch,
                                                        const double *__restrict__ r,
                                                        double *__restrict__ z,
                                                        const double *__restrict__ cp,
                                                        const double *__restrict__ bfp,
                                                        const double *__restrict__ Ky,
                                                        const double *__restrict__ Di, double ry) {
  const int nb = (ny + 3) / 4;
  const long n = (long)nx * nb;
  for (long i = blockIdx.x * (long)BLOCK + threadIdx.x; i < n; i += (long)gridDim.x * BLOCK) {
    const int x = (int)(i % nx), b = (int)(i / nx);
    const int bottom = 4 * b, top = min(bottom + 3, ny - 1);
    double dp[4], zl[4];
    long c = (long)bottom * pitch + x;
    dp[0] = r[c] / Di[c];
    for (int y = bottom + 1; y <= top; ++y) {
      c = (long)y * pitch + x;
      dp[y - bottom] = (r[c] - (-Ky[c] * ry) * dp[y - bottom - 1]) * bfp[c];
    }
    zl[top - bottom] = dp[top - bottom];
    for (int y = top - 1; y >= bottom; --y)
      zl[y - bottom] = dp[y - bottom] - cp[(long)y * pitch + x] * zl[y - bottom + 1];
    for (int y = bottom; y <= top; ++y) z[(long)y * pitch + x] = zl[y - bottom];
  }
}

// w = A u ; r = u - w   (kernels/tea_leaf_common_kernel.f90:162-174)
__global__ __launch_bounds__(BLOCK) void k_init_wr(Geo G, const double *__restrict__ u,
                                                    const double *__restrict__ Kx,
                                                    const double *__restrict__ Ky,
                                                    double *__restrict__ w, double *__restrict__ r,
                                                    double rx, double ry) {
  TILE_LOOP(G)
  dev::walk<PHYS>(G, x, tt.y0, tt.y1, [&](int xx, int yy) { return u[IDX(yy, xx)]; }, Kx, Ky,
                  [&](int y, double c, double l, double rr, double d, double up, double kxl,
                      double kxr, double kyd, double kyu) {
                    const double wv = dev::apply_a(c, l, rr, d, up, kxl, kxr, kyd, kyu, rx, ry);
                    w[IDX(y, x)] = wv;
                    r[IDX(y, x)] = c - wv;
                  });
  TILE_LOOP_END
}

// r = u0 - A u   (kernels/tea_leaf_common_kernel.f90:233-240)
__global__ __launch_bounds__(BLOCK) void k_residual(Geo G, const double *__restrict__ u,
                                                     const double *__restrict__ u0,
                                                     const double *__restrict__ Kx,
                                                     const double *__restrict__ Ky,
                                                     double *__restrict__ r, double rx, double ry) {
  TILE_LOOP(G)
  dev::walk<PHYS>(G, x, tt.y0, tt.y1, [&](int xx, int yy) { return u[IDX(yy, xx)]; }, Kx, Ky,
                  [&](int y, double c, double l, double rr, double d, double up, double kxl,
                      double kxr, double kyd, double kyu) {
                    r[IDX(y, x)] = u0[IDX(y, x)] -
                                   dev::apply_a(c, l, rr, d, up, kxl, kxr, kyd, kyu, rx, ry);
                  });
  TILE_LOOP_END
}

// sum a*b over the interior into partial slot `out` (2-norm, zrnorm, dots)
__global__ __launch_bounds__(BLOCK) void k_dot(Geo G, const double *__restrict__ a,
                                                const double *__restrict__ b,
                                                double *__restrict__ out) {
  __shared__ double lds[BLOCK / WAVE];
  double acc = 0.0;
  TILE_LOOP(G)
  for (int y = tt.y0; y < tt.y1; ++y) acc += a[IDX(y, x)] * b[IDX(y, x)];
  TILE_LOOP_END
  acc = dev::block_sum(acc, lds);
  if (threadIdx.x == 0) out[blockIdx.x] = acc;
}

// rrn = sum (r - r_store) z   (kernels/tea_leaf_ppcg_kernel.f90:553-557)
__global__ __launch_bounds__(BLOCK) void k_rrn(Geo G, const double *__restrict__ r,
                                                const double *__restrict__ rs,
                                                const double *__restrict__ z,
                                                double *__restrict__ out) {
  __shared__ double lds[BLOCK / WAVE];
  double acc = 0.0;
  TILE_LOOP(G)
  for (int y = tt.y0; y < tt.y1; ++y) acc += (r[IDX(y, x)] - rs[IDX(y, x)]) * z[IDX(y, x)];
  TILE_LOOP_END
  acc = dev::block_sum(acc, lds);
  if (threadIdx.x == 0) out[blockIdx.x] = acc;
}

// dst = a / b   (finalise: energy1 = u / density, kernels/tea_leaf_common_kernel.f90:198-202)
__global__ __launch_bounds__(BLOCK) void k_div(Geo G, const double *__restrict__ a,
                                                const double *__restrict__ b,
                                                double *__restrict__ dst) {
  TILE_LOOP(G)
  for (int y = tt.y0; y < tt.y1; ++y) dst[IDX(y, x)] = a[IDX(y, x)] / b[IDX(y, x)];
  TILE_LOOP_END
}

// dst = src on the interior
__global__ __launch_bounds__(BLOCK) void k_copy(Geo G, const double *__restrict__ src,
                                                 double *__restrict__ dst) {
  TILE_LOOP(G)
  for (int y = tt.y0; y < tt.y1; ++y) dst[IDX(y, x)] = src[IDX(y, x)];
  TILE_LOOP_END
}

// field summary, kernels/tea_leaf_common_kernel.f90 -> field_summary_kernel.f90:51-60
__global__ __launch_bounds__(BLOCK) void k_summary(Geo G, const double *__restrict__ vol,
                                                    const double *__restrict__ rho,
                                                    const double *__restrict__ e1,
                                                    const double *__restrict__ u,
                                                    double *__restrict__ part) {
  __shared__ double lds[BLOCK / WAVE];
  double sv = 0.0, sm = 0.0, se = 0.0, st = 0.0;
  TILE_LOOP(G)
  for (int y = tt.y0; y < tt.y1; ++y) {
    const long i = IDX(y, x);
    const double cv = vol[i], cm = cv * rho[i];
    sv += cv;
    sm += cm;
    se += cm * e1[i];
    st += cm * u[i];
  }
  TILE_LOOP_END
  sv = dev::block_sum(sv, lds);
  sm = dev::block_sum(sm, lds);
  se = dev::block_sum(se, lds);
  st = dev::block_sum(st, lds);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = sv;
    part[MAX_GRID + blockIdx.x] = sm;
    part[2 * MAX_GRID + blockIdx.x] = se;
    part[3 * MAX_GRID + blockIdx.x] = st;
  }
}

// ---------------------------------------------------------------------------
// reflective halo (kernels/update_halo_kernel.f90:136-195): left/right columns
// first, then bottom/top rows including the corners just written.
// ---------------------------------------------------------------------------
struct FieldList {
  double *f[11];
  int n;
};

__global__ __launch_bounds__(BLOCK) void k_halo_lr(FieldList fl, int nx, int ny, long pitch,
                                                    int depth) {
  const int rows = ny + 2 * depth;
  const long n = (long)fl.n * rows * depth;
  for (long i = blockIdx.x * (long)BLOCK + threadIdx.x; i < n; i += (long)gridDim.x * BLOCK) {
    const int fi = (int)(i / ((long)rows * depth));
    const int rem = (int)(i % ((long)rows * depth));
    const int y = -depth + rem / depth, j = rem % depth + 1;  // j = 1..depth
    double *m = fl.f[fi] + (long)y * pitch;
    m[-j] = m[j - 1];              // mesh(1-j,k) = mesh(0+j,k)
    m[nx - 1 + j] = m[nx - j];     // mesh(x_max+j,k) = mesh(x_max+1-j,k)
  }
}

__global__ __launch_bounds__(BLOCK) void k_halo_bt(FieldList fl, int nx, int ny, long pitch,
                                                    int depth) {
  const int cols = nx + 2 * depth;
  const long n = (long)fl.n * cols * depth;
  for (long i = blockIdx.x * (long)BLOCK + threadIdx.x; i < n; i += (long)gridDim.x * BLOCK) {
    const int fi = (int)(i / ((long)cols * depth));
    const int rem = (int)(i % ((long)cols * depth));
    const int x = -depth + rem % cols, kk = rem / cols + 1;  // kk = 1..depth
    double *m = fl.f[fi] + x;
    m[(long)(-kk) * pitch] = m[(long)(kk - 1) * pitch];
    m[(long)(ny - 1 + kk) * pitch] = m[(long)(ny - kk) * pitch];
  }
}

// ---------------------------------------------------------------------------
// CG kernels (kernels/tea_leaf_cg_kernel.f90)
// ---------------------------------------------------------------------------
// p = z = 0 ; z = M^-1 r (diag inline; block precomputed) ; p = z or r ; rro
__global__ __launch_bounds__(BLOCK) void k_cg_init(Geo G, const double *__restrict__ r,
                                                    double *__restrict__ p, double *__restrict__ z,
                                                    const double *__restrict__ Mi, int precond,
                                                    double *__restrict__ out) {
  __shared__ double lds[BLOCK / WAVE];
  double acc = 0.0;
  TILE_LOOP(G)
  for (int y = tt.y0; y < tt.y1; ++y) {
    const long i = IDX(y, x);
    double pv;
    if (precond == TL_PREC_NONE) {
      z[i] = 0.0;
      pv = r[i];
    } else if (precond == TL_PREC_JAC_DIAG) {
      pv = Mi[i] * r[i];
      z[i] = pv;
    } else {
      pv = z[i];
    }
    p[i] = pv;
    acc += r[i] * pv;
  }
  TILE_LOOP_END
  acc = dev::block_sum(acc, lds);
  if (threadIdx.x == 0) out[blockIdx.x] = acc;
}

// w = A p ; pw = sum w p   (kernels/tea_leaf_cg_kernel.f90:143-151)
__global__ __launch_bounds__(BLOCK) void k_cg_calc_w(Geo G, const double *__restrict__ p,
                                                      const double *__restrict__ Kx,
                                                      const double *__restrict__ Ky,
                                                      double *__restrict__ w, double rx, double ry,
                                                      double *__restrict__ out) {
  __shared__ double lds[BLOCK / WAVE];
  double acc = 0.0;
  TILE_LOOP(G)
  dev::walk<MIRROR>(G, x, tt.y0, tt.y1, [&](int xx, int yy) { return p[IDX(yy, xx)]; }, Kx, Ky,
                    [&](int y, double c, double l, double rr, double d, double up, double kxl,
                        double kxr, double kyd, double kyu) {
                      const double wv = dev::apply_a(c, l, rr, d, up, kxl, kxr, kyd, kyu, rx, ry);
                      w[IDX(y, x)] = wv;
                      acc += wv * c;
                    });
  TILE_LOOP_END
  acc = dev::block_sum(acc, lds);
  if (threadIdx.x == 0) out[blockIdx.x] = acc;
}

// u += alpha p ; r -= alpha w ; rrn = r.r or r.z   (:285-301, :237-254)
__global__ __launch_bounds__(BLOCK) void k_cg_calc_ur(Geo G, double *__restrict__ u,
                                                       const double *__restrict__ p,
                                                       double *__restrict__ r,
                                                       const double *__restrict__ w,
                                                       double *__restrict__ z,
                                                       const double *__restrict__ Mi, double alpha,
                                                       int precond, double *__restrict__ out) {
  __shared__ double lds[BLOCK / WAVE];
  double acc = 0.0;
  TILE_LOOP(G)
  for (int y = tt.y0; y < tt.y1; ++y) {
    const long i = IDX(y, x);
    u[i] = u[i] + alpha * p[i];
    const double rv = r[i] - alpha * w[i];
    r[i] = rv;
    if (precond == TL_PREC_NONE) {
      acc += rv * rv;
    } else if (precond == TL_PREC_JAC_DIAG) {
      const double zv = Mi[i] * rv;
      z[i] = zv;
      acc += rv * zv;
    }
  }
  TILE_LOOP_END
  acc = dev::block_sum(acc, lds);
  if (threadIdx.x == 0) out[blockIdx.x] = acc;
}

// p = src + beta p   (:328-344)
__global__ __launch_bounds__(BLOCK) void k_axpy_p(Geo G, const double *__restrict__ src,
                                                   double *__restrict__ p, double beta) {
  TILE_LOOP(G)
  for (int y = tt.y0; y < tt.y1; ++y) {
    const long i = IDX(y, x);
    p[i] = src[i] + beta * p[i];
  }
  TILE_LOOP_END
}

// ---------------------------------------------------------------------------
// Chebyshev kernels (kernels/tea_leaf_cheby_kernel.f90)
// ---------------------------------------------------------------------------
// w = A u ; r = u0 - w ; [z = Mi r] ; p = (z|r)/theta (init) or alpha p + beta (z|r) ;
// u_out = u_in + p.  u is ping-ponged; the output keeps the input's halo ring
// so the halo the next read sees is the one the reference would see.
// mode: 0 = init, 1 = iterate.  write_wr: also store w and r.
__global__ __launch_bounds__(BLOCK) void k_cheby(Geo G, const double *__restrict__ uin,
                                                  double *__restrict__ uout,
                                                  const double *__restrict__ u0,
                                                  const double *__restrict__ Kx,
                                                  const double *__restrict__ Ky,
                                                  double *__restrict__ p, double *__restrict__ w,
                                                  double *__restrict__ r, double *__restrict__ z,
                                                  const double *__restrict__ Mi, double rx,
                                                  double ry, int mode, double theta, double alpha,
                                                  double beta, int precond, int write_wr,
                                                  int want_norm, double *__restrict__ out) {
  __shared__ double lds[BLOCK / WAVE];
  double acc = 0.0;
  TILE_LOOP(G)
  dev::walk<PHYS>(G, x, tt.y0, tt.y1, [&](int xx, int yy) { return uin[IDX(yy, xx)]; }, Kx, Ky,
                  [&](int y, double c, double l, double rr, double d, double up, double kxl,
                      double kxr, double kyd, double kyu) {
                    const long i = IDX(y, x);
                    const double wv = dev::apply_a(c, l, rr, d, up, kxl, kxr, kyd, kyu, rx, ry);
                    const double rv = u0[i] - wv;
                    if (write_wr) {
                      w[i] = wv;
                      r[i] = rv;
                    }
                    acc += rv * rv;
                    if (mode < 0) return;  // stencil half only (block preconditioner)
                    double src = rv;
                    if (precond == TL_PREC_JAC_DIAG) {
                      src = Mi[i] * rv;
                      z[i] = src;
                    }
                    const double pv = (mode == 0) ? src / theta : alpha * p[i] + beta * src;
                    p[i] = pv;
                    uout[i] = c + pv;
                    if (x == 0) uout[i - 1] = uin[i - 1];
                    if (x == G.nx - 1) uout[i + 1] = uin[i + 1];
                    if (y == 0) uout[i - G.pitch] = uin[i - G.pitch];
                    if (y == G.ny - 1) uout[i + G.pitch] = uin[i + G.pitch];
                  });
  TILE_LOOP_END
  if (want_norm) {
    acc = dev::block_sum(acc, lds);
    if (threadIdx.x == 0) out[blockIdx.x] = acc;
  }
}

// pointwise tail of the Chebyshev step for the block preconditioner (z given):
// p = z/theta or alpha p + beta z ; u += p
__global__ __launch_bounds__(BLOCK) void k_cheby_pu(Geo G, double *__restrict__ u,
                                                     double *__restrict__ p,
                                                     const double *__restrict__ z, int mode,
                                                     double theta, double alpha, double beta) {
  TILE_LOOP(G)
  for (int y = tt.y0; y < tt.y1; ++y) {
    const long i = IDX(y, x);
    const double pv = (mode == 0) ? z[i] / theta : alpha * p[i] + beta * z[i];
    p[i] = pv;
    u[i] = u[i] + pv;
  }
  TILE_LOOP_END
}

// ---------------------------------------------------------------------------
// PPCG kernels (kernels/tea_leaf_ppcg_kernel.f90)
// ---------------------------------------------------------------------------
// sd = src * (1/theta) ; rtemp = r ; utemp = sd   (:62-85)
__global__ __launch_bounds__(BLOCK) void k_ppcg_init_sd(Geo G, const double *__restrict__ src,
                                                         const double *__restrict__ r,
                                                         double *__restrict__ sd,
                                                         double *__restrict__ rtemp,
                                                         double *__restrict__ utemp, double theta_r) {
  TILE_LOOP(G)
  for (int y = tt.y0; y < tt.y1; ++y) {
    const long i = IDX(y, x);
    const double s = src[i] * theta_r;
    sd[i] = s;
    rtemp[i] = r[i];
    utemp[i] = s;
  }
  TILE_LOOP_END
}

// ppcg_init step 1/2/3 (:134-199) after any block solve into z
__global__ __launch_bounds__(BLOCK) void k_ppcg_init(Geo G, const double *__restrict__ r,
                                                      double *__restrict__ p, double *__restrict__ z,
                                                      const double *__restrict__ Mi, int step,
                                                      int precond, int active,
                                                      double *__restrict__ out) {
  __shared__ double lds[BLOCK / WAVE];
  double acc = 0.0;
  TILE_LOOP(G)
  for (int y = tt.y0; y < tt.y1; ++y) {
    const long i = IDX(y, x);
    if (step == 1) {
      p[i] = 0.0;
      if (precond != TL_PREC_JAC_BLOCK) z[i] = 0.0;  // the block solve already wrote z
    } else if (step == 3) {
      p[i] = 0.0;
    }
    if (precond != TL_PREC_NONE || (active && step == 3)) {
      if ((step == 1 || step == 2) && precond == TL_PREC_JAC_DIAG) z[i] = Mi[i] * r[i];
      if (step == 1 || step == 3) p[i] = z[i];
    } else if (step == 1) {
      p[i] = r[i];
    }
    if (step == 1 || step == 3) acc += r[i] * p[i];
  }
  TILE_LOOP_END
  acc = dev::block_sum(acc, lds);
  if (threadIdx.x == 0) out[blockIdx.x] = acc;
}

// one inner sweep (:288-306 and :244-263): rtemp -= A sd ;
// sd' = alpha sd + beta (Mi) rtemp ; utemp += sd'.  sd is ping-ponged.
__global__ __launch_bounds__(BLOCK) void k_ppcg_sweep(Geo G, const double *__restrict__ sdin,
                                                       double *__restrict__ sdout,
                                                       double *__restrict__ rtemp,
                                                       double *__restrict__ utemp,
                                                       const double *__restrict__ Mi,
                                                       const double *__restrict__ Kx,
                                                       const double *__restrict__ Ky, double rx,
                                                       double ry, double alpha, double beta,
                                                       int diag_prec) {
  TILE_LOOP(G)
  dev::walk<MIRROR>(G, x, tt.y0, tt.y1, [&](int xx, int yy) { return sdin[IDX(yy, xx)]; }, Kx,
                    Ky,
                    [&](int y, double c, double l, double rr, double d, double up, double kxl,
                        double kxr, double kyd, double kyu) {
                      const long i = IDX(y, x);
                      const double rt =
                          rtemp[i] - dev::apply_a(c, l, rr, d, up, kxl, kxr, kyd, kyu, rx, ry);
                      rtemp[i] = rt;
                      const double s = diag_prec ? alpha * c + beta * Mi[i] * rt : alpha * c + beta * rt;
                      sdout[i] = s;
                      utemp[i] = utemp[i] + s;
                    });
  TILE_LOOP_END
}

// rtemp -= A sd (block-preconditioned variant, first half)
__global__ __launch_bounds__(BLOCK) void k_ppcg_stencil(Geo G, const double *__restrict__ sd,
                                                         double *__restrict__ rtemp,
                                                         const double *__restrict__ Kx,
                                                         const double *__restrict__ Ky, double rx,
                                                         double ry) {
  TILE_LOOP(G)
  dev::walk<MIRROR>(G, x, tt.y0, tt.y1, [&](int xx, int yy) { return sd[IDX(yy, xx)]; }, Kx, Ky,
                    [&](int y, double c, double l, double rr, double d, double up, double kxl,
                        double kxr, double kyd, double kyu) {
                      const long i = IDX(y, x);
                      rtemp[i] = rtemp[i] - dev::apply_a(c, l, rr, d, up, kxl, kxr, kyd, kyu, rx, ry);
                    });
  TILE_LOOP_END
}

// sd = alpha sd + beta z ; utemp += sd (block-preconditioned variant, second half)
__global__ __launch_bounds__(BLOCK) void k_ppcg_update(Geo G, double *__restrict__ sd,
                                                        const double *__restrict__ z,
                                                        double *__restrict__ utemp, double alpha,
                                                        double beta) {
  TILE_LOOP(G)
  for (int y = tt.y0; y < tt.y1; ++y) {
    const long i = IDX(y, x);
    const double s = alpha * sd[i] + beta * z[i];
    sd[i] = s;
    utemp[i] = utemp[i] + s;
  }
  TILE_LOOP_END
}

// ---------------------------------------------------------------------------
// Jacobi (kernels/tea_leaf_jacobi_kernel.f90:53-72); un is the previous u, its
// halo reads as zero (it is vector_r's, which nothing writes)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_jacobi(Geo G, const double *__restrict__ un,
                                                   double *__restrict__ uout,
                                                   const double *__restrict__ u0,
                                                   const double *__restrict__ Kx,
                                                   const double *__restrict__ Ky,
                                                   double *__restrict__ rcopy, double rx, double ry,
                                                   double *__restrict__ out) {
  __shared__ double lds[BLOCK / WAVE];
  double acc = 0.0;
  TILE_LOOP(G)
  dev::walk<ZERO>(G, x, tt.y0, tt.y1, [&](int xx, int yy) { return un[IDX(yy, xx)]; }, Kx, Ky,
                  [&](int y, double c, double l, double rr, double d, double up, double kxl,
                      double kxr, double kyd, double kyu) {
                    const long i = IDX(y, x);
                    const double v = (u0[i] + rx * (kxr * rr + kxl * l) + ry * (kyu * up + kyd * d)) /
                                     (1.0 + rx * (kxl + kxr) + ry * (kyd + kyu));
                    uout[i] = v;
                    if (rcopy) rcopy[i] = c;
                    acc += fabs(v - c);
                  });
  TILE_LOOP_END
  acc = dev::block_sum(acc, lds);
  if (threadIdx.x == 0) out[blockIdx.x] = acc;
}

// ===========================================================================
// host side
// ===========================================================================
namespace {

inline int launch_check() {
  TL_HIP(hipGetLastError());
  return TL_OK;
}

inline int flat_blocks(long n) {
  long b = (n + BLOCK - 1) / BLOCK;
  if (b > 8192) b = 8192;
  if (b < 1) b = 1;
  return (int)b;
}

int block_solve(const double *r, double *z) {
  State &s = S();
  const long n = (long)s.nx * ((s.ny + 3) / 4);
  hipLaunchKernelGGL(k_block_solve, dim3(flat_blocks(n)), dim3(BLOCK), 0, s.stream, s.nx, s.ny,
                     s.pitch, r, z, s.cp.p11, s.bfp.p11, s.Ky.p11, s.Di.p11, s.ry);
  return launch_check();
}

int reduce1(int slot, double *out) {
  int sl = slot;
  return finish_sums(&sl, 1, geo().blocks, out);
}

}  // namespace

#define G_LAUNCH(kern, ...)                                                               \
  do {                                                                                  \
    const Geo G__ = geo();                                                              \
    hipLaunchKernelGGL(kern, dim3(G__.blocks), dim3(BLOCK), 0, S().stream, G__, __VA_ARGS__); \
    TL_HIP(hipGetLastError());                                                          \
  } while (0)

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" int tl_common_init(double rx, double ry, const int zero_boundary[4]) {
  TL_REQUIRE_INIT();
  State &s = S();
  if (!zero_boundary) return fail(TL_ERR_ARG, "null zero_boundary");
  s.rx = rx;
  s.ry = ry;
  s.zb_all = zero_boundary[0] && zero_boundary[1] && zero_boundary[2] && zero_boundary[3];
  G_LAUNCH(k_init_u, s.energy1.p11, s.density.p11, s.u[s.ucur].p11, s.u0.p11);
  const long nk = (long)(s.nx + 2 * s.h - 1) * (s.ny + 2 * s.h - 1);
  hipLaunchKernelGGL(k_init_k, dim3(flat_blocks(nk)), dim3(BLOCK), 0, s.stream, s.nx, s.ny, s.h,
                     s.pitch, s.density.p11, s.Kx.p11, s.Ky.p11,
                     s.coefficient == TL_RECIP_CONDUCTIVITY ? 1 : 0, s.reflective,
                     make_int4(zero_boundary[0], zero_boundary[1], zero_boundary[2],
                               zero_boundary[3]));
  TL_HIP(hipGetLastError());
  const long nd = (long)(s.nx + 2 * s.h - 2) * (s.ny + 2 * s.h - 2);
  hipLaunchKernelGGL(k_init_di, dim3(flat_blocks(nd)), dim3(BLOCK), 0, s.stream, s.nx, s.ny, s.h,
                     s.pitch, s.Kx.p11, s.Ky.p11, s.Di.p11, s.Mi.p11, rx, ry,
                     s.precond == TL_PREC_JAC_DIAG ? 1 : 0);
  TL_HIP(hipGetLastError());
  if (s.precond == TL_PREC_JAC_BLOCK) {
    const long nb = (long)s.nx * ((s.ny + 3) / 4);
    hipLaunchKernelGGL(k_block_init, dim3(flat_blocks(nb)), dim3(BLOCK), 0, s.stream, s.nx, s.ny,
                       s.pitch, s.Ky.p11, s.Di.p11, s.cp.p11, s.bfp.p11, ry);
    TL_HIP(hipGetLastError());
  }
  G_LAUNCH(k_init_wr, s.u[s.ucur].p11, s.Kx.p11, s.Ky.p11, s.w.p11, s.r.p11, rx, ry);
  return TL_OK;
}

extern "C" int tl_calc_residual(void) {
  TL_REQUIRE_INIT();
  State &s = S();
  G_LAUNCH(k_residual, s.u[s.ucur].p11, s.u0.p11, s.Kx.p11, s.Ky.p11, s.r.p11, s.rx, s.ry);
  return TL_OK;
}

extern "C" int tl_calc_2norm(int which, double *norm) {
  TL_REQUIRE_INIT();
  State &s = S();
  if (!norm) return fail(TL_ERR_ARG, "null norm");
  if (which != 0 && which != 1)
    return fail(TL_ERR_ARG, "tea_leaf_common.f90: Invalid value for norm_array");
  const double *a = which == 0 ? s.u0.p11 : s.r.p11;
  G_LAUNCH(k_dot, a, a, part(P0));
  return reduce1(P0, norm);
}

extern "C" int tl_finalise(void) {
  TL_REQUIRE_INIT();
  State &s = S();
  G_LAUNCH(k_div, s.u[s.ucur].p11, s.density.p11, s.energy1.p11);
  return TL_OK;
}

extern "C" int tl_set_field(void) {
  TL_REQUIRE_INIT();
  State &s = S();
  G_LAUNCH(k_copy, s.energy0.p11, s.energy1.p11);
  return TL_OK;
}

extern "C" int tl_update_halo(const int fields[11], int depth) {
  TL_REQUIRE_INIT();
  State &s = S();
  if (!fields) return fail(TL_ERR_ARG, "null fields");
  if (depth < 1 || depth > s.h) return fail(TL_ERR_ARG, "update_halo: depth outside 1..halo");
  if (!s.reflective) return TL_OK;  // one chunk: tea_exchange and tile copies are no-ops
  FieldList fl;
  fl.n = 0;
  // FIELD_* order; the FIELD_R slot carries rtemp (update_halo.f90:79)
  double *slot[11] = {s.density.p11, s.energy0.p11, s.energy1.p11, s.u[s.ucur].p11,
                      s.p[s.pcur].p11, s.sd[s.sdcur].p11, s.rtemp.p11, s.z.p11,
                      s.Kx.p11, s.Ky.p11, s.Di.p11};
  for (int i = 0; i < 11; ++i)
    if (fields[i] == 1) fl.f[fl.n++] = slot[i];
  if (fl.n == 0) return TL_OK;
  const long nlr = (long)fl.n * (s.ny + 2 * depth) * depth;
  hipLaunchKernelGGL(k_halo_lr, dim3(flat_blocks(nlr)), dim3(BLOCK), 0, s.stream, fl, s.nx, s.ny,
                     s.pitch, depth);
  TL_HIP(hipGetLastError());
  const long nbt = (long)fl.n * (s.nx + 2 * depth) * depth;
  hipLaunchKernelGGL(k_halo_bt, dim3(flat_blocks(nbt)), dim3(BLOCK), 0, s.stream, fl, s.nx, s.ny,
                     s.pitch, depth);
  TL_HIP(hipGetLastError());
  return TL_OK;
}

extern "C" int tl_field_summary(double *vol, double *mass, double *ie, double *temp) {
  TL_REQUIRE_INIT();
  State &s = S();
  G_LAUNCH(k_summary, s.volume.p11, s.density.p11, s.energy1.p11, s.u[s.ucur].p11, part(P0));
  const int slots[4] = {P0, P1, P2, P3};
  double out[4];
  int rc = finish_sums(slots, 4, geo().blocks, out);
  if (rc) return rc;
  if (vol) *vol = out[0];
  if (mass) *mass = out[1];
  if (ie) *ie = out[2];
  if (temp) *temp = out[3];
  return TL_OK;
}

// ---- CG --------------------------------------------------------------------
extern "C" int tl_cg_init(double *rro) {
  TL_REQUIRE_INIT();
  State &s = S();
  if (!rro) return fail(TL_ERR_ARG, "null rro");
  if (s.precond == TL_PREC_JAC_BLOCK) {
    int rc = block_solve(s.r.p11, s.z.p11);
    if (rc) return rc;
  }
  G_LAUNCH(k_cg_init, s.r.p11, s.p[s.pcur].p11, s.z.p11, s.Mi.p11, s.precond, part(P0));
  return reduce1(P0, rro);
}

extern "C" int tl_cg_calc_w(double *pw) {
  TL_REQUIRE_INIT();
  State &s = S();
  if (!pw) return fail(TL_ERR_ARG, "null pw");
  hipEvent_t ev;
  prof_begin(0, &ev);
  G_LAUNCH(k_cg_calc_w, s.p[s.pcur].p11, s.Kx.p11, s.Ky.p11, s.w.p11, s.rx, s.ry, part(P0));
  prof_end(0, ev);
  return reduce1(P0, pw);
}

extern "C" int tl_cg_calc_ur(double alpha, double *rrn) {
  TL_REQUIRE_INIT();
  State &s = S();
  if (!rrn) return fail(TL_ERR_ARG, "null rrn");
  hipEvent_t ev;
  prof_begin(1, &ev);
  G_LAUNCH(k_cg_calc_ur, s.u[s.ucur].p11, s.p[s.pcur].p11, s.r.p11, s.w.p11, s.z.p11, s.Mi.p11,
           alpha, s.precond, part(P0));
  prof_end(1, ev);
  if (s.precond == TL_PREC_JAC_BLOCK) {
    int rc = block_solve(s.r.p11, s.z.p11);
    if (rc) return rc;
    G_LAUNCH(k_dot, s.r.p11, s.z.p11, part(P0));
  }
  return reduce1(P0, rrn);
}

extern "C" int tl_cg_calc_p(double beta) {
  TL_REQUIRE_INIT();
  State &s = S();
  const double *src = (s.precond != TL_PREC_NONE || s.ppcg_active) ? s.z.p11 : s.r.p11;
  G_LAUNCH(k_axpy_p, src, s.p[s.pcur].p11, beta);
  return TL_OK;
}

// ---- Chebyshev ---------------------------------------------------------------
static int cheby_step(int mode, double theta, double alpha, double beta, int want_norm,
                      double *norm) {
  State &s = S();
  const int in = s.ucur, outb = 1 - s.ucur;
  hipEvent_t ev;
  prof_begin(3, &ev);
  if (s.precond == TL_PREC_JAC_BLOCK) {
    // stencil part, then the tridiagonal solve, then the pointwise update
    G_LAUNCH(k_cheby, s.u[in].p11, s.u[in].p11, s.u0.p11, s.Kx.p11, s.Ky.p11, s.p[s.pcur].p11,
             s.w.p11, s.r.p11, s.z.p11, s.Mi.p11, s.rx, s.ry, -1, theta, alpha, beta,
             TL_PREC_NONE, 1, want_norm, part(P0));
    int rc = block_solve(s.r.p11, s.z.p11);
    if (rc) return rc;
    G_LAUNCH(k_cheby_pu, s.u[in].p11, s.p[s.pcur].p11, s.z.p11, mode, theta, alpha, beta);
  } else {
    G_LAUNCH(k_cheby, s.u[in].p11, s.u[outb].p11, s.u0.p11, s.Kx.p11, s.Ky.p11, s.p[s.pcur].p11,
             s.w.p11, s.r.p11, s.z.p11, s.Mi.p11, s.rx, s.ry, mode, theta, alpha, beta, s.precond,
             1, want_norm, part(P0));
    s.ucur = outb;
  }
  prof_end(3, ev);
  if (want_norm) return reduce1(P0, norm);
  return TL_OK;
}

extern "C" int tl_cheby_init(double theta) {
  TL_REQUIRE_INIT();
  return cheby_step(0, theta, 0.0, 0.0, 0, nullptr);
}

extern "C" int tl_cheby_iterate(double alpha, double beta) {
  TL_REQUIRE_INIT();
  return cheby_step(1, 0.0, alpha, beta, 0, nullptr);
}

extern "C" int tl_cheby_run(const double *ch_alphas, const double *ch_betas, int step_first,
                            int count, int want_norm, double *norm) {
  TL_REQUIRE_INIT();
  if (!ch_alphas || !ch_betas || step_first < 1 || count < 0)
    return fail(TL_ERR_ARG, "tl_cheby_run: bad arguments");
  if (want_norm && !norm) return fail(TL_ERR_ARG, "tl_cheby_run: null norm");
  const int fused = cheby_run_fused(ch_alphas, ch_betas, step_first, count, want_norm, norm);
  if (fused < 0) return -fused;
  if (!fused) {
    for (int i = 0; i < count; ++i) {
      const int st = step_first + i - 1;
      const int last = (i == count - 1) && want_norm;
      int rc = cheby_step(1, 0.0, ch_alphas[st], ch_betas[st], last, norm);
      if (rc) return rc;
    }
  }
  // the residual check of the run (tea_solve.f90:273) on a NaN or Inf norm
  if (want_norm && count > 0 && !std::isfinite(*norm))
    return fail(TL_ERR_NONFINITE, "Chebyshev: non-finite residual norm at step " +
                                      std::to_string(step_first + count - 1) +
                                      " (NaN or Inf in the fields)");
  return TL_OK;
}

// ---- PPCG ----------------------------------------------------------------------
extern "C" int tl_ppcg_init_sd(double theta) {
  TL_REQUIRE_INIT();
  State &s = S();
  const double *src = s.precond != TL_PREC_NONE ? s.z.p11 : s.r.p11;
  G_LAUNCH(k_ppcg_init_sd, src, s.r.p11, s.sd[s.sdcur].p11, s.rtemp.p11, s.utemp.p11,
           1.0 / theta);
  return TL_OK;
}

extern "C" int tl_ppcg_init(int step, double *rro) {
  TL_REQUIRE_INIT();
  State &s = S();
  if (!rro) return fail(TL_ERR_ARG, "null rro");
  if (step < 1 || step > 3) return fail(TL_ERR_ARG, "tl_ppcg_init: step must be 1, 2 or 3");
  const bool active = s.ppcg_active != 0;
  if ((s.precond != TL_PREC_NONE || (active && step == 3)) && (step == 1 || step == 2) &&
      s.precond == TL_PREC_JAC_BLOCK) {
    int rc = block_solve(s.r.p11, s.z.p11);
    if (rc) return rc;
  }
  G_LAUNCH(k_ppcg_init, s.r.p11, s.p[s.pcur].p11, s.z.p11, s.Mi.p11, step, s.precond,
           active ? 1 : 0, part(P0));
  if (step == 2) {
    *rro = 0.0;  // the wrapper's sum of zero tile contributions (tea_leaf_ppcg.f90:83)
    return TL_OK;
  }
  return reduce1(P0, rro);
}

extern "C" int tl_ppcg_inner(const double *alphas, const double *betas, int first, int count) {
  TL_REQUIRE_INIT();
  State &s = S();
  if (!alphas || !betas || first < 1 || count < 0)
    return fail(TL_ERR_ARG, "tl_ppcg_inner: bad arguments");
  for (int i = 0; i < count; ++i) {
    const double a = alphas[first - 1 + i], b = betas[first - 1 + i];
    hipEvent_t ev;
    prof_begin(2, &ev);
    if (s.precond == TL_PREC_JAC_BLOCK) {
      G_LAUNCH(k_ppcg_stencil, s.sd[s.sdcur].p11, s.rtemp.p11, s.Kx.p11, s.Ky.p11, s.rx, s.ry);
      int rc = block_solve(s.rtemp.p11, s.z.p11);
      if (rc) return rc;
      G_LAUNCH(k_ppcg_update, s.sd[s.sdcur].p11, s.z.p11, s.utemp.p11, a, b);
    } else {
      const int outb = 1 - s.sdcur;
      G_LAUNCH(k_ppcg_sweep, s.sd[s.sdcur].p11, s.sd[outb].p11, s.rtemp.p11, s.utemp.p11,
               s.Mi.p11, s.Kx.p11, s.Ky.p11, s.rx, s.ry, a, b,
               s.precond == TL_PREC_JAC_DIAG ? 1 : 0);
      s.sdcur = outb;
    }
    prof_end(2, ev);
  }
  return TL_OK;
}

extern "C" int tl_ppcg_store_r(void) {
  TL_REQUIRE_INIT();
  State &s = S();
  G_LAUNCH(k_copy, s.r.p11, s.r_store.p11);
  return TL_OK;
}

extern "C" int tl_ppcg_calc_zrnorm(double *rro) {
  TL_REQUIRE_INIT();
  State &s = S();
  if (!rro) return fail(TL_ERR_ARG, "null rro");
  const double *a = (s.precond != TL_PREC_NONE || s.ppcg_active) ? s.z.p11 : s.r.p11;
  G_LAUNCH(k_dot, a, s.r.p11, part(P0));
  return reduce1(P0, rro);
}

extern "C" int tl_ppcg_calc_rrn(double *rrn) {
  TL_REQUIRE_INIT();
  State &s = S();
  if (!rrn) return fail(TL_ERR_ARG, "null rrn");
  G_LAUNCH(k_rrn, s.r.p11, s.r_store.p11, s.z.p11, part(P0));
  return reduce1(P0, rrn);
}

extern "C" int tl_ppcg_update_z(void) {
  TL_REQUIRE_INIT();
  State &s = S();
  G_LAUNCH(k_copy, s.utemp.p11, s.z.p11);
  return TL_OK;
}

// ---- Jacobi --------------------------------------------------------------------
extern "C" int tl_jacobi_solve(double *error) {
  TL_REQUIRE_INIT();
  State &s = S();
  if (!error) return fail(TL_ERR_ARG, "null error");
  const int outb = 1 - s.ucur;
  hipEvent_t ev;
  prof_begin(4, &ev);
  G_LAUNCH(k_jacobi, s.u[s.ucur].p11, s.u[outb].p11, s.u0.p11, s.Kx.p11, s.Ky.p11, s.r.p11, s.rx,
           s.ry, part(P0));
  prof_end(4, ev);
  s.ucur = outb;
  return reduce1(P0, error);
}
