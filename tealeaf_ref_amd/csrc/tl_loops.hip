// Fused solver loops: consecutive iterations of the tea_leaf() DO loop
// (tea_solve.f90:152-414) with every scalar kept on the device.  The CG and
// PPCG kernels and their drivers are shared with the mixed-precision solver
// (tl_loopk.hpp); this file holds the FP64 entry points and the Jacobi loop.
#include <cmath>
#include <cstdlib>

#include "tl_loopk.hpp"
#include "tl_cheby2.hpp"
#include "tl_jacobi2.hpp"

namespace tl {
dev::Geo geo(int cols_per_tile);
dev::Geo geo_resident(int cols_per_tile, int blocks_per_cu);
}
using namespace tl;
using dev::Geo;

namespace {

#define TILE_LOOP(G)                                 \
  dev::TileIter it__((G).tiles, (G).blocks);         \
  for (long t__ = it__.t; t__ < it__.t_end; ++t__) { \
    const dev::Tile tt = dev::tile_of((G), t__);     \
    if (tt.x < (G).nx) {                             \
      const int x = tt.x;
#define TILE_LOOP_END \
  }                   \
  }
#define IDX(y, x) ((long)(y) * G.pitch + (x))

// ---------------------------------------------------------------------------
// Jacobi iteration n (tea_solve.f90:375-379): exit test of n-1, then the
// sweep from u[in] to u[out] with its |u - un| partials (parity-buffered).
// ---------------------------------------------------------------------------
// Exit test of iteration n-1 at the head of iteration n (none on the first):
// true when the loop stops (converged, non-finite error, or stopped before).
// herr[n-1] gets the error sqrt|sum|u - un||.
__device__ inline bool jacobi_head(int n, int first, double eps_abs,
                                   const double *__restrict__ prev_part, int nprev, Ctl *ctl,
                                   double *__restrict__ herr, double *lds) {
  if (first) {
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl->stop = 0;
    return false;
  }
  if (ctl->stop) return true;
  const double e = sqrt(fabs(dev::sum_partials(prev_part, nprev, lds)));
  const bool conv = fabs(e) < eps_abs;
  if (blockIdx.x == 0 && threadIdx.x == 0) herr[n - 1] = e;
  if (!conv && isfinite(e)) return false;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctl->stop = 1;
    ctl->reason = conv ? TL_LOOP_CONVERGED : TL_LOOP_NONFINITE;
    ctl->n_end = n - 1;
    ctl->error = e;
  }
  return true;
}

__global__ __launch_bounds__(BLOCK) void k_jl(Geo G, int n, int first, double eps_abs,
                                               const double *__restrict__ un,
                                               double *__restrict__ uout,
                                               const double *__restrict__ u0,
                                               const double *__restrict__ Kx,
                                               const double *__restrict__ Ky, double rx, double ry,
                                               const double *__restrict__ prev_part, int nprev,
                                               double *__restrict__ out_part, Ctl *ctl,
                                               double *__restrict__ herr) {
  __shared__ double lds[BLOCK / WAVE];
  if (jacobi_head(n, first, eps_abs, prev_part, nprev, ctl, herr, lds)) return;
  double acc = 0.0;
  TILE_LOOP(G)
  dev::walk<ZERO>(G, x, tt.y0, tt.y1, [&](int xx, int yy) { return un[IDX(yy, xx)]; }, Kx, Ky,
                  [&](int y, double c, double l, double rr, double d, double up, double kxl,
                      double kxr, double kyd, double kyu) {
                    const long i = IDX(y, x);
                    const double v = (u0[i] + rx * (kxr * rr + kxl * l) + ry * (kyu * up + kyd * d)) /
                                     (1.0 + rx * (kxl + kxr) + ry * (kyd + kyu));
                    uout[i] = v;
                    acc += fabs(v - c);
                  });
  TILE_LOOP_END
  acc = dev::block_sum(acc, lds);
  if (threadIdx.x == 0) out_part[blockIdx.x] = acc;
}

// The same iteration on the vector walk (two columns per lane, 16-byte loads,
// wave shuffles and LDS for the neighbours), for chunks with every face
// external.  Outside the mesh: ZE = false mirrors (the faces' coefficients
// are zero, so the value is multiplied by zero), ZE = true reads zero
// (reflective faces: the reference reads the halo of un, which it never
// writes, as the per-cell kernel's ZERO walk does).
template <bool NT, bool ZE>
__global__ __launch_bounds__(BLOCK) void k_jl_v(Geo G, int n, int first, double eps_abs,
                                                const double *__restrict__ un,
                                                double *__restrict__ uout,
                                                const double *__restrict__ u0,
                                                const double *__restrict__ Kx,
                                                const double *__restrict__ Ky, double rx,
                                                double ry, const double *__restrict__ prev_part,
                                                int nprev, double *__restrict__ out_part, Ctl *ctl,
                                                double *__restrict__ herr, int dir) {
  typedef typename dev::Vec<double, 2>::type VT;
  __shared__ double lds[BLOCK / WAVE];
  if (jacobi_head(n, first, eps_abs, prev_part, nprev, ctl, herr, lds)) return;
  const long P = G.pitch;
  double acc = 0.0;
  dev::TileIter it(G.tiles, G.blocks);
  for (long t = it.t; t < it.t_end; ++t) {
    const dev::TileV<2> tt = dev::tilev_of<2>(G, t);
    const int cc = tt.c <= G.nx ? tt.c : (G.nx / 2) * 2;
    auto ldop = [&](int col, int row) -> VT {
      return dev::ldv<NT, double, 2>(un + (long)row * P + col);
    };
    auto ld1 = [&](int col, int row) -> double { return un[(long)row * P + col]; };
    auto src = [&](int row) -> VT { return dev::ldv<NT, double, 2>(u0 + (long)row * P + cc); };
    // tea_leaf_jacobi_kernel.f90:60-66, in the reference's order of operations
    auto cell = [&](const VT &s0, int e, double c, double l, double r, double d, double u,
                    double kxl, double kxr, double kyd, double kyu) -> double {
      (void)c;
      return (s0[e] + rx * (kxr * r + kxl * l) + ry * (kyu * u + kyd * d)) /
             (1.0 + rx * (kxl + kxr) + ry * (kyd + kyu));
    };
    dev::walkv<NT, double, 2, false, TL_PF_W, true, ZE ? 1 : 0>(
        G, tt, Kx, Ky, rx, ry, ldop, ld1, src, [&](int y, VT v, VT w, const VT &) {
      if (tt.c >= G.nx) return;
      dev::stv<false, double, 2>(uout + (long)y * P, tt.c, G.nx, w);
      acc += fabs(w.x - v.x);
      if (tt.c + 1 < G.nx) acc += fabs(w.y - v.y);
    }, dir, cell);
  }
  acc = dev::block_sum(acc, lds);
  if (threadIdx.x == 0) out_part[blockIdx.x] = acc;
}

// Exit test of the two iterations of the previous pass (n-2, then n-1, as the
// reference tests after every iteration): partials of iteration n-2 in pa,
// of n-1 in pb (pa null: the previous launch did one iteration).  Returns
// true (and records the stop) when one of them has converged; with
// at_limit, records the stop at n-1 in any case.
__device__ inline bool jacobi_pair_exit(int n, double eps_abs, const double *__restrict__ pa,
                                        const double *__restrict__ pb, int nprev, Ctl *ctl,
                                        double *__restrict__ herr, double *lds, bool at_limit) {
  if (pa) {
    const double ea = sqrt(fabs(dev::sum_partials(pa, nprev, lds)));
    const bool conv_a = fabs(ea) < eps_abs;
    if (blockIdx.x == 0 && threadIdx.x == 0) herr[n - 2] = ea;
    if (conv_a || !isfinite(ea)) {
      if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl->stop = 1; ctl->reason = conv_a ? TL_LOOP_CONVERGED : TL_LOOP_NONFINITE;
        ctl->n_end = n - 2; ctl->error = ea;
      }
      return true;
    }
  }
  const double eb = sqrt(fabs(dev::sum_partials(pb, nprev, lds)));
  const bool conv = fabs(eb) < eps_abs, bad = !isfinite(eb);
  if (blockIdx.x == 0 && threadIdx.x == 0) herr[n - 1] = eb;
  if (conv || bad || at_limit) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      ctl->stop = 1;
      ctl->reason = conv ? TL_LOOP_CONVERGED : (bad ? TL_LOOP_NONFINITE : TL_LOOP_MAXITERS);
      ctl->n_end = n - 1;
      ctl->error = eb;
    }
    return true;
  }
  return false;
}

// Jacobi iterations n and n+1 in one pass (tl_jacobi2.hpp): exit test of the
// previous pass first; u of iteration n to umid, of n+1 to uout, the
// partials of sum |u_n - u_n-1| to part_a and of sum |u_n+1 - u_n| to part_b.
// Compiled for 4 waves per SIMD (5 registers spilled): 153.5 us per pass at
// 4096^2, against 176.5 with 3 waves (no spills) on 1024 tiles and 170.0 on
// 720.
#ifndef TL_JACOBI_PAIR_WAVES
#define TL_JACOBI_PAIR_WAVES 4
#endif

template <bool NT, bool ZE>
__global__ __launch_bounds__(BLOCK, TL_JACOBI_PAIR_WAVES) void k_jl_pair(
    Geo G, int n, int first, double eps_abs, const double *__restrict__ un,
    double *__restrict__ umid, double *__restrict__ uout, const double *__restrict__ u0,
    const double *__restrict__ Kx, const double *__restrict__ Ky, double rx, double ry,
    const double *__restrict__ prev_a, const double *__restrict__ prev_b, int nprev,
    double *__restrict__ part_a, double *__restrict__ part_b, Ctl *ctl,
    double *__restrict__ herr, int dir) {
  typedef typename dev::Vec<double, 2>::type VT;
  __shared__ double lds[BLOCK / WAVE];
  if (!first) {
    if (ctl->stop) return;
    if (jacobi_pair_exit(n, eps_abs, prev_a, prev_b, nprev, ctl, herr, lds, false)) return;
  } else if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctl->stop = 0;
  }
  const long P = G.pitch;
  double acc_a = 0.0, acc_b = 0.0;
  dev::TileIter it(G.tiles, G.blocks);
  for (long t = it.t; t < it.t_end; ++t) {
    const dev::TileV<2> tt = dev::tilev_of<2>(G, t);
    auto ldop = [&](int col, int row) -> VT {
      return dev::ldv<NT, double, 2>(un + (long)row * P + col);
    };
    auto ld1 = [&](int col, int row) -> double { return un[(long)row * P + col]; };
    auto src = [&](int col, int row) -> VT {
      return dev::ldv<NT, double, 2>(u0 + (long)row * P + col);
    };
    auto src1 = [&](int col, int row) -> double { return u0[(long)row * P + col]; };
    dev::walk_jacobi_pair<NT, 2, ZE>(G, tt, Kx, Ky, rx, ry, ldop, ld1, src, src1,
                                 [&](int y, VT v0, VT v1, VT v2) {
      if (tt.c >= G.nx) return;
      const long row = (long)y * P;
      dev::stv<false, double, 2>(umid + row, tt.c, G.nx, v1);
      dev::stv<false, double, 2>(uout + row, tt.c, G.nx, v2);
      acc_a += fabs(v1.x - v0.x);
      acc_b += fabs(v2.x - v1.x);
      if (tt.c + 1 < G.nx) {
        acc_a += fabs(v1.y - v0.y);
        acc_b += fabs(v2.y - v1.y);
      }
    }, dir);
  }
  acc_a = dev::block_sum(acc_a, lds);
  acc_b = dev::block_sum(acc_b, lds);
  if (threadIdx.x == 0) {
    part_a[blockIdx.x] = acc_a;
    part_b[blockIdx.x] = acc_b;
  }
}

// The exit test of the launch before iteration n, by one block: before a
// closing single iteration (at_limit 0), or after the last launch of the loop
// (at_limit 1: a stop in any case; pa null when that launch did one
// iteration).
__global__ void k_jl_pair_test(int n, double eps_abs, const double *__restrict__ pa,
                               const double *__restrict__ pb, int nprev, Ctl *ctl,
                               double *__restrict__ herr, int at_limit) {
  __shared__ double lds[BLOCK / WAVE];
  if (ctl->stop) return;
  jacobi_pair_exit(n, eps_abs, pa, pb, nprev, ctl, herr, lds, at_limit != 0);
}

__global__ void k_jl_tail(int n, double eps_abs, const double *__restrict__ prev_part, int nprev,
                          Ctl *ctl, double *__restrict__ herr) {
  __shared__ double lds[BLOCK / WAVE];
  if (ctl->stop) return;
  const double e = sqrt(fabs(dev::sum_partials(prev_part, nprev, lds)));
  if (threadIdx.x == 0) {
    herr[n - 1] = e;
    ctl->stop = 1;
    ctl->reason = fabs(e) < eps_abs ? TL_LOOP_CONVERGED
                                    : (isfinite(e) ? TL_LOOP_MAXITERS : TL_LOOP_NONFINITE);
    ctl->n_end = n - 1;
    ctl->error = e;
  }
}

// Columns per lane of the FP64 loop kernels (16-byte loads at 2).
#ifndef TL_V64
#define TL_V64 2
#endif

// ---------------------------------------------------------------------------
// Chebyshev iteration (kernels/tea_leaf_cheby_kernel.f90:111-200) on the
// vector walk: w = A u, r = u0 - w, (z = Mi r), p = alpha p + beta (r|z),
// u' = u + p into the other u buffer.  w, r (and z) are stored, and the
// partials of r.r written, only on the last iteration of a run: nothing reads
// them in between.  The halo ring of u is carried over to the new buffer, as
// the per-routine kernel does.  EDGE 0: chunks whose external faces are zero
// (not reflective), where the mirrored values outside the mesh are
// multiplied by zero; EDGE 2: reflective faces, whose nonzero coefficients
// take the halo cells of u as they are in memory, as the reference does (it
// refreshes only p's halo during the Chebyshev iterations, so u's halo
// stays as the first Chebyshev step left it, carried from buffer to buffer).
// ---------------------------------------------------------------------------
#ifndef TL_PF_CHEBY
#define TL_PF_CHEBY 0
#endif

#ifndef TL_CHEBY_WAVES
#define TL_CHEBY_WAVES 1
#endif

template <bool NT, int EDGE>
__global__ __launch_bounds__(BLOCK, TL_CHEBY_WAVES) void k_cheby_v(Geo G, const double *__restrict__ uin,
                                                    double *__restrict__ uout,
                                                    const double *__restrict__ u0,
                                                    const double *__restrict__ Kx,
                                                    const double *__restrict__ Ky,
                                                    double *__restrict__ p, double *__restrict__ w,
                                                    double *__restrict__ r, double *__restrict__ z,
                                                    const double *__restrict__ Mi, double rx,
                                                    double ry, double alpha, double beta, int diag,
                                                    int last, double *__restrict__ out_part,
                                                    int dir) {
  typedef typename dev::Vec<double, 2>::type VT;
  __shared__ double lds[BLOCK / WAVE];
  const long P = G.pitch;
  double acc = 0.0;
  dev::TileIter it(G.tiles, G.blocks);
  for (long t = it.t; t < it.t_end; ++t) {
    const dev::TileV<2> tt = dev::tilev_of<2>(G, t);
    const int cc = tt.c <= G.nx ? tt.c : (G.nx / 2) * 2;
    struct Side {
      VT u0, p, mi;
    };
    auto side = [&](int row) -> Side {
      const long i = (long)row * P + cc;
      Side e;
      e.u0 = dev::ldv<NT, double, 2>(u0 + i);
      e.p = dev::ldv<NT, double, 2>(p + i);
      e.mi = diag ? dev::ldv<NT, double, 2>(Mi + i) : VT(0);
      return e;
    };
    auto ldop = [&](int col, int row) -> VT {
      return dev::ldv<NT, double, 2>(uin + (long)row * P + col);
    };
    auto ld1 = [&](int col, int row) -> double { return uin[(long)row * P + col]; };
    dev::walkv<NT, double, 2, false, TL_PF_CHEBY, true, EDGE>(
        G, tt, Kx, Ky, rx, ry, ldop, ld1, side, [&](int y, VT v, VT wv, const Side &e) {
      if (tt.c >= G.nx) return;
      const long row = (long)y * P;
      const VT rv = e.u0 - wv;
      const VT src = diag ? e.mi * rv : rv;
      const VT pv = alpha * e.p + beta * src;
      dev::stv<false, double, 2>(p + row, tt.c, G.nx, pv);
      dev::stv<false, double, 2>(uout + row, tt.c, G.nx, v + pv);
      if (last) {
        dev::stv<false, double, 2>(w + row, tt.c, G.nx, wv);
        dev::stv<false, double, 2>(r + row, tt.c, G.nx, rv);
        if (diag) dev::stv<false, double, 2>(z + row, tt.c, G.nx, src);
        acc += rv.x * rv.x;
        if (tt.c + 1 < G.nx) acc += rv.y * rv.y;
      }
      // carry the halo ring of u to the new buffer
      if (tt.c == 0) uout[row - 1] = uin[row - 1];
      if (tt.c + 2 >= G.nx) uout[row + G.nx] = uin[row + G.nx];
      for (int side_row = 0; side_row < 2; ++side_row) {
        if (y != (side_row ? G.ny - 1 : 0)) continue;
        const long hr = side_row ? (long)G.ny * P : -P;
        uout[hr + tt.c] = uin[hr + tt.c];
        if (tt.c + 1 < G.nx) uout[hr + tt.c + 1] = uin[hr + tt.c + 1];
      }
    }, dir);
  }
  if (last) {
    acc = dev::block_sum(acc, lds);
    if (threadIdx.x == 0) out_part[blockIdx.x] = acc;
  }
}

// Chebyshev iterations k and k+1 in one pass (tl_cheby2.hpp): u and p are
// read from uin / pin and written to uout / pout.  On the last pass of a batch
// w, r (and z) of iteration k+1 are stored and the partials of r.r written.
// Compiled for 3 waves per SIMD (160 registers, no spills), on a grid of
// 3 blocks per CU so that every block is resident at once (cheby_geo): at
// 4096^2 179.6 us per pass, against 205.6 with 4 waves (29 registers
// spilled) on 1024 blocks and 213.7 with 3 waves on 1536.
#ifndef TL_CHEBY_PAIR_WAVES
#define TL_CHEBY_PAIR_WAVES 3
#endif

// DIAG (Jacobi-diagonal preconditioner) and LAST (the last pass of a run,
// which stores w, r, z and the r.r partials) are template parameters, so
// that the common pass holds no state it does not use.
template <bool NT, int DIAG, int LAST>
__global__ __launch_bounds__(BLOCK, TL_CHEBY_PAIR_WAVES) void k_cheby_pair(
    Geo G, const double *__restrict__ uin, double *__restrict__ uout,
    const double *__restrict__ u0, const double *__restrict__ Kx, const double *__restrict__ Ky,
    const double *__restrict__ pin, double *__restrict__ pout, double *__restrict__ w,
    double *__restrict__ r, double *__restrict__ z, const double *__restrict__ Mi, double rx,
    double ry, double a1, double b1, double a2, double b2, double *__restrict__ out_part,
    int dir) {
  constexpr bool diag = DIAG != 0, last = LAST != 0;
  typedef typename dev::Vec<double, 2>::type VT;
  __shared__ double lds[BLOCK / WAVE];
  const long P = G.pitch;
  double acc = 0.0;
  dev::TileIter it(G.tiles, G.blocks);
  for (long t = it.t; t < it.t_end; ++t) {
    const dev::TileV<2> tt = dev::tilev_of<2>(G, t);
    auto ldop = [&](int col, int row) -> VT {
      return dev::ldv<NT, double, 2>(uin + (long)row * P + col);
    };
    auto ld1 = [&](int col, int row) -> double { return uin[(long)row * P + col]; };
    auto side = [&](int col, int row) -> dev::ChebyIn<VT> {
      const long i = (long)row * P + col;
      dev::ChebyIn<VT> e;
      e.u0 = dev::ldv<NT, double, 2>(u0 + i);
      e.p = dev::ldv<NT, double, 2>(pin + i);
      e.mi = diag ? dev::ldv<NT, double, 2>(Mi + i) : VT(0);
      return e;
    };
    auto side1 = [&](int col, int row) -> dev::ChebyIn<double> {
      const long i = (long)row * P + col;
      dev::ChebyIn<double> e;
      e.u0 = u0[i];
      e.p = pin[i];
      e.mi = diag ? Mi[i] : 0.0;
      return e;
    };
    dev::walk_cheby_pair<NT, 2, DIAG>(
        G, tt, Kx, Ky, rx, ry, a1, b1, ldop, ld1, side, side1,
        [&](int y, VT u1, VT p1, VT wv, const dev::ChebyIn<VT> &e) {
      if (tt.c >= G.nx) return;
      const long row = (long)y * P;
      const VT rv = e.u0 - wv;
      const VT src = diag ? e.mi * rv : rv;
      const VT pv = a2 * p1 + b2 * src;
      dev::stv<false, double, 2>(pout + row, tt.c, G.nx, pv);
      dev::stv<false, double, 2>(uout + row, tt.c, G.nx, u1 + pv);
      if (last) {
        dev::stv<false, double, 2>(w + row, tt.c, G.nx, wv);
        dev::stv<false, double, 2>(r + row, tt.c, G.nx, rv);
        if (diag) dev::stv<false, double, 2>(z + row, tt.c, G.nx, src);
        acc += rv.x * rv.x;
        if (tt.c + 1 < G.nx) acc += rv.y * rv.y;
      }
      // carry the halo ring of u to the new buffer
      if (tt.c == 0) uout[row - 1] = uin[row - 1];
      if (tt.c + 2 >= G.nx) uout[row + G.nx] = uin[row + G.nx];
      for (int side_row = 0; side_row < 2; ++side_row) {
        if (y != (side_row ? G.ny - 1 : 0)) continue;
        const long hr = side_row ? (long)G.ny * P : -P;
        uout[hr + tt.c] = uin[hr + tt.c];
        if (tt.c + 1 < G.nx) uout[hr + tt.c + 1] = uin[hr + tt.c + 1];
      }
    }, dir);
  }
  if (last) {
    acc = dev::block_sum(acc, lds);
    if (threadIdx.x == 0) out_part[blockIdx.x] = acc;
  }
}

}  // namespace

namespace tl {
// One Chebyshev iteration with the block-Jacobi preconditioner in one pass
// (kernels/tea_leaf_cheby_kernel.f90:111-200 with tea_block_solve,
// kernels/tea_leaf_common_kernel.f90:378-468): w = A u and r = u0 - w row by
// row, the forward elimination of the 4-row Thomas solve as each row
// arrives, and once the rows of a preconditioner block have been walked the
// back substitution z = M^-1 r, p = alpha p + beta z and u' = u + p for the
// block's rows.  Tiles start on the blocks (multiples of 4 rows) and are
// walked upwards.  LAST (the last iteration of a run, before a residual
// check) also stores w, r and z and the partials of r.r.
template <bool NT, int LAST>
__global__ __launch_bounds__(BLOCK, TL_CHEBY_WAVES) void k_cheby_blk(
    Geo G, const double *__restrict__ uin, double *__restrict__ uout,
    const double *__restrict__ u0, const double *__restrict__ Kx, const double *__restrict__ Ky,
    double *__restrict__ p, double *__restrict__ w, double *__restrict__ r,
    double *__restrict__ z, const double *__restrict__ Di, const double *__restrict__ cp,
    const double *__restrict__ bfp, double rx, double ry, double alpha, double beta,
    double *__restrict__ out_part) {
  typedef typename dev::Vec<double, 2>::type VT;
  __shared__ double lds[BLOCK / WAVE];
  const long P = G.pitch;
  double acc = 0.0;
  dev::TileIter it(G.tiles, G.blocks);
  for (long t = it.t; t < it.t_end; ++t) {
    const dev::TileV<2> tt = dev::tilev_of<2>(G, t);
    const int cc = tt.c <= G.nx ? tt.c : (G.nx / 2) * 2;
    struct Side {
      VT u0, p, ky, bfp, cp, di;
    };
    auto side = [&](int row) -> Side {
      const long i = (long)row * P + cc;
      Side e;
      e.u0 = dev::ldv<NT, double, 2>(u0 + i);
      e.p = dev::ldv<NT, double, 2>(p + i);
      e.ky = dev::ldv<false, double, 2>(Ky + i);
      e.cp = dev::ldv<NT, double, 2>(cp + i);
      const bool head = (row & 3) == 0;  // first row of a block: Di, else bfp
      e.bfp = head ? VT(0) : dev::ldv<NT, double, 2>(bfp + i);
      e.di = head ? dev::ldv<NT, double, 2>(Di + i) : VT(0);
      return e;
    };
    auto ldop = [&](int col, int row) -> VT {
      return dev::ldv<NT, double, 2>(uin + (long)row * P + col);
    };
    auto ld1 = [&](int col, int row) -> double { return uin[(long)row * P + col]; };
    // the rows of the current block in walking order (shifted in), with
    // dp of the forward elimination, cp, p and u of each row
    VT dps[4], cps[4], pps[4], uus[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) dps[q] = cps[q] = pps[q] = uus[q] = VT(0);
    dev::walkv<NT, double, 2, false, TL_PF_CHEBY, true>(
        G, tt, Kx, Ky, rx, ry, ldop, ld1, side, [&](int y, VT v, VT wv, const Side &e) {
      if (tt.c >= G.nx) return;
      const long row = (long)y * P;
      const VT rv = e.u0 - wv;
      const int g0 = y & ~3;
      const int nq = G.ny - g0 < 4 ? G.ny - g0 : 4;
      // tea_block_solve's forward elimination, in its order of operations
      const VT dp = (y == g0) ? rv / e.di : (rv - (-e.ky * ry) * dps[3]) * e.bfp;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        dps[q] = dps[q + 1];
        cps[q] = cps[q + 1];
        pps[q] = pps[q + 1];
        uus[q] = uus[q + 1];
      }
      dps[3] = dp;
      cps[3] = e.cp;
      pps[3] = e.p;
      uus[3] = v;
      if (LAST) {
        dev::stv<false, double, 2>(w + row, tt.c, G.nx, wv);
        dev::stv<false, double, 2>(r + row, tt.c, G.nx, rv);
        acc += rv.x * rv.x;
        if (tt.c + 1 < G.nx) acc += rv.y * rv.y;
      }
      // carry the halo ring of u to the new buffer
      if (tt.c == 0) uout[row - 1] = uin[row - 1];
      if (tt.c + 2 >= G.nx) uout[row + G.nx] = uin[row + G.nx];
      for (int side_row = 0; side_row < 2; ++side_row) {
        if (y != (side_row ? G.ny - 1 : 0)) continue;
        const long hr = side_row ? (long)G.ny * P : -P;
        uout[hr + tt.c] = uin[hr + tt.c];
        if (tt.c + 1 < G.nx) uout[hr + tt.c + 1] = uin[hr + tt.c + 1];
      }
      if (y - g0 + 1 < nq) return;
      // the block is complete: its rows sit in slots 4 - nq .. 3; move them
      // to 0 .. nq - 1, then substitute back from the top row
      for (int k = nq; k < 4; ++k) {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          dps[q] = dps[q + 1];
          cps[q] = cps[q + 1];
          pps[q] = pps[q + 1];
          uus[q] = uus[q + 1];
        }
      }
      VT zn = VT(0);
#pragma unroll
      for (int q = 3; q >= 0; --q) {
        if (q >= nq) continue;
        const VT zq = (q == nq - 1) ? dps[q] : dps[q] - cps[q] * zn;
        const VT pv = alpha * pps[q] + beta * zq;
        const long ri = (long)(g0 + q) * P;
        dev::stv<false, double, 2>(p + ri, tt.c, G.nx, pv);
        dev::stv<false, double, 2>(uout + ri, tt.c, G.nx, uus[q] + pv);
        if (LAST) dev::stv<false, double, 2>(z + ri, tt.c, G.nx, zq);
        zn = zq;
      }
    }, 1);
  }
  acc = dev::block_sum(acc, lds);
  if (LAST && threadIdx.x == 0) out_part[blockIdx.x] = acc;
}

// Chebyshev iterations with the block preconditioner, one k_cheby_blk pass
// each (cheby_run_fused's contract).
int cheby_run_block(const double *ch_alphas, const double *ch_betas, int step_first, int count,
                    int want_norm, double *norm) {
  State &s = S();
  Geo G = geo(2 * BLOCK);
  // tiles start on the 4-row blocks of the preconditioner
  const int blocks = G.blocks < 0 ? -G.blocks : G.blocks;
  G.rows_per_tile = (G.rows_per_tile + 3) / 4 * 4;
  G.tiles = (long)G.tiles_x * ((G.ny + G.rows_per_tile - 1) / G.rows_per_tile);
  const int nb = (int)(G.tiles < blocks ? G.tiles : blocks);
  G.blocks = G.blocks < 0 ? -nb : nb;
  const bool nt = use_nontemporal();
  for (int i = 0; i < count; ++i) {
    const int st = step_first + i - 1;
    const int in = s.ucur, outb = 1 - s.ucur;
    const bool last = i + 1 == count;
    auto k = nt ? (last ? k_cheby_blk<true, 1> : k_cheby_blk<true, 0>)
                : (last ? k_cheby_blk<false, 1> : k_cheby_blk<false, 0>);
    hipEvent_t ev;
    prof_arm(3, &ev);
    klaunch(k, dim3(nb), dim3(BLOCK), G, s.u[in].p11, s.u[outb].p11,
            s.u0.p11, s.Kx.p11, s.Ky.p11, s.p[s.pcur].p11, s.w.p11, s.r.p11, s.z.p11,
            s.Di.p11, s.cp.p11, s.bfp.p11, s.rx, s.ry, ch_alphas[st], ch_betas[st],
            part(P0));
    prof_end(3, ev);
    if (hipGetLastError() != hipSuccess) return -fail(TL_ERR_HIP, "Chebyshev pass launch");
    s.ucur = outb;
  }
  if (want_norm && count > 0) {
    const int slot = P0;
    const int rc = finish_sums(&slot, 1, nb, norm);
    if (rc) return -rc;
  }
  return 1;
}

// Chebyshev iterations through the vector walk; returns 1 if it ran them,
// 0 if the configuration needs the per-routine kernel (internal faces, or
// reflective faces with the block preconditioner), or a negative error.
int cheby_run_fused(const double *ch_alphas, const double *ch_betas, int step_first, int count,
                    int want_norm, double *norm) {
  State &s = S();
  if (!s.zb_all) return 0;
  if (s.precond == TL_PREC_JAC_BLOCK)
    return s.reflective ? 0 : cheby_run_block(ch_alphas, ch_betas, step_first, count, want_norm,
                                              norm);
  const Geo G = geo_resident(2 * BLOCK, TL_CHEBY_PAIR_WAVES);
  const int nb = G.blocks < 0 ? -G.blocks : G.blocks;
  const int diag = s.precond == TL_PREC_JAC_DIAG;
  const bool nt = use_nontemporal();
  // reflective faces: one iteration per pass, reading u's halo cells
  const bool refl = s.reflective;
  auto k = nt ? (refl ? k_cheby_v<true, 2> : k_cheby_v<true, 0>)
              : (refl ? k_cheby_v<false, 2> : k_cheby_v<false, 0>);
  using K2 = decltype(&k_cheby_pair<true, 0, 0>);
  static const K2 k2tab[2][2][2] = {
      {{&k_cheby_pair<false, 0, 0>, &k_cheby_pair<false, 0, 1>},
       {&k_cheby_pair<false, 1, 0>, &k_cheby_pair<false, 1, 1>}},
      {{&k_cheby_pair<true, 0, 0>, &k_cheby_pair<true, 0, 1>},
       {&k_cheby_pair<true, 1, 0>, &k_cheby_pair<true, 1, 1>}}};
  const bool pairs = sweep_pairs() && !refl;
  for (int i = 0; i < count;) {
    const int st = step_first + i - 1;
    const int in = s.ucur, outb = 1 - s.ucur;
    hipEvent_t ev;
    prof_arm(3, &ev);
    if (pairs && i + 1 < count) {
      // two iterations per pass; p moves to the other buffer
      const int last = i + 2 == count;
      klaunch(k2tab[nt ? 1 : 0][diag][last], dim3(nb), dim3(BLOCK), G,
              s.u[in].p11, s.u[outb].p11, s.u0.p11, s.Kx.p11, s.Ky.p11,
              s.p[s.pcur].p11, s.p[1 - s.pcur].p11, s.w.p11, s.r.p11, s.z.p11,
              s.Mi.p11, s.rx, s.ry, ch_alphas[st], ch_betas[st], ch_alphas[st + 1],
              ch_betas[st + 1], part(P0), s.walk_dir);
      s.pcur = 1 - s.pcur;
      i += 2;
    } else {
      const int last = i + 1 == count;
      klaunch(k, dim3(nb), dim3(BLOCK), G, s.u[in].p11, s.u[outb].p11,
              s.u0.p11, s.Kx.p11, s.Ky.p11, s.p[s.pcur].p11, s.w.p11, s.r.p11,
              s.z.p11, s.Mi.p11, s.rx, s.ry, ch_alphas[st], ch_betas[st], diag, last,
              part(P0), s.walk_dir);
      i += 1;
    }
    s.walk_dir = -s.walk_dir;  // the next pass starts on the rows written last
    prof_end(3, ev);
    if (hipGetLastError() != hipSuccess) return -fail(TL_ERR_HIP, "Chebyshev pass launch");
    s.ucur = outb;
  }
  if (want_norm && count > 0) {
    const int slot = P0;
    const int rc = finish_sums(&slot, 1, nb, norm);
    if (rc) return -rc;
  }
  return 1;
}
}  // namespace tl

namespace {

// The FP64 fields as one buffer set of the shared loop kernels.
LoopBufs<double> fp64_bufs() {
  State &s = S();
  LoopBufs<double> b;
  b.G = geo(TL_V64 * BLOCK);
  b.u = s.u[s.ucur].p11;
  b.r = s.r.p11;
  b.w = s.w.p11;
  b.z = s.z.p11;
  b.r_store = s.r_store.p11;
  b.rtemp = s.rtemp.p11;
  b.utemp = s.utemp.p11;
  b.p[0] = s.p[0].p11;
  b.p[1] = s.p[1].p11;
  b.sd[0] = s.sd[0].p11;
  b.sd[1] = s.sd[1].p11;
  b.pcur = &s.pcur;
  b.sdcur = &s.sdcur;
  b.Mi = s.Mi.p11;
  b.Kx = s.Kx.p11;
  b.Ky = s.Ky.p11;
  b.rx = s.rx;
  b.ry = s.ry;
  b.diag = s.precond == TL_PREC_JAC_DIAG;
  b.block = s.precond == TL_PREC_JAC_BLOCK;
  b.Di = s.Di.p11;
  b.cp = s.cp.p11;
  b.bfp = s.bfp.p11;
  b.cg_src = (b.diag || b.block || s.ppcg_active) ? s.z.p11 : s.r.p11;
  if (b.block) {
    // tiles start on the 4-row blocks of the preconditioner
    Geo &G = b.G;
    const int blocks = G.blocks < 0 ? -G.blocks : G.blocks;
    G.rows_per_tile = (G.rows_per_tile + 3) / 4 * 4;
    G.tiles = (long)G.tiles_x * ((G.ny + G.rows_per_tile - 1) / G.rows_per_tile);
    const int nb = (int)(G.tiles < blocks ? G.tiles : blocks);
    G.blocks = G.blocks < 0 ? -nb : nb;
  }
  return b;
}

// A non-finite reduction on the per-routine path of a loop entry point.
int per_routine_nonfinite(tl_loop_state *st, const char *solver, int n, double v) {
  st->n_end = n;
  st->reason = TL_LOOP_NONFINITE;
  st->error = v;
  return nonfinite_fail(solver, n);
}

int check_state(tl_loop_state *st) {
  if (!st) return fail(TL_ERR_ARG, "null loop state");
  if (st->n_first < 1 || st->n_last < st->n_first)
    return fail(TL_ERR_ARG, "loop: need 1 <= n_first <= n_last");
  return TL_OK;
}

// The per-routine loop paths keep their exit-test errors on the host; they
// go to the device history, where tl_loop_errors finds the fused loops'.
template <typename F>
int with_errors(tl_loop_state *st, F run) {
  int rc = ensure_loop_buffers(st->n_last);
  if (rc) return rc;
  std::vector<double> errs;
  rc = run(errs);
  if (!errs.empty()) {
    State &s = S();
    TL_HIP(hipMemcpyAsync(loop_mem().herr + st->n_first, errs.data(),
                          sizeof(double) * errs.size(), hipMemcpyHostToDevice, s.stream));
    TL_HIP(hipStreamSynchronize(s.stream));
  }
  return rc;
}

// CG on a chunk with internal faces: the per-routine path, one iteration at
// a time; errs[n - n_first] gets the exit-test error of iteration n.
int cg_loop_per_routine(tl_loop_state *st, double *cg_alphas, double *cg_betas,
                        std::vector<double> &errs) {
  int rc;
  double rro = st->rro, old_error = st->old_error, err = 0.0;
  for (int n = st->n_first; n <= st->n_last; ++n) {
    if (n > st->n_first && st->check_switch && fabs(old_error) <= st->epslim &&
        n >= st->presteps) {
      st->n_end = n; st->reason = TL_LOOP_SWITCH; st->rro = rro; st->old_error = old_error;
      st->error = err;
      return TL_OK;
    }
    double pw, rrn;
    if ((rc = tl_cg_calc_w(&pw))) return rc;
    if (!std::isfinite(pw)) return per_routine_nonfinite(st, "CG", n, pw);
    const double alpha = rro / pw;
    if (cg_alphas) cg_alphas[n] = alpha;
    if ((rc = tl_cg_calc_ur(alpha, &rrn))) return rc;
    if (!std::isfinite(rrn)) return per_routine_nonfinite(st, "CG", n, rrn);
    const double beta = rrn / rro;
    if (cg_betas) cg_betas[n] = beta;
    if ((rc = tl_cg_calc_p(beta))) return rc;
    rro = rrn;
    err = sqrt(fabs(rrn));
    errs.push_back(err);
    if (fabs(err) < st->eps_abs) {
      st->n_end = n; st->reason = TL_LOOP_CONVERGED; st->rro = rro; st->error = err;
      return TL_OK;
    }
    old_error = err;
  }
  st->n_end = st->n_last; st->reason = TL_LOOP_MAXITERS; st->rro = rro; st->error = err;
  st->old_error = old_error;
  return TL_OK;
}

// PPCG without inner steps, or preconditioned with one or two: the
// per-routine path (tea_solve.f90:294-330), one outer iteration at a time.
int ppcg_loop_per_routine(tl_loop_state *st, const double *alphas, const double *betas, int inner,
                          double theta, int *inner_iters, std::vector<double> &errs) {
  int rc;
  double rro = st->rro, err = 0.0;
  for (int n = st->n_first; n <= st->n_last; ++n) {
    double pw, rrn, dummy;
    if ((rc = tl_cg_calc_w(&pw))) return rc;
    if ((rc = tl_ppcg_store_r())) return rc;
    if ((rc = tl_ppcg_calc_zrnorm(&rro))) return rc;
    if (!std::isfinite(pw) || !std::isfinite(rro))
      return per_routine_nonfinite(st, "PPCG", n, std::isfinite(pw) ? rro : pw);
    const double alpha = rro / pw;
    if ((rc = tl_cg_calc_ur(alpha, &dummy))) return rc;
    if (inner >= 0) {
      if ((rc = tl_ppcg_init_sd(theta))) return rc;
      if ((rc = tl_ppcg_inner(alphas, betas, 1, inner))) return rc;
      if ((rc = tl_ppcg_update_z())) return rc;
    }
    if (inner_iters) *inner_iters += inner;
    if ((rc = tl_ppcg_calc_rrn(&rrn))) return rc;
    if (!std::isfinite(rrn)) return per_routine_nonfinite(st, "PPCG", n, rrn);
    const double beta = rrn / rro;
    if ((rc = tl_cg_calc_p(beta))) return rc;
    rro = rrn;
    err = sqrt(fabs(rrn));
    errs.push_back(err);
    if (fabs(err) < st->eps_abs) {
      st->n_end = n; st->reason = TL_LOOP_CONVERGED; st->rro = rro; st->error = err;
      return TL_OK;
    }
    st->old_error = err;
  }
  st->n_end = st->n_last; st->reason = TL_LOOP_MAXITERS; st->rro = rro; st->error = err;
  return TL_OK;
}

}  // namespace

namespace tl {

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
// Nontemporal (streaming) accesses pay off once the working set of a
// loop (~6 fields) no longer fits the 256 MB Infinity Cache.
bool use_nontemporal() {
  static int forced = -2;
  if (forced == -2) {
    const char *e = getenv("TL_NT");
    forced = e ? atoi(e) : -1;
  }
  if (forced >= 0) return forced != 0;
  const State &s = S();
  return (double)s.nx * (double)s.ny * 8.0 * 6.0 > 256.0 * 1024 * 1024;
}

int batch_size() {
  static int b = 0;
  if (!b) {
    const char *e = getenv("TL_BATCH");
    b = e ? atoi(e) : 16;
    if (b < 1) b = 1;
    if (b > 4096) b = 4096;
  }
  return b;
}

int ensure_loop_buffers(long n_last) {
  State &s = S();
  const long need = n_last + 4;
  if (s.hist_len < need) {
    // forget the old buffer before allocating, so that a failed allocation
    // leaves no stale pointer or length behind
    double *old = s.hist;
    s.hist = nullptr;
    s.hist_len = 0;
    if (old) TL_HIP(hipFree(old));
    TL_HIP(hipMalloc((void **)&s.hist, sizeof(double) * 4 * need));
    s.hist_len = need;
  }
  return TL_OK;
}

LoopMem loop_mem() {
  State &s = S();
  LoopMem m;
  m.ctl = (Ctl *)s.ctl;
  m.hrro = s.hist;
  m.halpha = s.hist ? s.hist + s.hist_len : nullptr;
  m.hbeta = s.hist ? s.hist + 2 * s.hist_len : nullptr;
  m.herr = s.hist ? s.hist + 3 * s.hist_len : nullptr;
  m.parts = s.partials;
  return m;
}

int copy_hist(double *dst, const double *src_dev, int lo, int hi) {
  if (!dst || hi < lo) return TL_OK;
  State &s = S();
  TL_HIP(hipMemcpyAsync(dst + lo, src_dev + lo, sizeof(double) * (hi - lo + 1),
                        hipMemcpyDeviceToHost, s.stream));
  TL_HIP(hipStreamSynchronize(s.stream));
  return TL_OK;
}

}  // namespace tl

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" int tl_cg_loop(tl_loop_state *st, double *cg_alphas, double *cg_betas) {
  TL_REQUIRE_INIT();
  int rc = check_state(st);
  if (rc) return rc;
  State &s = S();
  if (!s.zb_all) return with_errors(st, [&](std::vector<double> &errs) {
    return cg_loop_per_routine(st, cg_alphas, cg_betas, errs);
  });
  LoopBufs<double> bb = fp64_bufs();
  return run_cg_loop<double, TL_V64>(bb, st, cg_alphas, cg_betas);
}

// ---------------------------------------------------------------------------
// tl_cg_loop_noalloc: the fused CG loop with its scalars (control block,
// per-iteration histories, partial-sum slots) in a file-scope device array
// instead of library allocations, for callers that cannot allocate.  Each
// device holds its own copy of the array until the process ends; every call
// on a device shares it, so calls that overlap on one device race.
// ---------------------------------------------------------------------------
constexpr int NOALLOC_MAX_ITERS = 16384;              // largest n_last accepted
constexpr long NOALLOC_HIST = NOALLOC_MAX_ITERS + 4;  // entries per history
constexpr size_t NOALLOC_CTL_BYTES = 256;
// 256 B + (4 x 16388 + 6 x 2048) doubles = 622,976 bytes per device
constexpr size_t NOALLOC_BYTES =
    NOALLOC_CTL_BYTES + sizeof(double) * (4 * NOALLOC_HIST + (size_t)NPART * MAX_GRID);
static_assert(sizeof(Ctl) <= NOALLOC_CTL_BYTES, "control block size");

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

namespace {

// Device address of `scratch` per device (the host-side name is not one).
constexpr int NOALLOC_MAX_DEVICES = 64;
void *scratch_addr[NOALLOC_MAX_DEVICES] = {nullptr};

int scratch_base(int device, void **out) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES)
    return fail(TL_ERR_ARG, "tl_cg_loop_noalloc: device index above 63");
  if (!scratch_addr[device]) {
    // the symbol's address on `device`: make it current for the query
    int cur = 0;
    TL_HIP(hipGetDevice(&cur));
    if (cur != device) TL_HIP(hipSetDevice(device));
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) TL_HIP(hipSetDevice(cur));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!p) return fail(TL_ERR_HIP, "hipGetSymbolAddress(scratch) returned null");
    scratch_addr[device] = p;
  }
  *out = scratch_addr[device];
  return TL_OK;
}

__global__ void k_zero_bytes(unsigned char *__restrict__ p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

}  // namespace

extern "C" int tl_cg_loop_noalloc(tl_loop_state *st, double *cg_alphas, double *cg_betas) {
  TL_REQUIRE_INIT();
  int rc = check_state(st);
  if (rc) return rc;
  if (st->n_last > NOALLOC_MAX_ITERS)
    return fail(TL_ERR_ARG, "tl_cg_loop_noalloc: n_last above " +
                                std::to_string(NOALLOC_MAX_ITERS) + " (its history size)");
  State &s = S();
  if (!s.zb_all)
    return fail(TL_ERR_ARG,
                "tl_cg_loop_noalloc: needs a chunk with all four faces external (fused path)");
  void *base = nullptr;
  if ((rc = scratch_base(s.device, &base))) return rc;
  LoopMem m;
  m.ctl = (Ctl *)base;
  m.hrro = (double *)((unsigned char *)base + NOALLOC_CTL_BYTES);
  m.halpha = m.hrro + NOALLOC_HIST;
  m.hbeta = m.halpha + NOALLOC_HIST;
  m.herr = m.hbeta + NOALLOC_HIST;
  m.parts = m.herr + NOALLOC_HIST;
  // the control block starts cleared (a kernel, not a memset node)
  hipLaunchKernelGGL(k_zero_bytes, dim3(1), dim3(64), 0, s.stream, (unsigned char *)m.ctl,
                     (int)NOALLOC_CTL_BYTES);
  TL_HIP(hipGetLastError());
  // kernel timing creates events on demand: off for this call
  const bool prof = s.prof;
  s.prof = false;
  LoopBufs<double> bb = fp64_bufs();
  rc = run_cg_loop<double, TL_V64>(bb, st, cg_alphas, cg_betas, nullptr, &m);
  s.prof = prof;
  return rc;
}

extern "C" int tl_ppcg_loop(tl_loop_state *st, const double *alphas, const double *betas,
                            int inner, double theta, int *inner_iters) {
  TL_REQUIRE_INIT();
  int rc = check_state(st);
  if (rc) return rc;
  if (!alphas || !betas) return fail(TL_ERR_ARG, "tl_ppcg_loop: null coefficients");
  State &s = S();
  // per routine: no inner steps, or a preconditioner with one or two inner
  // steps, where the fused pass that reads z (M^-1 r) as the first sweep's
  // operand would also write z as the last sweep's result, racing with the
  // neighbouring tiles' reads
  if (inner < 1 || (s.precond != TL_PREC_NONE && inner <= 2))
    return with_errors(st, [&](std::vector<double> &errs) {
      return ppcg_loop_per_routine(st, alphas, betas, inner, theta, inner_iters, errs);
    });
  // z.r entering the first iteration (tea_leaf_ppcg_calc_zrnorm); later
  // iterations take it from the last inner sweep of the one before
  double rro_first;
  if ((rc = tl_ppcg_calc_zrnorm(&rro_first))) return rc;
  LoopBufs<double> bb = fp64_bufs();
  return run_ppcg_loop<double, TL_V64>(bb, st, alphas, betas, inner, theta, rro_first, nullptr,
                                  inner_iters);
}

namespace {
// Jacobi iterations two per pass (k_jl_pair).  Launch j covers iterations
// n = n_first + 2j and n + 1, reading u from u[in_j] and writing iteration n
// to the spare p buffer ("mid") and n + 1 to u[1 - in_j]; an odd count ends
// with one k_jl_v iteration.  The partials of launch j go to P2/P3 (even j)
// or P4/P5 (odd j), so a launch never overwrites what the next one reads
// while an earlier block still reads it.  On exit at the first iteration of
// a pass, its u is copied from mid into the u buffer the loop leaves current.
int jacobi_loop_pairs(tl_loop_state *st) {
  State &s = S();
  const Geo G = geo(2 * BLOCK);
  const int nb = G.blocks < 0 ? -G.blocks : G.blocks;
  const int c0 = s.ucur;
  double *mid = s.p[1 - s.pcur].p11;  // not visible through TL_FIELD_P
  Ctl *ctl = (Ctl *)s.ctl;
  double *herr = loop_mem().herr;
  const bool nt = use_nontemporal();
  const bool ze = s.reflective;
  auto kp = nt ? (ze ? k_jl_pair<true, true> : k_jl_pair<true, false>)
               : (ze ? k_jl_pair<false, true> : k_jl_pair<false, false>);
  auto kv = nt ? (ze ? k_jl_v<true, true> : k_jl_v<true, false>)
               : (ze ? k_jl_v<false, true> : k_jl_v<false, false>);
  const int n_first = st->n_first, n_last = st->n_last;
  auto slot_a = [](int j) { return part((j & 1) ? P4 : P2); };
  auto slot_b = [](int j) { return part((j & 1) ? P5 : P3); };
  auto is_single = [&](int j) { return n_first + 2 * j == n_last; };
  auto launch = [&](int n, bool first) -> int {
    const int off = n - n_first;
    if (off & 1) return TL_OK;  // the second iteration of the pair at n - 1
    const int j = off / 2, in = c0 ^ (j & 1);
    const double *pa = first ? nullptr : slot_a(j - 1), *pb = first ? nullptr : slot_b(j - 1);
    const int dir = (j & 1) ? -1 : 1;
    // a single iteration after a pair: the exit test of the pair first
    if (is_single(j) && !first)
      hipLaunchKernelGGL(k_jl_pair_test, dim3(1), dim3(BLOCK), 0, s.stream, n, st->eps_abs, pa,
                         pb, nb, ctl, herr, 0);
    hipEvent_t ev;
    prof_arm(4, &ev);
    if (!is_single(j)) {
      klaunch(kp, dim3(nb), dim3(BLOCK), G, n, first ? 1 : 0, st->eps_abs, s.u[in].p11, mid,
              s.u[1 - in].p11, s.u0.p11, s.Kx.p11, s.Ky.p11, s.rx, s.ry, pa, pb, nb, slot_a(j),
              slot_b(j), ctl, herr, dir);
    } else if (first) {
      klaunch(kv, dim3(nb), dim3(BLOCK), G, n, 1, st->eps_abs, s.u[in].p11, s.u[1 - in].p11,
              s.u0.p11, s.Kx.p11, s.Ky.p11, s.rx, s.ry, slot_b(j), nb, slot_b(j), ctl, herr, dir);
    } else {
      klaunch(kv, dim3(nb), dim3(BLOCK), G, n, 0, st->eps_abs, s.u[in].p11, s.u[1 - in].p11,
              s.u0.p11, s.Kx.p11, s.Ky.p11, s.rx, s.ry, pb, nb, slot_b(j), ctl, herr, dir);
    }
    prof_end(4, ev);
    TL_HIP(hipGetLastError());
    return TL_OK;
  };
  auto tail = [&](int n) -> int {
    const int j = (n_last - n_first) / 2;  // the last launch
    hipLaunchKernelGGL(k_jl_pair_test, dim3(1), dim3(BLOCK), 0, s.stream, n, st->eps_abs,
                       is_single(j) ? nullptr : slot_a(j), slot_b(j), nb, ctl, herr, 1);
    TL_HIP(hipGetLastError());
    return TL_OK;
  };
  Ctl res;
  int rc;
  if ((rc = drive(n_first, n_last, launch, tail, &res, ctl))) return rc;
  st->n_end = res.n_end;
  st->reason = res.reason;
  st->error = res.error;
  if (res.reason == TL_LOOP_NONFINITE) return nonfinite_fail("Jacobi", res.n_end);
  const int off = res.n_end - n_first, j = off / 2, in = c0 ^ (j & 1);
  s.ucur = 1 - in;  // launch j wrote its last iteration to u[1 - in_j]
  if ((off & 1) == 0 && !is_single(j)) {
    // the loop ended at the first iteration of pair j, which is in mid
    TL_HIP(hipMemcpy2DAsync(s.u[s.ucur].p11, s.pitch * sizeof(double), mid,
                            s.pitch * sizeof(double), s.nx * sizeof(double), s.ny,
                            hipMemcpyDeviceToDevice, s.stream));
  }
  return TL_OK;
}
}  // namespace

extern "C" int tl_loop_errors(int lo, int hi, double *err) {
  TL_REQUIRE_INIT();
  State &s = S();
  if (!err) return fail(TL_ERR_ARG, "tl_loop_errors: null array");
  if (lo < 1 || hi < lo || !s.hist || hi + 1 >= s.hist_len)
    return fail(TL_ERR_ARG, "tl_loop_errors: iterations outside the last loop's history");
  return copy_hist(err, loop_mem().herr, lo, hi);
}

extern "C" int tl_jacobi_loop(tl_loop_state *st) {
  TL_REQUIRE_INIT();
  int rc = check_state(st);
  if (rc) return rc;
  if ((rc = ensure_loop_buffers(st->n_last))) return rc;
  State &s = S();
  // the vector walk where every face is external (reflective faces read zero
  // outside the mesh, as the reference reads the never-written halo of un);
  // the per-cell walk otherwise
  const bool vec = s.zb_all;
  if (vec && sweep_pairs()) return jacobi_loop_pairs(st);
  const Geo G = vec ? geo(2 * BLOCK) : geo(BLOCK);
  const int nb = G.blocks < 0 ? -G.blocks : G.blocks;
  const int u0c = s.ucur;
  Ctl *ctl = (Ctl *)s.ctl;
  double *herr = loop_mem().herr;
  const bool ze = s.reflective;
  auto kv = use_nontemporal() ? (ze ? k_jl_v<true, true> : k_jl_v<true, false>)
                              : (ze ? k_jl_v<false, true> : k_jl_v<false, false>);
  auto launch = [&](int n, bool first) -> int {
    const int in = u0c ^ ((n - st->n_first) & 1);
    hipEvent_t ev;
    prof_arm(4, &ev);
    if (vec) {
      // consecutive sweeps walk their rows in opposite directions
      const int dir = ((n - st->n_first) & 1) ? -1 : 1;
      klaunch(kv, dim3(nb), dim3(BLOCK), G, n, first ? 1 : 0,
              st->eps_abs, s.u[in].p11, s.u[1 - in].p11, s.u0.p11, s.Kx.p11, s.Ky.p11,
              s.rx, s.ry, part((n - 1) & 1 ? P4 : P5), nb, part(n & 1 ? P4 : P5), ctl,
              herr, dir);
    } else {
      klaunch(k_jl, dim3(nb), dim3(BLOCK), G, n, first ? 1 : 0,
              st->eps_abs, s.u[in].p11, s.u[1 - in].p11, s.u0.p11, s.Kx.p11,
              s.Ky.p11, s.rx, s.ry, part((n - 1) & 1 ? P4 : P5), nb,
              part(n & 1 ? P4 : P5), ctl, herr);
    }
    prof_end(4, ev);
    TL_HIP(hipGetLastError());
    return TL_OK;
  };
  auto tail = [&](int n) -> int {
    hipLaunchKernelGGL(k_jl_tail, dim3(1), dim3(BLOCK), 0, s.stream, n, st->eps_abs,
                       part((n - 1) & 1 ? P4 : P5), nb, ctl, herr);
    TL_HIP(hipGetLastError());
    return TL_OK;
  };
  Ctl res;
  if ((rc = drive(st->n_first, st->n_last, launch, tail, &res, ctl))) return rc;
  st->n_end = res.n_end;
  st->reason = res.reason;
  st->error = res.error;
  if (res.reason == TL_LOOP_NONFINITE) return nonfinite_fail("Jacobi", res.n_end);
  s.ucur = u0c ^ ((res.n_end - st->n_first + 1) & 1);
  return TL_OK;
}
