// Two consecutive Chebyshev iterations in one pass over the mesh (temporal
// blocking).  One iteration (kernels/tea_leaf_cheby_kernel.f90:111-200) is
//   w = A u;  r = u0 - w;  p' = alpha p + beta r (or beta Mi r);  u' = u + p'
// and the second needs u' at the neighbours of every cell, so a block
// computes the first iteration ("stage A") one row ahead of the second
// ("stage B") in its walking direction, on one ghost row beyond each end of
// its tile and, in the lanes at the block's edges, on the ghost column just
// outside the block.  Ghost results are used and discarded.  u and p are read
// from one buffer and written to another, because ghost cells read u and p
// that a neighbouring block rewrites.  Stage B stores what two separate
// iterations would store; a pass moves the bytes of one iteration.
#pragma once

#include "tl_sweep2.hpp"

namespace tl {
namespace dev {

// The per-cell inputs of an iteration other than u: u0, p and Mi.
template <typename X>
struct ChebyIn {
  X u0, p, mi;
};

// Walk one tile.  For each row yb of the tile (in walking order),
// body(yb, u1, p1, w, in) receives stage A's results on that row (u' and p'),
// A u' (for stage B) and the row's u0 and Mi (in.u0, in.mi).  opA(col, row)
// / opA1(col, row) load u (V values / one value); side(col, row) /
// side1(col, row) load u0, p and Mi (ChebyIn).  The body must not load from
// memory written by this pass.
template <bool NT, int V, int DIAG, typename OPA, typename OPA1, typename SIDE, typename SIDE1,
          typename BODY>
__device__ inline void walk_cheby_pair(const Geo &g, const TileV<V> &tt,
                                       const double *__restrict__ Kx,
                                       const double *__restrict__ Ky, double rx,
                                       double ry, double a1, double b1, OPA opA, OPA1 opA1,
                                       SIDE side, SIDE1 side1, BODY body, int dir) {
  typedef double T;
  typedef typename Vec<T, V>::type VT;
  constexpr int NW = BLOCK / WAVE;
  // per row parity and wave: stage A first and last value, stage A first Kx,
  // stage B first and last value
  __shared__ T xch[2][NW][5];
  const long P = g.pitch;
  const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
  const int c = tt.c, nx = g.nx, ny = g.ny;
  if (tt.y1 <= tt.y0) return;
  const bool fwd = dir > 0;
  const int cl = c <= nx ? c : (nx / V) * V;  // clamped column of this lane
  const int x0 = tt.x0 - wave * WAVE * V;     // first column of the block
  const int xe = x0 + NW * WAVE * V;          // first column right of the block
  const bool lane_l = wave == 0 && lane == 0, lane_r = wave == NW - 1 && lane == WAVE - 1;
  const bool ghost_l = lane_l && x0 > 0;      // ghost column x0 - 1
  const bool ghost_r = lane_r && xe < nx;     // ghost column xe
  const bool ghost = ghost_l || ghost_r;
  const bool kx_r = lane_r && xe <= nx;       // Kx face right of the block
  const int gx = ghost_l ? x0 - 1 : xe;       // this lane's ghost column
  auto in_mesh = [&](int y) { return y >= 0 && y < ny; };
  // stage B rows: the tile's rows; stage A rows: those plus the row beyond
  // each end, where it exists
  const int b_first = fwd ? tt.y0 : tt.y1 - 1;
  const int b_last = fwd ? tt.y1 - 1 : tt.y0;
  const int a_first = in_mesh(b_first - dir) ? b_first - dir : b_first;
  const int a_last = in_mesh(b_last + dir) ? b_last + dir : b_last;
  const int m = (a_last - a_first) * dir + 1;       // stage A rows
  const int iters = m + (a_last == b_last ? 1 : 0);  // + a stage-B-only step
  // Ky faces behind / ahead of row y in walking direction
  auto face_b = [&](int y) -> long { return (long)(fwd ? y : y + 1) * P; };
  auto face_a = [&](int y) -> long { return (long)(fwd ? y + 1 : y) * P; };

  // stage A state (row ya): u behind / centre, face behind
  VT sc = opA(cl, a_first);
  VT sb = in_mesh(a_first - dir) ? opA(cl, a_first - dir) : sc;
  VT kb = ldv<NT, T, V>(Ky + face_b(a_first) + cl);
  T gb = 0, gc = 0, gkb = 0;
  if (ghost) {
    gc = opA1(gx, a_first);
    gb = in_mesh(a_first - dir) ? opA1(gx, a_first - dir) : gc;
    gkb = Ky[face_b(a_first) + gx];
  }
  // stage B state (row yb): u' behind / at yb, p' at yb, and the faces and
  // inputs of yb
  VT u1b = VT(0), u1c = VT(0), p1c = VT(0);
  VT kxB = VT(0), kbB = VT(0), kaB = VT(0);
  ChebyIn<VT> inB = {VT(0), VT(0), VT(0)};
  T gu1c = 0, kxrB = 0;
  for (int k = 0; k < iters; ++k) {
    const bool doA = k < m;
    const int ya = a_first + k * dir;
    const int yb = ya - dir;
    const bool doB = k >= 1 && (yb - b_first) * dir >= 0 && (b_last - yb) * dir >= 0;
    // ---- loads of stage A's row
    VT sa = sc, kx = VT(0), ka = kb;
    ChebyIn<VT> in = {VT(0), VT(0), VT(0)};
    ChebyIn<T> gin = {0, 0, 0};
    T ga = gc, gka = gkb, gout = gc, gkl = 0, gkr = 0, ekr = 0;
    if (doA) {
      sa = in_mesh(ya + dir) ? opA(cl, ya + dir) : sc;
      kx = ldv<NT, T, V>(Kx + (long)ya * P + cl);
      ka = ldv<NT, T, V>(Ky + face_a(ya) + cl);
      in = side(cl, ya);
      if (ghost) {
        ga = in_mesh(ya + dir) ? opA1(gx, ya + dir) : gc;
        gka = Ky[face_a(ya) + gx];
        const int go = ghost_l ? gx - 1 : gx + 1;  // the ghost's outer neighbour
        gout = (go >= 0 && go < nx) ? opA1(go, ya) : gc;
        gkl = Kx[(long)ya * P + gx];
        gkr = Kx[(long)ya * P + gx + 1];
        gin = side1(gx, ya);
      }
      if (kx_r) ekr = Kx[(long)ya * P + xe];
    }
    // ---- exchange the edge values of both stages between the waves
    const int par = k & 1;
    if (lane == 0) {
      xch[par][wave][0] = sc[0];
      xch[par][wave][2] = kx[0];
      xch[par][wave][3] = u1c[0];
    }
    if (lane == WAVE - 1) {
      xch[par][wave][1] = sc[V - 1];
      xch[par][wave][4] = u1c[V - 1];
    }
    __syncthreads();
    T leftA = from_lower_lane(sc[V - 1]), rightA = from_upper_lane(sc[0]);
    T kxrA = from_upper_lane(kx[0]);
    T leftB = from_lower_lane(u1c[V - 1]), rightB = from_upper_lane(u1c[0]);
    if (lane == 0 && wave > 0) {
      leftA = xch[par][wave - 1][1];
      leftB = xch[par][wave - 1][4];
    }
    if (lane == WAVE - 1 && wave < NW - 1) {
      rightA = xch[par][wave + 1][0];
      kxrA = xch[par][wave + 1][2];
      rightB = xch[par][wave + 1][3];
    }
    if (kx_r) kxrA = ekr;
    // ---- stage A on row ya: own columns, then the ghost column
    VT u1a = u1c, p1a = p1c;  // stage A's results ahead of yb (the mirror when absent)
    T gu1a = gu1c;
    if (doA) {
      if (ghost_l) leftA = gc;
      if (ghost_r) rightA = gc;
      if (c == 0) leftA = sc[0];  // mirror at the left edge
      const VT &vd = fwd ? sb : sa, &vu = fwd ? sa : sb;
      const VT &kyd = fwd ? kb : ka, &kyu = fwd ? ka : kb;
      VT w;
#pragma unroll
      for (int e = 0; e < V; ++e) {
        const int col = c + e;
        const T l = (e == 0) ? leftA : sc[e - 1];
        T r;
        if (e + 1 < V) r = (col + 1 < nx) ? sc[e + 1] : sc[e];
        else r = (col + 1 < nx) ? rightA : sc[e];
        const T kr = (e + 1 < V) ? kx[e + 1] : kxrA;
        w[e] = op_cell<false>(sc[e], l, r, vd[e], vu[e], kx[e], kr, kyd[e], kyu[e], rx, ry);
      }
      const VT ra = in.u0 - w;
      p1a = DIAG ? a1 * in.p + b1 * (in.mi * ra) : a1 * in.p + b1 * ra;
      u1a = sc + p1a;
      if (ghost) {
        // inner neighbour: the block's edge cell; outer: gout
        const T inner = ghost_l ? sc[0] : sc[V - 1];
        const T gl = ghost_l ? gout : inner, gr = ghost_l ? inner : gout;
        const T gd = fwd ? gb : ga, gu = fwd ? ga : gb;
        const T gkd = fwd ? gkb : gka, gku = fwd ? gka : gkb;
        const T gw = op_cell<false>(gc, gl, gr, gd, gu, gkl, gkr, gkd, gku, rx, ry);
        const T gra = gin.u0 - gw;
        const T gp1 = DIAG ? a1 * gin.p + b1 * (gin.mi * gra) : a1 * gin.p + b1 * gra;
        gu1a = gc + gp1;
      }
    }
    // ---- stage B on row yb
    if (doB) {
      const VT &u1d = in_mesh(yb - dir) ? u1b : u1c;  // behind yb (mirrored)
      const VT &u1u = doA ? u1a : u1c;                // ahead of yb (mirrored)
      if (ghost_l) leftB = gu1c;
      if (ghost_r) rightB = gu1c;
      if (c == 0) leftB = u1c[0];
      const VT &vd = fwd ? u1d : u1u, &vu = fwd ? u1u : u1d;
      const VT &kyd = fwd ? kbB : kaB, &kyu = fwd ? kaB : kbB;
      VT w;
#pragma unroll
      for (int e = 0; e < V; ++e) {
        const int col = c + e;
        const T l = (e == 0) ? leftB : u1c[e - 1];
        T r;
        if (e + 1 < V) r = (col + 1 < nx) ? u1c[e + 1] : u1c[e];
        else r = (col + 1 < nx) ? rightB : u1c[e];
        const T kr = (e + 1 < V) ? kxB[e + 1] : kxrB;
        w[e] = op_cell<false>(u1c[e], l, r, vd[e], vu[e], kxB[e], kr, kyd[e], kyu[e], rx, ry);
      }
      body(yb, u1c, p1c, w, inB);
    }
    // ---- advance: stage A's row becomes stage B's row
    u1b = u1c;
    u1c = u1a;
    p1c = p1a;
    gu1c = gu1a;
    kxB = kx;
    kxrB = kxrA;
    kbB = kb;
    kaB = ka;
    inB = in;
    sb = sc;
    sc = sa;
    kb = ka;
    gb = gc;
    gc = ga;
    gkb = gka;
  }
}

}  // namespace dev
}  // namespace tl
