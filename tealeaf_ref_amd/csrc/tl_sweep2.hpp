// Two consecutive PPCG inner sweeps in one pass over the mesh (temporal
// blocking).  One sweep (kernels/tea_leaf_ppcg_kernel.f90:288-306) is
//   rtemp -= A sd;  sd' = alpha sd + beta (Mi) rtemp;  utemp += sd'
// and the second needs sd' at the neighbours of every cell, so a block
// computes the first sweep ("stage A") one row ahead of the second ("stage
// B") in its walking direction, on one ghost row beyond each end of its tile
// and, in the lanes at the block's edges, on the ghost column just outside
// the block.  Ghost results are used and discarded; rtemp is read from one
// buffer and written to another because ghost cells read rtemp that a
// neighbouring block rewrites.  Stage B stores exactly what two separate
// sweeps would store, bit for bit; a pass moves the bytes of one sweep.
#pragma once

#include "tl_stencilv.hpp"

namespace tl {
namespace dev {

// The operator at one cell from its centre, neighbours and faces.
template <bool DIFF, typename T>
__device__ inline T op_cell(T c, T l, T r, T d, T u, T kl, T kr, T kd, T ku, T rx, T ry) {
  return DIFF ? apply_diff<T>(c, l, r, d, u, kl, kr, kd, ku, rx, ry)
              : apply_t<T>(c, l, r, d, u, kl, kr, kd, ku, rx, ry);
}

// Walk one tile.  For each row yb of the tile (in walking order),
// body(yb, s1, r1, s0, w, mi, x) receives stage A's results on that row (sd'
// and rtemp'), stage A's operand there (sd), A sd' (for stage B), Mi (DIAG)
// and x = xtra(col, yb), the body's own loads of row yb (e.g. utemp).
// opA(col, row) / opA1(col, row) load stage A's operand (V values / one
// value), rtA / rtA1 its rtemp (r on the first sweep).  The body must not
// load from memory written by this pass.
// Every load of an iteration (stage A's row, the ghost cells, xtra of stage
// B's row) is issued before its barrier and arithmetic; with PF stage A's
// loads are issued one iteration earlier still and xtra two (a row of loads
// in flight behind the arithmetic, at the cost of the registers that hold
// it).
template <bool NT, typename T, int V, bool DIFF, int DIAG, bool PF, typename OPA, typename OPA1,
          typename RTA, typename RTA1, typename XTRA, typename BODY>
__device__ inline void walk_pair(const Geo &g, const TileV<V> &tt, const T *__restrict__ Kx,
                                 const T *__restrict__ Ky, const T *__restrict__ Mi, T rx, T ry,
                                 T a1, T b1, OPA opA, OPA1 opA1, RTA rtA, RTA1 rtA1, XTRA xtra,
                                 BODY body, int dir) {
  typedef typename Vec<T, V>::type VT;
  typedef decltype(xtra(0, 0)) X;
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
  // rows of stage B that exist: the tile's rows
  auto is_b = [&](int y) { return (y - b_first) * dir >= 0 && (b_last - y) * dir >= 0; };

  // the loads of one stage-A row ya (and xtra of the stage-B row it becomes)
  struct RowIn {
    VT sa, kx, ka, rt, mi;
    T ga, gka, gout, gkl, gkr, grt, gmi, ekr;
    X xb;
  };
  auto load_row = [&](int ya, const VT &sc_, T gc_, RowIn &o) {
    o.sa = in_mesh(ya + dir) ? opA(cl, ya + dir) : sc_;
    o.kx = ldv<NT, T, V>(Kx + (long)ya * P + cl);
    o.ka = ldv<NT, T, V>(Ky + face_a(ya) + cl);
    o.rt = rtA(cl, ya);
    if (DIAG) o.mi = ldv<NT, T, V>(Mi + (long)ya * P + cl);
    if (ghost) {
      o.ga = in_mesh(ya + dir) ? opA1(gx, ya + dir) : gc_;
      o.gka = Ky[face_a(ya) + gx];
      const int go = ghost_l ? gx - 1 : gx + 1;  // the ghost's outer neighbour
      o.gout = (go >= 0 && go < nx) ? opA1(go, ya) : gc_;
      o.gkl = Kx[(long)ya * P + gx];
      o.gkr = Kx[(long)ya * P + gx + 1];
      o.grt = rtA1(gx, ya);
      if (DIAG) o.gmi = Mi[(long)ya * P + gx];
    }
    if (kx_r) o.ekr = Kx[(long)ya * P + xe];
  };

  // stage A state (row ya): operand behind / centre, face behind
  VT sc = opA(cl, a_first);
  VT sb = in_mesh(a_first - dir) ? opA(cl, a_first - dir) : sc;
  VT kb = ldv<NT, T, V>(Ky + face_b(a_first) + cl);
  T gb = 0, gc = 0, gkb = 0;
  if (ghost) {
    gc = opA1(gx, a_first);
    gb = in_mesh(a_first - dir) ? opA1(gx, a_first - dir) : gc;
    gkb = Ky[face_b(a_first) + gx];
  }
  // stage B state (row yb): stage A's results behind / at yb, and the faces,
  // Mi, operand and xtra of stage A at yb
  VT s1b = VT(0), s1c = VT(0), r1c = VT(0), s0c = VT(0);
  VT kxB = VT(0), kbB = VT(0), kaB = VT(0), miB = VT(0);
  T gs1c = 0, kxrB = 0;
  X xB = X();
  RowIn nx_in;  // PF: the loads of the next iteration's row, in flight
  if (PF) {
    nx_in.ekr = 0;
    load_row(a_first, sc, gc, nx_in);
    if (is_b(a_first)) nx_in.xb = xtra(cl, a_first);
  }
  for (int k = 0; k < iters; ++k) {
    const bool doA = k < m;
    const int ya = a_first + k * dir;
    const int yb = ya - dir;
    const bool doB = k >= 1 && is_b(yb);
    // ---- loads of stage A's row (issued now, or one iteration ago with PF)
    RowIn in;
    in.sa = sc; in.kx = VT(0); in.ka = kb; in.rt = VT(0); in.mi = VT(0);
    in.ga = gc; in.gka = gkb; in.gout = gc; in.gkl = 0; in.gkr = 0; in.grt = 0; in.gmi = 0;
    in.ekr = 0;
    in.xb = X();
    if (PF) {
      if (doA) in = nx_in;
      if (k + 1 < m) {
        nx_in.ekr = 0;
        load_row(ya + dir, in.sa, in.ga, nx_in);
        if (is_b(ya + dir)) nx_in.xb = xtra(cl, ya + dir);
      }
    } else {
      if (doA) load_row(ya, sc, gc, in);
      if (doB) xB = xtra(cl, yb);
    }
    // ---- exchange the edge values of both stages between the waves
    const int par = k & 1;
    if (lane == 0) {
      xch[par][wave][0] = sc[0];
      xch[par][wave][2] = in.kx[0];
      xch[par][wave][3] = s1c[0];
    }
    if (lane == WAVE - 1) {
      xch[par][wave][1] = sc[V - 1];
      xch[par][wave][4] = s1c[V - 1];
    }
    __syncthreads();
    T leftA = from_lower_lane(sc[V - 1]), rightA = from_upper_lane(sc[0]);
    T kxrA = from_upper_lane(in.kx[0]);
    T leftB = from_lower_lane(s1c[V - 1]), rightB = from_upper_lane(s1c[0]);
    if (lane == 0 && wave > 0) {
      leftA = xch[par][wave - 1][1];
      leftB = xch[par][wave - 1][4];
    }
    if (lane == WAVE - 1 && wave < NW - 1) {
      rightA = xch[par][wave + 1][0];
      kxrA = xch[par][wave + 1][2];
      rightB = xch[par][wave + 1][3];
    }
    if (kx_r) kxrA = in.ekr;
    // ---- stage A on row ya: own columns, then the ghost column
    VT s1a = s1c;  // stage A's result ahead of yb (the mirror when absent)
    VT r1a = r1c;
    T gs1a = gs1c;
    if (doA) {
      if (ghost_l) leftA = gc;
      if (ghost_r) rightA = gc;
      if (c == 0) leftA = sc[0];  // mirror at the left edge
      const VT &vd = fwd ? sb : in.sa, &vu = fwd ? in.sa : sb;
      const VT &kyd = fwd ? kb : in.ka, &kyu = fwd ? in.ka : kb;
      VT w;
#pragma unroll
      for (int e = 0; e < V; ++e) {
        const int col = c + e;
        const T l = (e == 0) ? leftA : sc[e - 1];
        T r;
        if (e + 1 < V) r = (col + 1 < nx) ? sc[e + 1] : sc[e];
        else r = (col + 1 < nx) ? rightA : sc[e];
        const T kr = (e + 1 < V) ? in.kx[e + 1] : kxrA;
        w[e] = op_cell<DIFF>(sc[e], l, r, vd[e], vu[e], in.kx[e], kr, kyd[e], kyu[e], rx, ry);
      }
      r1a = in.rt - w;
      s1a = DIAG ? a1 * sc + (b1 * in.mi) * r1a : a1 * sc + b1 * r1a;
      if (ghost) {
        // inner neighbour: the block's edge cell; outer: gout
        const T inner = ghost_l ? sc[0] : sc[V - 1];
        const T gl = ghost_l ? in.gout : inner, gr = ghost_l ? inner : in.gout;
        const T gd = fwd ? gb : in.ga, gu = fwd ? in.ga : gb;
        const T gkd = fwd ? gkb : in.gka, gku = fwd ? in.gka : gkb;
        const T gw = op_cell<DIFF>(gc, gl, gr, gd, gu, in.gkl, in.gkr, gkd, gku, rx, ry);
        const T grt1 = in.grt - gw;
        gs1a = DIAG ? a1 * gc + (b1 * in.gmi) * grt1 : a1 * gc + b1 * grt1;
      }
    }
    // ---- stage B on row yb
    if (doB) {
      const VT &s1d = in_mesh(yb - dir) ? s1b : s1c;  // behind yb (mirrored)
      const VT &s1u = doA ? s1a : s1c;                // ahead of yb (mirrored)
      if (ghost_l) leftB = gs1c;
      if (ghost_r) rightB = gs1c;
      if (c == 0) leftB = s1c[0];
      const VT &vd = fwd ? s1d : s1u, &vu = fwd ? s1u : s1d;
      const VT &kyd = fwd ? kbB : kaB, &kyu = fwd ? kaB : kbB;
      VT w;
#pragma unroll
      for (int e = 0; e < V; ++e) {
        const int col = c + e;
        const T l = (e == 0) ? leftB : s1c[e - 1];
        T r;
        if (e + 1 < V) r = (col + 1 < nx) ? s1c[e + 1] : s1c[e];
        else r = (col + 1 < nx) ? rightB : s1c[e];
        const T kr = (e + 1 < V) ? kxB[e + 1] : kxrB;
        w[e] = op_cell<DIFF>(s1c[e], l, r, vd[e], vu[e], kxB[e], kr, kyd[e], kyu[e], rx, ry);
      }
      body(yb, s1c, r1c, s0c, w, miB, xB);
    }
    // ---- advance: stage A's row becomes stage B's row
    s1b = s1c;
    s1c = s1a;
    r1c = r1a;
    gs1c = gs1a;
    s0c = sc;
    kxB = in.kx;
    kxrB = kxrA;
    kbB = kb;
    kaB = in.ka;
    miB = in.mi;
    if (PF) xB = in.xb;
    sb = sc;
    sc = in.sa;
    kb = in.ka;
    gb = gc;
    gc = in.ga;
    gkb = in.gka;
  }
}

}  // namespace dev
}  // namespace tl
