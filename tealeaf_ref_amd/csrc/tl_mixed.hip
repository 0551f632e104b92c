// Mixed precision: FP32 inner solves under FP64 iterative refinement.
//
// Outer loop (FP64):  r = u0 - A u;  stop when ||r|| <= rel_tol ||r_start||;
//                     solve A d = r in FP32;  u += d.
// Inner solve (FP32 fields and stencil, FP64 reductions and scalars): the
// same CG / PPCG loop kernels as the FP64 solver, four columns per lane.
// The PPCG polynomial preconditioner uses Chebyshev coefficients from FP64
// eigenvalue bounds of the Lanczos tridiagonal of an FP32 CG solve, as the
// reference derives them from its CG presteps (tea_leaf_cheby.f90:142-205,
// scaled as tea_solve.f90:193-194).
//
// The FP64 refinement pass reads u and d (5-point), u0, Kx, Ky and writes
// u' = u + d into the other u buffer and r as FP32: 48 B per cell.
//
// Both levels apply the operator in difference form (tl_stencilv.hpp): with
// Di x - sum K x_n, FP32 rounding limits each refinement to a residual
// reduction of about cond(A) * 2^-24, which is 1e-2 for the 4096^2 decks.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "tl_loopk.hpp"

namespace tl {
dev::Geo geo(int cols_per_tile);
}
using namespace tl;
using dev::Geo;

namespace {

constexpr int PADL32 = 64;  // floats before interior column 1 (256 B)

// FP32 field slots inside S().f32
enum F32 { D = 0, R, W, Z, RS, RT, UT, P_0, P_1, SD_0, SD_1, KX, KY, NF32 };

float *f32(int k) {
  State &s = S();
  return s.f32 + (long)k * s.field32 + s.pitch32 + PADL32;
}

int ensure_f32() {
  State &s = S();
  if (s.f32) return TL_OK;
  // columns up to nx+3 are loaded by the 4-wide vectors; one halo row on
  // each side (row ny is read as the upper face of Ky)
  s.pitch32 = ((long)PADL32 + s.nx + 4 + 63) / 64 * 64;
  s.field32 = (size_t)(s.pitch32 * (s.ny + 2) + 64);
  const size_t bytes = sizeof(float) * s.field32 * NF32;
  TL_HIP(hipMalloc((void **)&s.f32, bytes));
  s.f32_bytes = bytes;
  // columns and rows outside the interior stay zero for good
  TL_HIP(hipMemsetAsync(s.f32, 0, bytes, s.stream));
  return TL_OK;
}

Geo geo32() {
  State &s = S();
  static int target = -1, maxrows = -1;
  if (target < 0) {
    const char *e = getenv("TL_TILES32");
    target = e ? atoi(e) : 1024;
    e = getenv("TL_MAXROWS32");
    maxrows = e ? atoi(e) : 64;
  }
  const Grid g = grid_for(s.nx, s.ny, 4 * BLOCK, target, maxrows);
  Geo G;
  G.nx = s.nx;
  G.ny = s.ny;
  G.pitch = s.pitch32;
  G.tiles_x = g.tiles_x;
  G.rows_per_tile = g.rows_per_tile;
  G.tiles = g.tiles;
  G.blocks = g.blocks;
  return G;
}

// Kx, Ky -> FP32 on columns [0, nx] and rows [0, ny] (every face the
// stencil reads; zero faces stay exactly zero)
__global__ void k_mx_coef(int nx, int ny, long pitch, long pitch32, const double *__restrict__ Kx,
                          const double *__restrict__ Ky, float *__restrict__ Kx32,
                          float *__restrict__ Ky32) {
  const long n = (long)(nx + 1) * (ny + 1);
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const int x = (int)(i % (nx + 1)), y = (int)(i / (nx + 1));
    Kx32[(long)y * pitch32 + x] = (float)Kx[(long)y * pitch + x];
    Ky32[(long)y * pitch32 + x] = (float)Ky[(long)y * pitch + x];
  }
}

// u' = u + d (WITH_D), r = u0 - A u' in FP64 with the operator in
// difference form, r stored as FP32 with the partials of r.r in FP64
// (slot A) and of r32.r32 (slot B).
template <bool NT, bool WITH_D>
__global__ __launch_bounds__(BLOCK) void k_mx_refine(Geo G, long pitch32,
                                                      const double *__restrict__ u,
                                                      const float *__restrict__ d,
                                                      double *__restrict__ uout,
                                                      const double *__restrict__ u0,
                                                      const double *__restrict__ Kx,
                                                      const double *__restrict__ Ky, double rx,
                                                      double ry, float *__restrict__ r32,
                                                      double *__restrict__ part_a,
                                                      double *__restrict__ part_b) {
  typedef typename dev::Vec<double, 2>::type VD;
  typedef typename dev::Vec<float, 2>::type VF;
  __shared__ double lds[BLOCK / WAVE];
  const long P = G.pitch;
  double aa = 0.0, ab = 0.0;
  dev::TileIter it(G.tiles, G.blocks);
  for (long t = it.t; t < it.t_end; ++t) {
    const dev::TileV<2> tt = dev::tilev_of<2>(G, t);
    auto ldop = [&](int col, int row) -> VD {
      VD v = dev::ldv<NT, double, 2>(u + (long)row * P + col);
      if (WITH_D) v += __builtin_convertvector(dev::ldv<NT, float, 2>(d + (long)row * pitch32 + col), VD);
      return v;
    };
    auto ld1 = [&](int col, int row) -> double {
      double v = u[(long)row * P + col];
      if (WITH_D) v += (double)d[(long)row * pitch32 + col];
      return v;
    };
    const int cc = tt.c <= G.nx ? tt.c : (G.nx / 2) * 2;
    auto side = [&](int row) -> VD { return dev::ldv<NT, double, 2>(u0 + (long)row * P + cc); };
    dev::walkv<NT, double, 2, true, true>(
        G, tt, Kx, Ky, rx, ry, ldop, ld1, side, [&](int y, VD v, VD wv, VD u0v) {
      if (tt.c >= G.nx) return;
      const VD r = u0v - wv;
      const VF rf = __builtin_convertvector(r, VF);
      if (WITH_D) dev::stv<(NT && TL_ST_NT_LOOP), double, 2>(uout + (long)y * P, tt.c, G.nx, v);
      dev::stv<(NT && TL_ST_NT_LOOP), float, 2>(r32 + (long)y * pitch32, tt.c, G.nx, rf);
      aa += r.x * r.x;
      ab += (double)rf.x * (double)rf.x;
      if (tt.c + 1 < G.nx) {
        aa += r.y * r.y;
        ab += (double)rf.y * (double)rf.y;
      }
    });
  }
  aa = dev::block_sum(aa, lds);
  ab = dev::block_sum(ab, lds);
  if (threadIdx.x == 0) {
    part_a[blockIdx.x] = aa;
    part_b[blockIdx.x] = ab;
  }
}

// ---------------------------------------------------------------------------
// eigenvalue bounds (FP64, host)
// ---------------------------------------------------------------------------
// Eigenvalues of the symmetric tridiagonal matrix (diagonal d[0..n-1],
// sub-diagonal e[1..n-1]) by the implicit QL method with Wilkinson shifts
// (Bowdler, Martin, Reinsch and Wilkinson, Numer. Math. 11 (1968) 293-306),
// at most 30 sweeps per eigenvalue as the reference allows
// (tea_leaf_cheby.f90:90-140).  Returns false when a sweep limit is hit.
bool ql_eigenvalues(std::vector<double> &d, std::vector<double> e) {
  const int n = (int)d.size();
  if (n < 1) return true;
  // shift so that e[i] couples d[i] and d[i+1]
  for (int i = 0; i + 1 < n; ++i) e[i] = e[i + 1];
  e[n - 1] = 0.0;
  for (int l = 0; l < n; ++l) {
    int sweeps = 0;
    for (;;) {
      int m = l;
      while (m < n - 1) {
        const double tol = fabs(d[m]) + fabs(d[m + 1]);
        if (fabs(e[m]) + tol == tol) break;
        ++m;
      }
      if (m == l) break;
      if (sweeps++ == 30) return false;
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = sqrt(g * g + 1.0);
      g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? fabs(r) : -fabs(r)));
      double s = 1.0, c = 1.0, p = 0.0;
      bool deflated = false;
      for (int i = m - 1; i >= l; --i) {
        const double f = s * e[i], b = c * e[i];
        r = sqrt(f * f + g * g);
        e[i + 1] = r;
        if (r == 0.0) {
          d[i + 1] -= p;
          e[m] = 0.0;
          deflated = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0 * c * b;
        p = s * r;
        d[i + 1] = g + p;
        g = c * r - b;
      }
      if (deflated) continue;
      d[l] -= p;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  return true;
}

// Extreme eigenvalues of the Lanczos tridiagonal of CG iterations 1..n:
// diagonal 1/alpha_i + beta_{i-1}/alpha_{i-1}, off-diagonal sqrt(beta_i)/alpha_i
// (tea_leaf_cheby.f90:142-176); alphas/betas indexed by iteration.
bool lanczos_bounds(const std::vector<double> &al, const std::vector<double> &be, int n,
                    double *emin, double *emax) {
  std::vector<double> d(n, 0.0), e(n, 0.0);
  for (int i = 1; i <= n; ++i) {
    d[i - 1] = 1.0 / al[i];
    if (i > 1) d[i - 1] += be[i - 1] / al[i - 1];
    if (i < n) e[i] = sqrt(be[i]) / al[i];
  }
  if (!ql_eigenvalues(d, e)) return false;
  double lo = d[0], hi = d[0];
  for (double v : d) {
    lo = v < lo ? v : lo;
    hi = v > hi ? v : hi;
  }
  if (!(lo > 0.0) || !(hi > lo)) return false;
  *emin = lo;
  *emax = hi;
  return true;
}

// Chebyshev coefficients of the polynomial preconditioner
// (tea_leaf_cheby.f90:178-205)
void cheby_coefs(double emin, double emax, int n, std::vector<double> &ca,
                 std::vector<double> &cb, double *theta) {
  *theta = (emax + emin) / 2.0;
  const double delta = (emax - emin) / 2.0, sigma = *theta / delta;
  double rho_old = 1.0 / sigma;
  ca.assign(n, 0.0);
  cb.assign(n, 0.0);
  for (int i = 0; i < n; ++i) {
    const double rho_new = 1.0 / (2.0 * sigma - rho_old);
    ca[i] = rho_new * rho_old;
    cb[i] = 2.0 * rho_new / delta;
    rho_old = rho_new;
  }
}

// ---------------------------------------------------------------------------
// driver pieces
// ---------------------------------------------------------------------------
LoopBufs<float> f32_bufs(int &pcur, int &sdcur) {
  State &s = S();
  LoopBufs<float> b;
  b.G = geo32();
  b.u = f32(D);
  b.r = f32(R);
  b.w = f32(W);
  b.z = f32(Z);
  b.r_store = f32(RS);
  b.rtemp = f32(RT);
  b.utemp = f32(UT);
  b.p[0] = f32(P_0);
  b.p[1] = f32(P_1);
  b.sd[0] = f32(SD_0);
  b.sd[1] = f32(SD_1);
  b.pcur = &pcur;
  b.sdcur = &sdcur;
  b.Mi = nullptr;
  b.Kx = f32(KX);
  b.Ky = f32(KY);
  b.rx = (float)s.rx;
  b.ry = (float)s.ry;
  b.diag = 0;
  b.cg_src = f32(R);
  return b;
}

// One refinement pass; returns ||r||^2 (FP64) and r32.r32 on the host.
int refine(bool with_d, double *rr64, double *rr32) {
  State &s = S();
  const Geo G = geo(2 * BLOCK);
  const int nb = G.blocks < 0 ? -G.blocks : G.blocks;
  const bool nt = use_nontemporal();
  const double *uin = s.u[s.ucur].p11;
  double *uout = s.u[1 - s.ucur].p11;
  hipEvent_t ev;
  prof_arm(7, &ev);
  if (with_d) {
    auto k = nt ? k_mx_refine<true, true> : k_mx_refine<false, true>;
    klaunch(k, dim3(nb), dim3(BLOCK), G, s.pitch32, uin, f32(D), uout,
            s.u0.p11, s.Kx.p11, s.Ky.p11, s.rx, s.ry, f32(R), part(P4), part(P5));
  } else {
    auto k = nt ? k_mx_refine<true, false> : k_mx_refine<false, false>;
    klaunch(k, dim3(nb), dim3(BLOCK), G, s.pitch32, uin, f32(D), uout,
            s.u0.p11, s.Kx.p11, s.Ky.p11, s.rx, s.ry, f32(R), part(P4), part(P5));
  }
  prof_end(7, ev);
  TL_HIP(hipGetLastError());
  const int slots[2] = {P4, P5};
  double out[2];
  int rc = finish_sums(slots, 2, nb, out);
  if (rc) return rc;
  // a non-finite result leaves u as it was (the new u went to the other buffer)
  if (!std::isfinite(out[0]) || !std::isfinite(out[1]))
    return fail(TL_ERR_NONFINITE, "mixed precision: non-finite residual");
  if (with_d) s.ucur = 1 - s.ucur;
  *rr64 = out[0];
  *rr32 = out[1];
  return TL_OK;
}

int zero_d() {
  State &s = S();
  TL_HIP(hipMemsetAsync(s.f32 + (long)D * s.field32, 0, sizeof(float) * s.field32, s.stream));
  return TL_OK;
}

}  // namespace

extern "C" int tl_mixed_solve(tl_mixed_state *st) {
  TL_REQUIRE_INIT();
  if (!st) return fail(TL_ERR_ARG, "tl_mixed_solve: null state");
  State &s = S();
  if (st->solver != TL_MIXED_CG && st->solver != TL_MIXED_PPCG)
    return fail(TL_ERR_ARG, "tl_mixed_solve: solver must be TL_MIXED_CG or TL_MIXED_PPCG");
  if (s.precond != TL_PREC_NONE)
    return fail(TL_ERR_ARG, "tl_mixed_solve: needs tl_preconditioner_type none");
  if (!s.zb_all)
    return fail(TL_ERR_ARG, "tl_mixed_solve: needs a chunk with all four faces external");
  if (st->max_outer < 1 || st->max_inner < 1 || !(st->rel_tol > 0.0) ||
      !(st->inner_tol > 0.0 && st->inner_tol < 1.0) ||
      (st->solver == TL_MIXED_PPCG && (st->inner_steps < 1 || st->presteps < 2)))
    return fail(TL_ERR_ARG, "tl_mixed_solve: bad limits or tolerances");
  int rc;
  if ((rc = ensure_f32())) return rc;
  if ((rc = ensure_loop_buffers(st->max_inner > st->presteps ? st->max_inner : st->presteps)))
    return rc;
  {
    const long n = (long)(s.nx + 1) * (s.ny + 1);
    const int blocks = (int)((n + BLOCK - 1) / BLOCK < 4096 ? (n + BLOCK - 1) / BLOCK : 4096);
    hipLaunchKernelGGL(k_mx_coef, dim3(blocks), dim3(BLOCK), 0, s.stream, s.nx, s.ny, s.pitch,
                       s.pitch32, s.Kx.p11, s.Ky.p11, f32(KX), f32(KY));
    TL_HIP(hipGetLastError());
  }
  st->outer = 0;
  st->inner_iters = 0;
  st->sweeps = 0;
  st->eig_iters = 0;
  double rr64, rr32;
  if ((rc = refine(false, &rr64, &rr32))) return rc;
  const double r0 = sqrt(rr64);
  st->initial_residual = r0;
  st->final_residual = r0;
  st->reason = TL_LOOP_CONVERGED;
  const double target = st->rel_tol * r0;
  if (r0 == 0.0) return TL_OK;

  bool have_eig = st->solver == TL_MIXED_PPCG && st->eigmin > 0.0 && st->eigmax > st->eigmin;
  std::vector<double> ca, cb;
  double theta = 0.0;
  if (have_eig) cheby_coefs(st->eigmin, st->eigmax, st->inner_steps, ca, cb, &theta);
  int pcur = 0, sdcur = 0;
  LoopBufs<float> b = f32_bufs(pcur, sdcur);
  double res = r0;
  // The true residual's rounding floor: a target below it (e.g. relative to
  // a small initial residual) cannot be met.  A refinement whose inner solve
  // reduced its own residual at least 4-fold while the true residual did not
  // halve has reached it: the solve ends there (another refinement on a
  // residual of rounding noise costs hundreds of inner iterations for
  // nothing).  A refinement whose inner solve gained less, such as the
  // eigenvalue-estimate CG of PPCG's first refinement, which stops when the
  // estimate settles, counts as a stall; two in a row end the solve too.
  int stalls = 0;  // refinements in a row that did not halve the true residual
  bool floor = false;
  while (res > target && st->outer < st->max_outer && stalls < 2 && !floor) {
    double inner_gain = 0.0;  // the inner solve's start / end residual
    if ((rc = zero_d())) return rc;
    // ask an inner solve for no more than the remaining reduction (with a
    // margin), and never less than inner_tol
    double rel = 0.25 * target / res;
    rel = rel > st->inner_tol ? rel : st->inner_tol;
    rel = rel < 0.5 ? rel : 0.5;
    if (st->solver == TL_MIXED_CG || !have_eig) {
      // FP32 CG from d = 0, p = r.  For PPCG this first solve's coefficients
      // give the eigenvalue estimate (the Lanczos tridiagonal).  Its largest
      // eigenvalue converges from below, and the x1.05 margin of
      // tea_solve.f90:193-194 must cover what is left (a bound below the
      // spectrum makes the polynomial preconditioner amplify the top modes),
      // so the CG runs in doubling segments from `presteps` iterations until
      // the largest estimate moved by under 1 % over the last segment, at
      // most 64 x presteps (or max_inner) iterations in all.
      const bool estimate = st->solver == TL_MIXED_PPCG;
      const int cap = estimate ? std::min(st->max_inner, 64 * st->presteps) : st->max_inner;
      std::vector<double> al(estimate ? cap + 2 : 0, 0.0), be(estimate ? cap + 2 : 0, 0.0);
      tl_loop_state ls;
      std::memset(&ls, 0, sizeof(ls));
      ls.n_first = 1;
      ls.n_last = estimate ? std::min(st->presteps, cap) : cap;
      ls.eps_abs = rel * sqrt(rr32);
      ls.rro = rr32;
      double emin = 0.0, emax = 0.0, emax_prev = 0.0;
      bool eig_ok = false;
      for (;;) {
        // the first segment starts from p = r; later ones continue the loop
        if ((rc = run_cg_loop<float, 4>(b, &ls, estimate ? al.data() : nullptr,
                                        estimate ? be.data() : nullptr,
                                        ls.n_first == 1 ? b.r : nullptr)))
          return rc;
        st->inner_iters += ls.n_end - ls.n_first + 1;
        inner_gain = sqrt(rr32) / std::max(ls.error, 1e-300);
        if (!estimate) break;
        eig_ok = ls.n_end >= 2 && lanczos_bounds(al, be, ls.n_end, &emin, &emax);
        st->eig_iters = ls.n_end;
        if (!eig_ok || ls.reason != TL_LOOP_MAXITERS || ls.n_end >= cap) break;
        if (emax_prev > 0.0 && emax - emax_prev <= 0.01 * emax) break;
        emax_prev = emax;
        ls.n_first = ls.n_end + 1;
        ls.n_last = std::min(2 * ls.n_end, cap);
      }
      if (estimate) {
        if (eig_ok) {
          st->eigmin = emin * 0.95;  // tea_solve.f90:193-194
          st->eigmax = emax * 1.05;
          cheby_coefs(st->eigmin, st->eigmax, st->inner_steps, ca, cb, &theta);
          have_eig = true;
        } else if (ls.n_end >= 2) {
          return fail(TL_ERR_EIGEN, "tl_mixed_solve: eigenvalue estimate failed");
        }
      }
    } else {
      // FP32 PPCG from d = 0: z = poly(A) r, p = z
      if ((rc = run_ppcg_precondition<float, 4>(b, ca.data(), cb.data(), st->inner_steps, theta)))
        return rc;
      const int slot = P2;
      double zr;
      if ((rc = finish_sums(&slot, 1, b.G.blocks, &zr))) return rc;
      st->sweeps += st->inner_steps;
      if (!(zr > 0.0)) return fail(TL_ERR_NONFINITE, "tl_mixed_solve: z.r is not positive");
      tl_loop_state ls;
      std::memset(&ls, 0, sizeof(ls));
      ls.n_first = 1;
      ls.n_last = st->max_inner;
      ls.eps_abs = rel * sqrt(zr);
      int sweeps = 0;
      if ((rc = run_ppcg_loop<float, 4>(b, &ls, ca.data(), cb.data(), st->inner_steps, theta, zr,
                                        b.z, &sweeps)))
        return rc;
      st->inner_iters += ls.n_end;
      st->sweeps += sweeps;
      inner_gain = sqrt(zr) / std::max(ls.error, 1e-300);
    }
    if ((rc = refine(true, &rr64, &rr32))) return rc;
    const double prev = res;
    res = sqrt(rr64);
    st->outer += 1;
    if (res > 0.5 * prev) {
      stalls += 1;
      floor = inner_gain >= 4.0;
    } else {
      stalls = 0;
    }
    st->final_residual = res;
  }
  st->reason = res <= target ? TL_LOOP_CONVERGED
                             : ((floor || stalls >= 2) ? TL_LOOP_FLOOR : TL_LOOP_MAXITERS);
  return TL_OK;
}
