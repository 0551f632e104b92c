// Fused solver loops shared by the FP64 entry points (tl_loops.hip) and the
// FP32 inner solves of the mixed-precision solver (tl_mixed.hip).
//
// Every reduction is finished redundantly by every block of the *next*
// kernel: block b of kernel k+1 sums the per-block partials kernel k wrote,
// in one fixed order, so all blocks see the same bits.  That removes the host
// round trip per dot product and needs no atomics or cross-XCD fences;
// kernel boundaries order everything.  The loop control block (stop flag,
// reason, iteration) is written by block 0 of the kernel that decides; the
// kernels of an in-flight batch that follow see the flag and return.
#pragma once

#include <cmath>
#include <cstdlib>

#include "tl_internal.hpp"
#include "tl_stencil.hpp"
#include "tl_stencilv.hpp"
#include "tl_sweep2.hpp"

namespace tl {

// stop: 0 while the loop runs; otherwise nonzero, and for a stop decided at
// the head of the first kernel of iteration n it is stop_at(n), so that the
// blocks of that same launch (which still apply the last p update) tell it
// from a stop of an earlier launch.
struct Ctl {
  int stop, reason, n_end, pad;
  double error, old_error, rro;
};
__host__ __device__ constexpr int stop_at(int n) { return n + 1; }
constexpr int STOP_UPDATE = -1;  // a stop recorded by an update pass
// tl_init allocates 64 device bytes for the control block and 256 pinned
// host bytes for its copies (two batches in flight and the final read)
static_assert(sizeof(Ctl) <= 64 && 3 * sizeof(Ctl) <= 256, "control block size");

// Device memory a fused loop keeps its scalars in: the control block, the
// per-iteration histories indexed by iteration n (rro entering n, alpha and
// beta of n, the exit-test error sqrt|rrn| of n) and the partial-sum slots.
struct LoopMem {
  Ctl *ctl;
  double *hrro, *halpha, *hbeta, *herr;
  double *parts;  // NPART slots of MAX_GRID doubles
  double *part(int slot) const { return parts + (long)slot * MAX_GRID; }
};

struct LoopParams {
  double eps_abs;
  double epslim;
  int check_switch;
  int presteps;
};

bool use_nontemporal();
int batch_size();
// allocates the library's loop history for iterations up to n_last
int ensure_loop_buffers(long n_last);
// the library's own loop memory (after ensure_loop_buffers)
LoopMem loop_mem();
int copy_hist(double *dst, const double *src_dev, int lo, int hi);

// TL_ERR_NONFINITE for a loop that stopped on a NaN or Inf reduction
inline int nonfinite_fail(const char *solver, int n) {
  return fail(TL_ERR_NONFINITE, std::string(solver) + ": non-finite reduction at iteration " +
                                    std::to_string(n) + " (NaN or Inf in the fields)");
}

// Drive a loop: launch(n, first) queues iteration n; tail(n) queues the exit
// test of iteration n-1.  Batches of iterations are queued two deep and the
// control block `dctl` is polled once per batch.  Allocates nothing.
template <typename LAUNCH, typename TAIL>
int drive(int n_first, int n_last, LAUNCH launch, TAIL tail, Ctl *out, const Ctl *dctl) {
  State &s = S();
  Ctl *hc = (Ctl *)s.hctl;
  const int B = batch_size();
  hipEvent_t *ev = s.loop_ev;
  int n = n_first, k = 0, rc = TL_OK;
  bool stopped = false;
  while (n <= n_last) {
    const int m = (n + B - 1 < n_last) ? n + B - 1 : n_last;
    for (int it = n; it <= m; ++it) {
      rc = launch(it, it == n_first);
      if (rc) return rc;
    }
    if (hipMemcpyAsync(&hc[k & 1], dctl, sizeof(Ctl), hipMemcpyDeviceToHost, s.stream) !=
            hipSuccess ||
        hipEventRecord(ev[k & 1], s.stream) != hipSuccess)
      return fail(TL_ERR_HIP, "queueing the loop control block copy");
    if (k >= 1) {
      if (hipEventSynchronize(ev[(k - 1) & 1]) != hipSuccess)
        return fail(TL_ERR_HIP, "hipEventSynchronize");
      if (hc[(k - 1) & 1].stop) {
        stopped = true;
        break;
      }
    }
    n = m + 1;
    ++k;
  }
  if (!stopped) {
    rc = tail(n_last + 1);
    if (rc) return rc;
  }
  if (hipMemcpyAsync(&hc[2], dctl, sizeof(Ctl), hipMemcpyDeviceToHost, s.stream) != hipSuccess ||
      hipStreamSynchronize(s.stream) != hipSuccess)
    return fail(TL_ERR_HIP, "reading the loop control block");
  *out = hc[2];
  return TL_OK;
}

// ---------------------------------------------------------------------------
// kernels (one copy per translation unit)
// ---------------------------------------------------------------------------
namespace {

using dev::Geo;

// FP32 fields belong to the mixed-precision solver, which applies the
// operator in difference form; FP64 fields use the reference's Di form.
template <typename T>
constexpr bool diff_form() {
  return sizeof(T) == 4;
}

// Row prefetch of the stencil walks (see walkv) per kernel.  The w pass
// keeps two rows of loads in flight: 3 % faster than one at 4096^2 (110
// registers, 4 waves per SIMD, as many as its 1024 tiles put on a CU);
// three rows need more than 128 registers and measured no faster.
#ifndef TL_PF_W
#define TL_PF_W 2
#endif
#ifndef TL_PF_SWEEP
#define TL_PF_SWEEP 0
#endif
// Stores: written fields are re-read by the next kernel, so plain stores that
// leave them in the 256 MB Infinity Cache beat streaming stores (measured at
// 4096^2).  Consecutive passes walk their rows in opposite directions (the
// CG and PPCG updates in reverse, the PPCG sweeps alternately), so that each
// pass reads first what the one before wrote last.
// Streaming hints of the CG and PPCG passes (measured at 4096^2).  u is
// touched only by the update pass, so its store streams past the Infinity
// Cache (TL_ST_NT_U, TL_ST_NT_PPU), which then still holds the r and p the
// next w pass reads: the CG w pass drops from 141 to about 126 us (FP64
// only: the FP32 fields of the mixed solver are half the size and its CG
// measured 2 % slower with the streaming store).  Plain
// loads of the w pass's operand (TL_LD_NT_OPND=0) or of p and w in the
// update pass (TL_LD_NT_UR=0), and streaming r_store (TL_ST_NT_RSTORE),
// measured slower or no faster.
#ifndef TL_ST_NT_U
#define TL_ST_NT_U 1
#endif
#ifndef TL_ST_NT_PPU
#define TL_ST_NT_PPU 1
#endif
#ifndef TL_ST_NT_RSTORE
#define TL_ST_NT_RSTORE 0
#endif
#ifndef TL_LD_NT_OPND
#define TL_LD_NT_OPND 1
#endif
#ifndef TL_LD_NT_UR
#define TL_LD_NT_UR 1
#endif
#ifndef TL_ST_NT_LOOP
#define TL_ST_NT_LOOP 0
#endif
// utemp of the FP64 PPCG sweeps is stored streaming, which leaves the
// Infinity Cache to the sd and rtemp the next pass reads (sweep pass 210 ->
// 199 us, PPCG step 0.489 -> 0.467 s at 4096^2)
#ifndef TL_ST_NT_UTEMP
#define TL_ST_NT_UTEMP 1
#endif
#ifndef TL_ST_NT_SWEEP
#define TL_ST_NT_SWEEP 0
#endif
#ifndef TL_LDSX_W
#define TL_LDSX_W 1
#endif
#ifndef TL_LDSX_SWEEP
#define TL_LDSX_SWEEP 1
#endif
// minimum waves per SIMD the w pass is compiled for
#ifndef TL_W_WAVES
#define TL_W_WAVES 1
#endif
// minimum waves per SIMD the two-sweep pass is compiled for: FP64 at 4
// spills 14 registers and runs 210.5 us per pass at 4096^2, against 260.8
// with 3 waves (no spills) on 1024 tiles and 220.3 on 720
#ifndef TL_PAIR_WAVES64
#define TL_PAIR_WAVES64 4
#endif
#ifndef TL_PAIR_WAVES32
#define TL_PAIR_WAVES32 4
#endif

// Scalar head of the first kernel of an iteration: beta of iteration n-1,
// exit test (tea_solve.f90:402-410) and, for CG, the switch test (:167).
struct Head {
  double beta;
  bool stop;
};

__device__ inline Head loop_head(int n, int first, int tail, const LoopParams &lp,
                                 double rro_in, const double *__restrict__ prev_part, int nprev,
                                 Ctl *ctl, double *__restrict__ hrro,
                                 double *__restrict__ hbeta, double *__restrict__ herr,
                                 double *lds) {
  Head h;
  h.beta = 0.0;
  h.stop = false;
  if (!first) {
    const double rrn = dev::sum_partials(prev_part, nprev, lds);
    const double rro = hrro[n - 1];
    h.beta = rrn / rro;
    const double err = sqrt(fabs(rrn));
    // a NaN or Inf reduction (e.g. a zero density) stops the loop instead of
    // iterating to the limit
    const bool bad = !isfinite(rrn);
    const bool conv = !bad && fabs(err) < lp.eps_abs;
    const bool sw = !bad && !conv && !tail && lp.check_switch && fabs(err) <= lp.epslim &&
                    n >= lp.presteps;
    h.stop = bad || conv || sw || tail;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      hbeta[n - 1] = h.beta;
      herr[n - 1] = err;
      if (h.stop) {
        ctl->stop = stop_at(n);
        ctl->reason = bad ? TL_LOOP_NONFINITE
                          : (conv ? TL_LOOP_CONVERGED : (sw ? TL_LOOP_SWITCH : TL_LOOP_MAXITERS));
        ctl->n_end = sw ? n : n - 1;
        ctl->error = err;
        ctl->old_error = err;
        ctl->rro = rrn;
      } else {
        hrro[n] = rrn;
      }
    }
  } else if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctl->stop = 0;
    hrro[n] = rro_in;
  }
  return h;
}

// A non-finite p.w (or z.r) in the update pass of iteration n: block 0
// records the stop; every block computed the same value and returns.
__device__ inline bool nonfinite_stop(int n, double v, double v2, Ctl *ctl) {
  if (isfinite(v) && isfinite(v2)) return false;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctl->stop = STOP_UPDATE;
    ctl->reason = TL_LOOP_NONFINITE;
    ctl->n_end = n;
    ctl->error = isfinite(v) ? v2 : v;
  }
  return true;
}

// First kernel of a CG or PPCG iteration n: p = src + beta p (of n-1; src is
// r for CG, z for PPCG and preconditioned CG), w = A p, partials of p.w.
// On the launch that stops the loop only the p update is done, so p ends as
// the reference leaves it.
template <bool NT, typename T, int V>
__global__ __launch_bounds__(BLOCK, TL_W_WAVES) void k_loop_w(Geo G, int n, int first, int tail,
                                                   LoopParams lp, double rro_in,
                                                   const T *__restrict__ src,
                                                   const T *__restrict__ pin,
                                                   T *__restrict__ pout,
                                                   const T *__restrict__ Kx,
                                                   const T *__restrict__ Ky, T *__restrict__ w,
                                                   T rx, T ry,
                                                   const double *__restrict__ prev_part,
                                                   int nprev, double *__restrict__ out_part,
                                                   Ctl *ctl, double *__restrict__ hrro,
                                                   double *__restrict__ hbeta,
                                                   double *__restrict__ herr, int dir) {
  typedef typename dev::Vec<T, V>::type VT;
  __shared__ double lds[BLOCK / WAVE];
  // launches queued past the stop return at once; a block of the launch
  // whose head stops the loop may already see that stop, and still runs
  if (!first) {
    const int st = ctl->stop;
    if (st && st != stop_at(n)) return;
  }
  const Head h =
      loop_head(n, first, tail, lp, rro_in, prev_part, nprev, ctl, hrro, hbeta, herr, lds);
  const T beta = (T)h.beta;
  const bool upd = !first;
  const long P = G.pitch;
  double acc = 0.0;
  dev::TileIter it(G.tiles, G.blocks);
  for (long t = it.t; t < it.t_end; ++t) {
    const dev::TileV<V> tt = dev::tilev_of<V>(G, t);
    if (h.stop) {
      if (tt.c < G.nx)
        for (int y = tt.y0; y < tt.y1; ++y) {
          const long i = (long)y * P + tt.c;
          const VT v = dev::ldv<NT, T, V>(src + i) + beta * dev::ldv<NT, T, V>(pin + i);
          dev::stv<(NT && TL_ST_NT_LOOP), T, V>(pout + (long)y * P, tt.c, G.nx, v);
        }
      continue;
    }
    auto ldop = [&](int col, int row) -> VT {
      const long i = (long)row * P + col;
      const VT pv = dev::ldv<(NT && TL_LD_NT_OPND), T, V>(pin + i);
      return upd ? dev::ldv<(NT && TL_LD_NT_OPND), T, V>(src + i) + beta * pv : pv;
    };
    auto ld1 = [&](int col, int row) -> T {
      const long i = (long)row * P + col;
      return upd ? src[i] + beta * pin[i] : pin[i];
    };
    auto none = [](int) { return dev::NoExtra(); };
    dev::walkv<NT, T, V, diff_form<T>(), TL_PF_W, TL_LDSX_W>(
        G, tt, Kx, Ky, rx, ry, ldop, ld1, none, [&](int y, VT v, VT wv, dev::NoExtra) {
      if (tt.c < G.nx) {
        dev::stv<(NT && TL_ST_NT_LOOP), T, V>(w + (long)y * P, tt.c, G.nx, wv);
        dev::stv<(NT && TL_ST_NT_LOOP), T, V>(pout + (long)y * P, tt.c, G.nx, v);
#pragma unroll
        for (int e = 0; e < V; ++e)
          if (tt.c + e < G.nx) acc += (double)wv[e] * (double)v[e];
      }
    }, dir);
  }
  if (h.stop) return;
  acc = dev::block_sum(acc, lds);
  if (threadIdx.x == 0) out_part[blockIdx.x] = acc;
}

// Rows per group of the unpreconditioned CG update pass (k_cg_ur): 8 rows
// of loads in flight per lane made the pass 3 % faster than one row at
// 4096^2 and 8192^2.
#ifndef TL_UR_ROWS
#define TL_UR_ROWS 8
#endif
constexpr int UR_ROWS = TL_UR_ROWS;

// CG iteration n, second pass: alpha = rro/pw; u += alpha p; r -= alpha w;
// (z = Mi r); partials of r.r or r.z (kernels/tea_leaf_cg_kernel.f90:237-300)
template <bool NT, typename T, int V>
__global__ __launch_bounds__(BLOCK) void k_cg_ur(Geo G, int n, T *__restrict__ u,
                                                  const T *__restrict__ p, T *__restrict__ r,
                                                  const T *__restrict__ w, T *__restrict__ z,
                                                  const T *__restrict__ Mi, int diag,
                                                  const double *__restrict__ prev_part, int nprev,
                                                  double *__restrict__ out_part, Ctl *ctl,
                                                  const double *__restrict__ hrro,
                                                  double *__restrict__ halpha, int rev) {
  typedef typename dev::Vec<T, V>::type VT;
  __shared__ double lds[BLOCK / WAVE];
  if (ctl->stop) return;
  const double pw = dev::sum_partials(prev_part, nprev, lds);
  if (nonfinite_stop(n, pw, 0.0, ctl)) return;
  const double alpha = hrro[n] / pw;
  if (blockIdx.x == 0 && threadIdx.x == 0) halpha[n] = alpha;
  const T a = (T)alpha;
  const long P = G.pitch;
  double acc = 0.0;
  dev::TileIter it(G.tiles, G.blocks);
  for (long t = it.t; t < it.t_end; ++t) {
    const dev::TileV<V> tt = dev::tilev_of<V>(G, t);
    if (tt.c >= G.nx) continue;
    int k = tt.y0;
    if (!diag) {
      // groups of UR_ROWS rows with every load issued before the first
      // store, so that each lane has 4 x UR_ROWS loads in flight; the r.r
      // partial still adds the rows in walking order
      for (; k + UR_ROWS <= tt.y1; k += UR_ROWS) {
        VT uu[UR_ROWS], pp[UR_ROWS], rr[UR_ROWS], ww[UR_ROWS];
#pragma unroll
        for (int q = 0; q < UR_ROWS; ++q) {
          const int y = rev ? tt.y1 - 1 - (k + q - tt.y0) : k + q;
          const long i = (long)y * P + tt.c;
          uu[q] = dev::ldv<NT, T, V>(u + i);
          pp[q] = dev::ldv<(NT && TL_LD_NT_UR), T, V>(p + i);
          rr[q] = dev::ldv<NT, T, V>(r + i);
          ww[q] = dev::ldv<(NT && TL_LD_NT_UR), T, V>(w + i);
        }
#pragma unroll
        for (int q = 0; q < UR_ROWS; ++q) {
          const int y = rev ? tt.y1 - 1 - (k + q - tt.y0) : k + q;
          const VT uv = uu[q] + a * pp[q];
          const VT rv = rr[q] - a * ww[q];
          dev::stv<(NT && (TL_ST_NT_LOOP || (TL_ST_NT_U && sizeof(T) == 8))), T, V>(
              u + (long)y * P, tt.c, G.nx, uv);
          dev::stv<(NT && TL_ST_NT_LOOP), T, V>(r + (long)y * P, tt.c, G.nx, rv);
#pragma unroll
          for (int e = 0; e < V; ++e)
            if (tt.c + e < G.nx) acc += (double)rv[e] * (double)rv[e];
        }
      }
    }
    for (; k < tt.y1; ++k) {
      const int y = rev ? tt.y1 - 1 - (k - tt.y0) : k;
      const long i = (long)y * P + tt.c;
      const VT uv = dev::ldv<NT, T, V>(u + i) + a * dev::ldv<(NT && TL_LD_NT_UR), T, V>(p + i);
      const VT rv = dev::ldv<NT, T, V>(r + i) - a * dev::ldv<(NT && TL_LD_NT_UR), T, V>(w + i);
      dev::stv<(NT && (TL_ST_NT_LOOP || (TL_ST_NT_U && sizeof(T) == 8))), T, V>(
          u + (long)y * P, tt.c, G.nx, uv);
      dev::stv<(NT && TL_ST_NT_LOOP), T, V>(r + (long)y * P, tt.c, G.nx, rv);
      if (diag) {
        const VT zv = dev::ldv<NT, T, V>(Mi + i) * rv;
        dev::stv<(NT && TL_ST_NT_LOOP), T, V>(z + (long)y * P, tt.c, G.nx, zv);
#pragma unroll
        for (int e = 0; e < V; ++e)
          if (tt.c + e < G.nx) acc += (double)rv[e] * (double)zv[e];
      } else {
#pragma unroll
        for (int e = 0; e < V; ++e)
          if (tt.c + e < G.nx) acc += (double)rv[e] * (double)rv[e];
      }
    }
  }
  acc = dev::block_sum(acc, lds);
  if (threadIdx.x == 0) out_part[blockIdx.x] = acc;
}

// CG iteration n, second pass with the block-Jacobi preconditioner:
// u += alpha p, r -= alpha w over groups of 4 rows, then z = M^-1 r by the
// 4-row Thomas solve of tea_block_solve (kernels/tea_leaf_common_kernel.f90:
// 378-468) and the partials of r.z.  Tiles start on multiples of 4 rows.
template <bool NT, typename T, int V>
__global__ __launch_bounds__(BLOCK) void k_cg_ur_blk(
    Geo G, int n, T *__restrict__ u, const T *__restrict__ p, T *__restrict__ r,
    const T *__restrict__ w, T *__restrict__ z, const T *__restrict__ Di,
    const T *__restrict__ Ky, const T *__restrict__ cp, const T *__restrict__ bfp, T ry,
    const double *__restrict__ prev_part, int nprev, double *__restrict__ out_part, Ctl *ctl,
    const double *__restrict__ hrro, double *__restrict__ halpha, int rev) {
  typedef typename dev::Vec<T, V>::type VT;
  __shared__ double lds[BLOCK / WAVE];
  if (ctl->stop) return;
  const double pw = dev::sum_partials(prev_part, nprev, lds);
  if (nonfinite_stop(n, pw, 0.0, ctl)) return;
  const double alpha = hrro[n] / pw;
  if (blockIdx.x == 0 && threadIdx.x == 0) halpha[n] = alpha;
  const T a = (T)alpha;
  const long P = G.pitch;
  double acc = 0.0;
  dev::TileIter it(G.tiles, G.blocks);
  for (long t = it.t; t < it.t_end; ++t) {
    const dev::TileV<V> tt = dev::tilev_of<V>(G, t);
    if (tt.c >= G.nx) continue;
    const int ngroups = (tt.y1 - tt.y0 + 3) / 4;
    for (int m = 0; m < ngroups; ++m) {
      const int g0 = tt.y0 + 4 * (rev ? ngroups - 1 - m : m);
      const int nq = min(4, G.ny - g0);
      VT rr[4], dp[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (q >= nq) break;
        const long i = (long)(g0 + q) * P + tt.c;
        const VT uv = dev::ldv<NT, T, V>(u + i) + a * dev::ldv<NT, T, V>(p + i);
        rr[q] = dev::ldv<NT, T, V>(r + i) - a * dev::ldv<NT, T, V>(w + i);
        dev::stv<false, T, V>(u + i - tt.c, tt.c, G.nx, uv);
        dev::stv<false, T, V>(r + i - tt.c, tt.c, G.nx, rr[q]);
      }
      const long i0 = (long)g0 * P + tt.c;
      dp[0] = rr[0] / dev::ldv<NT, T, V>(Di + i0);
#pragma unroll
      for (int q = 1; q < 4; ++q) {
        if (q >= nq) break;
        const long i = i0 + (long)q * P;
        dp[q] = (rr[q] - (-dev::ldv<NT, T, V>(Ky + i) * ry) * dp[q - 1]) *
                dev::ldv<NT, T, V>(bfp + i);
      }
      VT zn = dp[nq - 1];
      for (int q = nq - 1; q >= 0; --q) {
        const long i = i0 + (long)q * P;
        const VT zq = q == nq - 1 ? dp[q] : dp[q] - dev::ldv<NT, T, V>(cp + i) * zn;
        dev::stv<false, T, V>(z + i - tt.c, tt.c, G.nx, zq);
#pragma unroll
        for (int e = 0; e < V; ++e)
          if (tt.c + e < G.nx) acc += (double)rr[q][e] * (double)zq[e];
        zn = zq;
      }
    }
  }
  acc = dev::block_sum(acc, lds);
  if (threadIdx.x == 0) out_part[blockIdx.x] = acc;
}

// PPCG outer iteration n, second pass (tea_solve.f90:300-309): rro = z.r
// (partials from the previous iteration's last sweep, or rro_in on the
// first), alpha = rro/pw; u += alpha p; r_store = r; r -= alpha w; (z = Mi r)
template <bool NT, typename T, int V>
__global__ __launch_bounds__(BLOCK) void k_pp_ur(Geo G, int n, int first, double rro_in,
                                                  T *__restrict__ u, const T *__restrict__ p,
                                                  T *__restrict__ r, T *__restrict__ r_store,
                                                  const T *__restrict__ w, T *__restrict__ z,
                                                  const T *__restrict__ Mi, int diag,
                                                  const double *__restrict__ part_pw,
                                                  const double *__restrict__ part_zr, int nprev,
                                                  Ctl *ctl, double *__restrict__ hrro,
                                                  double *__restrict__ halpha, int rev) {
  typedef typename dev::Vec<T, V>::type VT;
  __shared__ double lds[BLOCK / WAVE];
  if (ctl->stop) return;
  const double pw = dev::sum_partials(part_pw, nprev, lds);
  const double rro = first ? rro_in : dev::sum_partials(part_zr, nprev, lds);
  if (nonfinite_stop(n, pw, rro, ctl)) return;
  const double alpha = rro / pw;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    hrro[n] = rro;
    halpha[n] = alpha;
  }
  const T a = (T)alpha;
  const long P = G.pitch;
  dev::TileIter it(G.tiles, G.blocks);
  for (long t = it.t; t < it.t_end; ++t) {
    const dev::TileV<V> tt = dev::tilev_of<V>(G, t);
    if (tt.c >= G.nx) continue;
    int k = tt.y0;
    if (sizeof(T) == 8 && !diag) {
      // FP64: groups of UR_ROWS rows with every load issued before the first
      // store, as in k_cg_ur (728 against 767 us at 8192^2, the same at
      // 4096^2; in FP32 2 % faster at 4096^2 and 2 % slower at 8192^2)
      for (; k + UR_ROWS <= tt.y1; k += UR_ROWS) {
        VT uu[UR_ROWS], pp[UR_ROWS], rr[UR_ROWS], ww[UR_ROWS];
#pragma unroll
        for (int q = 0; q < UR_ROWS; ++q) {
          const int y = rev ? tt.y1 - 1 - (k + q - tt.y0) : k + q;
          const long i = (long)y * P + tt.c;
          uu[q] = dev::ldv<NT, T, V>(u + i);
          pp[q] = dev::ldv<NT, T, V>(p + i);
          rr[q] = dev::ldv<NT, T, V>(r + i);
          ww[q] = dev::ldv<NT, T, V>(w + i);
        }
#pragma unroll
        for (int q = 0; q < UR_ROWS; ++q) {
          const int y = rev ? tt.y1 - 1 - (k + q - tt.y0) : k + q;
          const VT uv = uu[q] + a * pp[q];
          const VT rv = rr[q] - a * ww[q];
          dev::stv<(NT && (TL_ST_NT_LOOP || (TL_ST_NT_PPU && sizeof(T) == 8))), T, V>(
              u + (long)y * P, tt.c, G.nx, uv);
          dev::stv<(NT && (TL_ST_NT_LOOP || TL_ST_NT_RSTORE)), T, V>(r_store + (long)y * P,
                                                                      tt.c, G.nx, rr[q]);
          dev::stv<(NT && TL_ST_NT_LOOP), T, V>(r + (long)y * P, tt.c, G.nx, rv);
        }
      }
    }
    for (; k < tt.y1; ++k) {
      const int y = rev ? tt.y1 - 1 - (k - tt.y0) : k;
      const long i = (long)y * P + tt.c;
      const VT uv = dev::ldv<NT, T, V>(u + i) + a * dev::ldv<NT, T, V>(p + i);
      const VT ro = dev::ldv<NT, T, V>(r + i);
      const VT rv = ro - a * dev::ldv<NT, T, V>(w + i);
      dev::stv<(NT && (TL_ST_NT_LOOP || (TL_ST_NT_PPU && sizeof(T) == 8))), T, V>(
          u + (long)y * P, tt.c, G.nx, uv);
      dev::stv<(NT && (TL_ST_NT_LOOP || TL_ST_NT_RSTORE)), T, V>(r_store + (long)y * P, tt.c,
                                                                  G.nx, ro);
      dev::stv<(NT && TL_ST_NT_LOOP), T, V>(r + (long)y * P, tt.c, G.nx, rv);
      if (diag) dev::stv<(NT && TL_ST_NT_LOOP), T, V>(z + (long)y * P, tt.c, G.nx, dev::ldv<NT, T, V>(Mi + i) * rv);
    }
  }
}

// PPCG outer iteration n, second pass with the block-Jacobi preconditioner:
// as k_pp_ur, with z = M^-1 r by the 4-row Thomas solve of tea_block_solve
// (kernels/tea_leaf_common_kernel.f90:378-468) over groups of 4 rows (tiles
// start on multiples of 4 rows).
template <bool NT, typename T, int V>
__global__ __launch_bounds__(BLOCK) void k_pp_ur_blk(
    Geo G, int n, int first, double rro_in, T *__restrict__ u, const T *__restrict__ p,
    T *__restrict__ r, T *__restrict__ r_store, const T *__restrict__ w, T *__restrict__ z,
    const T *__restrict__ Di, const T *__restrict__ Ky, const T *__restrict__ cp,
    const T *__restrict__ bfp, T ry, const double *__restrict__ part_pw,
    const double *__restrict__ part_zr, int nprev, Ctl *ctl, double *__restrict__ hrro,
    double *__restrict__ halpha, int rev) {
  typedef typename dev::Vec<T, V>::type VT;
  __shared__ double lds[BLOCK / WAVE];
  if (ctl->stop) return;
  const double pw = dev::sum_partials(part_pw, nprev, lds);
  const double rro = first ? rro_in : dev::sum_partials(part_zr, nprev, lds);
  if (nonfinite_stop(n, pw, rro, ctl)) return;
  const double alpha = rro / pw;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    hrro[n] = rro;
    halpha[n] = alpha;
  }
  const T a = (T)alpha;
  const long P = G.pitch;
  dev::TileIter it(G.tiles, G.blocks);
  for (long t = it.t; t < it.t_end; ++t) {
    const dev::TileV<V> tt = dev::tilev_of<V>(G, t);
    if (tt.c >= G.nx) continue;
    const int ngroups = (tt.y1 - tt.y0 + 3) / 4;
    for (int m = 0; m < ngroups; ++m) {
      const int g0 = tt.y0 + 4 * (rev ? ngroups - 1 - m : m);
      const int nq = min(4, G.ny - g0);
      VT rr[4], dp[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (q >= nq) break;
        const long i = (long)(g0 + q) * P + tt.c;
        const VT uv = dev::ldv<NT, T, V>(u + i) + a * dev::ldv<NT, T, V>(p + i);
        const VT ro = dev::ldv<NT, T, V>(r + i);
        rr[q] = ro - a * dev::ldv<NT, T, V>(w + i);
        dev::stv<false, T, V>(u + i - tt.c, tt.c, G.nx, uv);
        dev::stv<false, T, V>(r_store + i - tt.c, tt.c, G.nx, ro);
        dev::stv<false, T, V>(r + i - tt.c, tt.c, G.nx, rr[q]);
      }
      const long i0 = (long)g0 * P + tt.c;
      dp[0] = rr[0] / dev::ldv<NT, T, V>(Di + i0);
#pragma unroll
      for (int q = 1; q < 4; ++q) {
        if (q >= nq) break;
        const long i = i0 + (long)q * P;
        dp[q] = (rr[q] - (-dev::ldv<NT, T, V>(Ky + i) * ry) * dp[q - 1]) *
                dev::ldv<NT, T, V>(bfp + i);
      }
      VT zn = dp[nq - 1];
      for (int q = nq - 1; q >= 0; --q) {
        const long i = i0 + (long)q * P;
        const VT zq = q == nq - 1 ? dp[q] : dp[q] - dev::ldv<NT, T, V>(cp + i) * zn;
        dev::stv<false, T, V>(z + i - tt.c, tt.c, G.nx, zq);
        zn = zq;
      }
    }
  }
}

// One inner sweep with the block-Jacobi preconditioner
// (kernels/tea_leaf_ppcg_kernel.f90:269-285): rtemp -= A sd; z = M^-1 rtemp
// (tea_block_solve); sd' = alpha sd + beta z; utemp += sd'.  The rows are
// walked upwards on tiles that start on the 4-row blocks; the forward
// elimination runs as each row arrives and, once the rows of a block have
// been walked, the back substitution and the sd / utemp updates of the block
// follow.  FIRST forms sd = src/theta, rtemp = r, utemp = sd on the fly
// (tea_leaf_ppcg_init_sd); LAST writes z = utemp (update_z) and the
// partials of (r - r_store).z and z.r, and nothing else (sd, rtemp and utemp
// are dead after the last sweep of a fused solve).
template <bool NT, typename T, int V, int FIRST, int LAST>
__global__ __launch_bounds__(BLOCK) void k_pp_sweep_blk(
    Geo G, int want_rrn, T theta_r, T alpha, T beta, const T *__restrict__ src,
    const T *__restrict__ sdin, T *__restrict__ sdout, T *__restrict__ rtemp,
    T *__restrict__ utemp, const T *__restrict__ r, const T *__restrict__ r_store,
    T *__restrict__ z, const T *__restrict__ Di, const T *__restrict__ cp,
    const T *__restrict__ bfp, const T *__restrict__ Kx, const T *__restrict__ Ky, T rx, T ry,
    double *__restrict__ part_rrn, double *__restrict__ part_zr, const Ctl *ctl) {
  typedef typename dev::Vec<T, V>::type VT;
  __shared__ double lds[BLOCK / WAVE];
  if (ctl && ctl->stop) return;
  const long P = G.pitch;
  double arrn = 0.0, azr = 0.0;
  dev::TileIter it(G.tiles, G.blocks);
  for (long t = it.t; t < it.t_end; ++t) {
    const dev::TileV<V> tt = dev::tilev_of<V>(G, t);
    auto ldop = [&](int col, int row) -> VT {
      const long i = (long)row * P + col;
      return FIRST ? dev::ldv<NT, T, V>(src + i) * theta_r : dev::ldv<NT, T, V>(sdin + i);
    };
    auto ld1 = [&](int col, int row) -> T {
      const long i = (long)row * P + col;
      return FIRST ? src[i] * theta_r : sdin[i];
    };
    struct Side {
      VT rv, rt, ut, rs, ky, bfp, cp, di;
    };
    const int cc = tt.c <= G.nx ? tt.c : (G.nx / V) * V;
    auto side = [&](int row) -> Side {
      const long i = (long)row * P + cc;
      const bool head = (row & 3) == 0;  // first row of a block: Di, else bfp
      Side e;
      e.rv = (FIRST || LAST) ? dev::ldv<NT, T, V>(r + i) : VT(0);
      e.rt = FIRST ? VT(0) : dev::ldv<NT, T, V>(rtemp + i);
      e.ut = FIRST ? VT(0) : dev::ldv<NT, T, V>(utemp + i);
      e.rs = (LAST && want_rrn) ? dev::ldv<NT, T, V>(r_store + i) : VT(0);
      e.ky = dev::ldv<false, T, V>(Ky + i);
      e.cp = dev::ldv<NT, T, V>(cp + i);
      e.bfp = head ? VT(0) : dev::ldv<NT, T, V>(bfp + i);
      e.di = head ? dev::ldv<NT, T, V>(Di + i) : VT(0);
      return e;
    };
    // the rows of the current block in walking order (shifted in): dp of
    // the forward elimination, cp, sd and utemp before the sweep, r, r_store
    VT dps[4], cps[4], sds[4], uts[4], rvs[4], rss[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) dps[q] = cps[q] = sds[q] = uts[q] = rvs[q] = rss[q] = VT(0);
    dev::walkv<NT, T, V, diff_form<T>(), TL_PF_SWEEP, TL_LDSX_SWEEP>(
        G, tt, Kx, Ky, rx, ry, ldop, ld1, side, [&](int y, VT v, VT wv, const Side &e) {
      if (tt.c >= G.nx) return;
      const VT rt = (FIRST ? e.rv : e.rt) - wv;
      if (!LAST) dev::stv<false, T, V>(rtemp + (long)y * P, tt.c, G.nx, rt);
      const int g0 = y & ~3;
      const int nq = G.ny - g0 < 4 ? G.ny - g0 : 4;
      const VT dp = (y == g0) ? rt / e.di : (rt - (-e.ky * ry) * dps[3]) * e.bfp;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        dps[q] = dps[q + 1];
        cps[q] = cps[q + 1];
        sds[q] = sds[q + 1];
        uts[q] = uts[q + 1];
        rvs[q] = rvs[q + 1];
        rss[q] = rss[q + 1];
      }
      dps[3] = dp;
      cps[3] = e.cp;
      sds[3] = v;
      uts[3] = FIRST ? v : e.ut;
      rvs[3] = e.rv;
      rss[3] = e.rs;
      if (y - g0 + 1 < nq) return;
      for (int k = nq; k < 4; ++k) {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          dps[q] = dps[q + 1];
          cps[q] = cps[q + 1];
          sds[q] = sds[q + 1];
          uts[q] = uts[q + 1];
          rvs[q] = rvs[q + 1];
          rss[q] = rss[q + 1];
        }
      }
      VT zn = VT(0);
#pragma unroll
      for (int q = 3; q >= 0; --q) {
        if (q >= nq) continue;
        const VT zq = (q == nq - 1) ? dps[q] : dps[q] - cps[q] * zn;
        zn = zq;
        const VT sv = alpha * sds[q] + beta * zq;
        const VT ut = uts[q] + sv;
        const long ri = (long)(g0 + q) * P;
        if (!LAST) {
          dev::stv<false, T, V>(sdout + ri, tt.c, G.nx, sv);
          dev::stv<false, T, V>(utemp + ri, tt.c, G.nx, ut);
        } else {
          dev::stv<false, T, V>(z + ri, tt.c, G.nx, ut);
#pragma unroll
          for (int e = 0; e < V; ++e)
            if (tt.c + e < G.nx) {
              if (want_rrn) arrn += ((double)rvs[q][e] - (double)rss[q][e]) * (double)ut[e];
              azr += (double)ut[e] * (double)rvs[q][e];
            }
        }
      }
    }, 1);
  }
  if (LAST) {
    arrn = dev::block_sum(arrn, lds);
    azr = dev::block_sum(azr, lds);
    if (threadIdx.x == 0) {
      part_rrn[blockIdx.x] = arrn;
      part_zr[blockIdx.x] = azr;
    }
  }
}

// One inner sweep of the polynomial preconditioner (kernels/tea_leaf_ppcg_kernel.f90:288-306):
// rtemp -= A sd; sd' = alpha sd + beta (Mi) rtemp; utemp += sd'.  The first
// sweep forms sd = src/theta, rtemp = r, utemp = sd on the fly
// (tea_leaf_ppcg_init_sd); the last writes z = utemp
// (tea_leaf_ppcg_update_z) and the partials of (r - r_store).z
// (tea_leaf_ppcg_calc_rrn) and of z.r (the next iteration's zrnorm).
template <bool NT, typename T, int V>
__global__ __launch_bounds__(BLOCK) void k_pp_sweep(
    Geo G, int first_sweep, int last_sweep, int want_rrn, T theta_r, T alpha, T beta, int diag,
    const T *__restrict__ src, const T *__restrict__ sdin, T *__restrict__ sdout,
    T *__restrict__ rtemp, T *__restrict__ utemp, const T *__restrict__ r,
    const T *__restrict__ r_store, T *__restrict__ z, const T *__restrict__ Mi,
    const T *__restrict__ Kx, const T *__restrict__ Ky, T rx, T ry,
    double *__restrict__ part_rrn, double *__restrict__ part_zr, const Ctl *ctl, int dir) {
  typedef typename dev::Vec<T, V>::type VT;
  __shared__ double lds[BLOCK / WAVE];
  if (ctl && ctl->stop) return;  // no control block: sweeps outside a loop
  const long P = G.pitch;
  double arrn = 0.0, azr = 0.0;
  dev::TileIter it(G.tiles, G.blocks);
  for (long t = it.t; t < it.t_end; ++t) {
    const dev::TileV<V> tt = dev::tilev_of<V>(G, t);
    auto ldop = [&](int col, int row) -> VT {
      const long i = (long)row * P + col;
      return first_sweep ? dev::ldv<NT, T, V>(src + i) * theta_r : dev::ldv<NT, T, V>(sdin + i);
    };
    auto ld1 = [&](int col, int row) -> T {
      const long i = (long)row * P + col;
      return first_sweep ? src[i] * theta_r : sdin[i];
    };
    // per-row loads besides the operand: r, rtemp, utemp, r_store, Mi
    struct Side {
      VT rv, rt, ut, rs, mi;
    };
    const int cc = tt.c <= G.nx ? tt.c : (G.nx / V) * V;
    auto side = [&](int row) -> Side {
      const long i = (long)row * P + cc;
      Side e;
      e.rv = (first_sweep || last_sweep) ? dev::ldv<NT, T, V>(r + i) : VT(0);
      e.rt = first_sweep ? VT(0) : dev::ldv<NT, T, V>(rtemp + i);
      e.ut = first_sweep ? VT(0) : dev::ldv<NT, T, V>(utemp + i);
      e.rs = (last_sweep && want_rrn) ? dev::ldv<NT, T, V>(r_store + i) : VT(0);
      e.mi = diag ? dev::ldv<NT, T, V>(Mi + i) : VT(0);
      return e;
    };
    dev::walkv<NT, T, V, diff_form<T>(), TL_PF_SWEEP, TL_LDSX_SWEEP>(
        G, tt, Kx, Ky, rx, ry, ldop, ld1, side, [&](int y, VT v, VT wv, const Side &e) {
      if (tt.c >= G.nx) return;
      const VT rv = e.rv;
      const VT rt0 = first_sweep ? rv : e.rt;
      const VT ut0 = first_sweep ? v : e.ut;
      const VT rt = rt0 - wv;
      const VT sv = diag ? alpha * v + (beta * e.mi) * rt : alpha * v + beta * rt;
      const VT ut = ut0 + sv;
      dev::stv<(NT && TL_ST_NT_SWEEP && sizeof(T) == 8), T, V>(rtemp + (long)y * P, tt.c, G.nx, rt);
      dev::stv<(NT && TL_ST_NT_SWEEP && sizeof(T) == 8), T, V>(sdout + (long)y * P, tt.c, G.nx, sv);
      dev::stv<(NT && (TL_ST_NT_SWEEP || TL_ST_NT_UTEMP) && sizeof(T) == 8), T, V>(
          utemp + (long)y * P, tt.c, G.nx, ut);
      if (last_sweep) {
        dev::stv<(NT && TL_ST_NT_SWEEP && sizeof(T) == 8), T, V>(z + (long)y * P, tt.c, G.nx, ut);
        const VT rs = e.rs;
#pragma unroll
        for (int e = 0; e < V; ++e)
          if (tt.c + e < G.nx) {
            if (want_rrn) arrn += ((double)rv[e] - (double)rs[e]) * (double)ut[e];
            azr += (double)ut[e] * (double)rv[e];
          }
      }
    }, dir);
  }
  if (last_sweep) {
    arrn = dev::block_sum(arrn, lds);
    azr = dev::block_sum(azr, lds);
    if (threadIdx.x == 0) {
      part_rrn[blockIdx.x] = arrn;
      part_zr[blockIdx.x] = azr;
    }
  }
}

// Inner sweeps i and i+1 in one pass (tl_sweep2.hpp): reads sdin, rtin
// (r on the first sweep), utemp; writes sdout, rtout, utemp.  FIRST / LAST
// as first_sweep / last_sweep in k_pp_sweep, applying to sweep i / sweep
// i+1, except that the LAST pass writes z alone: it runs only inside the
// fused solves, where sd, rtemp and utemp are dead after the sweeps (unlike
// the per-routine tl_ppcg_inner, whose caller reads utemp in update_z).
// DIAG: the Jacobi-diagonal preconditioner.  They are template
// parameters, so that each of the eight kernels holds only the pointers and
// registers it uses (the middle passes of an outer iteration need neither r,
// r_store, z nor Mi).
#ifndef TL_PF_PAIR
#define TL_PF_PAIR 0
#endif

#ifndef TL_SWEEP_FIRST_PLAIN
#define TL_SWEEP_FIRST_PLAIN 1
#endif

template <typename VT>
struct SweepX {
  VT ut, rv, rs;  // utemp, and r and r_store on the last sweep
};

template <bool NT, typename T, int V, int FIRST, int LAST, int DIAG>
__global__ __launch_bounds__(BLOCK, sizeof(T) == 4 ? TL_PAIR_WAVES32 : TL_PAIR_WAVES64) void k_pp_sweep2(
    Geo G, int want_rrn, T theta_r, T a1, T b1, T a2, T b2, const T *__restrict__ src,
    const T *__restrict__ sdin, T *__restrict__ sdout, const T *__restrict__ rtin,
    T *__restrict__ rtout, T *__restrict__ utemp, const T *__restrict__ r,
    const T *__restrict__ r_store, T *__restrict__ z, const T *__restrict__ Mi,
    const T *__restrict__ Kx, const T *__restrict__ Ky, T rx, T ry, double *__restrict__ part_rrn,
    double *__restrict__ part_zr, const Ctl *ctl, int dir) {
  typedef typename dev::Vec<T, V>::type VT;
  typedef SweepX<VT> X;
  constexpr bool ST = TL_ST_NT_SWEEP && sizeof(T) == 8;
  __shared__ double lds[BLOCK / WAVE];
  if (ctl && ctl->stop) return;
  const long P = G.pitch;
  double arrn = 0.0, azr = 0.0;
  dev::TileIter it(G.tiles, G.blocks);
  for (long t = it.t; t < it.t_end; ++t) {
    const dev::TileV<V> tt = dev::tilev_of<V>(G, t);
    // first pass without a preconditioner: src is r, read here one row
    // ahead of rtA's read of the same row.  In FP32 plain loads let the
    // second come from the L2 (mixed PPCG step 1.4 % faster at 4096^2); in
    // FP64 the streaming loads measured faster (1.9 %)
    constexpr bool NTF =
        NT && !(FIRST && !DIAG && sizeof(T) == 4 && TL_SWEEP_FIRST_PLAIN);
    auto opA = [&](int col, int row) -> VT {
      const long i = (long)row * P + col;
      return FIRST ? dev::ldv<NTF, T, V>(src + i) * theta_r : dev::ldv<NT, T, V>(sdin + i);
    };
    auto opA1 = [&](int col, int row) -> T {
      const long i = (long)row * P + col;
      return FIRST ? src[i] * theta_r : sdin[i];
    };
    auto rtA = [&](int col, int row) -> VT {
      const long i = (long)row * P + col;
      return FIRST ? dev::ldv<NTF, T, V>(r + i) : dev::ldv<NT, T, V>(rtin + i);
    };
    auto rtA1 = [&](int col, int row) -> T {
      const long i = (long)row * P + col;
      return FIRST ? r[i] : rtin[i];
    };
    auto xtra = [&](int col, int row) -> X {
      const long i = (long)row * P + col;
      X x;
      x.ut = FIRST ? VT(0) : dev::ldv<NT, T, V>(utemp + i);
      x.rv = LAST ? dev::ldv<NT, T, V>(r + i) : VT(0);
      x.rs = (LAST && want_rrn) ? dev::ldv<NT, T, V>(r_store + i) : VT(0);
      return x;
    };
    dev::walk_pair<NT, T, V, diff_form<T>(), DIAG, (TL_PF_PAIR != 0)>(
        G, tt, Kx, Ky, Mi, rx, ry, a1, b1, opA, opA1, rtA, rtA1, xtra,
        [&](int y, const VT &s1, const VT &r1, const VT &s0, const VT &w, const VT &mi,
            const X &x) {
          if (tt.c >= G.nx) return;
          const long row = (long)y * P;
          const VT r2 = r1 - w;
          const VT s2 = DIAG ? a2 * s1 + (b2 * mi) * r2 : a2 * s1 + b2 * r2;
          const VT ut0 = FIRST ? s0 : x.ut;
          const VT ut = (ut0 + s1) + s2;
          if (!LAST) {
            dev::stv<(NT && ST), T, V>(rtout + row, tt.c, G.nx, r2);
            dev::stv<(NT && ST), T, V>(sdout + row, tt.c, G.nx, s2);
            dev::stv<(NT && (ST || (TL_ST_NT_UTEMP && sizeof(T) == 8))), T, V>(utemp + row, tt.c,
                                                                                G.nx, ut);
          } else {
            // the last pass of a solve's sweeps: only z = utemp (update_z)
            // is read afterwards; the next first pass starts sd, rtemp and
            // utemp afresh from r
            dev::stv<(NT && ST), T, V>(z + row, tt.c, G.nx, ut);
#pragma unroll
            for (int e = 0; e < V; ++e)
              if (tt.c + e < G.nx) {
                if (want_rrn) arrn += ((double)x.rv[e] - (double)x.rs[e]) * (double)ut[e];
                azr += (double)ut[e] * (double)x.rv[e];
              }
          }
        },
        dir);
  }
  if (LAST) {
    arrn = dev::block_sum(arrn, lds);
    azr = dev::block_sum(azr, lds);
    if (threadIdx.x == 0) {
      part_rrn[blockIdx.x] = arrn;
      part_zr[blockIdx.x] = azr;
    }
  }
}

// The two-sweep pass for a (first, last, diag) combination.
template <bool NT, typename T, int V>
auto sweep2_kernel(int first, int last, int diag) {
  using K = decltype(&k_pp_sweep2<NT, T, V, 0, 0, 0>);
  static const K tab[8] = {
      &k_pp_sweep2<NT, T, V, 0, 0, 0>, &k_pp_sweep2<NT, T, V, 0, 0, 1>,
      &k_pp_sweep2<NT, T, V, 0, 1, 0>, &k_pp_sweep2<NT, T, V, 0, 1, 1>,
      &k_pp_sweep2<NT, T, V, 1, 0, 0>, &k_pp_sweep2<NT, T, V, 1, 0, 1>,
      &k_pp_sweep2<NT, T, V, 1, 1, 0>, &k_pp_sweep2<NT, T, V, 1, 1, 1>};
  return tab[(first ? 4 : 0) + (last ? 2 : 0) + (diag ? 1 : 0)];
}

// ---------------------------------------------------------------------------
// host loops over one set of buffers (FP64 fields or the FP32 inner fields)
// ---------------------------------------------------------------------------
template <typename T>
struct LoopBufs {
  Geo G;                  // geometry; G.pitch in elements of T
  T *u, *r, *w, *z, *r_store, *rtemp, *utemp;
  T *p[2];
  T *sd[2];
  int *pcur, *sdcur;
  const T *Mi, *Kx, *Ky;
  T rx, ry;
  int diag;               // Jacobi-diagonal preconditioner
  const T *cg_src;        // r, or z when preconditioned: the CG direction source
  // block-Jacobi preconditioner (FP64 CG and PPCG): Di, Thomas coefficients;
  // G's tiles then start on multiples of 4 rows
  int block = 0;
  const T *Di = nullptr, *cp = nullptr, *bfp = nullptr;
};

inline int abs_blocks(const Geo &G) { return G.blocks < 0 ? -G.blocks : G.blocks; }

// Two inner sweeps per pass unless TL_SWEEP_PAIRS=0 (read on every call).
inline bool sweep_pairs() {
  const char *e = getenv("TL_SWEEP_PAIRS");
  return !e || atoi(e) != 0;
}

// Queue the `inner` sweeps of one PPCG outer iteration (or of a preconditioner
// application when ctl is null): pairs of sweeps in one pass each, a single
// pass for an odd last one.  sd ping-pongs per pass; rtemp alternates
// between b.rtemp and b.w (w is dead between the update and the next w pass).
// Passes walk in alternating directions starting with `dir`; without pairs,
// both sweeps of what would be a pair take that pair's direction, so the two
// modes produce the same bits.  The last sweep leaves the partials of
// (r - r_store).z in slot P3 and of z.r in slot P2 of `parts`.
template <typename T, int V>
int queue_sweeps(const LoopBufs<T> &b, const double *alphas, const double *betas, int inner,
                 T theta_r, int want_rrn, const Ctl *ctl, const LoopMem &m, int &sd, T *&rt_cur,
                 int dir) {
  const Geo G = b.G;
  const int nb = abs_blocks(G);
  const bool nt = use_nontemporal();
  if constexpr (sizeof(T) == 8) {
    if (b.block) {
      // block preconditioner: one sweep per pass, walked upwards; the first
      // reads z (M^-1 r from the update pass) and the last writes z, so they
      // must be different passes (tl_ppcg_loop routes solves of one or two
      // preconditioned sweeps to the per-routine path)
      using KB = decltype(&k_pp_sweep_blk<true, T, V, 0, 0>);
      static const KB kb[2][2][2] = {
          {{&k_pp_sweep_blk<false, T, V, 0, 0>, &k_pp_sweep_blk<false, T, V, 0, 1>},
           {&k_pp_sweep_blk<false, T, V, 1, 0>, &k_pp_sweep_blk<false, T, V, 1, 1>}},
          {{&k_pp_sweep_blk<true, T, V, 0, 0>, &k_pp_sweep_blk<true, T, V, 0, 1>},
           {&k_pp_sweep_blk<true, T, V, 1, 0>, &k_pp_sweep_blk<true, T, V, 1, 1>}}};
      for (int i = 1; i <= inner; ++i) {
        hipEvent_t ev;
        prof_arm(2, &ev);
        klaunch(kb[nt ? 1 : 0][i == 1 ? 1 : 0][i == inner ? 1 : 0], dim3(nb),
                dim3(BLOCK), G, want_rrn, theta_r, (T)alphas[i - 1],
                (T)betas[i - 1], b.z, b.sd[sd], b.sd[1 - sd], rt_cur, b.utemp, b.r,
                b.r_store, b.z, b.Di, b.cp, b.bfp, b.Kx, b.Ky, b.rx, b.ry,
                m.part(P3), m.part(P2), ctl);
        prof_end(2, ev);
        sd = 1 - sd;
      }
      TL_HIP(hipGetLastError());
      return TL_OK;
    }
  }
  auto k1 = nt ? k_pp_sweep<true, T, V> : k_pp_sweep<false, T, V>;
  const T *src = b.diag ? b.z : b.r;
  const bool pairs = sweep_pairs();
  for (int i = 1; i <= inner;) {
    hipEvent_t ev;
    prof_arm(2, &ev);
    if (pairs && i + 1 <= inner) {
      T *rt_next = rt_cur == b.rtemp ? b.w : b.rtemp;
      auto k2 = nt ? sweep2_kernel<true, T, V>(i == 1, i + 1 == inner, b.diag)
                   : sweep2_kernel<false, T, V>(i == 1, i + 1 == inner, b.diag);
      klaunch(k2, dim3(nb), dim3(BLOCK), G, want_rrn, theta_r,
              (T)alphas[i - 1], (T)betas[i - 1], (T)alphas[i], (T)betas[i], src,
              b.sd[sd], b.sd[1 - sd], rt_cur, rt_next, b.utemp, b.r, b.r_store, b.z,
              b.Mi, b.Kx, b.Ky, b.rx, b.ry, m.part(P3), m.part(P2), ctl, dir);
      rt_cur = rt_next;
      i += 2;
    } else {
      klaunch(k1, dim3(nb), dim3(BLOCK), G, i == 1 ? 1 : 0,
              i == inner ? 1 : 0, want_rrn, theta_r, (T)alphas[i - 1], (T)betas[i - 1],
              b.diag, src, b.sd[sd], b.sd[1 - sd], rt_cur, b.utemp, b.r, b.r_store,
              b.z, b.Mi, b.Kx, b.Ky, b.rx, b.ry, m.part(P3), m.part(P2), ctl, dir);
      if (pairs || (i & 1) == 0) dir = -dir;
      i += 1;
      prof_end(2, ev);
      sd = 1 - sd;
      continue;
    }
    prof_end(2, ev);
    sd = 1 - sd;
    dir = -dir;
  }
  TL_HIP(hipGetLastError());
  return TL_OK;
}

// CG iterations (tea_solve.f90:341-369); see tl_cg_loop.  pin_first, when
// given, replaces p[pcur] as the direction of the first iteration.  The
// scalars live in `mem` when given (which must hold histories for
// iterations up to st->n_last), else in the library's loop memory.
template <typename T, int V>
int run_cg_loop(const LoopBufs<T> &b, tl_loop_state *st, double *cg_alphas, double *cg_betas,
                const T *pin_first = nullptr, const LoopMem *mem = nullptr) {
  State &s = S();
  int rc;
  if (!mem && (rc = ensure_loop_buffers(st->n_last))) return rc;
  const LoopMem m = mem ? *mem : loop_mem();
  const Geo G = b.G;
  const int nb = abs_blocks(G);
  const bool nt = use_nontemporal();
  auto kw = nt ? k_loop_w<true, T, V> : k_loop_w<false, T, V>;
  auto kur = nt ? k_cg_ur<true, T, V> : k_cg_ur<false, T, V>;
  auto kub = nt ? k_cg_ur_blk<true, T, V> : k_cg_ur_blk<false, T, V>;
  const int c0 = *b.pcur;
  LoopParams lp;
  lp.eps_abs = st->eps_abs;
  lp.epslim = st->epslim;
  lp.check_switch = st->check_switch;
  lp.presteps = st->presteps;
  auto launch = [&](int n, bool first) -> int {
    const int cin = c0 ^ ((n - st->n_first) & 1);
    const T *pin = (first && pin_first) ? pin_first : b.p[cin];
    hipEvent_t ev;
    prof_arm(0, &ev);
    klaunch(kw, dim3(nb), dim3(BLOCK), G, n, first ? 1 : 0, 0, lp,
            st->rro, b.cg_src, pin, b.p[1 - cin], b.Kx, b.Ky, b.w, b.rx, b.ry,
            m.part(P2), nb, m.part(P1), m.ctl, m.hrro, m.hbeta, m.herr, 1);
    prof_end(0, ev);
    prof_arm(1, &ev);
    if (b.block)
      klaunch(kub, dim3(nb), dim3(BLOCK), G, n, b.u, b.p[1 - cin], b.r,
              b.w, b.z, b.Di, b.Ky, b.cp, b.bfp, b.ry, m.part(P1), nb, m.part(P2),
              m.ctl, m.hrro, m.halpha, 1);
    else
      klaunch(kur, dim3(nb), dim3(BLOCK), G, n, b.u, b.p[1 - cin], b.r,
              b.w, b.z, b.Mi, b.diag, m.part(P1), nb, m.part(P2), m.ctl, m.hrro,
              m.halpha, 1);
    prof_end(1, ev);
    TL_HIP(hipGetLastError());
    return TL_OK;
  };
  auto tail = [&](int n) -> int {
    const int cin = c0 ^ ((n - st->n_first) & 1);
    hipLaunchKernelGGL(kw, dim3(nb), dim3(BLOCK), 0, s.stream, G, n, 0, 1, lp, st->rro, b.cg_src,
                       b.p[cin], b.p[1 - cin], b.Kx, b.Ky, b.w, b.rx, b.ry, m.part(P2), nb,
                       m.part(P1), m.ctl, m.hrro, m.hbeta, m.herr, 1);
    TL_HIP(hipGetLastError());
    return TL_OK;
  };
  Ctl res;
  if ((rc = drive(st->n_first, st->n_last, launch, tail, &res, m.ctl))) return rc;
  st->n_end = res.n_end;
  st->reason = res.reason;
  st->error = res.error;
  st->rro = res.rro;
  st->old_error = res.old_error;
  if (res.reason == TL_LOOP_NONFINITE) return nonfinite_fail("CG", res.n_end);
  // the launch that stopped the loop applied the last p update
  const int n_det = res.reason == TL_LOOP_SWITCH ? res.n_end : res.n_end + 1;
  *b.pcur = c0 ^ ((n_det - st->n_first + 1) & 1);
  const int last_done = res.reason == TL_LOOP_SWITCH ? res.n_end - 1 : res.n_end;
  if ((rc = copy_hist(cg_alphas, m.halpha, st->n_first, last_done))) return rc;
  if ((rc = copy_hist(cg_betas, m.hbeta, st->n_first, last_done))) return rc;
  return TL_OK;
}

// PPCG outer iterations after the switch (tea_solve.f90:289-331); see
// tl_ppcg_loop.  rro_first is z.r entering n_first; pin_first, when given,
// replaces p[pcur] as the direction of the first iteration.
template <typename T, int V>
int run_ppcg_loop(const LoopBufs<T> &b, tl_loop_state *st, const double *alphas,
                  const double *betas, int inner, double theta, double rro_first,
                  const T *pin_first, int *inner_iters) {
  State &s = S();
  int rc;
  if ((rc = ensure_loop_buffers(st->n_last))) return rc;
  const LoopMem m = loop_mem();
  const Geo G = b.G;
  const int nb = abs_blocks(G);
  const bool nt = use_nontemporal();
  auto kw = nt ? k_loop_w<true, T, V> : k_loop_w<false, T, V>;
  auto kur = nt ? k_pp_ur<true, T, V> : k_pp_ur<false, T, V>;
  const int c0 = *b.pcur;
  int sd = *b.sdcur;
  LoopParams lp;
  lp.eps_abs = st->eps_abs;
  lp.epslim = 0.0;
  lp.check_switch = 0;
  lp.presteps = 0;
  const T theta_r = (T)(1.0 / theta);
  // Row directions: the w pass walks d, the update -d, the sweep passes
  // alternate from d; with an odd number of sweep passes the next outer
  // iteration flips d, so every pass starts on the rows written last.
  const int passes = inner / 2 + inner % 2;
  auto launch = [&](int n, bool first) -> int {
    const int cin = c0 ^ ((n - st->n_first) & 1);
    const T *pin = (first && pin_first) ? pin_first : b.p[cin];
    const int d = ((passes & 1) && ((n - st->n_first) & 1)) ? -1 : 1;
    hipEvent_t ev;
    prof_arm(5, &ev);
    klaunch(kw, dim3(nb), dim3(BLOCK), G, n, first ? 1 : 0, 0, lp,
            rro_first, b.z, pin, b.p[1 - cin], b.Kx, b.Ky, b.w, b.rx, b.ry,
            m.part(P3), nb, m.part(P1), m.ctl, m.hrro, m.hbeta, m.herr, d);
    prof_end(5, ev);
    prof_arm(6, &ev);
    bool block_ur = false;
    if constexpr (sizeof(T) == 8) {
      if (b.block) {
        auto kub = nt ? k_pp_ur_blk<true, T, V> : k_pp_ur_blk<false, T, V>;
        klaunch(kub, dim3(nb), dim3(BLOCK), G, n, first ? 1 : 0,
                rro_first, b.u, b.p[1 - cin], b.r, b.r_store, b.w, b.z, b.Di, b.Ky,
                b.cp, b.bfp, b.ry, m.part(P1), m.part(P2), nb, m.ctl, m.hrro,
                m.halpha, d > 0 ? 1 : 0);
        block_ur = true;
      }
    }
    if (!block_ur)
      klaunch(kur, dim3(nb), dim3(BLOCK), G, n, first ? 1 : 0,
              rro_first, b.u, b.p[1 - cin], b.r, b.r_store, b.w, b.z, b.Mi, b.diag,
              m.part(P1), m.part(P2), nb, m.ctl, m.hrro, m.halpha, d > 0 ? 1 : 0);
    prof_end(6, ev);
    T *rt_cur = b.rtemp;
    return queue_sweeps<T, V>(b, alphas, betas, inner, theta_r, 1, m.ctl, m, sd, rt_cur, d);
  };
  auto tail = [&](int n) -> int {
    const int cin = c0 ^ ((n - st->n_first) & 1);
    hipLaunchKernelGGL(kw, dim3(nb), dim3(BLOCK), 0, s.stream, G, n, 0, 1, lp, rro_first, b.z,
                       b.p[cin], b.p[1 - cin], b.Kx, b.Ky, b.w, b.rx, b.ry, m.part(P3), nb,
                       m.part(P1), m.ctl, m.hrro, m.hbeta, m.herr, 1);
    TL_HIP(hipGetLastError());
    return TL_OK;
  };
  Ctl res;
  if ((rc = drive(st->n_first, st->n_last, launch, tail, &res, m.ctl))) return rc;
  st->n_end = res.n_end;
  st->reason = res.reason;
  st->error = res.error;
  st->rro = res.rro;
  st->old_error = res.old_error;
  if (res.reason == TL_LOOP_NONFINITE) return nonfinite_fail("PPCG", res.n_end);
  const int executed = res.n_end - st->n_first + 1;  // outer iterations that ran
  if (inner_iters) *inner_iters += executed * inner;
  *b.pcur = c0 ^ ((executed + 1) & 1);               // + the stopping launch's p update
  const int swaps = sweep_pairs() ? passes : inner;
  *b.sdcur = *b.sdcur ^ ((executed * swaps) & 1);
  return TL_OK;
}

// Sweeps only (the polynomial preconditioner applied to r into z, from
// scratch): tea_leaf_ppcg_init_sd + the inner steps + update_z, with the
// partials of z.r left in slot P2.  Used to start an inner PPCG solve.
template <typename T, int V>
int run_ppcg_precondition(const LoopBufs<T> &b, const double *alphas, const double *betas,
                          int inner, double theta) {
  int sd = *b.sdcur;
  T *rt_cur = b.rtemp;
  const LoopMem m = loop_mem();
  const int rc = queue_sweeps<T, V>(b, alphas, betas, inner, (T)(1.0 / theta), 0,
                                    (const Ctl *)nullptr, m, sd, rt_cur, 1);
  if (rc) return rc;
  *b.sdcur = sd;
  return TL_OK;
}

}  // namespace
}  // namespace tl
