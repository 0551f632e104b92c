// Fused solver loops: consecutive iterations of the tea_leaf() DO loop
// (tea_solve.f90:152-414) with every scalar kept on the device.
//
// Each reduction is finished redundantly by every block of the *next* kernel:
// block b of kernel k+1 sums the per-block partials kernel k wrote, in one
// fixed order, so all blocks (and the host, later) see the same bits.  That
// removes the host round trip per dot product and needs no atomics or
// cross-XCD fences; kernel boundaries order everything.  The loop control
// block (stop flag, reason, iteration) is written by block 0 of the kernel
// that decides; later kernels of an in-flight batch see the flag and return.
#include <cmath>
#include <cstdlib>

#include "tl_internal.hpp"
#include "tl_stencil.hpp"
#include "tl_stencil2.hpp"

namespace tl {
dev::Geo geo(int cols_per_tile);
}
using namespace tl;
using dev::Geo;

namespace {

struct Ctl {
  int stop, reason, n_end, pad;
  double error, old_error, rro;
};

struct LoopParams {
  double eps_abs;
  double epslim;
  int check_switch;
  int presteps;
};

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
// CG: iteration n = K1(n) [beta of n-1, exit/switch test, p = src + beta p,
// w = A p, pw partials] then K2(n) [alpha = rro/pw, u += alpha p,
// r -= alpha w, (z = Mi r), rrn partials].
// ---------------------------------------------------------------------------
template <bool NT>
__global__ __launch_bounds__(BLOCK) void k_cgl_w(Geo G, int n, int first, int tail,
                                                  LoopParams lp, double rro_in,
                                                  const double *__restrict__ src,
                                                  const double *__restrict__ pin,
                                                  double *__restrict__ pout,
                                                  const double *__restrict__ Kx,
                                                  const double *__restrict__ Ky,
                                                  double *__restrict__ w, double rx, double ry,
                                                  const double *__restrict__ prev_part, int nprev,
                                                  double *__restrict__ out_part, Ctl *ctl,
                                                  double *__restrict__ hrro,
                                                  double *__restrict__ hbeta) {
  using dev::d2;
  __shared__ double lds[BLOCK / WAVE];
  double beta = 0.0;
  bool stop = false;
  if (!first) {
    if (ctl->stop) return;
    const double rrn = dev::sum_partials(prev_part, nprev, lds);
    const double rro = hrro[n - 1];
    beta = rrn / rro;                                   // tea_solve.f90:362
    const double err = sqrt(fabs(rrn));                 // :368,402
    const bool conv = fabs(err) < lp.eps_abs;           // :410
    const bool sw = !conv && !tail && lp.check_switch && fabs(err) <= lp.epslim &&
                    n >= lp.presteps;                   // :167 at the top of iteration n
    stop = conv || sw || tail;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      hbeta[n - 1] = beta;
      if (stop) {
        ctl->stop = 1;
        ctl->reason = conv ? TL_LOOP_CONVERGED : (sw ? TL_LOOP_SWITCH : TL_LOOP_MAXITERS);
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
  // p = src + beta p of iteration n-1 (tea_solve.f90:366) is applied even when
  // the loop stops here, so p ends exactly as the reference leaves it.
  const bool upd = !first;
  const long P = G.pitch;
  double acc = 0.0;
  dev::TileIter it(G.tiles, G.blocks);
  for (long t = it.t; t < it.t_end; ++t) {
    const dev::Tile2 tt = dev::tile2_of(G, t);
    if (stop) {
      if (tt.c < G.nx)
        for (int y = tt.y0; y < tt.y1; ++y) {
          const long i = (long)y * P + tt.c;
          const d2 v = dev::ld2<NT>(src + i) + beta * dev::ld2<NT>(pin + i);
          dev::st_pair<NT>(pout + (long)y * P, tt.c, G.nx, v);
        }
      continue;
    }
    auto ldp = [&](int col, int row) -> d2 {
      const long i = (long)row * P + col;
      const d2 pv = dev::ld2<NT>(pin + i);
      return upd ? dev::ld2<NT>(src + i) + beta * pv : pv;
    };
    auto ld1p = [&](int col, int row) -> double {
      const long i = (long)row * P + col;
      return upd ? src[i] + beta * pin[i] : pin[i];
    };
    dev::walk2<NT>(G, tt, Kx, Ky, rx, ry, ldp, ld1p, [&](int y, d2 v, d2 wv) {
      if (tt.c < G.nx) {
        dev::st_pair<NT>(w + (long)y * P, tt.c, G.nx, wv);
        dev::st_pair<NT>(pout + (long)y * P, tt.c, G.nx, v);
        acc += wv.x * v.x;
        if (tt.c + 1 < G.nx) acc += wv.y * v.y;
      }
    });
  }
  if (stop) return;
  acc = dev::block_sum(acc, lds);
  if (threadIdx.x == 0) out_part[blockIdx.x] = acc;
}

template <bool NT>
__global__ __launch_bounds__(BLOCK) void k_cgl_ur(Geo G, int n, double *__restrict__ u,
                                                   const double *__restrict__ p,
                                                   double *__restrict__ r,
                                                   const double *__restrict__ w,
                                                   double *__restrict__ z,
                                                   const double *__restrict__ Mi, int diag,
                                                   const double *__restrict__ prev_part, int nprev,
                                                   double *__restrict__ out_part, Ctl *ctl,
                                                   const double *__restrict__ hrro,
                                                   double *__restrict__ halpha) {
  using dev::d2;
  __shared__ double lds[BLOCK / WAVE];
  if (ctl->stop) return;
  const double pw = dev::sum_partials(prev_part, nprev, lds);
  const double alpha = hrro[n] / pw;                    // tea_solve.f90:352
  if (blockIdx.x == 0 && threadIdx.x == 0) halpha[n] = alpha;
  const long P = G.pitch;
  double acc = 0.0;
  dev::TileIter it(G.tiles, G.blocks);
  for (long t = it.t; t < it.t_end; ++t) {
    const dev::Tile2 tt = dev::tile2_of(G, t);
    if (tt.c >= G.nx) continue;
    const bool pair = tt.c + 1 < G.nx;
    for (int y = tt.y0; y < tt.y1; ++y) {
      const long i = (long)y * P + tt.c;
      const d2 uv = dev::ld2<NT>(u + i) + alpha * dev::ld2<NT>(p + i);
      const d2 rv = dev::ld2<NT>(r + i) - alpha * dev::ld2<NT>(w + i);
      dev::st_pair<NT>(u + (long)y * P, tt.c, G.nx, uv);
      dev::st_pair<NT>(r + (long)y * P, tt.c, G.nx, rv);
      if (diag) {
        const d2 zv = dev::ld2<NT>(Mi + i) * rv;
        dev::st_pair<NT>(z + (long)y * P, tt.c, G.nx, zv);
        acc += rv.x * zv.x;
        if (pair) acc += rv.y * zv.y;
      } else {
        acc += rv.x * rv.x;
        if (pair) acc += rv.y * rv.y;
      }
    }
  }
  acc = dev::block_sum(acc, lds);
  if (threadIdx.x == 0) out_part[blockIdx.x] = acc;
}

// ---------------------------------------------------------------------------
// PPCG outer iteration n (tea_solve.f90:289-331), z aliased into the z field:
//  Ka(n): beta = rrn/rro of n-1, exit test, p = z + beta p, w = A p,
//         partials of p.w and z.r
//  Kb(n): alpha = rro/pw, u += alpha p, r_store = r, r -= alpha w, (z = Mi r)
//  S(n,i): inner sweep i; the first forms sd = src/theta on the fly, the last
//         writes z = utemp and the partials of (r - r_store).z
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_ppl_w(Geo G, int n, int first, int tail,
                                                  double eps_abs, const double *__restrict__ z,
                                                  const double *__restrict__ r,
                                                  const double *__restrict__ pin,
                                                  double *__restrict__ pout,
                                                  const double *__restrict__ Kx,
                                                  const double *__restrict__ Ky,
                                                  double *__restrict__ w, double rx, double ry,
                                                  const double *__restrict__ prev_part, int nprev,
                                                  double *__restrict__ part_pw,
                                                  double *__restrict__ part_zr, Ctl *ctl,
                                                  const double *__restrict__ hrro,
                                                  double *__restrict__ hbeta) {
  __shared__ double lds[BLOCK / WAVE];
  double beta = 0.0;
  bool stop = false;
  if (!first) {
    if (ctl->stop) return;
    const double rrn = dev::sum_partials(prev_part, nprev, lds);
    beta = rrn / hrro[n - 1];                           // :325
    const double err = sqrt(fabs(rrn));                 // :329,402
    const bool conv = fabs(err) < eps_abs;
    stop = conv || tail;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      hbeta[n - 1] = beta;
      if (stop) {
        ctl->stop = 1;
        ctl->reason = conv ? TL_LOOP_CONVERGED : TL_LOOP_MAXITERS;
        ctl->n_end = n - 1;
        ctl->error = err;
        ctl->old_error = err;
        ctl->rro = rrn;
      }
    }
  } else if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctl->stop = 0;
  }
  const bool upd = !first;
  double apw = 0.0, azr = 0.0;
  TILE_LOOP(G)
  if (stop) {
    for (int y = tt.y0; y < tt.y1; ++y) {
      const long i = IDX(y, x);
      pout[i] = z[i] + beta * pin[i];                   // cg_calc_p with tl_ppcg_active
    }
  } else {
    dev::walk<MIRROR>(
        G, x, tt.y0, tt.y1,
        [&](int xx, int yy) {
          const long i = IDX(yy, xx);
          return upd ? z[i] + beta * pin[i] : pin[i];
        },
        Kx, Ky,
        [&](int y, double c, double l, double rr, double d, double up, double kxl, double kxr,
            double kyd, double kyu) {
          const long i = IDX(y, x);
          const double wv = dev::apply_a(c, l, rr, d, up, kxl, kxr, kyd, kyu, rx, ry);
          w[i] = wv;
          pout[i] = c;
          apw += wv * c;
          azr += z[i] * r[i];
        });
  }
  TILE_LOOP_END
  if (stop) return;
  apw = dev::block_sum(apw, lds);
  azr = dev::block_sum(azr, lds);
  if (threadIdx.x == 0) {
    part_pw[blockIdx.x] = apw;
    part_zr[blockIdx.x] = azr;
  }
}

__global__ __launch_bounds__(BLOCK) void k_ppl_ur(Geo G, int n, double *__restrict__ u,
                                                   const double *__restrict__ p,
                                                   double *__restrict__ r,
                                                   double *__restrict__ r_store,
                                                   const double *__restrict__ w,
                                                   double *__restrict__ z,
                                                   const double *__restrict__ Mi, int diag,
                                                   const double *__restrict__ part_pw,
                                                   const double *__restrict__ part_zr, int nprev,
                                                   Ctl *ctl, double *__restrict__ hrro,
                                                   double *__restrict__ halpha) {
  __shared__ double lds[BLOCK / WAVE];
  if (ctl->stop) return;
  const double pw = dev::sum_partials(part_pw, nprev, lds);
  const double rro = dev::sum_partials(part_zr, nprev, lds);
  const double alpha = rro / pw;                        // :307
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    hrro[n] = rro;
    halpha[n] = alpha;
  }
  TILE_LOOP(G)
  for (int y = tt.y0; y < tt.y1; ++y) {
    const long i = IDX(y, x);
    u[i] = u[i] + alpha * p[i];
    const double ro = r[i];
    r_store[i] = ro;
    const double rv = ro - alpha * w[i];
    r[i] = rv;
    if (diag) z[i] = Mi[i] * rv;
  }
  TILE_LOOP_END
}

__global__ __launch_bounds__(BLOCK) void k_ppl_sweep(
    Geo G, int first_sweep, int last_sweep, double theta_r, double alpha, double beta, int diag,
    const double *__restrict__ src, const double *__restrict__ sdin, double *__restrict__ sdout,
    double *__restrict__ rtemp, double *__restrict__ utemp, const double *__restrict__ r,
    const double *__restrict__ r_store, double *__restrict__ z, const double *__restrict__ Mi,
    const double *__restrict__ Kx, const double *__restrict__ Ky, double rx, double ry,
    double *__restrict__ part_rrn, Ctl *ctl) {
  __shared__ double lds[BLOCK / WAVE];
  if (ctl->stop) return;
  double acc = 0.0;
  TILE_LOOP(G)
  dev::walk<MIRROR>(
      G, x, tt.y0, tt.y1,
      [&](int xx, int yy) {
        const long i = IDX(yy, xx);
        return first_sweep ? src[i] * theta_r : sdin[i];
      },
      Kx, Ky,
      [&](int y, double c, double l, double rr, double d, double up, double kxl, double kxr,
          double kyd, double kyu) {
        const long i = IDX(y, x);
        const double rt0 = first_sweep ? r[i] : rtemp[i];
        const double ut0 = first_sweep ? c : utemp[i];
        const double rt = rt0 - dev::apply_a(c, l, rr, d, up, kxl, kxr, kyd, kyu, rx, ry);
        const double s = diag ? alpha * c + beta * Mi[i] * rt : alpha * c + beta * rt;
        const double ut = ut0 + s;
        rtemp[i] = rt;
        sdout[i] = s;
        utemp[i] = ut;
        if (last_sweep) {
          z[i] = ut;                                    // tea_leaf_ppcg_update_z
          acc += (r[i] - r_store[i]) * ut;              // tea_leaf_ppcg_calc_rrn
        }
      });
  TILE_LOOP_END
  if (last_sweep) {
    acc = dev::block_sum(acc, lds);
    if (threadIdx.x == 0) part_rrn[blockIdx.x] = acc;
  }
}

// ---------------------------------------------------------------------------
// Jacobi iteration n (tea_solve.f90:375-379): exit test of n-1, then the
// sweep from u[in] to u[out] with its |u - un| partials (parity-buffered).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_jl(Geo G, int n, int first, double eps_abs,
                                               const double *__restrict__ un,
                                               double *__restrict__ uout,
                                               const double *__restrict__ u0,
                                               const double *__restrict__ Kx,
                                               const double *__restrict__ Ky, double rx, double ry,
                                               const double *__restrict__ prev_part, int nprev,
                                               double *__restrict__ out_part, Ctl *ctl) {
  __shared__ double lds[BLOCK / WAVE];
  if (!first) {
    if (ctl->stop) return;
    const double e = sqrt(fabs(dev::sum_partials(prev_part, nprev, lds)));
    const bool conv = fabs(e) < eps_abs;
    if (conv) {
      if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl->stop = 1; ctl->reason = TL_LOOP_CONVERGED; ctl->n_end = n - 1; ctl->error = e;
      }
      return;
    }
  } else if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctl->stop = 0;
  }
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

__global__ void k_jl_tail(int n, double eps_abs, const double *__restrict__ prev_part, int nprev,
                          Ctl *ctl) {
  __shared__ double lds[BLOCK / WAVE];
  if (ctl->stop) return;
  const double e = sqrt(fabs(dev::sum_partials(prev_part, nprev, lds)));
  if (threadIdx.x == 0) {
    ctl->stop = 1;
    ctl->reason = fabs(e) < eps_abs ? TL_LOOP_CONVERGED : TL_LOOP_MAXITERS;
    ctl->n_end = n - 1;
    ctl->error = e;
  }
}

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
  if (!s.ctl) {
    TL_HIP(hipMalloc(&s.ctl, sizeof(Ctl)));
    TL_HIP(hipMemsetAsync(s.ctl, 0, sizeof(Ctl), s.stream));
    TL_HIP(hipHostMalloc(&s.hctl, 4 * sizeof(Ctl), hipHostMallocDefault));
  }
  const long need = n_last + 4;
  if (s.hist_len < need) {
    if (s.hist) TL_HIP(hipFree(s.hist));
    TL_HIP(hipMalloc((void **)&s.hist, sizeof(double) * 3 * need));
    s.hist_len = need;
  }
  return TL_OK;
}

double *h_rro() { return S().hist; }
double *h_alpha() { return S().hist + S().hist_len; }
double *h_beta() { return S().hist + 2 * S().hist_len; }

// Drive a loop: `launch(n, first)` queues iteration n; `tail(n)` queues the
// exit test of iteration n-1.  Batches of iterations are queued two deep and
// the control block is polled once per batch.
template <typename LAUNCH, typename TAIL>
int drive(int n_first, int n_last, LAUNCH launch, TAIL tail, Ctl *out) {
  State &s = S();
  Ctl *hc = (Ctl *)s.hctl;
  const int B = batch_size();
  hipEvent_t ev[2];
  TL_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
  TL_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
  int n = n_first, k = 0;
  bool stopped = false;
  int rc = TL_OK;
  while (n <= n_last) {
    const int m = (n + B - 1 < n_last) ? n + B - 1 : n_last;
    for (int it = n; it <= m; ++it) {
      rc = launch(it, it == n_first);
      if (rc) goto done;
    }
    if ((rc = hipMemcpyAsync(&hc[k & 1], s.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s.stream)) !=
        hipSuccess) {
      rc = hip_fail((hipError_t)rc, "hipMemcpyAsync(ctl)");
      goto done;
    }
    if (hipEventRecord(ev[k & 1], s.stream) != hipSuccess) {
      rc = fail(TL_ERR_HIP, "hipEventRecord");
      goto done;
    }
    if (k >= 1) {
      if (hipEventSynchronize(ev[(k - 1) & 1]) != hipSuccess) {
        rc = fail(TL_ERR_HIP, "hipEventSynchronize");
        goto done;
      }
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
    if (rc) goto done;
  }
  if (hipMemcpyAsync(&hc[2], s.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s.stream) != hipSuccess ||
      hipStreamSynchronize(s.stream) != hipSuccess) {
    rc = fail(TL_ERR_HIP, "reading the loop control block");
    goto done;
  }
  *out = hc[2];
done:
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  return rc;
}

int copy_hist(double *dst, const double *src_dev, int lo, int hi) {
  if (!dst || hi < lo) return TL_OK;
  State &s = S();
  TL_HIP(hipMemcpyAsync(dst + lo, src_dev + lo, sizeof(double) * (hi - lo + 1),
                        hipMemcpyDeviceToHost, s.stream));
  TL_HIP(hipStreamSynchronize(s.stream));
  return TL_OK;
}

int check_state(tl_loop_state *st) {
  if (!st) return fail(TL_ERR_ARG, "null loop state");
  if (st->n_first < 1 || st->n_last < st->n_first)
    return fail(TL_ERR_ARG, "loop: need 1 <= n_first <= n_last");
  return TL_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" int tl_cg_loop(tl_loop_state *st, double *cg_alphas, double *cg_betas) {
  TL_REQUIRE_INIT();
  int rc = check_state(st);
  if (rc) return rc;
  State &s = S();
  if (s.precond == TL_PREC_JAC_BLOCK || !s.zb_all) {
    // block-tridiagonal preconditioner, or a chunk with internal faces: the
    // per-routine path, one iteration at a time
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
      const double alpha = rro / pw;
      if (cg_alphas) cg_alphas[n] = alpha;
      if ((rc = tl_cg_calc_ur(alpha, &rrn))) return rc;
      const double beta = rrn / rro;
      if (cg_betas) cg_betas[n] = beta;
      if ((rc = tl_cg_calc_p(beta))) return rc;
      rro = rrn;
      err = sqrt(fabs(rrn));
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
  if ((rc = ensure_loop_buffers(st->n_last))) return rc;
  const Geo G = geo(dev::COLS_PER_BLOCK);
  const int nb = G.blocks < 0 ? -G.blocks : G.blocks;
  const int diag = s.precond == TL_PREC_JAC_DIAG;
  const double *src = (diag || s.ppcg_active) ? s.z.p11 : s.r.p11;
  const bool nt = use_nontemporal();
  auto kw = nt ? k_cgl_w<true> : k_cgl_w<false>;
  auto kur = nt ? k_cgl_ur<true> : k_cgl_ur<false>;
  const int c0 = s.pcur;
  LoopParams lp;
  lp.eps_abs = st->eps_abs;
  lp.epslim = st->epslim;
  lp.check_switch = st->check_switch;
  lp.presteps = st->presteps;
  Ctl *ctl = (Ctl *)s.ctl;
  auto launch = [&](int n, bool first) -> int {
    const int cin = c0 ^ ((n - st->n_first) & 1);
    hipEvent_t ev;
    prof_begin(0, &ev);
    hipLaunchKernelGGL(kw, dim3(nb), dim3(BLOCK), 0, s.stream, G, n, first ? 1 : 0, 0, lp,
                       st->rro, src, s.p[cin].p11, s.p[1 - cin].p11, s.Kx.p11, s.Ky.p11, s.w.p11,
                       s.rx, s.ry, part(P2), nb, part(P1), ctl, h_rro(), h_beta());
    prof_end(0, ev);
    prof_begin(1, &ev);
    hipLaunchKernelGGL(kur, dim3(nb), dim3(BLOCK), 0, s.stream, G, n, s.u[s.ucur].p11,
                       s.p[1 - cin].p11, s.r.p11, s.w.p11, s.z.p11, s.Mi.p11, diag, part(P1), nb,
                       part(P2), ctl, h_rro(), h_alpha());
    prof_end(1, ev);
    TL_HIP(hipGetLastError());
    return TL_OK;
  };
  auto tail = [&](int n) -> int {
    const int cin = c0 ^ ((n - st->n_first) & 1);
    hipLaunchKernelGGL(kw, dim3(nb), dim3(BLOCK), 0, s.stream, G, n, 0, 1, lp, st->rro, src,
                       s.p[cin].p11, s.p[1 - cin].p11, s.Kx.p11, s.Ky.p11, s.w.p11, s.rx, s.ry,
                       part(P2), nb, part(P1), ctl, h_rro(), h_beta());
    TL_HIP(hipGetLastError());
    return TL_OK;
  };
  Ctl res;
  if ((rc = drive(st->n_first, st->n_last, launch, tail, &res))) return rc;
  st->n_end = res.n_end;
  st->reason = res.reason;
  st->error = res.error;
  st->rro = res.rro;
  st->old_error = res.old_error;
  // the launch that stopped the loop applied the last p update
  const int n_det = res.reason == TL_LOOP_SWITCH ? res.n_end : res.n_end + 1;
  s.pcur = c0 ^ ((n_det - st->n_first + 1) & 1);
  const int last_done = res.reason == TL_LOOP_SWITCH ? res.n_end - 1 : res.n_end;
  if ((rc = copy_hist(cg_alphas, h_alpha(), st->n_first, last_done))) return rc;
  if ((rc = copy_hist(cg_betas, h_beta(), st->n_first, last_done))) return rc;
  return TL_OK;
}

extern "C" int tl_ppcg_loop(tl_loop_state *st, const double *alphas, const double *betas,
                            int inner, double theta, int *inner_iters) {
  TL_REQUIRE_INIT();
  int rc = check_state(st);
  if (rc) return rc;
  if (!alphas || !betas) return fail(TL_ERR_ARG, "tl_ppcg_loop: null coefficients");
  State &s = S();
  if (s.precond == TL_PREC_JAC_BLOCK || inner < 1) {
    // per-routine path (tea_solve.f90:294-330), one outer iteration at a time
    double rro = st->rro, err = 0.0;
    for (int n = st->n_first; n <= st->n_last; ++n) {
      double pw, rrn, dummy;
      if ((rc = tl_cg_calc_w(&pw))) return rc;
      if ((rc = tl_ppcg_store_r())) return rc;
      if ((rc = tl_ppcg_calc_zrnorm(&rro))) return rc;
      const double alpha = rro / pw;
      if ((rc = tl_cg_calc_ur(alpha, &dummy))) return rc;
      if (inner >= 0) {
        if ((rc = tl_ppcg_init_sd(theta))) return rc;
        if ((rc = tl_ppcg_inner(alphas, betas, 1, inner))) return rc;
        if ((rc = tl_ppcg_update_z())) return rc;
      }
      if (inner_iters) *inner_iters += inner;
      if ((rc = tl_ppcg_calc_rrn(&rrn))) return rc;
      const double beta = rrn / rro;
      if ((rc = tl_cg_calc_p(beta))) return rc;
      rro = rrn;
      err = sqrt(fabs(rrn));
      if (fabs(err) < st->eps_abs) {
        st->n_end = n; st->reason = TL_LOOP_CONVERGED; st->rro = rro; st->error = err;
        return TL_OK;
      }
      st->old_error = err;
    }
    st->n_end = st->n_last; st->reason = TL_LOOP_MAXITERS; st->rro = rro; st->error = err;
    return TL_OK;
  }
  if ((rc = ensure_loop_buffers(st->n_last))) return rc;
  const Geo G = geo(BLOCK);
  const int nb = G.blocks;
  const int diag = s.precond == TL_PREC_JAC_DIAG;
  const int c0 = s.pcur;
  Ctl *ctl = (Ctl *)s.ctl;
  int sd = s.sdcur;
  const double theta_r = 1.0 / theta;
  auto launch = [&](int n, bool first) -> int {
    const int cin = c0 ^ ((n - st->n_first) & 1);
    hipEvent_t ev;
    prof_begin(5, &ev);
    hipLaunchKernelGGL(k_ppl_w, dim3(nb), dim3(BLOCK), 0, s.stream, G, n, first ? 1 : 0, 0,
                       st->eps_abs, s.z.p11, s.r.p11, s.p[cin].p11, s.p[1 - cin].p11, s.Kx.p11,
                       s.Ky.p11, s.w.p11, s.rx, s.ry, part(P3), nb, part(P1), part(P2), ctl,
                       h_rro(), h_beta());
    prof_end(5, ev);
    prof_begin(6, &ev);
    hipLaunchKernelGGL(k_ppl_ur, dim3(nb), dim3(BLOCK), 0, s.stream, G, n, s.u[s.ucur].p11,
                       s.p[1 - cin].p11, s.r.p11, s.r_store.p11, s.w.p11, s.z.p11, s.Mi.p11, diag,
                       part(P1), part(P2), nb, ctl, h_rro(), h_alpha());
    prof_end(6, ev);
    const double *src = diag ? s.z.p11 : s.r.p11;
    for (int i = 1; i <= inner; ++i) {
      prof_begin(2, &ev);
      hipLaunchKernelGGL(k_ppl_sweep, dim3(nb), dim3(BLOCK), 0, s.stream, G, i == 1 ? 1 : 0,
                         i == inner ? 1 : 0, theta_r, alphas[i - 1], betas[i - 1], diag, src,
                         s.sd[sd].p11, s.sd[1 - sd].p11, s.rtemp.p11, s.utemp.p11, s.r.p11,
                         s.r_store.p11, s.z.p11, s.Mi.p11, s.Kx.p11, s.Ky.p11, s.rx, s.ry,
                         part(P3), ctl);
      prof_end(2, ev);
      sd = 1 - sd;
    }
    TL_HIP(hipGetLastError());
    return TL_OK;
  };
  auto tail = [&](int n) -> int {
    const int cin = c0 ^ ((n - st->n_first) & 1);
    hipLaunchKernelGGL(k_ppl_w, dim3(nb), dim3(BLOCK), 0, s.stream, G, n, 0, 1, st->eps_abs,
                       s.z.p11, s.r.p11, s.p[cin].p11, s.p[1 - cin].p11, s.Kx.p11, s.Ky.p11,
                       s.w.p11, s.rx, s.ry, part(P3), nb, part(P1), part(P2), ctl, h_rro(),
                       h_beta());
    TL_HIP(hipGetLastError());
    return TL_OK;
  };
  Ctl res;
  if ((rc = drive(st->n_first, st->n_last, launch, tail, &res))) return rc;
  st->n_end = res.n_end;
  st->reason = res.reason;
  st->error = res.error;
  st->rro = res.rro;
  st->old_error = res.old_error;
  const int executed = res.n_end - st->n_first + 1;   // outer iterations that ran
  if (inner_iters) *inner_iters += executed * inner;
  s.pcur = c0 ^ ((executed + 1) & 1);                 // + the p update of the stopping launch
  s.sdcur = s.sdcur ^ ((executed * inner) & 1);
  return TL_OK;
}

extern "C" int tl_jacobi_loop(tl_loop_state *st) {
  TL_REQUIRE_INIT();
  int rc = check_state(st);
  if (rc) return rc;
  if ((rc = ensure_loop_buffers(st->n_last))) return rc;
  State &s = S();
  const Geo G = geo(BLOCK);
  const int nb = G.blocks;
  const int u0c = s.ucur;
  Ctl *ctl = (Ctl *)s.ctl;
  auto launch = [&](int n, bool first) -> int {
    const int in = u0c ^ ((n - st->n_first) & 1);
    hipEvent_t ev;
    prof_begin(4, &ev);
    hipLaunchKernelGGL(k_jl, dim3(nb), dim3(BLOCK), 0, s.stream, G, n, first ? 1 : 0, st->eps_abs,
                       s.u[in].p11, s.u[1 - in].p11, s.u0.p11, s.Kx.p11, s.Ky.p11, s.rx, s.ry,
                       part((n - 1) & 1 ? P4 : P5), nb, part(n & 1 ? P4 : P5), ctl);
    prof_end(4, ev);
    TL_HIP(hipGetLastError());
    return TL_OK;
  };
  auto tail = [&](int n) -> int {
    hipLaunchKernelGGL(k_jl_tail, dim3(1), dim3(BLOCK), 0, s.stream, n, st->eps_abs,
                       part((n - 1) & 1 ? P4 : P5), nb, ctl);
    TL_HIP(hipGetLastError());
    return TL_OK;
  };
  Ctl res;
  if ((rc = drive(st->n_first, st->n_last, launch, tail, &res))) return rc;
  st->n_end = res.n_end;
  st->reason = res.reason;
  st->error = res.error;
  s.ucur = u0c ^ ((res.n_end - st->n_first + 1) & 1);
  return TL_OK;
}
