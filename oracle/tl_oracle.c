/*
 * TEST INFRASTRUCTURE ONLY -- the checker, never the product.
 *
 * Plain-C restatement of the reference TeaLeaf v1.403 for one chunk and one
 * tile: deck reader, mesh/state generation, every kernel on the implicit-solve
 * path, the tea_leaf() step driver and the diffuse() step loop.  Each function
 * names the reference file:line it restates.  Loop order, operation order and
 * reduction order follow the reference, so that with gcc -O2 (no FMA) it agrees
 * bit for bit with the reference built by oracle/ref.mk and run on one thread.
 *
 * How it is pinned (tests/test_oracle_*.py): the shipped goldens
 * Benchmarks/tea_bm_{1,2}.out (step rows, iteration counts), the QA constants of
 * field_summary.f90:110-115, fixtures captured from oracle/_ref for every solver
 * and option, and closed forms (tqli on tridiag(-1,2,-1), Chebyshev recurrence,
 * operator symmetry, A*const = const, dense direct solves on tiny grids).
 */
#include "tl_oracle.h"

#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* indexing: reference bounds x_min-h:x_max+h with x_min = y_min = 1         */
/* ------------------------------------------------------------------------ */
#define W(c) ((c)->nx + 2 * (c)->h)
#define IX(c, j, k) ((long)((j) + (c)->h - 1) + (long)((k) + (c)->h - 1) * (long)W(c))
#define F(c, a, j, k) ((a)[IX(c, j, k)])
/* volume: x_min-2:x_max+2 (kernels/field_summary_kernel.f90:36) */
#define VOL(c, j, k) ((c)->volume[(long)((j) + 1) + (long)((k) + 1) * (long)((c)->nx + 4)])
/* cp, bfp: x_min:x_max, y_min:y_max */
#define T(c, a, j, k) ((a)[(long)((j) - 1) + (long)((k) - 1) * (long)(c)->nx])
/* vertexx: x_min-2:x_max+3 ; cellx: x_min-2:x_max+2 */
#define VX(c, j) ((c)->vertexx[(j) + 2])
#define VY(c, k) ((c)->vertexy[(k) + 2])
#define CX(c, j) ((c)->cellx[(j) + 2])
#define CY(c, k) ((c)->celly[(k) + 2])

static const double g_small = (double)1.0e-16f; /* data.f90:30 (default-real literal) */

/* ------------------------------------------------------------------------ */
/* chunk storage (build_field.f90:43-243: every array allocated and zeroed)  */
/* ------------------------------------------------------------------------ */
tlo_chunk *tlo_chunk_create(int nx, int ny, int h) {
  tlo_chunk *c = (tlo_chunk *)calloc(1, sizeof(tlo_chunk));
  long n;
  c->nx = nx; c->ny = ny; c->h = h;
  n = (long)(nx + 2 * h) * (long)(ny + 2 * h);
#define AL(f) c->f = (double *)calloc((size_t)n, sizeof(double))
  AL(density); AL(energy0); AL(energy1); AL(u); AL(u0); AL(p); AL(r);
  AL(r_store); AL(Mi); AL(w); AL(z); AL(utemp); AL(rtemp); AL(Di); AL(Kx);
  AL(Ky); AL(sd);
#undef AL
  c->cp = (double *)calloc((size_t)nx * ny, sizeof(double));
  c->bfp = (double *)calloc((size_t)nx * ny, sizeof(double));
  c->volume = (double *)calloc((size_t)(nx + 4) * (ny + 4), sizeof(double));
  c->vertexx = (double *)calloc((size_t)nx + 6, sizeof(double));
  c->vertexy = (double *)calloc((size_t)ny + 6, sizeof(double));
  c->cellx = (double *)calloc((size_t)nx + 5, sizeof(double));
  c->celly = (double *)calloc((size_t)ny + 5, sizeof(double));
  return c;
}

void tlo_chunk_destroy(tlo_chunk *c) {
  if (!c) return;
  free(c->density); free(c->energy0); free(c->energy1); free(c->u); free(c->u0);
  free(c->p); free(c->r); free(c->r_store); free(c->Mi); free(c->w); free(c->z);
  free(c->utemp); free(c->rtemp); free(c->Di); free(c->Kx); free(c->Ky);
  free(c->sd); free(c->cp); free(c->bfp); free(c->volume); free(c->vertexx);
  free(c->vertexy); free(c->cellx); free(c->celly);
  free(c);
}

long tlo_field_size(const tlo_chunk *c) { return (long)W(c) * (long)(c->ny + 2 * c->h); }

/* ids 1..11 as data.f90:44-55; 12.. are the remaining haloed vectors */
double *tlo_field(tlo_chunk *c, int id) {
  switch (id) {
    case 1: return c->density;  case 2: return c->energy0; case 3: return c->energy1;
    case 4: return c->u;        case 5: return c->p;       case 6: return c->sd;
    case 7: return c->r;        case 8: return c->z;       case 9: return c->Kx;
    case 10: return c->Ky;      case 11: return c->Di;     case 12: return c->u0;
    case 13: return c->w;       case 14: return c->r_store; case 15: return c->Mi;
    case 16: return c->utemp;   case 17: return c->rtemp;
    default: return NULL;
  }
}

/* ------------------------------------------------------------------------ */
/* setup kernels                                                             */
/* ------------------------------------------------------------------------ */
static void chunk_dxdy(const tlo_params *p, double *dx, double *dy) {
  /* initialise_chunk.f90:35-36: REAL(grid%x_cells) is default (single) real */
  *dx = (p->xmax - p->xmin) / (double)(float)p->x_cells;
  *dy = (p->ymax - p->ymin) / (double)(float)p->y_cells;
}

/* kernels/initialise_chunk_kernel.f90:58-130 (vertex/cell coords, volume) */
void tlo_initialise_chunk(tlo_chunk *c, const tlo_params *p) {
  int j, k;
  double dx, dy;
  chunk_dxdy(p, &dx, &dy);
  c->dx = dx; c->dy = dy;
  for (j = -1; j <= c->nx + 3; ++j) VX(c, j) = p->xmin + dx * (double)(float)(j - 1);
  for (k = -1; k <= c->ny + 3; ++k) VY(c, k) = p->ymin + dy * (double)(float)(k - 1);
  for (j = -1; j <= c->nx + 2; ++j) CX(c, j) = 0.5 * (VX(c, j) + VX(c, j + 1));
  for (k = -1; k <= c->ny + 2; ++k) CY(c, k) = 0.5 * (VY(c, k) + VY(c, k + 1));
  for (k = -1; k <= c->ny + 2; ++k)
    for (j = -1; j <= c->nx + 2; ++j) VOL(c, j, k) = dx * dy;
}

/* kernels/generate_chunk_kernel.f90:82-138; the u0 dummy receives field%u
 * (generate_chunk.f90), so u = energy0*density over the haloed range. */
void tlo_generate_chunk(tlo_chunk *c, const tlo_params *p) {
  int j, k, s, h = c->h;
  for (k = 1 - h; k <= c->ny + h; ++k)
    for (j = 1 - h; j <= c->nx + h; ++j) {
      F(c, c->energy0, j, k) = p->states[1].energy;
      F(c, c->density, j, k) = p->states[1].density;
    }
  for (s = 2; s <= p->number_of_states; ++s) {
    const tlo_state *st = &p->states[s];
    double xc = st->xmin, yc = st->ymin;
    for (k = 1 - h; k <= c->ny + h; ++k)
      for (j = 1 - h; j <= c->nx + h; ++j) {
        int hit = 0;
        if (st->geometry == TLO_G_RECT) {
          if (j >= 1 && j <= c->nx && k >= 1 && k <= c->ny)
            if (VX(c, j + 1) >= st->xmin && VX(c, j) < st->xmax)
              if (VY(c, k + 1) >= st->ymin && VY(c, k) < st->ymax) hit = 1;
        } else if (st->geometry == TLO_G_CIRC) {
          double rad = sqrt((CX(c, j) - xc) * (CX(c, j) - xc) + (CY(c, k) - yc) * (CY(c, k) - yc));
          if (rad <= st->radius) hit = 1;
        } else if (st->geometry == TLO_G_POINT) {
          if (VX(c, j) == xc && VY(c, k) == yc) hit = 1;
        }
        if (hit) {
          F(c, c->energy0, j, k) = st->energy;
          F(c, c->density, j, k) = st->density;
        }
      }
  }
  for (k = 1 - h; k <= c->ny + h; ++k)
    for (j = 1 - h; j <= c->nx + h; ++j)
      F(c, c->u, j, k) = F(c, c->energy0, j, k) * F(c, c->density, j, k);
}

/* kernels/set_field_kernel.f90:40-46 */
void tlo_set_field(tlo_chunk *c) {
  int j, k;
  for (k = 1; k <= c->ny; ++k)
    for (j = 1; j <= c->nx; ++j) F(c, c->energy1, j, k) = F(c, c->energy0, j, k);
}

/* ------------------------------------------------------------------------ */
/* preconditioners: kernels/tea_leaf_common_kernel.f90:276-468               */
/* ------------------------------------------------------------------------ */
void tlo_diag_init(tlo_chunk *c) { /* :276-307 */
  int j, k, h = c->h;
  for (k = 1 - h + 1; k <= c->ny + h - 1; ++k)
    for (j = 1 - h + 1; j <= c->nx + h - 1; ++j) {
      double di = F(c, c->Di, j, k);
      F(c, c->Mi, j, k) = (di != 0.0) ? 1.0 / di : 0.0;
    }
}

void tlo_diag_solve(tlo_chunk *c, const double *r, double *z, int depth) { /* :309-335 */
  int j, k;
  for (k = 1 - depth; k <= c->ny + depth; ++k)
    for (j = 1 - depth; j <= c->nx + depth; ++j) F(c, z, j, k) = F(c, c->Mi, j, k) * F(c, r, j, k);
}

void tlo_block_init(tlo_chunk *c) { /* :337-376, jac_block_size = 4 */
  int j, k, ko, bottom, top;
  double ry = c->ry;
  for (ko = 1; ko <= c->ny; ko += 4) {
    bottom = ko;
    top = (ko + 3 < c->ny) ? ko + 3 : c->ny;
    for (j = 1; j <= c->nx; ++j) {
      k = bottom;
      T(c, c->cp, j, k) = (-F(c, c->Ky, j, k + 1) * ry) / F(c, c->Di, j, k);
      for (k = bottom + 1; k <= top; ++k) {
        T(c, c->bfp, j, k) = 1.0 / (F(c, c->Di, j, k) - (-F(c, c->Ky, j, k) * ry) * T(c, c->cp, j, k - 1));
        T(c, c->cp, j, k) = (-F(c, c->Ky, j, k + 1) * ry) * T(c, c->bfp, j, k);
      }
    }
  }
}

static void block_solve_rows(tlo_chunk *c, const double *r, double *z, int bottom, int top) {
  int j, k;
  double dp_l[4], z_l[4], ry = c->ry;
  for (j = 1; j <= c->nx; ++j) {
    k = bottom;
    dp_l[0] = F(c, r, j, k) / F(c, c->Di, j, k);
    for (k = bottom + 1; k <= top; ++k)
      dp_l[k - bottom] = (F(c, r, j, k) - (-F(c, c->Ky, j, k) * ry) * dp_l[k - bottom - 1]) * T(c, c->bfp, j, k);
    k = top;
    z_l[k - bottom] = dp_l[k - bottom];
    for (k = top - 1; k >= bottom; --k)
      z_l[k - bottom] = dp_l[k - bottom] - T(c, c->cp, j, k) * z_l[k - bottom + 1];
    for (k = bottom; k <= top; ++k) F(c, z, j, k) = z_l[k - bottom];
  }
}

void tlo_block_solve(tlo_chunk *c, const double *r, double *z) { /* :378-468 */
  int ko, ki, k_extra = c->ny - (c->ny % 4);
  for (ko = 1; ko <= k_extra; ko += 4) block_solve_rows(c, r, z, ko, ko + 3);
  for (ki = k_extra + 1; ki <= c->ny; ki += 4) {
    int bottom = ki < c->ny ? ki : c->ny;
    int top = (ki + 3 < c->ny) ? ki + 3 : c->ny;
    block_solve_rows(c, r, z, bottom, top);
  }
}

static void apply_precond(tlo_chunk *c, int precond, const double *r, double *z) {
  if (precond == TLO_PREC_JAC_BLOCK) tlo_block_solve(c, r, z);
  else if (precond == TLO_PREC_JAC_DIAG) tlo_diag_solve(c, r, z, 0);
}

/* (A x)_{j,k} with the precomputed Di, kernels/tea_leaf_cg_kernel.f90:145-147 */
static inline double smvp_di(const tlo_chunk *c, const double *x, int j, int k) {
  return F(c, c->Di, j, k) * F(c, x, j, k)
       - c->ry * (F(c, c->Ky, j, k + 1) * F(c, x, j, k + 1) + F(c, c->Ky, j, k) * F(c, x, j, k - 1))
       - c->rx * (F(c, c->Kx, j + 1, k) * F(c, x, j + 1, k) + F(c, c->Kx, j, k) * F(c, x, j - 1, k));
}

/* ------------------------------------------------------------------------ */
/* common kernels: kernels/tea_leaf_common_kernel.f90                         */
/* ------------------------------------------------------------------------ */
void tlo_common_init(tlo_chunk *c, const int zb[4], int reflective, int coef, int precond) {
  int j, k, h = c->h;
  double rx = c->rx, ry = c->ry;
  for (k = 1; k <= c->ny; ++k) /* :67-72 */
    for (j = 1; j <= c->nx; ++j) {
      F(c, c->u, j, k) = F(c, c->energy1, j, k) * F(c, c->density, j, k);
      F(c, c->u0, j, k) = F(c, c->energy1, j, k) * F(c, c->density, j, k);
    }
  for (k = 1 - h; k <= c->ny + h; ++k) /* :75-92, w used as temporary */
    for (j = 1 - h; j <= c->nx + h; ++j)
      F(c, c->w, j, k) = (coef == TLO_RECIP_CONDUCTIVITY) ? 1.0 / F(c, c->density, j, k)
                                                          : F(c, c->density, j, k);
  for (k = 1 - h + 1; k <= c->ny + h; ++k) /* :94-101 */
    for (j = 1 - h + 1; j <= c->nx + h; ++j) {
      F(c, c->Kx, j, k) = (F(c, c->w, j - 1, k) + F(c, c->w, j, k)) / (2.0 * F(c, c->w, j - 1, k) * F(c, c->w, j, k));
      F(c, c->Ky, j, k) = (F(c, c->w, j, k - 1) + F(c, c->w, j, k)) / (2.0 * F(c, c->w, j, k - 1) * F(c, c->w, j, k));
    }
  if (!reflective) { /* :104-141 */
    if (zb[0]) for (k = 1 - h; k <= c->ny + h; ++k) for (j = 1 - h; j <= 1; ++j) F(c, c->Kx, j, k) = 0.0;
    if (zb[1]) for (k = 1 - h; k <= c->ny + h; ++k) for (j = c->nx + 1; j <= c->nx + h; ++j) F(c, c->Kx, j, k) = 0.0;
    if (zb[2]) for (k = 1 - h; k <= 1; ++k) for (j = 1 - h; j <= c->nx + h; ++j) F(c, c->Ky, j, k) = 0.0;
    if (zb[3]) for (k = c->ny + 1; k <= c->ny + h; ++k) for (j = 1 - h; j <= c->nx + h; ++j) F(c, c->Ky, j, k) = 0.0;
  }
  for (k = 1 - h + 1; k <= c->ny + h - 1; ++k) /* :144-152 */
    for (j = 1 - h + 1; j <= c->nx + h - 1; ++j)
      F(c, c->Di, j, k) = (1.0 + ry * (F(c, c->Ky, j, k + 1) + F(c, c->Ky, j, k))
                               + rx * (F(c, c->Kx, j + 1, k) + F(c, c->Kx, j, k)));
  if (precond == TLO_PREC_JAC_BLOCK) tlo_block_init(c); /* :154-160 */
  else if (precond == TLO_PREC_JAC_DIAG) tlo_diag_init(c);
  for (k = 1; k <= c->ny; ++k) /* :162-174 */
    for (j = 1; j <= c->nx; ++j) {
      F(c, c->w, j, k) = smvp_di(c, c->u, j, k);
      F(c, c->r, j, k) = F(c, c->u, j, k) - F(c, c->w, j, k);
    }
}

void tlo_finalise(tlo_chunk *c) { /* :179-206 */
  int j, k;
  for (k = 1; k <= c->ny; ++k)
    for (j = 1; j <= c->nx; ++j) F(c, c->energy1, j, k) = F(c, c->u, j, k) / F(c, c->density, j, k);
}

void tlo_calc_residual(tlo_chunk *c) { /* :208-244 */
  int j, k;
  for (k = 1; k <= c->ny; ++k)
    for (j = 1; j <= c->nx; ++j) F(c, c->r, j, k) = F(c, c->u0, j, k) - smvp_di(c, c->u, j, k);
}

double tlo_calc_2norm(tlo_chunk *c, int which) { /* :246-274; 0 = u0, 1 = r (tea_leaf_common.f90:112-130) */
  const double *a = which == 0 ? c->u0 : c->r;
  double norm = 0.0;
  int j, k;
  for (k = 1; k <= c->ny; ++k)
    for (j = 1; j <= c->nx; ++j) norm = norm + F(c, a, j, k) * F(c, a, j, k);
  return norm;
}

/* ------------------------------------------------------------------------ */
/* CG kernels: kernels/tea_leaf_cg_kernel.f90                                 */
/* ------------------------------------------------------------------------ */
double tlo_cg_init(tlo_chunk *c, int precond) { /* :31-112 */
  int j, k;
  double rro = 0.0;
  for (k = 1; k <= c->ny; ++k)
    for (j = 1; j <= c->nx; ++j) { F(c, c->p, j, k) = 0.0; F(c, c->z, j, k) = 0.0; }
  if (precond != TLO_PREC_NONE) {
    apply_precond(c, precond, c->r, c->z);
    for (k = 1; k <= c->ny; ++k) for (j = 1; j <= c->nx; ++j) F(c, c->p, j, k) = F(c, c->z, j, k);
  } else {
    for (k = 1; k <= c->ny; ++k) for (j = 1; j <= c->nx; ++j) F(c, c->p, j, k) = F(c, c->r, j, k);
  }
  for (k = 1; k <= c->ny; ++k)
    for (j = 1; j <= c->nx; ++j) rro = rro + F(c, c->r, j, k) * F(c, c->p, j, k);
  return rro;
}

double tlo_cg_calc_w(tlo_chunk *c) { /* :114-156 */
  int j, k;
  double pw = 0.0;
  for (k = 1; k <= c->ny; ++k) {
    for (j = 1; j <= c->nx; ++j) F(c, c->w, j, k) = smvp_di(c, c->p, j, k);
    for (j = 1; j <= c->nx; ++j) pw = pw + F(c, c->w, j, k) * F(c, c->p, j, k);
  }
  return pw;
}

double tlo_cg_calc_ur(tlo_chunk *c, double alpha, int precond) { /* :198-304 */
  int j, k;
  double rrn = 0.0;
  for (k = 1; k <= c->ny; ++k)
    for (j = 1; j <= c->nx; ++j) F(c, c->u, j, k) = F(c, c->u, j, k) + alpha * F(c, c->p, j, k);
  if (precond == TLO_PREC_JAC_DIAG) {
    for (k = 1; k <= c->ny; ++k)
      for (j = 1; j <= c->nx; ++j) {
        F(c, c->r, j, k) = F(c, c->r, j, k) - alpha * F(c, c->w, j, k);
        F(c, c->z, j, k) = F(c, c->Mi, j, k) * F(c, c->r, j, k);
        rrn = rrn + F(c, c->r, j, k) * F(c, c->z, j, k);
      }
  } else if (precond == TLO_PREC_JAC_BLOCK) {
    for (k = 1; k <= c->ny; ++k)
      for (j = 1; j <= c->nx; ++j) F(c, c->r, j, k) = F(c, c->r, j, k) - alpha * F(c, c->w, j, k);
    tlo_block_solve(c, c->r, c->z);
    for (k = 1; k <= c->ny; ++k)
      for (j = 1; j <= c->nx; ++j) rrn = rrn + F(c, c->r, j, k) * F(c, c->z, j, k);
  } else {
    for (k = 1; k <= c->ny; ++k)
      for (j = 1; j <= c->nx; ++j) {
        F(c, c->r, j, k) = F(c, c->r, j, k) - alpha * F(c, c->w, j, k);
        rrn = rrn + F(c, c->r, j, k) * F(c, c->r, j, k);
      }
  }
  return rrn;
}

void tlo_cg_calc_p(tlo_chunk *c, double beta, int precond, int ppcg_active) { /* :306-347 */
  int j, k;
  const double *src = (precond != TLO_PREC_NONE || ppcg_active) ? c->z : c->r;
  for (k = 1; k <= c->ny; ++k)
    for (j = 1; j <= c->nx; ++j) F(c, c->p, j, k) = F(c, src, j, k) + beta * F(c, c->p, j, k);
}

/* ------------------------------------------------------------------------ */
/* Chebyshev kernels: kernels/tea_leaf_cheby_kernel.f90                       */
/* ------------------------------------------------------------------------ */
void tlo_cheby_init(tlo_chunk *c, double theta, int precond) { /* :30-109 */
  int j, k;
  for (k = 1; k <= c->ny; ++k)
    for (j = 1; j <= c->nx; ++j) {
      F(c, c->w, j, k) = smvp_di(c, c->u, j, k);
      F(c, c->r, j, k) = F(c, c->u0, j, k) - F(c, c->w, j, k);
    }
  if (precond != TLO_PREC_NONE) {
    apply_precond(c, precond, c->r, c->z);
    for (k = 1; k <= c->ny; ++k) for (j = 1; j <= c->nx; ++j) F(c, c->p, j, k) = F(c, c->z, j, k) / theta;
  } else {
    for (k = 1; k <= c->ny; ++k) for (j = 1; j <= c->nx; ++j) F(c, c->p, j, k) = F(c, c->r, j, k) / theta;
  }
  for (k = 1; k <= c->ny; ++k)
    for (j = 1; j <= c->nx; ++j) F(c, c->u, j, k) = F(c, c->u, j, k) + F(c, c->p, j, k);
}

void tlo_cheby_iterate(tlo_chunk *c, double alpha, double beta, int precond) { /* :111-200 */
  int j, k;
  double rx = c->rx, ry = c->ry;
  for (k = 1; k <= c->ny; ++k)
    for (j = 1; j <= c->nx; ++j) {
      F(c, c->w, j, k) = (1.0 + ry * (F(c, c->Ky, j, k + 1) + F(c, c->Ky, j, k))
                              + rx * (F(c, c->Kx, j + 1, k) + F(c, c->Kx, j, k))) * F(c, c->u, j, k)
          - ry * (F(c, c->Ky, j, k + 1) * F(c, c->u, j, k + 1) + F(c, c->Ky, j, k) * F(c, c->u, j, k - 1))
          - rx * (F(c, c->Kx, j + 1, k) * F(c, c->u, j + 1, k) + F(c, c->Kx, j, k) * F(c, c->u, j - 1, k));
      F(c, c->r, j, k) = F(c, c->u0, j, k) - F(c, c->w, j, k);
    }
  if (precond != TLO_PREC_NONE) {
    apply_precond(c, precond, c->r, c->z);
    for (k = 1; k <= c->ny; ++k)
      for (j = 1; j <= c->nx; ++j) F(c, c->p, j, k) = alpha * F(c, c->p, j, k) + beta * F(c, c->z, j, k);
  } else {
    for (k = 1; k <= c->ny; ++k)
      for (j = 1; j <= c->nx; ++j) F(c, c->p, j, k) = alpha * F(c, c->p, j, k) + beta * F(c, c->r, j, k);
  }
  for (k = 1; k <= c->ny; ++k)
    for (j = 1; j <= c->nx; ++j) F(c, c->u, j, k) = F(c, c->u, j, k) + F(c, c->p, j, k);
}

/* ------------------------------------------------------------------------ */
/* PPCG kernels: kernels/tea_leaf_ppcg_kernel.f90                             */
/* ------------------------------------------------------------------------ */
void tlo_ppcg_init_sd(tlo_chunk *c, double theta, int precond) { /* :31-88 */
  int j, k;
  double theta_r = 1.0 / theta;
  const double *src = precond != TLO_PREC_NONE ? c->z : c->r;
  for (k = 1; k <= c->ny; ++k)
    for (j = 1; j <= c->nx; ++j) {
      F(c, c->sd, j, k) = F(c, src, j, k) * theta_r;
      F(c, c->rtemp, j, k) = F(c, c->r, j, k);
      F(c, c->utemp, j, k) = F(c, c->sd, j, k);
    }
}

double tlo_ppcg_init(tlo_chunk *c, int step, int precond, int ppcg_active) { /* :90-201 */
  int j, k;
  double rro = 0.0;
  if (step == 1) {
    for (k = 1; k <= c->ny; ++k) for (j = 1; j <= c->nx; ++j) { F(c, c->p, j, k) = 0.0; F(c, c->z, j, k) = 0.0; }
  } else if (step == 3) {
    for (k = 1; k <= c->ny; ++k) for (j = 1; j <= c->nx; ++j) F(c, c->p, j, k) = 0.0;
  }
  if (precond != TLO_PREC_NONE || (ppcg_active && step == 3)) {
    if (step == 1 || step == 2) apply_precond(c, precond, c->r, c->z);
    if (step == 1 || step == 3)
      for (k = 1; k <= c->ny; ++k) for (j = 1; j <= c->nx; ++j) F(c, c->p, j, k) = F(c, c->z, j, k);
  } else if (step == 1) {
    for (k = 1; k <= c->ny; ++k) for (j = 1; j <= c->nx; ++j) F(c, c->p, j, k) = F(c, c->r, j, k);
  }
  if (step == 1 || step == 3)
    for (k = 1; k <= c->ny; ++k)
      for (j = 1; j <= c->nx; ++j) rro = rro + F(c, c->r, j, k) * F(c, c->p, j, k);
  return rro; /* the wrapper returns 0 for step 2 (tea_leaf_ppcg.f90:83,110) */
}

void tlo_ppcg_inner(tlo_chunk *c, double alpha, double beta, int precond) { /* :204-310, interior bounds */
  int j, k;
  for (k = 1; k <= c->ny; ++k)
    for (j = 1; j <= c->nx; ++j)
      F(c, c->rtemp, j, k) = F(c, c->rtemp, j, k) - smvp_di(c, c->sd, j, k);
  if (precond == TLO_PREC_JAC_DIAG) {
    for (k = 1; k <= c->ny; ++k)
      for (j = 1; j <= c->nx; ++j) {
        F(c, c->sd, j, k) = alpha * F(c, c->sd, j, k) + beta * F(c, c->Mi, j, k) * F(c, c->rtemp, j, k);
        F(c, c->utemp, j, k) = F(c, c->utemp, j, k) + F(c, c->sd, j, k);
      }
  } else if (precond == TLO_PREC_JAC_BLOCK) {
    tlo_block_solve(c, c->rtemp, c->z);
    for (k = 1; k <= c->ny; ++k)
      for (j = 1; j <= c->nx; ++j) {
        F(c, c->sd, j, k) = alpha * F(c, c->sd, j, k) + beta * F(c, c->z, j, k);
        F(c, c->utemp, j, k) = F(c, c->utemp, j, k) + F(c, c->sd, j, k);
      }
  } else {
    for (k = 1; k <= c->ny; ++k)
      for (j = 1; j <= c->nx; ++j) {
        F(c, c->sd, j, k) = alpha * F(c, c->sd, j, k) + beta * F(c, c->rtemp, j, k);
        F(c, c->utemp, j, k) = F(c, c->utemp, j, k) + F(c, c->sd, j, k);
      }
  }
}

double tlo_ppcg_calc_zrnorm(tlo_chunk *c, int precond, int ppcg_active) { /* :411-451 */
  int j, k;
  double norm = 0.0;
  const double *a = (precond != TLO_PREC_NONE || ppcg_active) ? c->z : c->r;
  for (k = 1; k <= c->ny; ++k)
    for (j = 1; j <= c->nx; ++j) norm = norm + F(c, a, j, k) * F(c, c->r, j, k);
  return norm;
}

void tlo_ppcg_update_z(tlo_chunk *c) { /* :453-477 */
  int j, k;
  for (k = 1; k <= c->ny; ++k) for (j = 1; j <= c->nx; ++j) F(c, c->z, j, k) = F(c, c->utemp, j, k);
}

void tlo_ppcg_store_r(tlo_chunk *c) { /* :507-530 */
  int j, k;
  for (k = 1; k <= c->ny; ++k) for (j = 1; j <= c->nx; ++j) F(c, c->r_store, j, k) = F(c, c->r, j, k);
}

double tlo_ppcg_calc_rrn(tlo_chunk *c) { /* :532-561 */
  int j, k;
  double rrn = 0.0;
  for (k = 1; k <= c->ny; ++k)
    for (j = 1; j <= c->nx; ++j)
      rrn = rrn + (F(c, c->r, j, k) - F(c, c->r_store, j, k)) * F(c, c->z, j, k);
  return rrn;
}

/* ------------------------------------------------------------------------ */
/* Jacobi: kernels/tea_leaf_jacobi_kernel.f90:26-77; un is vector_r          */
/* (tea_leaf_jacobi.f90:38)                                                   */
/* ------------------------------------------------------------------------ */
double tlo_jacobi_solve(tlo_chunk *c) {
  int j, k;
  double err = 0.0, rx = c->rx, ry = c->ry;
  double *un = c->r;
  for (k = 1; k <= c->ny; ++k) for (j = 1; j <= c->nx; ++j) F(c, un, j, k) = F(c, c->u, j, k);
  for (k = 1; k <= c->ny; ++k)
    for (j = 1; j <= c->nx; ++j) {
      F(c, c->u, j, k) = (F(c, c->u0, j, k)
            + rx * (F(c, c->Kx, j + 1, k) * F(c, un, j + 1, k) + F(c, c->Kx, j, k) * F(c, un, j - 1, k))
            + ry * (F(c, c->Ky, j, k + 1) * F(c, un, j, k + 1) + F(c, c->Ky, j, k) * F(c, un, j, k - 1)))
          / (1.0 + rx * (F(c, c->Kx, j, k) + F(c, c->Kx, j + 1, k))
                 + ry * (F(c, c->Ky, j, k) + F(c, c->Ky, j, k + 1)));
      err = err + fabs(F(c, c->u, j, k) - F(c, un, j, k));
    }
  return err;
}

/* ------------------------------------------------------------------------ */
/* halo: update_halo.f90:49-91 -> kernels/update_halo_kernel.f90:49-195      */
/* (single chunk, single tile: only the reflective external update exists)   */
/* ------------------------------------------------------------------------ */
static void halo_cell(tlo_chunk *c, double *m, int depth) {
  int j, k, nx = c->nx, ny = c->ny;
  for (k = 1 - depth; k <= ny + depth; ++k) for (j = 1; j <= depth; ++j) F(c, m, 1 - j, k) = F(c, m, 0 + j, k);
  for (k = 1 - depth; k <= ny + depth; ++k) for (j = 1; j <= depth; ++j) F(c, m, nx + j, k) = F(c, m, nx + 1 - j, k);
  for (k = 1; k <= depth; ++k) for (j = 1 - depth; j <= nx + depth; ++j) F(c, m, j, 1 - k) = F(c, m, j, 0 + k);
  for (k = 1; k <= depth; ++k) for (j = 1 - depth; j <= nx + depth; ++j) F(c, m, j, ny + k) = F(c, m, j, ny + 1 - k);
}

void tlo_update_halo(tlo_chunk *c, const int fields[11], int depth, int reflective) {
  /* update_boundary passes vector_rtemp in the FIELD_R slot (update_halo.f90:79) */
  double *slot[11];
  int i;
  if (!reflective) return;
  slot[0] = c->density; slot[1] = c->energy0; slot[2] = c->energy1; slot[3] = c->u;
  slot[4] = c->p; slot[5] = c->sd; slot[6] = c->rtemp; slot[7] = c->z; slot[8] = c->Kx;
  slot[9] = c->Ky; slot[10] = c->Di;
  for (i = 0; i < 11; ++i)
    if (fields[i] == 1) halo_cell(c, slot[i], depth);
}

/* kernels/field_summary_kernel.f90:51-60 */
void tlo_field_summary(tlo_chunk *c, double out[4]) {
  int j, k;
  double vol = 0.0, mass = 0.0, ie = 0.0, temp = 0.0;
  for (k = 1; k <= c->ny; ++k)
    for (j = 1; j <= c->nx; ++j) {
      double cell_vol = VOL(c, j, k);
      double cell_mass = cell_vol * F(c, c->density, j, k);
      vol = vol + cell_vol;
      mass = mass + cell_mass;
      ie = ie + cell_mass * F(c, c->energy1, j, k);
      temp = temp + cell_mass * F(c, c->u, j, k);
    }
  out[0] = vol; out[1] = mass; out[2] = ie; out[3] = temp;
}

/* ------------------------------------------------------------------------ */
/* host scalar code: tea_leaf_cheby.f90:90-205                                */
/* ------------------------------------------------------------------------ */
static double fsign(double a, double b) { return b >= 0.0 ? fabs(a) : -fabs(a); }

/* implicit-QL eigenvalues of a symmetric tridiagonal (tea_leaf_cheby.f90:90-140);
 * d[0..n-1] diagonal, e[0..n-1] sub-diagonal stored from e[1] (eoshift'ed here) */
int tlo_tqli(double *d, double *e, int n) {
  int i, iter, l, m;
  double b, c, dd, f, g, p, r, s;
  for (i = 0; i < n - 1; ++i) e[i] = e[i + 1]; /* eoshift(e,1) */
  if (n > 0) e[n - 1] = 0.0;
  for (l = 0; l < n; ++l) {
    iter = 0;
    for (;;) {
      for (m = l; m < n - 1; ++m) {
        dd = fabs(d[m]) + fabs(d[m + 1]);
        if (fabs(e[m]) + dd == dd) break;
      }
      if (m == l) break;
      if (iter == 30) return 1;
      iter = iter + 1;
      g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      r = sqrt(g * g + 1.0);
      g = d[m] - d[l] + e[l] / (g + fsign(r, g));
      s = 1.0; c = 1.0; p = 0.0;
      {
        int restart = 0;
        for (i = m - 1; i >= l; --i) {
          f = s * e[i];
          b = c * e[i];
          r = sqrt(f * f + g * g);
          e[i + 1] = r;
          if (r == 0.0) {
            d[i + 1] = d[i + 1] - p;
            e[m] = 0.0;
            restart = 1;
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
        if (restart) continue;
      }
      d[l] = d[l] - p;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  return 0;
}

/* tea_leaf_cheby.f90:142-176; cg_* are 1-based (index 0 unused) */
int tlo_calc_eigenvalues(const double *cg_alphas, const double *cg_betas, int n,
                         double *eigmin, double *eigmax) {
  double *diag = (double *)calloc((size_t)n + 1, sizeof(double));
  double *off = (double *)calloc((size_t)n + 1, sizeof(double));
  int i, info;
  for (i = 1; i <= n; ++i) {
    diag[i - 1] = 1.0 / cg_alphas[i];
    if (i > 1) diag[i - 1] = diag[i - 1] + cg_betas[i - 1] / cg_alphas[i - 1];
    if (i < n) off[i] = sqrt(cg_betas[i]) / cg_alphas[i];
  }
  info = tlo_tqli(diag, off, n);
  if (info == 0) {
    double mn = diag[0], mx = diag[0];
    for (i = 1; i < n; ++i) { if (diag[i] < mn) mn = diag[i]; if (diag[i] > mx) mx = diag[i]; }
    *eigmin = mn; *eigmax = mx;
    if (mn < 0.0 || mx < 0.0) info = 1;
  }
  free(diag); free(off);
  return info;
}

/* tea_leaf_cheby.f90:178-205; alphas/betas 1-based */
void tlo_calc_ch_coefs(double *alphas, double *betas, double eigmin, double eigmax,
                       double *theta_out, int n) {
  double theta = (eigmax + eigmin) / 2.0, delta = (eigmax - eigmin) / 2.0;
  double sigma = theta / delta, rho_old = 1.0 / sigma, rho_new;
  int i;
  for (i = 1; i <= n; ++i) {
    rho_new = 1.0 / (2.0 * sigma - rho_old);
    alphas[i] = rho_new * rho_old;
    betas[i] = 2.0 * rho_new / delta;
    rho_old = rho_new;
  }
  *theta_out = theta;
}

/* ------------------------------------------------------------------------ */
/* the step driver: tea_solve.f90:37-503                                      */
/* ------------------------------------------------------------------------ */
typedef struct {
  const tlo_params *P;
  tlo_chunk *c;
  double dt;
  /* SAVEd across calls (tea_solve.f90:55,58) */
  double eigmin, eigmax, theta, cn;
  int first;
} tlo_solver;

static void set_fields(int f[11], int a, int b) {
  memset(f, 0, 11 * sizeof(int));
  if (a) f[a - 1] = 1;
  if (b) f[b - 1] = 1;
}

/* tea_leaf_ppcg.f90:211-431 on one chunk (bounds are never extended) */
static void run_ppcg_inner_steps(tlo_solver *S, const double *a, const double *b, int inner) {
  tlo_chunk *c = S->c;
  const tlo_params *P = S->P;
  int fields[11], cur, bx, h = P->halo_depth;
  if (inner < 0) return;
  tlo_ppcg_init_sd(c, S->theta, P->precond);
  for (cur = 1; cur <= inner; cur += h) {
    set_fields(fields, TLO_F_SD, TLO_F_R);
    tlo_update_halo(c, fields, h, P->reflective);
    set_fields(fields, TLO_F_SD, 0);
    for (bx = h - 1; bx >= 0; --bx) {
      int idx = cur + h - 1 - bx;
      tlo_ppcg_inner(c, a[idx], b[idx], P->precond);
      if (P->reflective) tlo_update_halo(c, fields, 1, 1);
      if (idx == inner) break;
    }
  }
  tlo_ppcg_update_z(c);
}

static int tea_leaf_step(tlo_solver *S, tlo_step_record *rec, char *msg, int msglen) {
  tlo_chunk *c = S->c;
  const tlo_params *P = S->P;
  int fields[11], zb[4] = {1, 1, 1, 1};
  int n, est_itc = 0, cheby_calc_steps = 0, max_cheby_iters = 0, ppcg_inner_iters = 0;
  int ch_switch_check = 0, ppcg_active = 0, exited = 0;
  double old_error, error = 0.0, initial_residual, rro = 0.0, pw, rrn, alpha, beta;
  int maxi = P->max_iters;
  double *cg_alphas = (double *)calloc((size_t)maxi + 3, sizeof(double));
  double *cg_betas = (double *)calloc((size_t)maxi + 3, sizeof(double));
  int ncoef = (maxi + 3 > P->tl_ppcg_inner_steps + 3) ? maxi + 3 : P->tl_ppcg_inner_steps + 3;
  double *ch_alphas = (double *)calloc((size_t)ncoef, sizeof(double));
  double *ch_betas = (double *)calloc((size_t)ncoef, sizeof(double));

  rec->switch_n = 0; rec->est_itc = 0; rec->eigmin = rec->eigmax = 0.0;

  set_fields(fields, TLO_F_ENERGY1, TLO_F_DENSITY);            /* :89-94 */
  tlo_update_halo(c, fields, P->halo_depth, P->reflective);
  c->rx = S->dt / (c->dx * c->dx);                              /* tea_leaf_common.f90:24-25 */
  c->ry = S->dt / (c->dy * c->dy);
  tlo_common_init(c, zb, P->reflective, P->coefficient, P->precond);
  set_fields(fields, TLO_F_U, 0);                               /* :99-103 */
  tlo_update_halo(c, fields, 1, P->reflective);
  tlo_calc_residual(c);                                         /* :106-110 */
  initial_residual = tlo_calc_2norm(c, 1);
  old_error = initial_residual;                                 /* :113 (before the sqrt) */
  initial_residual = sqrt(fabs(initial_residual));
  rec->initial_residual = initial_residual;

  if (P->use_cg || P->use_cheby || P->use_ppcg) {               /* :123-142 */
    rro = tlo_cg_init(c, P->precond);
    set_fields(fields, TLO_F_U, TLO_F_P);
    tlo_update_halo(c, fields, 1, P->reflective);
    set_fields(fields, TLO_F_P, 0);
  } else if (P->use_jacobi) {
    set_fields(fields, TLO_F_U, 0);
  }

  for (n = 1; n <= maxi; ++n) {                                 /* :152 */
    if (!ch_switch_check) {                                     /* :160-174 */
      if (cheby_calc_steps > 0) ch_switch_check = 1;
      else if (!S->first && P->use_ppcg && n > 1) ch_switch_check = 1;
      else if (fabs(old_error) <= P->tl_ch_cg_epslim && n >= P->tl_ch_cg_presteps) ch_switch_check = 1;
      else ch_switch_check = 0;
    }
    if ((P->use_cheby || P->use_ppcg) && ch_switch_check) {
      if (cheby_calc_steps == 0) {                              /* :183-253 */
        max_cheby_iters = maxi - n + 2;
        if (S->first) {
          int info = tlo_calc_eigenvalues(cg_alphas, cg_betas, n - 1, &S->eigmin, &S->eigmax);
          S->first = 0;
          if (info != 0) {
            snprintf(msg, msglen, "tea_leaf: Error in calculating eigenvalues");
            free(cg_alphas); free(cg_betas); free(ch_alphas); free(ch_betas);
            return -2;
          }
          S->eigmin = S->eigmin * 0.95;
          S->eigmax = S->eigmax * 1.05;
        }
        if (P->use_cheby) {
          tlo_calc_ch_coefs(ch_alphas, ch_betas, S->eigmin, S->eigmax, &S->theta, max_cheby_iters);
          set_fields(fields, TLO_F_U, 0);
        } else {
          tlo_calc_ch_coefs(ch_alphas, ch_betas, S->eigmin, S->eigmax, &S->theta, P->tl_ppcg_inner_steps);
          ppcg_active = 1;
        }
        S->cn = S->eigmax / S->eigmin;
        rec->switch_n = n; rec->eigmin = S->eigmin; rec->eigmax = S->eigmax;
        rro = tlo_ppcg_init(c, 2, P->precond, ppcg_active);
        run_ppcg_inner_steps(S, ch_alphas, ch_betas, P->tl_ppcg_inner_steps);
        ppcg_inner_iters = ppcg_inner_iters + P->tl_ppcg_inner_steps;
        rro = tlo_ppcg_init(c, 3, P->precond, ppcg_active);
        /* :246-250 reset fields to P for both solvers, so from here on the
         * Chebyshev iterations refresh p's halo, not u's */
        set_fields(fields, TLO_F_P, 0);
        tlo_update_halo(c, fields, 1, P->reflective);
      }
      if (P->use_cheby) {                                       /* :259-280 */
        if (cheby_calc_steps == 0) {
          /* tea_leaf_cheby_first_step (tea_leaf_cheby.f90:208-252); its error
           * dummy is old_error here (tea_solve.f90:261) */
          double bb = tlo_calc_2norm(c, 0), it_alpha, gamm;
          tlo_cheby_init(c, S->theta, P->precond);
          tlo_update_halo(c, fields, 1, P->reflective);
          tlo_cheby_iterate(c, ch_alphas[1], ch_betas[1], P->precond);
          old_error = tlo_calc_2norm(c, 1);
          it_alpha = P->eps / 2.0 * sqrt(bb / old_error);
          gamm = (sqrt(S->cn) - 1.0) / (sqrt(S->cn) + 1.0);
          est_itc = (int)lround(log(it_alpha) / log(gamm));
          rec->est_itc = est_itc;
          cheby_calc_steps = 1;
        } else {
          tlo_cheby_iterate(c, ch_alphas[cheby_calc_steps], ch_betas[cheby_calc_steps], P->precond);
          if (n >= est_itc && n % 10 == 0) error = tlo_calc_2norm(c, 1);
        }
      } else if (P->use_ppcg) {                                 /* :289-331 */
        pw = tlo_cg_calc_w(c);
        tlo_ppcg_store_r(c);
        rro = tlo_ppcg_calc_zrnorm(c, P->precond, ppcg_active);
        alpha = rro / pw;
        (void)tlo_cg_calc_ur(c, alpha, P->precond);
        run_ppcg_inner_steps(S, ch_alphas, ch_betas, P->tl_ppcg_inner_steps);
        ppcg_inner_iters = ppcg_inner_iters + P->tl_ppcg_inner_steps;
        rrn = tlo_ppcg_calc_rrn(c);
        beta = rrn / rro;
        tlo_cg_calc_p(c, beta, P->precond, ppcg_active);
        error = rrn;
        rro = rrn;
      }
      cheby_calc_steps = cheby_calc_steps + 1;
    } else if (P->use_cg || P->use_cheby || P->use_ppcg) {      /* :341-369 */
      fields[TLO_F_P - 1] = 1;
      pw = tlo_cg_calc_w(c);
      alpha = rro / pw;
      cg_alphas[n] = alpha;
      rrn = tlo_cg_calc_ur(c, alpha, P->precond);
      beta = rrn / rro;
      cg_betas[n] = beta;
      tlo_cg_calc_p(c, beta, P->precond, ppcg_active);
      error = rrn;
      rro = rrn;
    } else if (P->use_jacobi) {                                 /* :375-379 */
      error = tlo_jacobi_solve(c);
    }
    tlo_update_halo(c, fields, 1, P->reflective);               /* :384 */
    error = sqrt(fabs(error));                                  /* :402 */
    if (fabs(error) < P->eps * initial_residual) { exited = 1; break; }
    old_error = error;
  }
  if (!exited) n = maxi + 1;

  if (P->tl_check_result) {                                     /* :416-431 */
    set_fields(fields, TLO_F_U, 0);
    tlo_update_halo(c, fields, 1, P->reflective);
    tlo_calc_residual(c);
    rec->exact_error = sqrt(tlo_calc_2norm(c, 1)) / initial_residual;
  }
  rec->conduction_error = error / initial_residual;             /* :438-453 */
  rec->iterations = n - 1;
  rec->ppcg_inner_iters = ppcg_inner_iters;

  tlo_finalise(c);                                              /* :461-467 */
  set_fields(fields, TLO_F_ENERGY1, 0);
  tlo_update_halo(c, fields, 1, P->reflective);

  free(cg_alphas); free(cg_betas); free(ch_alphas); free(ch_betas);
  return 0;
}

static void summary_into(tlo_chunk *c, tlo_step_record *r) {
  double s[4];
  tlo_field_summary(c, s);
  r->has_summary = 1;
  r->summary[0] = s[0]; r->summary[1] = s[1]; r->summary[2] = s[1] / s[0];
  r->summary[3] = s[2]; r->summary[4] = s[3];
}

/* initialise.f90 + start.f90:24-133 + diffuse.f90:23-105 */
int tlo_run(const tlo_params *P, tlo_step_record *rec, int max_records, int max_steps,
            char *msg, int msglen) {
  tlo_solver S;
  int fields[11], step = 0, nrec = 0, rc;
  double time = 0.0;
  tlo_chunk *c = tlo_chunk_create(P->x_cells, P->y_cells, P->halo_depth);
  memset(&S, 0, sizeof(S));
  S.P = P; S.c = c; S.first = 1; S.dt = P->dtinit;
  tlo_initialise_chunk(c, P);
  tlo_generate_chunk(c, P);
  memset(fields, 0, sizeof(fields));
  fields[TLO_F_DENSITY - 1] = fields[TLO_F_ENERGY0 - 1] = fields[TLO_F_ENERGY1 - 1] = 1;
  tlo_update_halo(c, fields, P->halo_depth, P->reflective);
  tlo_set_field(c);
  if (nrec < max_records) {
    memset(&rec[nrec], 0, sizeof(rec[nrec]));
    rec[nrec].step = 0;
    summary_into(c, &rec[nrec]);
    nrec++;
  }
  for (;;) {
    tlo_step_record r;
    memset(&r, 0, sizeof(r));
    step = step + 1;
    S.dt = P->dtinit;                                           /* timestep.f90 / calc_dt.f90:34 */
    rc = tea_leaf_step(&S, &r, msg, msglen);
    if (rc != 0) { tlo_chunk_destroy(c); return rc; }
    time = time + S.dt;
    r.step = step;
    if (P->summary_frequency != 0 && step % P->summary_frequency == 0) summary_into(c, &r);
    if (time + g_small > P->end_time || step >= P->end_step || (max_steps > 0 && step >= max_steps)) {
      summary_into(c, &r);
      if (nrec < max_records) rec[nrec++] = r;
      break;
    }
    if (nrec < max_records) rec[nrec++] = r;
  }
  tlo_chunk_destroy(c);
  return nrec;
}

/* ------------------------------------------------------------------------ */
/* deck reader: initialise.f90:103-111, parse.f90:143-354, read_input.f90     */
/* ------------------------------------------------------------------------ */
void tlo_default_params(tlo_params *p) { /* read_input.f90:39-101 */
  memset(p, 0, sizeof(*p));
  p->xmin = 0.0; p->ymin = 0.0; p->xmax = 100.0; p->ymax = 100.0;
  p->x_cells = 10; p->y_cells = 10;
  p->end_time = 10.0; p->end_step = 640000; /* g_ibig, data.f90:28 */
  p->visit_frequency = 0; p->summary_frequency = 10;
  p->dtinit = (double)0.1f;                  /* default-real literal */
  p->max_iters = 1000; p->eps = (double)1.0e-10f;
  p->coefficient = TLO_CONDUCTIVITY;
  p->tl_ch_cg_presteps = 25; p->tl_ch_cg_epslim = 1.0;
  p->precond = TLO_PREC_NONE; p->tl_ppcg_inner_steps = -1;
  p->use_jacobi = 1; p->halo_depth = 1;
}

typedef struct {
  char (*lines)[128];
  int nlines, pos;
  char line[128];   /* current remaining words */
} deck_reader;

static void clean_line(const char *src, int len, char *out) {
  /* a100 read; chars outside 32..128 -> blank; TRIM(ADJUSTL); cut at ! and ; */
  char tmp[128];
  int i, n = len < 100 ? len : 100, s = 0, e;
  for (i = 0; i < n; ++i) {
    unsigned char ch = (unsigned char)src[i];
    tmp[i] = (ch < 32 || ch > 128) ? ' ' : (char)ch;
  }
  tmp[n] = 0;
  for (i = 0; i < n; ++i) if (tmp[i] == '!' || tmp[i] == ';') { tmp[i] = 0; break; }
  while (tmp[s] == ' ') s++;
  e = (int)strlen(tmp);
  while (e > s && tmp[e - 1] == ' ') e--;
  memcpy(out, tmp + s, (size_t)(e - s));
  out[e - s] = 0;
}

/* parse_getline for mask '*tea': returns 0 with d->line set, or -1 at the end */
static int deck_getline(deck_reader *d, char *here, char *msg, int msglen) {
  while (d->pos < d->nlines) {
    char l[128];
    int i;
    strcpy(l, d->lines[d->pos++]);
    for (i = 0; l[i]; ++i) {
      l[i] = (char)tolower((unsigned char)l[i]);
      if (l[i] == '=' || l[i] == ',') l[i] = ' ';
    }
    if (strncmp(l, "*select:", 8) == 0) continue;
    if (l[0] == '*') {
      char nugget[128];
      int s = 0;
      while (l[s] && l[s] != ' ') s++;
      memcpy(nugget, l, (size_t)s); nugget[s] = 0;
      if (strncmp(nugget + 1, "end", 3) == 0) {
        char want[130];
        char *last = strrchr(here, '*');
        snprintf(want, sizeof(want), "*%s", nugget + 4);
        if (!last || strcmp(last, want) != 0) {
          snprintf(msg, msglen, "parse_getline: Unmatched */*end pair.");
          return -2;
        }
        if (strcmp(here, "*tea") == 0) { *last = 0; return -1; }
        *last = 0;
      } else {
        strcat(here, nugget);
      }
      continue;
    }
    if (strcmp(here, "*tea") == 0 && l[0]) { strcpy(d->line, l); return 0; }
  }
  return -1;
}

static int deck_getword(deck_reader *d, int wrap, char *here, char *word, char *msg, int msglen) {
  for (;;) {
    char *l = d->line;
    int s = 0, e;
    while (l[s] == ' ') s++;
    e = s;
    while (l[e] && l[e] != ' ') e++;
    memcpy(word, l + s, (size_t)(e - s));
    word[e - s] = 0;
    if (strcmp(word, "\\") == 0 || (word[0] == 0 && wrap)) {
      int st = deck_getline(d, here, msg, msglen);
      if (st == -2) return -2;
      if (st != 0) { word[0] = 0; return 0; }
      continue;
    }
    memmove(l, l + e, strlen(l + e) + 1);
    return 0;
  }
}

static int parse_ival(const char *w, int *out, char *msg, int msglen) { /* parse.f90:279-294, I7 */
  char buf[8], *end;
  long v;
  int n = (int)strlen(w);
  if (n > 7) n = 7;
  memcpy(buf, w, (size_t)n); buf[n] = 0;
  v = strtol(buf, &end, 10);
  while (*end == ' ') end++;
  if (end == buf || *end) {
    snprintf(msg, msglen, "parse_getival: Error attempting to convert to integer:%s", w);
    return -3;
  }
  *out = (int)v;
  return 0;
}

static int parse_rval(const char *w, double *out, char *msg, int msglen) { /* parse.f90:319-338, E27.20 */
  char buf[64], *end;
  int i;
  snprintf(buf, sizeof(buf), "%s", w);
  if (!strchr(buf, '.')) strncat(buf, ".0", sizeof(buf) - strlen(buf) - 1);
  buf[27] = 0;
  for (i = 0; buf[i]; ++i) if (buf[i] == 'd' || buf[i] == 'D') buf[i] = 'e';
  *out = strtod(buf, &end);
  if (end == buf || *end) {
    snprintf(msg, msglen, "parse_getrval: Error attempting to convert to real:%s", w);
    return -3;
  }
  return 0;
}

int tlo_parse_deck(const char *text, tlo_params *p, char *msg, int msglen) {
  deck_reader d;
  char here[256], word[128], val[128];
  int pass, nl = 0, cap = 64, rc = 0;
  const char *s = text;
  tlo_default_params(p);
  memset(&d, 0, sizeof(d));
  d.lines = malloc(sizeof(*d.lines) * (size_t)cap);
  while (*s) { /* initialise.f90:103-111: keep the non-blank cleaned lines */
    const char *e = strchr(s, '\n');
    int len = e ? (int)(e - s) : (int)strlen(s);
    char out[128];
    if (len > 0 && s[len - 1] == '\r') len--;
    clean_line(s, len, out);
    if (out[0]) {
      if (nl == cap) { cap *= 2; d.lines = realloc(d.lines, sizeof(*d.lines) * (size_t)cap); }
      strcpy(d.lines[nl++], out);
    }
    s = e ? e + 1 : s + strlen(s);
  }
  d.nlines = nl;

  for (pass = 0; pass < 2 && rc == 0; ++pass) {
    int state_max = 0;
    d.pos = 0; here[0] = 0; d.line[0] = 0;
    for (;;) {
      int st = deck_getline(&d, here, msg, msglen);
      if (st == -2) { rc = -2; break; }
      if (st != 0) break;
      for (;;) {
        if ((rc = deck_getword(&d, 0, here, word, msg, msglen))) break;
        if (!word[0]) break;
        if (pass == 0) { /* read_input.f90:106-120: count states */
          if (strcmp(word, "state") == 0) {
            int sn;
            if ((rc = deck_getword(&d, 1, here, val, msg, msglen))) break;
            if ((rc = parse_ival(val, &sn, msg, msglen))) break;
            if (sn > state_max) state_max = sn;
            break;
          }
          continue;
        }
#define GETI(dst) do { if ((rc = deck_getword(&d, 1, here, val, msg, msglen))) goto done; \
                       if ((rc = parse_ival(val, &(dst), msg, msglen))) goto done; } while (0)
#define GETR(dst) do { if ((rc = deck_getword(&d, 1, here, val, msg, msglen))) goto done; \
                       if ((rc = parse_rval(val, &(dst), msg, msglen))) goto done; } while (0)
        if (!strcmp(word, "initial_timestep")) GETR(p->dtinit);
        else if (!strcmp(word, "end_time")) GETR(p->end_time);
        else if (!strcmp(word, "end_step")) GETI(p->end_step);
        else if (!strcmp(word, "xmin")) GETR(p->xmin);
        else if (!strcmp(word, "xmax")) GETR(p->xmax);
        else if (!strcmp(word, "ymin")) GETR(p->ymin);
        else if (!strcmp(word, "ymax")) GETR(p->ymax);
        else if (!strcmp(word, "x_cells")) GETI(p->x_cells);
        else if (!strcmp(word, "y_cells")) GETI(p->y_cells);
        else if (!strcmp(word, "visit_frequency")) GETI(p->visit_frequency);
        else if (!strcmp(word, "summary_frequency")) GETI(p->summary_frequency);
        else if (!strcmp(word, "tiles_per_task") || !strcmp(word, "sub_tiles_per_tile") ||
                 !strcmp(word, "tiles_per_problem")) { int dummy; GETI(dummy); (void)dummy; }
        else if (!strcmp(word, "tl_ch_cg_presteps")) GETI(p->tl_ch_cg_presteps);
        else if (!strcmp(word, "tl_ppcg_inner_steps")) GETI(p->tl_ppcg_inner_steps);
        else if (!strcmp(word, "tl_ch_cg_epslim")) GETR(p->tl_ch_cg_epslim);
        else if (!strcmp(word, "tl_check_result")) p->tl_check_result = 1;
        else if (!strcmp(word, "tl_preconditioner_type")) {
          for (;;) {
            if ((rc = deck_getword(&d, 0, here, val, msg, msglen))) goto done;
            if (!val[0]) break;
            if (!strcmp(val, "none")) p->precond = TLO_PREC_NONE;
            else if (!strcmp(val, "jac_diag")) p->precond = TLO_PREC_JAC_DIAG;
            else if (!strcmp(val, "jac_block")) p->precond = TLO_PREC_JAC_BLOCK;
            else {
              snprintf(msg, msglen, "read_input: Invalid preconditioner type specified: %s", val);
              rc = -4; goto done;
            }
          }
        }
        else if (!strcmp(word, "use_fortran_kernels")) { }
        else if (!strcmp(word, "verbose_on")) p->verbose_on = 1;
        else if (!strcmp(word, "tl_use_jacobi")) { p->use_cheby = 0; p->use_cg = 0; p->use_ppcg = 0; p->use_jacobi = 1; }
        else if (!strcmp(word, "tl_use_cg")) { p->use_cheby = 0; p->use_cg = 1; p->use_ppcg = 0; p->use_jacobi = 0; }
        else if (!strcmp(word, "tl_use_ppcg")) { p->use_cheby = 0; p->use_cg = 0; p->use_ppcg = 1; p->use_jacobi = 0; }
        else if (!strcmp(word, "tl_use_chebyshev")) { p->use_cheby = 1; p->use_cg = 0; p->use_ppcg = 0; p->use_jacobi = 0; }
        else if (!strcmp(word, "reflective_boundary")) p->reflective = 1;
        else if (!strcmp(word, "profiler_on")) p->profiler_on = 1;
        else if (!strcmp(word, "halo_depth")) {
          GETI(p->halo_depth);
          if (p->halo_depth < 1) {
            snprintf(msg, msglen, "read_input: Invalid halo exchange depth specified");
            rc = -4; goto done;
          }
        }
        else if (!strcmp(word, "tl_max_iters")) GETI(p->max_iters);
        else if (!strcmp(word, "tl_eps")) GETR(p->eps);
        else if (!strcmp(word, "tl_coefficient_density")) p->coefficient = TLO_CONDUCTIVITY;
        else if (!strcmp(word, "tl_coefficient_inverse_density")) p->coefficient = TLO_RECIP_CONDUCTIVITY;
        else if (!strcmp(word, "test_problem")) GETI(p->test_problem);
        else if (!strcmp(word, "state")) {
          int sn;
          tlo_state *st;
          GETI(sn);
          if (sn < 1 || sn > TLO_MAX_STATES) {
            snprintf(msg, msglen, "read_input: state number out of range");
            rc = -4; goto done;
          }
          st = &p->states[sn];
          if (st->defined) {
            snprintf(msg, msglen, "read_input: State defined twice.");
            rc = -4; goto done;
          }
          st->defined = 1;
          for (;;) {
            if ((rc = deck_getword(&d, 0, here, word, msg, msglen))) goto done;
            if (!word[0]) break;
            if (!strcmp(word, "xmin")) GETR(st->xmin);
            else if (!strcmp(word, "ymin")) GETR(st->ymin);
            else if (!strcmp(word, "xmax")) GETR(st->xmax);
            else if (!strcmp(word, "ymax")) GETR(st->ymax);
            else if (!strcmp(word, "radius")) GETR(st->radius);
            else if (!strcmp(word, "density")) GETR(st->density);
            else if (!strcmp(word, "energy")) GETR(st->energy);
            else if (!strcmp(word, "geometry")) {
              if ((rc = deck_getword(&d, 1, here, val, msg, msglen))) goto done;
              if (!strcmp(val, "rectangle")) st->geometry = TLO_G_RECT;
              else if (!strcmp(val, "circle")) st->geometry = TLO_G_CIRC;
              else if (!strcmp(val, "point")) st->geometry = TLO_G_POINT;
            }
          }
        }
#undef GETI
#undef GETR
      }
      if (rc) break;
    }
    if (pass == 0 && rc == 0) {
      p->number_of_states = state_max;
      if (state_max < 1) {
        snprintf(msg, msglen, "read_input: No states defined.");
        rc = -4;
      } else if (state_max > TLO_MAX_STATES) {
        snprintf(msg, msglen, "read_input: too many states");
        rc = -4;
      }
    }
  }
done:
  free(d.lines);
  if (rc) return rc;
  if (p->tl_ppcg_inner_steps == -1 && p->use_ppcg) /* read_input.f90:323-327, single-precision REAL */
    p->tl_ppcg_inner_steps = 4 * (int)sqrtf(sqrtf((float)(p->x_cells * p->y_cells)));
  if (p->halo_depth > 1 && p->precond == TLO_PREC_JAC_BLOCK) {
    snprintf(msg, msglen, "read_input: Unable to use nonstandard halo depth with block jacobi preconditioner");
    return -4;
  }
  { /* read_input.f90:346-359 */
    double dx = (p->xmax - p->xmin) / (double)(float)p->x_cells;
    double dy = (p->ymax - p->ymin) / (double)(float)p->y_cells;
    int n;
    for (n = 2; n <= p->number_of_states; ++n) {
      p->states[n].xmin = p->states[n].xmin + (dx / 100.0);
      p->states[n].ymin = p->states[n].ymin + (dy / 100.0);
      p->states[n].xmax = p->states[n].xmax - (dx / 100.0);
      p->states[n].ymax = p->states[n].ymax - (dy / 100.0);
    }
  }
  return 0;
}
