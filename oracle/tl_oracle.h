/*
 * TEST INFRASTRUCTURE ONLY.  CPU restatement of the reference TeaLeaf hot path
 * (single chunk, single tile), used by tests/, __graft_entry__.smoke() and the
 * CPU-baseline leg of bench.py as the checker.  The product path never links,
 * loads or calls anything here.
 *
 * Layout: every haloed field is (nx+2h)*(ny+2h) doubles, x fastest, exactly
 * the reference's Fortran layout x_min-h:x_max+h, y_min-h:y_max+h
 * (FTNREF2D, ftocmacros.h:13).  cp/bfp are interior only (nx*ny); volume has
 * the fixed 2-deep halo of kernels/field_summary_kernel.f90:36.
 *
 * Built with gcc at -O2 and no -march, so no FMA contraction happens; with
 * sequential (single-thread) reductions it reproduces the reference built by
 * oracle/ref.mk bit for bit.
 */
#ifndef TL_ORACLE_H
#define TL_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { TLO_PREC_NONE = 1, TLO_PREC_JAC_DIAG = 2, TLO_PREC_JAC_BLOCK = 3 };
enum { TLO_CONDUCTIVITY = 1, TLO_RECIP_CONDUCTIVITY = 2 };
enum { TLO_G_RECT = 1, TLO_G_CIRC = 2, TLO_G_POINT = 3 };
/* field ids of data.f90:44-55 (1-based, as in the reference) */
enum { TLO_F_DENSITY = 1, TLO_F_ENERGY0, TLO_F_ENERGY1, TLO_F_U, TLO_F_P,
       TLO_F_SD, TLO_F_R, TLO_F_Z, TLO_F_KX, TLO_F_KY, TLO_F_DI };

#define TLO_MAX_STATES 64

typedef struct {
  int defined, geometry;
  double density, energy, xmin, xmax, ymin, ymax, radius;
} tlo_state;

/* Deck parameters with the defaults of read_input.f90:39-101. */
typedef struct {
  double xmin, ymin, xmax, ymax;
  int x_cells, y_cells;
  double end_time, dtinit, eps, tl_ch_cg_epslim;
  int end_step, visit_frequency, summary_frequency, max_iters;
  int tl_ch_cg_presteps, tl_ppcg_inner_steps, tl_check_result;
  int precond, coefficient, reflective, halo_depth, test_problem;
  int use_cg, use_cheby, use_ppcg, use_jacobi, profiler_on, verbose_on;
  int number_of_states;
  tlo_state states[TLO_MAX_STATES + 1]; /* 1-based */
} tlo_params;

typedef struct {
  int nx, ny, h;
  double dx, dy, rx, ry;
  double *density, *energy0, *energy1, *u, *u0, *p, *r, *r_store, *Mi, *w,
         *z, *utemp, *rtemp, *Di, *Kx, *Ky, *sd;
  double *cp, *bfp;          /* interior only */
  double *volume;            /* 2-deep halo */
  double *vertexx, *vertexy, *cellx, *celly; /* 1D, offsets as reference */
} tlo_chunk;

/* Per-timestep record produced by the driver. */
typedef struct {
  int step;
  int iterations;            /* the printed "Iteration count" (n-1) */
  int ppcg_inner_iters;
  int switch_n;              /* iteration of the CG->Cheby/PPCG switch, 0 if none */
  int est_itc;
  double conduction_error;   /* error/initial_residual as printed */
  double initial_residual;
  double exact_error;        /* only with tl_check_result */
  double eigmin, eigmax;
  int has_summary;
  double summary[5];         /* vol, mass, mass/vol, ie, temp */
} tlo_step_record;

/* ---- deck ---- */
void tlo_default_params(tlo_params *p);
/* returns 0, or a nonzero code with msg filled (report_error cases) */
int tlo_parse_deck(const char *text, tlo_params *p, char *msg, int msglen);

/* ---- chunk ---- */
tlo_chunk *tlo_chunk_create(int nx, int ny, int h);
void tlo_chunk_destroy(tlo_chunk *c);
double *tlo_field(tlo_chunk *c, int field_id);   /* ids as data.f90 + extras */
long tlo_field_size(const tlo_chunk *c);          /* (nx+2h)*(ny+2h) */
void tlo_initialise_chunk(tlo_chunk *c, const tlo_params *p);
void tlo_generate_chunk(tlo_chunk *c, const tlo_params *p);
void tlo_set_field(tlo_chunk *c);

/* ---- kernels (the reference kernels directory) ---- */
void tlo_common_init(tlo_chunk *c, const int zero_boundary[4], int reflective,
                     int coef, int precond);
void tlo_finalise(tlo_chunk *c);
void tlo_calc_residual(tlo_chunk *c);
double tlo_calc_2norm(tlo_chunk *c, int which);
void tlo_diag_init(tlo_chunk *c);
void tlo_diag_solve(tlo_chunk *c, const double *r, double *z, int depth);
void tlo_block_init(tlo_chunk *c);
void tlo_block_solve(tlo_chunk *c, const double *r, double *z);
double tlo_cg_init(tlo_chunk *c, int precond);
double tlo_cg_calc_w(tlo_chunk *c);
double tlo_cg_calc_ur(tlo_chunk *c, double alpha, int precond);
void tlo_cg_calc_p(tlo_chunk *c, double beta, int precond, int ppcg_active);
void tlo_cheby_init(tlo_chunk *c, double theta, int precond);
void tlo_cheby_iterate(tlo_chunk *c, double alpha, double beta, int precond);
void tlo_ppcg_init_sd(tlo_chunk *c, double theta, int precond);
double tlo_ppcg_init(tlo_chunk *c, int step, int precond, int ppcg_active);
void tlo_ppcg_inner(tlo_chunk *c, double alpha, double beta, int precond);
double tlo_ppcg_calc_zrnorm(tlo_chunk *c, int precond, int ppcg_active);
void tlo_ppcg_update_z(tlo_chunk *c);
void tlo_ppcg_store_r(tlo_chunk *c);
double tlo_ppcg_calc_rrn(tlo_chunk *c);
double tlo_jacobi_solve(tlo_chunk *c);
void tlo_update_halo(tlo_chunk *c, const int fields[11], int depth, int reflective);
void tlo_field_summary(tlo_chunk *c, double out[4]);

/* ---- host scalar routines (tea_leaf_cheby.f90, tea_leaf_ppcg.f90) ---- */
int tlo_tqli(double *d, double *e, int n);
int tlo_calc_eigenvalues(const double *cg_alphas, const double *cg_betas,
                         int n, double *eigmin, double *eigmax);
void tlo_calc_ch_coefs(double *alphas, double *betas, double eigmin,
                       double eigmax, double *theta, int n);

/* ---- whole program: initialise + diffuse loop (tea_leaf.f90) ----
 * Runs the deck; fills up to max_records step records (one per timestep,
 * plus record 0 for the step-0 summary).  Returns the number of records, or
 * a negative error code (msg filled).  max_steps > 0 caps the step loop. */
int tlo_run(const tlo_params *p, tlo_step_record *rec, int max_records,
            int max_steps, char *msg, int msglen);

#ifdef __cplusplus
}
#endif
#endif
