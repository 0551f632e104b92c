/*
 * tealeaf_hip.h -- C ABI of the MI355X (gfx950) TeaLeaf solver library
 * (libtealeaf_hip.so).  Plain C types only: ints, doubles and host pointers.
 *
 * One process drives one GPU holding the whole mesh as a single chunk.  The
 * library owns every device buffer from tl_init() to tl_finalize(); no device
 * pointer crosses this boundary.  Every entry point returns 0 on success or a
 * TL_ERR_* code (message in tl_last_error()).  Entry points that return a
 * scalar through a pointer synchronise the library's stream; all others are
 * asynchronous and ordered on that stream.  Not thread safe.
 *
 * Host arrays use the reference layout: x fastest, bounds
 * (x_min-halo : x_max+halo, y_min-halo : y_max+halo) with x_min = y_min = 1
 * (FTNREF2D, ftocmacros.h:13; e.g. kernels/tea_leaf_cg_kernel.f90:131-132).
 *
 * Each declaration names the reference routine it replaces.  The reference
 * plugs kernels in at the `IF (use_fortran_kernels)` branch of every driver
 * wrapper (e.g. tea_leaf_cg.f90:61); a host binds these functions there.
 */
#ifndef TEALEAF_HIP_H
#define TEALEAF_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define TL_OK 0
#define TL_ERR_HIP 1          /* a HIP runtime call failed                  */
#define TL_ERR_STATE 2        /* not initialised / already initialised     */
#define TL_ERR_ARG 3          /* bad argument (field id, size, option)     */
#define TL_ERR_NONFINITE 4    /* a reduction produced NaN or Inf           */
#define TL_ERR_EIGEN 5        /* eigenvalue estimate failed (tqli)         */

/* ---- option values (data.f90:73-79) ----------------------------------- */
#define TL_CONDUCTIVITY 1          /* tl_coefficient_density             */
#define TL_RECIP_CONDUCTIVITY 2    /* tl_coefficient_inverse_density     */
#define TL_PREC_NONE 1
#define TL_PREC_JAC_DIAG 2
#define TL_PREC_JAC_BLOCK 3

/* ---- field ids: FIELD_* of data.f90:44-55, then the other arrays of
 *      field_type (definitions.f90:136-174) ------------------------------ */
#define TL_FIELD_DENSITY 1
#define TL_FIELD_ENERGY0 2
#define TL_FIELD_ENERGY1 3
#define TL_FIELD_U 4
#define TL_FIELD_P 5
#define TL_FIELD_SD 6
#define TL_FIELD_R 7
#define TL_FIELD_Z 8
#define TL_FIELD_KX 9
#define TL_FIELD_KY 10
#define TL_FIELD_DI 11
#define TL_FIELD_U0 12
#define TL_FIELD_W 13
#define TL_FIELD_R_STORE 14
#define TL_FIELD_MI 15
#define TL_FIELD_UTEMP 16
#define TL_FIELD_RTEMP 17
#define TL_FIELD_VOLUME 18
#define TL_FIELD_CP 19
#define TL_FIELD_BFP 20
#define TL_NUM_FIELDS 20

/* ---- lifecycle ---------------------------------------------------------- */

/* Allocates and zeroes every field for an nx*ny chunk with `halo` halo
 * cells (build_field.f90:43-243) on HIP device `device`.
 * Replaces: build_field() + the ALLOCATEs of chunk%tiles(t)%field. */
int tl_init(int nx, int ny, int halo, int device);

/* Frees everything.  (The reference never deallocates, tea_leaf.f90:70.) */
int tl_finalize(void);

/* Last error message (static storage, valid until the next call). */
const char *tl_last_error(void);

/* Solver options that the reference keeps in definitions_module globals:
 * coefficient (tea_solve.f90:75), tl_preconditioner_type, reflective_boundary
 * (read_input.f90:207-233), chunk%halo_exchange_depth. */
int tl_set_options(int coefficient, int preconditioner, int reflective);

/* tl_ppcg_active (definitions.f90:71), toggled by tea_solve.f90:79,209. */
int tl_set_ppcg_active(int active);

/* Copy a host array with `host_halo` halo cells to/from a device field.
 * Only the cells inside both halos are copied.  volume uses host_halo = 2
 * (kernels/field_summary_kernel.f90:36); cp and bfp use host_halo = 0.
 * The copy goes through a device staging buffer of its size, which the
 * library keeps until tl_finalize (the snapshot copies share it). */
int tl_upload(int field_id, const double *host, int host_halo);
int tl_download(int field_id, double *host, int host_halo);

/* The interior ny x nx cells of a field to/from an array in device memory
 * of the library's GPU, row stride ld doubles (ld >= nx): a device-to-device
 * copy, done when the call returns.  No counterpart in the reference: for
 * callers whose data already lives on the GPU (a torch tensor), so that it
 * does not travel through the host.  The caller finishes any work writing
 * the array (e.g. synchronizes its stream) first.  A pointer that is not
 * device memory of this GPU is refused with TL_ERR_ARG. */
int tl_upload_device(int field_id, const double *dev, long ld);
int tl_download_device(int field_id, double *dev, long ld);

/* Block until all queued work is done. */
int tl_synchronize(void);

/* ---- set-up on the device (initialise_chunk + generate_chunk) ----------- */
#define TL_GEOM_RECTANGLE 1   /* g_rect, g_circ, g_point of data.f90:66-68 */
#define TL_GEOM_CIRCLE 2
#define TL_GEOM_POINT 3

typedef struct {
  int geometry;               /* TL_GEOM_*; ignored for the first state; any
                               * other value covers no cell                  */
  double density, energy;
  double xmin, xmax, ymin, ymax, radius;  /* as read_input.f90 leaves them   */
} tl_state;

/* density, energy0 and u = energy0*density over the haloed chunk from the
 * deck's states (states[0] = state 1 fills everything, later states
 * overwrite where their geometry hits), and volume = dx*dy, computed on the
 * device with the reference's vertex and cell-centre coordinates
 * x = xmin + dx*(j-1) and 0.5*(x_j + x_{j+1}) rounded as there.
 * Replaces initialise_chunk (kernels/initialise_chunk_kernel.f90:26-132)
 * and generate_chunk (kernels/generate_chunk_kernel.f90:30-142). */
int tl_generate(double xmin, double ymin, double dx, double dy, int nstates,
                const tl_state *states);

/* ---- per-step kernels (tea_leaf_common.f90) ------------------------------ */

/* u = u0 = energy1*density; Kx, Ky from density or 1/density; zero-flux
 * faces unless reflective; Di; preconditioner set-up; w = A u; r = u - w.
 * Replaces tea_leaf_init_common (tea_leaf_common.f90:12-61) ->
 * tea_leaf_common_init_kernel (kernels/tea_leaf_common_kernel.f90:26-177). */
int tl_common_init(double rx, double ry, const int zero_boundary[4]);

/* r = u0 - A u.  Replaces tea_leaf_calc_residual (tea_leaf_common.f90:63-92). */
int tl_calc_residual(void);

/* norm = sum(arr^2), arr = u0 (which = 0) or r (which = 1).
 * Replaces tea_leaf_calc_2norm (tea_leaf_common.f90:94-138). */
int tl_calc_2norm(int which, double *norm);

/* energy1 = u/density.  Replaces tea_leaf_finalise (tea_leaf_common.f90:140-163). */
int tl_finalise(void);

/* energy1 = energy0 on the interior.  Replaces set_field (set_field.f90). */
int tl_set_field(void);

/* Reflective external-halo update of the fields flagged in fields[0..10]
 * (FIELD_* order) to `depth`; no-op unless reflective.  The FIELD_R slot
 * updates rtemp, as update_boundary does (update_halo.f90:79).
 * Replaces update_halo (update_halo.f90:32-47) on one chunk. */
int tl_update_halo(const int fields[11], int depth);

/* vol = sum(volume), mass = sum(volume*density), ie = sum(mass*energy1),
 * temp = sum(mass*u).  Replaces field_summary_kernel
 * (kernels/field_summary_kernel.f90:26-64) and the tea_sum calls of
 * field_summary.f90:88-91. */
int tl_field_summary(double *vol, double *mass, double *ie, double *temp);

/* ---- CG (tea_leaf_cg.f90) ---------------------------------------------- */
int tl_cg_init(double *rro);                  /* tea_leaf_cg_init  :11-50   */
int tl_cg_calc_w(double *pw);                 /* tea_leaf_cg_calc_w :52-88  */
int tl_cg_calc_ur(double alpha, double *rrn); /* tea_leaf_cg_calc_ur :90-131 */
int tl_cg_calc_p(double beta);                /* tea_leaf_cg_calc_p :133-158 */

/* ---- Chebyshev (tea_leaf_cheby.f90) --------------------------------------- */
int tl_cheby_init(double theta);              /* tea_leaf_cheby_init :13-49 */
/* One iteration with coefficients ch_alphas(step), ch_betas(step).
 * Replaces tea_leaf_cheby_iterate (tea_leaf_cheby.f90:51-88). */
int tl_cheby_iterate(double alpha, double beta);

/* ---- PPCG (tea_leaf_ppcg.f90) -------------------------------------------- */
int tl_ppcg_init(int step, double *rro);      /* tea_leaf_ppcg_init :72-116 */
int tl_ppcg_init_sd(double theta);            /* tea_leaf_ppcg_init_sd :33-68 */
/* Inner sweeps first..first+count-1 with coefficients alphas[i-1], betas[i-1]
 * (1-based sweep index i).  Replaces the sweep loop of
 * tea_leaf_run_ppcg_inner_steps (tea_leaf_ppcg.f90:231-424), including the
 * reflective halo refresh of sd after every sweep (:316-339). */
int tl_ppcg_inner(const double *alphas, const double *betas, int first, int count);
int tl_ppcg_store_r(void);                    /* tea_leaf_ppcg_store_r :433-454 */
int tl_ppcg_calc_zrnorm(double *rro);         /* tea_leaf_ppcg_calc_zrnorm :118-148 */
int tl_ppcg_calc_rrn(double *rrn);            /* tea_leaf_ppcg_calc_rrn :456-486 */
int tl_ppcg_update_z(void);                   /* tea_leaf_ppcg_update_z :150-172 */

/* ---- Jacobi (tea_leaf_jacobi.f90) -------------------------------------- */
int tl_jacobi_solve(double *error);           /* tea_leaf_jacobi_solve :11-46 */

/* ---- fused solver loops (device-resident alpha/beta and exit test) -------
 * Each runs consecutive iterations of the tea_leaf() DO loop
 * (tea_solve.f90:152-414) for one solver branch, keeping the reductions,
 * alpha = rro/pw, beta = rrn/rro and the exit test |sqrt(rrn)| < eps*r0 on
 * the device, so the host synchronises once per batch instead of twice per
 * iteration.  Iteration counts, the recorded CG alphas/betas and the final
 * error are those of the reference loop. */
typedef struct {
  /* in */
  int n_first;          /* first iteration number n to execute (1-based)      */
  int n_last;           /* last iteration allowed (tl_max_iters)              */
  double eps_abs;       /* eps * initial_residual (tea_solve.f90:410)          */
  int check_switch;     /* test the CG->Cheby/PPCG switch (tea_solve.f90:167)  */
  double epslim;        /* tl_ch_cg_epslim                                     */
  int presteps;         /* tl_ch_cg_presteps                                   */
  /* in/out */
  double rro;           /* rro entering n_first / after the last iteration    */
  double old_error;     /* old_error entering n_first / on exit               */
  /* out */
  int n_end;            /* iteration at which the loop stopped                */
  int reason;           /* TL_LOOP_* below                                    */
  double error;         /* sqrt|rrn| (or sqrt|error|) of iteration n_end      */
} tl_loop_state;

#define TL_LOOP_CONVERGED 1  /* iteration n_end met the exit test            */
#define TL_LOOP_SWITCH 2     /* switch condition true at the top of n_end    */
#define TL_LOOP_MAXITERS 3   /* n_end == n_last without meeting the test     */
#define TL_LOOP_NONFINITE 4  /* a reduction (p.w, r.r, z.r or the Jacobi
                              * error) of iteration n_end was NaN or Inf; the
                              * call returns TL_ERR_NONFINITE.  The reference
                              * does not check (it iterates to max_iters); a
                              * zero density gives such values through Kx/Ky
                              * (kernels/tea_leaf_common_kernel.f90:97-98). */
#define TL_LOOP_FLOOR 5      /* tl_mixed_solve only: the true residual has
                              * reached its rounding floor above rel_tol (a
                              * refinement whose inner solve gained 4x, or
                              * two in a row, failed to halve it)            */

/* CG branch (tea_solve.f90:341-369).  cg_alphas/cg_betas are host arrays
 * indexed by n (entry n written for every executed iteration), as
 * cg_alphas(n)/cg_betas(n) at :353,363.  Either may be NULL. */
int tl_cg_loop(tl_loop_state *st, double *cg_alphas, double *cg_betas);

/* tl_cg_loop without allocating: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc): the loop's control
 * block, per-iteration histories and partial sums live in a file-scope
 * device array of 622,976 bytes that each device holds until the process
 * ends, instead of the library's loop buffers.  Same kernels and results as
 * tl_cg_loop, bit for bit.  Limits: n_last <= 16384 (TL_ERR_ARG otherwise), a
 * chunk with all four faces external, kernel timing off for the call.
 * Every call on a device shares that device's array: calls that overlap on
 * one device race.  tl_loop_errors does not see its histories. */
int tl_cg_loop_noalloc(tl_loop_state *st, double *cg_alphas, double *cg_betas);

/* PPCG outer iterations after the switch (tea_solve.f90:289-331): calc_w,
 * store_r, zrnorm, calc_ur, `inner` sweeps with coefficients alphas/betas
 * (length inner), calc_rrn, calc_p.  inner_iters is incremented by `inner`
 * per executed outer iteration. */
int tl_ppcg_loop(tl_loop_state *st, const double *alphas, const double *betas,
                 int inner, double theta, int *inner_iters);

/* Jacobi branch (tea_solve.f90:375-379). */
int tl_jacobi_loop(tl_loop_state *st);

/* The exit-test error of every iteration lo..hi the last tl_cg_loop,
 * tl_ppcg_loop or tl_jacobi_loop call executed: err[n] = sqrt|rrn| (CG, PPCG)
 * or sqrt|error| (Jacobi) of iteration n, the value the reference prints as
 * "Residual" after each iteration with verbose_on (tea_solve.f90:402-408).
 * err is a host array indexed by n, as cg_alphas. */
int tl_loop_errors(int lo, int hi, double *err);

/* Chebyshev iterations n_first..n_last with coefficient index
 * step_first.. (tea_solve.f90:265-279), queued without synchronisation;
 * the residual norm is computed only on the last one when want_norm != 0. */
int tl_cheby_run(const double *ch_alphas, const double *ch_betas, int step_first,
                 int count, int want_norm, double *norm);

/* ---- mixed precision: FP32 inner solves under FP64 iterative refinement ---
 * No counterpart in the reference: it replaces the whole iteration loop of
 * tea_leaf() (tea_solve.f90:152-414) for a deck that asks for it.  Solves
 * A u = u0 for the current step after tl_common_init, starting from the u
 * that tl_common_init set (u = u0).  Each refinement computes the true
 * residual r = u0 - A u in FP64, solves A d = r with FP32 fields and stencil
 * (CG, or PPCG with the Chebyshev polynomial preconditioner of the
 * reference) whose dot products, alpha/beta and eigenvalue bounds are FP64,
 * and adds u += d in FP64.  It stops when ||u0 - A u|| <= rel_tol times the
 * residual of the starting u (a true residual, unlike the recurrence
 * residual of tea_solve.f90:402-410), after max_outer refinements, or when
 * the true residual has reached its rounding floor (reason
 * TL_LOOP_FLOOR): a refinement whose inner solve reduced its own residual
 * at least 4-fold, or two in a row, failed to halve it.  Requires
 * tl_preconditioner_type none and a chunk with all four faces external. */
#define TL_MIXED_CG 1
#define TL_MIXED_PPCG 2

typedef struct {
  /* in */
  int solver;             /* TL_MIXED_CG or TL_MIXED_PPCG                        */
  int max_outer;          /* refinement limit                                    */
  int max_inner;          /* iteration limit of one inner solve                  */
  int inner_steps;        /* PPCG polynomial degree (tl_ppcg_inner_steps)        */
  int presteps;           /* FP32 CG iterations that estimate the eigenvalues   */
  double rel_tol;         /* true relative residual to reach                     */
  double inner_tol;       /* residual reduction asked of an inner solve; one
                           * that needs less to reach rel_tol is asked for a
                           * quarter of what remains                             */
  /* in/out: PPCG eigenvalue bounds; > 0 on entry reuses them, else they are
   * estimated and returned.  Pass back only bounds of an earlier step of the
   * same chunk, mesh and rx, ry (the matrix is then the same): bounds below
   * the top of the spectrum make the polynomial preconditioner indefinite */
  double eigmin, eigmax;
  /* out */
  int outer;              /* refinements done                                    */
  int inner_iters;        /* CG or PPCG outer iterations summed over inner solves */
  int sweeps;             /* PPCG polynomial sweeps summed                       */
  int reason;             /* TL_LOOP_CONVERGED, TL_LOOP_FLOOR or TL_LOOP_MAXITERS */
  double initial_residual;/* ||u0 - A u|| at entry                               */
  double final_residual;  /* ||u0 - A u|| on exit (FP64)                         */
  int eig_iters;          /* PPCG: FP32 CG iterations whose Lanczos tridiagonal
                           * gave the eigenvalue bounds (presteps, doubled until
                           * the largest estimate moves by under 1 %); 0 when
                           * the bounds were passed in                           */
} tl_mixed_state;

int tl_mixed_solve(tl_mixed_state *st);

/* ---- adjoint of the implicit step ----------------------------------------
 * No counterpart in the reference: the building blocks of a reverse-mode
 * gradient of a functional of u after N steps with respect to the per-cell
 * log-density.  A step solves A(rho) u_n = b_n (b_1 = rho e_0, then
 * b_n = u_{n-1}); A is symmetric, so the adjoint of a solve is another solve
 * with the same operator (tl_cg_loop or tl_mixed_solve on the right-hand side
 * placed in u0), and the dependence of A on rho sits in the face
 * coefficients.  Snapshot slots are device arrays laid out like a field: the
 * checkpoints, the adjoint seed and the gradient stay in HBM. */

/* Reserve at least `count` snapshot slots (zeroed; kept until tl_finalize). */
int tl_snapshot_reserve(int count);
/* slot <- field, field <- slot (the whole allocation, halos included). */
int tl_snapshot_save(int slot, int field_id);
int tl_snapshot_load(int slot, int field_id);
/* Every value of a slot, halos included, set to `value`. */
int tl_snapshot_fill(int slot, double value);
/* Host <-> slot copies, as tl_upload / tl_download. */
int tl_snapshot_upload(int slot, const double *host, int host_halo);
int tl_snapshot_download(int slot, double *host, int host_halo);
/* Device array <-> slot copies of the interior cells, as tl_upload_device. */
int tl_snapshot_upload_device(int slot, const double *dev, long ld);
int tl_snapshot_download_device(int slot, double *dev, long ld);
/* grad += dJ/dlog(rho) of one step: for every cell c and each of its internal
 * faces f (neighbour b), -r_f (lambda_c - lambda_b)(u_c - u_b) dK_f/dlog(rho_c)
 * with r_f = rx or ry of the last tl_common_init and dK_f/dlog(rho_c) =
 * -1/(2 rho_c) (tl_coefficient_density) or rho_c/2 (inverse density); plus
 * lambda_c rho_c e0_c when e0_slot >= 0 (the first step's b_1 = rho e_0).
 * lambda is read from a field, u_n, e_0 and grad from slots. */
int tl_adjoint_face_grad(int lambda_field, int u_slot, int e0_slot, int grad_slot);

/* Device memory the library holds (fields, FP32 fields, snapshot slots,
 * reduction buffers), in bytes.  No counterpart in the reference. */
int tl_device_bytes(double *bytes);

/* Achievable HBM bandwidth of the device, STREAM style: the best copy
 * (c = a, 16 B per element) and triad (c = a + s b, 24 B per element) rate
 * in GB/s of hand-written kernels over three FP64 arrays of n elements (n
 * even; allocated for the call and freed), each timed over `reps` launches.
 * No counterpart in the reference: the ceiling the solver kernels' rates are
 * compared with.  Synchronises. */
int tl_hbm_stream(long n, int reps, double *copy_gbs, double *triad_gbs);

/* ---- timing of the dominant kernels (HIP events on the library stream) --- */
/* on = 1 times every launch of the solver kernels with HIP events, on = N > 1
 * every N-th launch of each class (events between launches slow a loop by a
 * few per cent; sampling keeps that out of a timed run), on = 0 stops. */
int tl_profile_enable(int on);
/* Returns timed launch count and their total milliseconds for a kernel class:
 * 0 = CG calc_w, 1 = CG calc_ur, 2 = PPCG inner sweep, 3 = Chebyshev iterate,
 * 4 = Jacobi, 5 = PPCG calc_w, 6 = PPCG calc_ur, 7 = mixed-precision FP64
 * refinement (u += d, true residual).  Launches that returned at once (queued
 * past a loop's exit: shorter than a quarter of the median timed launch of
 * the class) are left out and counted by tl_profile_idle. */
int tl_profile_get(int kernel_class, long *launches, double *total_ms);
/* Timed launches of a kernel class left out of tl_profile_get as idle. */
int tl_profile_idle(int kernel_class, long *launches);
int tl_profile_reset(void);

#ifdef __cplusplus
}
#endif

#endif /* TEALEAF_HIP_H */
