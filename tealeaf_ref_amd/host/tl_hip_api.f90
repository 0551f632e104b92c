!> ISO_C_BINDING interface to libtealeaf_hip.so (include/tealeaf_hip.h).
!> One interface per C entry point; scalars come back through intent(out)
!> arguments, exactly as the reference kernels return them.
MODULE tl_hip_api

  USE ISO_C_BINDING
  IMPLICIT NONE

  INTEGER(c_int), PARAMETER :: TL_OK = 0
  INTEGER(c_int), PARAMETER :: TL_LOOP_CONVERGED = 1, TL_LOOP_SWITCH = 2, TL_LOOP_MAXITERS = 3, &
                                TL_LOOP_NONFINITE = 4, TL_LOOP_FLOOR = 5

  INTEGER(c_int), PARAMETER :: TLF_DENSITY = 1, TLF_ENERGY0 = 2, TLF_ENERGY1 = 3, TLF_U = 4, &
                               TLF_P = 5, TLF_SD = 6, TLF_R = 7, TLF_Z = 8, TLF_KX = 9,     &
                               TLF_KY = 10, TLF_DI = 11, TLF_U0 = 12, TLF_W = 13,           &
                               TLF_VOLUME = 18

  TYPE, BIND(C) :: tl_loop_state
    INTEGER(c_int) :: n_first, n_last
    REAL(c_double) :: eps_abs
    INTEGER(c_int) :: check_switch
    REAL(c_double) :: epslim
    INTEGER(c_int) :: presteps
    REAL(c_double) :: rro, old_error
    INTEGER(c_int) :: n_end, reason
    REAL(c_double) :: error
  END TYPE tl_loop_state

  INTEGER(c_int), PARAMETER :: TL_MIXED_CG = 1, TL_MIXED_PPCG = 2

  TYPE, BIND(C) :: tl_state
    INTEGER(c_int) :: geometry
    REAL(c_double) :: density, energy, xmin, xmax, ymin, ymax, radius
  END TYPE tl_state

  TYPE, BIND(C) :: tl_mixed_state
    INTEGER(c_int) :: solver, max_outer, max_inner, inner_steps, presteps
    REAL(c_double) :: rel_tol, inner_tol, eigmin, eigmax
    INTEGER(c_int) :: outer, inner_iters, sweeps, reason
    REAL(c_double) :: initial_residual, final_residual
    INTEGER(c_int) :: eig_iters
  END TYPE tl_mixed_state

  INTERFACE
    INTEGER(c_int) FUNCTION tl_init(nx, ny, halo, device) BIND(C, name="tl_init")
      IMPORT :: c_int
      INTEGER(c_int), VALUE :: nx, ny, halo, device
    END FUNCTION
    INTEGER(c_int) FUNCTION tl_finalize() BIND(C, name="tl_finalize")
      IMPORT :: c_int
    END FUNCTION
    TYPE(c_ptr) FUNCTION tl_last_error() BIND(C, name="tl_last_error")
      IMPORT :: c_ptr
    END FUNCTION
    INTEGER(c_int) FUNCTION tl_set_options(coef, prec, refl) BIND(C, name="tl_set_options")
      IMPORT :: c_int
      INTEGER(c_int), VALUE :: coef, prec, refl
    END FUNCTION
    INTEGER(c_int) FUNCTION tl_set_ppcg_active(active) BIND(C, name="tl_set_ppcg_active")
      IMPORT :: c_int
      INTEGER(c_int), VALUE :: active
    END FUNCTION
    INTEGER(c_int) FUNCTION tl_upload(fid, host, halo) BIND(C, name="tl_upload")
      IMPORT :: c_int, c_double
      INTEGER(c_int), VALUE :: fid, halo
      REAL(c_double), INTENT(IN) :: host(*)
    END FUNCTION
    INTEGER(c_int) FUNCTION tl_download(fid, host, halo) BIND(C, name="tl_download")
      IMPORT :: c_int, c_double
      INTEGER(c_int), VALUE :: fid, halo
      REAL(c_double), INTENT(OUT) :: host(*)
    END FUNCTION
    INTEGER(c_int) FUNCTION tl_synchronize() BIND(C, name="tl_synchronize")
      IMPORT :: c_int
    END FUNCTION
    INTEGER(c_int) FUNCTION tl_common_init(rx, ry, zb) BIND(C, name="tl_common_init")
      IMPORT :: c_int, c_double
      REAL(c_double), VALUE :: rx, ry
      INTEGER(c_int), INTENT(IN) :: zb(4)
    END FUNCTION
    INTEGER(c_int) FUNCTION tl_calc_residual() BIND(C, name="tl_calc_residual")
      IMPORT :: c_int
    END FUNCTION
    INTEGER(c_int) FUNCTION tl_calc_2norm(which, norm) BIND(C, name="tl_calc_2norm")
      IMPORT :: c_int, c_double
      INTEGER(c_int), VALUE :: which
      REAL(c_double), INTENT(OUT) :: norm
    END FUNCTION
    INTEGER(c_int) FUNCTION tl_finalise() BIND(C, name="tl_finalise")
      IMPORT :: c_int
    END FUNCTION
    INTEGER(c_int) FUNCTION tl_set_field() BIND(C, name="tl_set_field")
      IMPORT :: c_int
    END FUNCTION
    INTEGER(c_int) FUNCTION tl_update_halo(fields, depth) BIND(C, name="tl_update_halo")
      IMPORT :: c_int
      INTEGER(c_int), INTENT(IN) :: fields(11)
      INTEGER(c_int), VALUE :: depth
    END FUNCTION
    INTEGER(c_int) FUNCTION tl_field_summary(vol, mass, ie, temp) BIND(C, name="tl_field_summary")
      IMPORT :: c_int, c_double
      REAL(c_double), INTENT(OUT) :: vol, mass, ie, temp
    END FUNCTION
    INTEGER(c_int) FUNCTION tl_cg_init(rro) BIND(C, name="tl_cg_init")
      IMPORT :: c_int, c_double
      REAL(c_double), INTENT(OUT) :: rro
    END FUNCTION
    INTEGER(c_int) FUNCTION tl_cheby_init(theta) BIND(C, name="tl_cheby_init")
      IMPORT :: c_int, c_double
      REAL(c_double), VALUE :: theta
    END FUNCTION
    INTEGER(c_int) FUNCTION tl_cheby_iterate(a, b) BIND(C, name="tl_cheby_iterate")
      IMPORT :: c_int, c_double
      REAL(c_double), VALUE :: a, b
    END FUNCTION
    INTEGER(c_int) FUNCTION tl_cheby_run(a, b, step_first, cnt, want, norm) BIND(C, name="tl_cheby_run")
      IMPORT :: c_int, c_double
      REAL(c_double), INTENT(IN) :: a(*), b(*)
      INTEGER(c_int), VALUE :: step_first, cnt, want
      REAL(c_double), INTENT(OUT) :: norm
    END FUNCTION
    INTEGER(c_int) FUNCTION tl_ppcg_init(step, rro) BIND(C, name="tl_ppcg_init")
      IMPORT :: c_int, c_double
      INTEGER(c_int), VALUE :: step
      REAL(c_double), INTENT(OUT) :: rro
    END FUNCTION
    INTEGER(c_int) FUNCTION tl_ppcg_init_sd(theta) BIND(C, name="tl_ppcg_init_sd")
      IMPORT :: c_int, c_double
      REAL(c_double), VALUE :: theta
    END FUNCTION
    INTEGER(c_int) FUNCTION tl_ppcg_inner(a, b, first, cnt) BIND(C, name="tl_ppcg_inner")
      IMPORT :: c_int, c_double
      REAL(c_double), INTENT(IN) :: a(*), b(*)
      INTEGER(c_int), VALUE :: first, cnt
    END FUNCTION
    INTEGER(c_int) FUNCTION tl_ppcg_update_z() BIND(C, name="tl_ppcg_update_z")
      IMPORT :: c_int
    END FUNCTION
    INTEGER(c_int) FUNCTION tl_cg_loop(st, al, be) BIND(C, name="tl_cg_loop")
      IMPORT :: c_int, c_double, tl_loop_state
      TYPE(tl_loop_state), INTENT(INOUT) :: st
      REAL(c_double), INTENT(INOUT) :: al(*), be(*)
    END FUNCTION
    INTEGER(c_int) FUNCTION tl_ppcg_loop(st, a, b, inner, theta, inner_iters) BIND(C, name="tl_ppcg_loop")
      IMPORT :: c_int, c_double, tl_loop_state
      TYPE(tl_loop_state), INTENT(INOUT) :: st
      REAL(c_double), INTENT(IN) :: a(*), b(*)
      INTEGER(c_int), VALUE :: inner
      REAL(c_double), VALUE :: theta
      INTEGER(c_int), INTENT(INOUT) :: inner_iters
    END FUNCTION
    INTEGER(c_int) FUNCTION tl_generate(xmin, ymin, dx, dy, nstates, states) BIND(C, name="tl_generate")
      IMPORT :: c_int, c_double, tl_state
      REAL(c_double), VALUE :: xmin, ymin, dx, dy
      INTEGER(c_int), VALUE :: nstates
      TYPE(tl_state), INTENT(IN) :: states(*)
    END FUNCTION
    INTEGER(c_int) FUNCTION tl_mixed_solve(st) BIND(C, name="tl_mixed_solve")
      IMPORT :: c_int, tl_mixed_state
      TYPE(tl_mixed_state), INTENT(INOUT) :: st
    END FUNCTION
    INTEGER(c_int) FUNCTION tl_jacobi_loop(st) BIND(C, name="tl_jacobi_loop")
      IMPORT :: c_int, tl_loop_state
      TYPE(tl_loop_state), INTENT(INOUT) :: st
    END FUNCTION
    INTEGER(c_int) FUNCTION tl_cg_calc_w(pw) BIND(C, name="tl_cg_calc_w")
      IMPORT :: c_int, c_double
      REAL(c_double), INTENT(OUT) :: pw
    END FUNCTION
    INTEGER(c_int) FUNCTION tl_cg_calc_ur(alpha, rrn) BIND(C, name="tl_cg_calc_ur")
      IMPORT :: c_int, c_double
      REAL(c_double), VALUE :: alpha
      REAL(c_double), INTENT(OUT) :: rrn
    END FUNCTION
    INTEGER(c_int) FUNCTION tl_cg_calc_p(beta) BIND(C, name="tl_cg_calc_p")
      IMPORT :: c_int, c_double
      REAL(c_double), VALUE :: beta
    END FUNCTION
    INTEGER(c_int) FUNCTION tl_ppcg_store_r() BIND(C, name="tl_ppcg_store_r")
      IMPORT :: c_int
    END FUNCTION
    INTEGER(c_int) FUNCTION tl_ppcg_calc_zrnorm(rro) BIND(C, name="tl_ppcg_calc_zrnorm")
      IMPORT :: c_int, c_double
      REAL(c_double), INTENT(OUT) :: rro
    END FUNCTION
    INTEGER(c_int) FUNCTION tl_ppcg_calc_rrn(rrn) BIND(C, name="tl_ppcg_calc_rrn")
      IMPORT :: c_int, c_double
      REAL(c_double), INTENT(OUT) :: rrn
    END FUNCTION
    INTEGER(c_int) FUNCTION tl_jacobi_solve(err) BIND(C, name="tl_jacobi_solve")
      IMPORT :: c_int, c_double
      REAL(c_double), INTENT(OUT) :: err
    END FUNCTION
    INTEGER(c_int) FUNCTION tl_loop_errors(lo, hi, err) BIND(C, name="tl_loop_errors")
      IMPORT :: c_int, c_double
      INTEGER(c_int), VALUE :: lo, hi
      REAL(c_double), INTENT(INOUT) :: err(*)
    END FUNCTION
  END INTERFACE

END MODULE tl_hip_api
