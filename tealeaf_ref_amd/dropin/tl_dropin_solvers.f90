!> The reference's driver-wrapper modules, re-implemented on the C ABI: the
!> module and routine names and argument lists are the reference's (they
!> are what its unmodified tea_solve.f90, diffuse.f90 and start.f90 call);
!> each body calls the library routine that replaces the wrapper's kernel
!> (include/tealeaf_hip.h names the reference routine of each).  One chunk
!> of one tile: the per-tile loops and the MPI sums of the reference are
!> identities here.
!>
!>   update_halo_module      update_halo.f90:32-92      -> tl_update_halo
!>   set_field_module        set_field.f90:26-57        -> tl_set_field
!>   tea_leaf_common_module  tea_leaf_common.f90:12-163 -> tl_common_init, ...
!>   tea_leaf_cg_module      tea_leaf_cg.f90:11-158     -> tl_cg_*
!>   tea_leaf_cheby_module   tea_leaf_cheby.f90:13-252  -> tl_cheby_*, tl_eigen
!>   tea_leaf_ppcg_module    tea_leaf_ppcg.f90:33-486   -> tl_ppcg_*
!>   tea_leaf_jacobi_module  tea_leaf_jacobi.f90:11-46  -> tl_jacobi_solve

MODULE update_halo_module

  USE tea_module
  USE tl_dropin

  IMPLICIT NONE

CONTAINS

  !> Reflective external halos of the flagged fields (one chunk: no
  !> exchange between chunks or tiles).
  SUBROUTINE update_halo(fields, depth)
    INTEGER :: fields(NUM_FIELDS), depth
    INTEGER(c_int) :: f(NUM_FIELDS)
    REAL(KIND=8) :: timer, halo_time
    IF (profiler_on) halo_time = timer()
    IF (reflective_boundary .AND. ANY(chunk%chunk_neighbours == EXTERNAL_FACE)) THEN
      f = INT(fields, c_int)
      CALL tl_check(tl_update_halo(f, INT(depth, c_int)), 'update_halo')
    ENDIF
    IF (profiler_on) profiler%halo_update = profiler%halo_update + (timer() - halo_time)
  END SUBROUTINE update_halo

END MODULE update_halo_module


MODULE set_field_module

CONTAINS

  SUBROUTINE set_field()
    USE tea_module
    USE tl_dropin
    IMPLICIT NONE
    REAL(KIND=8) :: kernel_time, timer
    IF (profiler_on) kernel_time = timer()
    CALL tl_check(tl_set_field(), 'set_field')
    IF (profiler_on) profiler%set_field = profiler%set_field + (timer() - kernel_time)
  END SUBROUTINE set_field

END MODULE set_field_module


MODULE tea_leaf_common_module

  USE definitions_module
  USE report_module
  USE tl_dropin

  IMPLICIT NONE

CONTAINS

  !> rx, ry from the cell widths (tea_leaf_common.f90:24-25), the faces of
  !> the chunk that are external to the mesh, then the library's set-up.
  SUBROUTINE tea_leaf_init_common()
    INTEGER(c_int) :: zb(4)
    INTEGER :: t
    t = 1
    chunk%tiles(t)%field%rx = dt / (chunk%tiles(t)%field%celldx(chunk%tiles(t)%field%x_min)**2)
    chunk%tiles(t)%field%ry = dt / (chunk%tiles(t)%field%celldy(chunk%tiles(t)%field%y_min)**2)
    zb = MERGE(1_c_int, 0_c_int, chunk%tiles(t)%tile_neighbours == EXTERNAL_FACE .AND. &
                                 chunk%chunk_neighbours == EXTERNAL_FACE)
    CALL tl_check(tl_common_init(chunk%tiles(t)%field%rx, chunk%tiles(t)%field%ry, zb), &
                  'tea_leaf_init_common')
  END SUBROUTINE tea_leaf_init_common

  SUBROUTINE tea_leaf_calc_residual()
    CALL tl_check(tl_calc_residual(), 'tea_leaf_calc_residual')
  END SUBROUTINE tea_leaf_calc_residual

  SUBROUTINE tea_leaf_calc_2norm(norm_array, norm)
    INTEGER :: norm_array
    REAL(KIND=8) :: norm
    IF (norm_array /= 0 .AND. norm_array /= 1) &
      CALL report_error("tea_leaf_common.f90", "Invalid value for norm_array")
    CALL tl_check(tl_calc_2norm(INT(norm_array, c_int), norm), 'tea_leaf_calc_2norm')
  END SUBROUTINE tea_leaf_calc_2norm

  SUBROUTINE tea_leaf_finalise()
    CALL tl_check(tl_finalise(), 'tea_leaf_finalise')
  END SUBROUTINE tea_leaf_finalise

END MODULE tea_leaf_common_module


MODULE tea_leaf_cg_module

  USE definitions_module
  USE tl_dropin

  IMPLICIT NONE

CONTAINS

  SUBROUTINE tea_leaf_cg_init(rro)
    REAL(KIND=8) :: rro
    CALL tl_sync_ppcg_active(tl_ppcg_active)
    CALL tl_check(tl_cg_init(rro), 'tea_leaf_cg_init')
  END SUBROUTINE tea_leaf_cg_init

  SUBROUTINE tea_leaf_cg_calc_w(pw)
    REAL(KIND=8) :: pw
    CALL tl_check(tl_cg_calc_w(pw), 'tea_leaf_cg_calc_w')
  END SUBROUTINE tea_leaf_cg_calc_w

  SUBROUTINE tea_leaf_cg_calc_ur(alpha, rrn)
    REAL(KIND=8) :: alpha, rrn
    CALL tl_check(tl_cg_calc_ur(alpha, rrn), 'tea_leaf_cg_calc_ur')
  END SUBROUTINE tea_leaf_cg_calc_ur

  SUBROUTINE tea_leaf_cg_calc_p(beta)
    REAL(KIND=8) :: beta
    CALL tl_sync_ppcg_active(tl_ppcg_active)
    CALL tl_check(tl_cg_calc_p(beta), 'tea_leaf_cg_calc_p')
  END SUBROUTINE tea_leaf_cg_calc_p

END MODULE tea_leaf_cg_module


MODULE tea_leaf_cheby_module

  USE definitions_module
  USE tea_leaf_common_module
  USE update_halo_module
  USE tl_eigen
  USE tl_dropin

  IMPLICIT NONE

CONTAINS

  SUBROUTINE tea_leaf_cheby_init(theta)
    REAL(KIND=8) :: theta
    CALL tl_check(tl_cheby_init(theta), 'tea_leaf_cheby_init')
  END SUBROUTINE tea_leaf_cheby_init

  !> One Chebyshev iteration with the coefficients of step cheby_calc_steps.
  SUBROUTINE tea_leaf_cheby_iterate(ch_alphas, ch_betas, max_cheby_iters, cheby_calc_steps)
    INTEGER :: cheby_calc_steps, max_cheby_iters
    REAL(KIND=8), DIMENSION(:) :: ch_alphas, ch_betas
    IF (cheby_calc_steps < 1 .OR. cheby_calc_steps > max_cheby_iters) &
      CALL report_error('tea_leaf_cheby_iterate', 'coefficient index out of range')
    CALL tl_check(tl_cheby_iterate(ch_alphas(cheby_calc_steps), ch_betas(cheby_calc_steps)), &
                  'tea_leaf_cheby_iterate')
  END SUBROUTINE tea_leaf_cheby_iterate

  !> Eigenvalue bounds from the first tl_ch_cg_presteps CG coefficients
  !> (the Lanczos tridiagonal; tl_eigen).
  SUBROUTINE tea_calc_eigenvalues(cg_alphas, cg_betas, eigmin, eigmax, max_iters, &
                                  tl_ch_cg_presteps, info)
    INTEGER :: tl_ch_cg_presteps, max_iters, info
    REAL(KIND=8), DIMENSION(max_iters) :: cg_alphas, cg_betas
    REAL(KIND=8) :: eigmin, eigmax
    REAL(KIND=8), ALLOCATABLE :: al(:), be(:)
    ALLOCATE(al(0:tl_ch_cg_presteps), be(0:tl_ch_cg_presteps))
    al(0) = 0.0_8
    be(0) = 0.0_8
    al(1:tl_ch_cg_presteps) = cg_alphas(1:tl_ch_cg_presteps)
    be(1:tl_ch_cg_presteps) = cg_betas(1:tl_ch_cg_presteps)
    CALL calc_eigenvalues(al, be, tl_ch_cg_presteps, eigmin, eigmax, info)
  END SUBROUTINE tea_calc_eigenvalues

  SUBROUTINE tea_calc_ch_coefs(ch_alphas, ch_betas, eigmin, eigmax, theta, max_cheby_iters)
    INTEGER :: max_cheby_iters
    REAL(KIND=8), DIMENSION(max_cheby_iters) :: ch_alphas, ch_betas
    REAL(KIND=8) :: eigmin, eigmax, theta
    CALL calc_ch_coefs(ch_alphas, ch_betas, eigmin, eigmax, theta, max_cheby_iters)
  END SUBROUTINE tea_calc_ch_coefs

  !> tea_leaf_cheby.f90:208-252: ||u0||, p from the residual, the first
  !> iteration, ||r|| into `error`, and the iteration estimate.
  SUBROUTINE tea_leaf_cheby_first_step(ch_alphas, ch_betas, fields, error, theta, cn, &
                                       max_cheby_iters, est_itc, solve_time)
    INTEGER :: est_itc, max_cheby_iters
    INTEGER, DIMENSION(:) :: fields
    REAL(KIND=8) :: error, theta, cn, solve_time
    REAL(KIND=8), DIMENSION(:) :: ch_alphas, ch_betas
    REAL(KIND=8) :: it_alpha, gamm, bb, halo_time, timer
    CALL tea_leaf_calc_2norm(0, bb)
    CALL tea_leaf_cheby_init(theta)
    IF (profiler_on) halo_time = timer()
    CALL update_halo(fields, 1)
    IF (profiler_on) solve_time = solve_time + (timer() - halo_time)
    CALL tea_leaf_cheby_iterate(ch_alphas, ch_betas, max_cheby_iters, 1)
    CALL tea_leaf_calc_2norm(1, error)
    it_alpha = eps / 2.0_8 * SQRT(bb / error)
    gamm = (SQRT(cn) - 1.0_8) / (SQRT(cn) + 1.0_8)
    est_itc = NINT(LOG(it_alpha) / (LOG(gamm)))
    IF (parallel%boss) THEN
      WRITE(g_out, '(a11)') "est itc"
      WRITE(g_out, '(11i11)') est_itc
      WRITE(0, '(a11)') "est itc"
      WRITE(0, '(11i11)') est_itc
    ENDIF
  END SUBROUTINE tea_leaf_cheby_first_step

END MODULE tea_leaf_cheby_module


MODULE tea_leaf_ppcg_module

  USE definitions_module
  USE tea_leaf_cheby_module
  USE update_halo_module
  USE tl_dropin

  IMPLICIT NONE

CONTAINS

  SUBROUTINE tea_leaf_ppcg_init_sd(theta)
    REAL(KIND=8) :: theta
    CALL tl_sync_ppcg_active(tl_ppcg_active)
    CALL tl_check(tl_ppcg_init_sd(theta), 'tea_leaf_ppcg_init_sd')
  END SUBROUTINE tea_leaf_ppcg_init_sd

  SUBROUTINE tea_leaf_ppcg_init(ppcg_inner_iters, ch_alphas, ch_betas, theta, solve_time, step, rro)
    INTEGER :: ppcg_inner_iters, step
    REAL(KIND=8) :: rro, theta, solve_time
    REAL(KIND=8), DIMENSION(:) :: ch_alphas, ch_betas
    CALL tl_sync_ppcg_active(tl_ppcg_active)
    CALL tl_check(tl_ppcg_init(INT(step, c_int), rro), 'tea_leaf_ppcg_init')
  END SUBROUTINE tea_leaf_ppcg_init

  SUBROUTINE tea_leaf_ppcg_calc_zrnorm(rrn)
    REAL(KIND=8) :: rrn
    CALL tl_sync_ppcg_active(tl_ppcg_active)
    CALL tl_check(tl_ppcg_calc_zrnorm(rrn), 'tea_leaf_ppcg_calc_zrnorm')
  END SUBROUTINE tea_leaf_ppcg_calc_zrnorm

  SUBROUTINE tea_leaf_ppcg_update_z()
    CALL tl_sync_ppcg_active(tl_ppcg_active)
    CALL tl_check(tl_ppcg_update_z(), 'tea_leaf_ppcg_update_z')
  END SUBROUTINE tea_leaf_ppcg_update_z

  SUBROUTINE tea_calc_ls_coefs(ch_alphas, ch_betas, eigmin, eigmax, theta, ppcg_inner_steps)
    INTEGER :: ppcg_inner_steps
    REAL(KIND=8), DIMENSION(ppcg_inner_steps) :: ch_alphas, ch_betas
    REAL(KIND=8) :: eigmin, eigmax, theta
    CALL tea_calc_ch_coefs(ch_alphas, ch_betas, eigmin, eigmax, theta, ppcg_inner_steps)
  END SUBROUTINE tea_calc_ls_coefs

  !> tea_leaf_ppcg.f90:211-431: sd from the residual, the halo of sd and r,
  !> the inner sweeps (with the reflective refresh of sd after each, inside
  !> tl_ppcg_inner), z = utemp.
  SUBROUTINE tea_leaf_run_ppcg_inner_steps(ch_alphas, ch_betas, theta, ppcg_inner_steps, &
                                           solve_time)
    INTEGER :: ppcg_inner_steps
    REAL(KIND=8) :: theta, solve_time
    REAL(KIND=8), DIMENSION(MAX(ppcg_inner_steps, 0)) :: ch_alphas, ch_betas
    INTEGER :: fields(NUM_FIELDS)
    REAL(KIND=8) :: halo_time, timer
    IF (ppcg_inner_steps < 0) RETURN
    CALL tea_leaf_ppcg_init_sd(theta)
    fields = 0
    fields(FIELD_SD) = 1
    fields(FIELD_R) = 1
    IF (profiler_on) halo_time = timer()
    CALL update_halo(fields, chunk%halo_exchange_depth)
    IF (profiler_on) solve_time = solve_time + (timer() - halo_time)
    IF (ppcg_inner_steps > 0) &
      CALL tl_check(tl_ppcg_inner(ch_alphas, ch_betas, 1_c_int, INT(ppcg_inner_steps, c_int)), &
                    'tea_leaf_run_ppcg_inner_steps')
    CALL tea_leaf_ppcg_update_z()
  END SUBROUTINE tea_leaf_run_ppcg_inner_steps

  SUBROUTINE tea_leaf_ppcg_store_r()
    CALL tl_check(tl_ppcg_store_r(), 'tea_leaf_ppcg_store_r')
  END SUBROUTINE tea_leaf_ppcg_store_r

  SUBROUTINE tea_leaf_ppcg_calc_rrn(rrn)
    REAL(KIND=8) :: rrn
    CALL tl_sync_ppcg_active(tl_ppcg_active)
    CALL tl_check(tl_ppcg_calc_rrn(rrn), 'tea_leaf_ppcg_calc_rrn')
  END SUBROUTINE tea_leaf_ppcg_calc_rrn

END MODULE tea_leaf_ppcg_module


MODULE tea_leaf_jacobi_module

  USE definitions_module
  USE tl_dropin

  IMPLICIT NONE

CONTAINS

  SUBROUTINE tea_leaf_jacobi_solve(error)
    REAL(KIND=8) :: error
    CALL tl_check(tl_jacobi_solve(error), 'tea_leaf_jacobi_solve')
  END SUBROUTINE tea_leaf_jacobi_solve

END MODULE tea_leaf_jacobi_module
