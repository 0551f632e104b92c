!> Support module of the drop-in driver modules: the reference's own
!> tea_solve.f90, diffuse.f90 and set-up (start.f90, read_input.f90, ...)
!> compiled unmodified and linked with the modules of this directory, whose
!> kernel branches call libtealeaf_hip.so instead of the reference's kernels.
!> The library holds every field of the (single) chunk on the GPU from the
!> call of generate_chunk on; the reference's host arrays are allocated by
!> its build_field but not used for the solve.
MODULE tl_dropin

  USE ISO_C_BINDING
  USE tl_hip_api
  USE report_module

  IMPLICIT NONE

  LOGICAL :: tl_dropin_ready = .FALSE.
  ! the library's copy of the reference's tl_ppcg_active (definitions.f90:73)
  INTEGER :: tl_dropin_ppcg_active = -1

CONTAINS

  !> The reference switches PPCG preconditioning on and off through the
  !> global tl_ppcg_active (tea_solve.f90:79,209), which its kernels read
  !> (kernels/tea_leaf_cg_kernel.f90:328, tea_leaf_ppcg_kernel.f90:157,432);
  !> the library keeps its own flag: set it before each call that reads it.
  SUBROUTINE tl_sync_ppcg_active(active)
    LOGICAL, INTENT(IN) :: active
    INTEGER :: a
    a = MERGE(1, 0, active)
    IF (a == tl_dropin_ppcg_active) RETURN
    CALL tl_check(tl_set_ppcg_active(INT(a, c_int)), 'tl_ppcg_active')
    tl_dropin_ppcg_active = a
  END SUBROUTINE tl_sync_ppcg_active

  !> A non-zero status from the library ends the run through the
  !> reference's report_error, with the library's message.
  SUBROUTINE tl_check(rc, where)
    INTEGER(c_int), INTENT(IN) :: rc
    CHARACTER(LEN=*), INTENT(IN) :: where
    CHARACTER(KIND=c_char), POINTER :: cmsg(:)
    CHARACTER(LEN=512) :: msg
    INTEGER :: i
    IF (rc == TL_OK) RETURN
    CALL C_F_POINTER(tl_last_error(), cmsg, [512])
    msg = ' '
    DO i = 1, 512
      IF (cmsg(i) == C_NULL_CHAR) EXIT
      msg(i:i) = cmsg(i)
    ENDDO
    CALL report_error(where, TRIM(msg))
  END SUBROUTINE tl_check

END MODULE tl_dropin
