!> The reference's external set-up and summary routines on the C ABI (same
!> names and argument lists as generate_chunk.f90:22-80 and
!> field_summary.f90:27-135, which the unmodified start.f90 and diffuse.f90
!> call).

!> Puts the chunk on the GPU: the library is initialised for the chunk's
!> cells and halo depth with the deck's options, and the states are
!> generated on the device (initialise_chunk + generate_chunk geometry,
!> vertex and cell-centre coordinates as the reference rounds them).
SUBROUTINE generate_chunk()

  USE tea_module
  USE tl_dropin

  IMPLICIT NONE

  TYPE(tl_state), ALLOCATABLE :: ts(:)
  INTEGER :: s
  REAL(KIND=8) :: dx, dy

  IF (tiles_per_task /= 1 .OR. parallel%max_task /= 1) &
    CALL report_error('generate_chunk', 'the GPU library holds one chunk of one tile')
  IF (.NOT. tl_dropin_ready) THEN
    CALL tl_check(tl_init(INT(chunk%tiles(1)%x_cells, c_int), INT(chunk%tiles(1)%y_cells, c_int), &
                          INT(chunk%halo_exchange_depth, c_int), 0_c_int), 'generate_chunk')
    tl_dropin_ready = .TRUE.
  ENDIF
  CALL tl_check(tl_set_options(INT(coefficient, c_int), INT(tl_preconditioner_type, c_int), &
                               MERGE(1_c_int, 0_c_int, reflective_boundary)), 'generate_chunk')
  ! the cell widths initialise_chunk computed (initialise_chunk.f90:34-35):
  ! start.f90 has since reset grid%x_cells to the chunk count
  dx = chunk%tiles(1)%field%celldx(chunk%tiles(1)%field%x_min)
  dy = chunk%tiles(1)%field%celldy(chunk%tiles(1)%field%y_min)
  ALLOCATE(ts(number_of_states))
  DO s = 1, number_of_states
    ts(s)%geometry = INT(states(s)%geometry, c_int)
    ts(s)%density = states(s)%density
    ts(s)%energy = states(s)%energy
    ts(s)%xmin = states(s)%xmin
    ts(s)%xmax = states(s)%xmax
    ts(s)%ymin = states(s)%ymin
    ts(s)%ymax = states(s)%ymax
    ts(s)%radius = states(s)%radius
  ENDDO
  CALL tl_check(tl_generate(grid%xmin, grid%ymin, dx, dy, INT(number_of_states, c_int), ts), &
                'generate_chunk')

END SUBROUTINE generate_chunk


!> Volume, mass, internal energy and temperature of the chunk from the
!> device, printed as the reference prints them, with its QA check of the
!> six benchmark problems at the end of the run.
SUBROUTINE field_summary()

  USE tea_module
  USE tl_dropin

  IMPLICIT NONE

  REAL(KIND=8) :: vol, mass, ie, temp, qa_diff, kernel_time, timer
  REAL(KIND=8), PARAMETER :: expected(6) = [157.55084183279294_8, 106.27221178646569_8, &
      99.955877498324000_8, 97.277332050749976_8, 95.462351583362249_8, 95.174738768320850_8]

  IF (parallel%boss) THEN
    WRITE(g_out, *)
    WRITE(g_out, *) 'Time ', time
    WRITE(g_out, '(a13,5a26)') '           ', 'Volume', 'Mass', 'Density', 'Energy', 'U'
  ENDIF
  IF (profiler_on) kernel_time = timer()
  CALL tl_check(tl_field_summary(vol, mass, ie, temp), 'field_summary')
  IF (profiler_on) profiler%summary = profiler%summary + (timer() - kernel_time)
  IF (parallel%boss) THEN
    WRITE(g_out, '(a6,i7,5e26.17)') ' step:', step, vol, mass, mass / vol, ie, temp
    WRITE(g_out, *)
    FLUSH(g_out)
  ENDIF
  IF (complete .AND. parallel%boss .AND. test_problem >= 1 .AND. test_problem <= 6) THEN
    qa_diff = ABS((100.0_8 * (temp / expected(test_problem))) - 100.0_8)
    WRITE(*, '(a,i4,a,e16.7,a)') "Test problem", test_problem, " is within", qa_diff, &
      "% of the expected solution"
    WRITE(g_out, '(a,i4,a,e16.7,a)') "Test problem", test_problem, " is within", qa_diff, &
      "% of the expected solution"
    IF (qa_diff < 0.001) THEN
      WRITE(*, *) "This test is considered PASSED"
      WRITE(g_out, *) "This test is considered PASSED"
    ELSE
      WRITE(*, *) "This test is considered NOT PASSED"
      WRITE(g_out, *) "This is test is considered NOT PASSED"
    ENDIF
  ENDIF

END SUBROUTINE field_summary
