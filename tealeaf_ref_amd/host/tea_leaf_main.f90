!> Drop-in for the reference executable: reads ./tea.in, writes ./tea.out
!> (tea_leaf.f90:35-73), with the field work on the GPU.
PROGRAM tea_leaf
  USE tea_host
  IMPLICIT NONE
  CALL initialise('tea.in', 'tea.out', 0)
  WRITE(*, *) 'Output file tea.out opened. All output will go there.'
  DO
    IF (diffuse_step()) EXIT
  ENDDO
  CALL shutdown()
END PROGRAM tea_leaf
