!> Host scalar routines of the Chebyshev and PPCG solvers (tea_leaf_cheby.f90):
!> eigenvalue bounds from the CG Lanczos tridiagonal and the Chebyshev
!> recurrence coefficients.  Used by the native host (tea_host.f90) and by
!> the drop-in driver modules (tealeaf_ref_amd/dropin).
MODULE tl_eigen

  IMPLICIT NONE

CONTAINS

  !> Eigenvalues of the symmetric tridiagonal (d, e) by the implicit QL
  !> method with Wilkinson shifts: the published routine `tqli` of
  !> W. H. Press et al., Numerical Recipes (2nd ed., section 11.3), after
  !> Bowdler, Martin, Reinsch and Wilkinson's tql1 (Handbook for Automatic
  !> Computation II, 1971), eigenvalues only.  The reference uses the same
  !> routine (tea_leaf_cheby.f90:90-140); its order of operations decides the
  !> printed eigenvalues and the Chebyshev iteration counts, so it is kept.
  !> e(2:n) holds the sub-diagonal on entry.  info = 1 after 30 sweeps
  !> without convergence of an eigenvalue.
  SUBROUTINE ql_eigenvalues(d, e, n, info)
    INTEGER, INTENT(IN) :: n
    REAL(KIND=8), INTENT(INOUT) :: d(n), e(n)
    INTEGER, INTENT(OUT) :: info
    INTEGER :: l, m, i, sweeps
    REAL(KIND=8) :: g, r, s, c, p, f, b, tol
    LOGICAL :: underflow
    info = 0
    IF (n < 1) RETURN
    e(1:n-1) = e(2:n)
    e(n) = 0.0_8
    DO l = 1, n
      sweeps = 0
      DO
        ! find the first negligible sub-diagonal at or after l
        m = l
        DO WHILE (m < n)
          tol = ABS(d(m)) + ABS(d(m + 1))
          IF (ABS(e(m)) + tol == tol) EXIT
          m = m + 1
        ENDDO
        IF (m == l) EXIT
        IF (sweeps == 30) THEN
          info = 1
          RETURN
        ENDIF
        sweeps = sweeps + 1
        g = (d(l + 1) - d(l)) / (2.0_8 * e(l))
        r = SQRT(g * g + 1.0_8)
        g = d(m) - d(l) + e(l) / (g + SIGN(r, g))
        s = 1.0_8
        c = 1.0_8
        p = 0.0_8
        underflow = .FALSE.
        DO i = m - 1, l, -1
          f = s * e(i)
          b = c * e(i)
          r = SQRT(f * f + g * g)
          e(i + 1) = r
          IF (r == 0.0_8) THEN
            d(i + 1) = d(i + 1) - p
            e(m) = 0.0_8
            underflow = .TRUE.
            EXIT
          ENDIF
          s = f / r
          c = g / r
          g = d(i + 1) - p
          r = (d(i) - g) * s + 2.0_8 * c * b
          p = s * r
          d(i + 1) = g + p
          g = c * r - b
        ENDDO
        IF (underflow) CYCLE
        d(l) = d(l) - p
        e(l) = g
        e(m) = 0.0_8
      ENDDO
    ENDDO
  END SUBROUTINE ql_eigenvalues

  !> Lanczos tridiagonal from the CG coefficients and its extreme eigenvalues
  !> (tea_leaf_cheby.f90:142-176).
  SUBROUTINE calc_eigenvalues(al, be, nsteps, emin, emax, info)
    INTEGER, INTENT(IN) :: nsteps
    REAL(KIND=8), INTENT(IN) :: al(0:), be(0:)
    REAL(KIND=8), INTENT(OUT) :: emin, emax
    INTEGER, INTENT(OUT) :: info
    REAL(KIND=8), ALLOCATABLE :: d(:), e(:)
    INTEGER :: i
    ALLOCATE(d(MAX(nsteps, 1)), e(MAX(nsteps, 1)))
    d = 0.0_8
    e = 0.0_8
    DO i = 1, nsteps
      d(i) = 1.0_8 / al(i)
      IF (i > 1) d(i) = d(i) + be(i - 1) / al(i - 1)
      IF (i < nsteps) e(i + 1) = SQRT(be(i)) / al(i)
    ENDDO
    CALL ql_eigenvalues(d, e, nsteps, info)
    emin = 0.0_8
    emax = 0.0_8
    IF (info /= 0) RETURN
    emin = MINVAL(d(1:nsteps))
    emax = MAXVAL(d(1:nsteps))
    IF (emin < 0.0_8 .OR. emax < 0.0_8) info = 1
  END SUBROUTINE calc_eigenvalues

  !> Chebyshev recurrence coefficients (tea_leaf_cheby.f90:178-205)
  SUBROUTINE calc_ch_coefs(ca, cb, emin, emax, th, n)
    INTEGER, INTENT(IN) :: n
    REAL(KIND=8), INTENT(OUT) :: ca(:), cb(:), th
    REAL(KIND=8), INTENT(IN) :: emin, emax
    REAL(KIND=8) :: delta, sigma, rho_old, rho_new
    INTEGER :: i
    th = (emax + emin) / 2
    delta = (emax - emin) / 2
    sigma = th / delta
    rho_old = 1.0_8 / sigma
    DO i = 1, n
      rho_new = 1.0_8 / (2.0_8 * sigma - rho_old)
      ca(i) = rho_new * rho_old
      cb(i) = 2.0_8 * rho_new / delta
      rho_old = rho_new
    ENDDO
  END SUBROUTINE calc_ch_coefs

END MODULE tl_eigen
