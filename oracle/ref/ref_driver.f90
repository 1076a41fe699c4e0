! Test infrastructure only: C-callable driver around the reference's own
! self-contained Fortran files, compiled from their sources in the reference
! tree (Constants.f90, MathSplines.f90, MathDCsrch.f, MathDCstep.f,
! MathRKsuite.f, MathLNsrch.f90 with the serial OutputFiles.f90 and
! MPI_Functions.f90 it uses) into oracle/_ref/libprofess_ref.so by
! oracle/ref/Makefile.
!
! Exports (BIND(C)):
!   ref_spline_set / ref_spline_val  -> MathSplines spline_cubic_set/_val
!   ref_dcsrch                       -> MathDCsrch dcsrch (task as an int code)
!   ref_lnsrch                       -> MathLNsrch LnSrch (task as 70 characters,
!                                       the module's points(2,3) copied out)
!   ref_wgc_eta_table                -> restatement of makeKernel2 + the WGC ODE
!                                       right-hand side (KEDF_KernelODE.f90:254-537)
!                                       integrated by the reference's RKSUITE
MODULE ref_wgc_ode
  IMPLICIT NONE
  REAL(KIND=8) :: ode_beta = 0.d0, ode_gamma = 0.d0
CONTAINS
  ! Right-hand side restated from KEDF_KernelODE.f90:481-537 (same arithmetic).
  SUBROUTINE wgc_rhs(T, Y, YP)
    REAL(KIND=8) :: T, Y(*), YP(*)
    REAL(KIND=8) :: eta, flind, bb, gg
    eta = T
    bb = ode_beta
    gg = ode_gamma
    YP(1) = Y(2)
    IF (eta == 1.d0) THEN
      flind = 2.d0
    ELSE IF (eta == 0.0d0) THEN
      flind = 1.d0
    ELSE
      flind = 1.d0/(1.d0/2.d0+(1.d0-eta**2)/(4.d0*eta)*LOG(ABS((1.d0+eta)/(1.d0-eta))))
    END IF
    YP(2) = 20.d0 * (flind - 3.d0*eta**2 - 1.d0) - &
            (gg + 1.d0 - 6.d0*5.d0/3.d0)*eta*Y(2) - &
            36.d0 * (5.d0/3.d0 - bb) * bb * Y(1)
    YP(2) = YP(2)/eta**2
  END SUBROUTINE wgc_rhs
END MODULE ref_wgc_ode


SUBROUTINE ref_spline_set(n, t, y, ibcbeg, ybcbeg, ibcend, ybcend, ypp) BIND(C, name="ref_spline_set")
  USE ISO_C_BINDING
  USE MathSplines, ONLY: spline_cubic_set
  IMPLICIT NONE
  INTEGER(C_INT), VALUE :: n, ibcbeg, ibcend
  REAL(C_DOUBLE), VALUE :: ybcbeg, ybcend
  REAL(C_DOUBLE) :: t(n), y(n), ypp(n)
  CALL spline_cubic_set(INT(n), t, y, INT(ibcbeg), ybcbeg, INT(ibcend), ybcend, ypp)
END SUBROUTINE ref_spline_set


SUBROUTINE ref_spline_val(n, t, y, ypp, m, tval, yval) BIND(C, name="ref_spline_val")
  USE ISO_C_BINDING
  USE MathSplines, ONLY: spline_cubic_val
  IMPLICIT NONE
  INTEGER(C_INT), VALUE :: n, m
  REAL(C_DOUBLE) :: t(n), y(n), ypp(n), tval(m), yval(m)
  REAL(C_DOUBLE) :: yp, ypp2
  INTEGER :: i
  DO i = 1, m
    CALL spline_cubic_val(INT(n), t, y, ypp, tval(i), yval(i), yp, ypp2)
  END DO
END SUBROUTINE ref_spline_val


! task codes: 0 START, 1 FG, 2 CONVERGENCE, 3 WARNING, 4 ERROR
SUBROUTINE ref_dcsrch(stp, f, g, ftol, gtol, xtol, taskcode, stpmin, stpmax, isave, dsave) &
    BIND(C, name="ref_dcsrch")
  USE ISO_C_BINDING
  IMPLICIT NONE
  REAL(C_DOUBLE) :: stp
  REAL(C_DOUBLE), VALUE :: f, g, ftol, gtol, xtol, stpmin, stpmax
  INTEGER(C_INT) :: taskcode
  INTEGER(C_INT) :: isave(2)
  REAL(C_DOUBLE) :: dsave(13)
  CHARACTER(LEN=60) :: task
  INTEGER :: is(2)
  EXTERNAL dcsrch
  SELECT CASE (taskcode)
    CASE (0)
      task = 'START'
    CASE DEFAULT
      task = 'FG'
  END SELECT
  is = isave
  CALL dcsrch(stp, f, g, ftol, gtol, xtol, task, stpmin, stpmax, is, dsave)
  isave = is
  IF (task(1:2) == 'FG') THEN
    taskcode = 1
  ELSE IF (task(1:4) == 'CONV') THEN
    taskcode = 2
  ELSE IF (task(1:4) == 'WARN') THEN
    taskcode = 3
  ELSE
    taskcode = 4
  END IF
END SUBROUTINE ref_dcsrch


! Restatement of makeKernel2 (KEDF_KernelODE.f90:254-478) driving the
! reference's RKSUITE (SETUP_RK/UT_RK, MathRKsuite.f:42,465): knots
! eta in [1e-3,3] step 1e-3 then [3.001,100] step 1e-2; start at eta=100 with
! w = wInf, w' = 0 and integrate toward 0.9*eta(1), TOL 5e-5, METHOD 2.
! Outputs w, w' (=Y(2)), w'' (=YP(2)) at every knot, zeros at the last knot.
! Returns the number of knots written, or -L on an RKSUITE failure at knot L.
FUNCTION ref_wgc_eta_table(alpha, beta, gamma, nmax, eta_out, w_out, w1_out, w2_out) &
    BIND(C, name="ref_wgc_eta_table") RESULT(nout)
  USE ISO_C_BINDING
  USE ref_wgc_ode
  IMPLICIT NONE
  REAL(C_DOUBLE), VALUE :: alpha, beta, gamma
  INTEGER(C_INT), VALUE :: nmax
  REAL(C_DOUBLE) :: eta_out(nmax), w_out(nmax), w1_out(nmax), w2_out(nmax)
  INTEGER(C_INT) :: nout
  INTEGER, PARAMETER :: NEQ = 2, LENWRK = 32*NEQ, METHOD = 2
  REAL(KIND=8) :: HSTART, T, TEND, TOL, TSTART, TWANT, winf
  INTEGER :: L, UFLAG, i, npoints(2), ntot
  LOGICAL :: ERRASS, MESAGE
  REAL(KIND=8) :: THRES(NEQ), WORK(LENWRK), Y(NEQ), YMAX(NEQ), YP(NEQ), YSTART(NEQ)
  REAL(KIND=8) :: t_start(2), t_end(2), t_step(2)
  EXTERNAL SETUP_RK, UT_RK

  ode_beta = beta
  ode_gamma = gamma
  winf = -1.6d0 * 20.d0/(36.d0*alpha*beta)
  t_start(1) = 1.d-3
  t_end(1) = 3.d0
  t_step(1) = 1d-3
  t_start(2) = t_end(1) + t_step(1)
  t_end(2) = 100.d0
  t_step(2) = t_step(1)*10.d0
  npoints(:) = FLOOR((t_end(:)-t_start(:))/t_step(:)) - 1
  ntot = SUM(npoints)
  IF (ntot > nmax) THEN
    nout = 0
    RETURN
  END IF
  DO i = 1, npoints(1)
    eta_out(i) = t_start(1) + REAL(i-1,kind=8)*t_step(1)
  END DO
  DO i = npoints(1)+1, ntot
    eta_out(i) = t_start(2) + REAL(i-npoints(1)-1,kind=8)*t_step(2)
  END DO
  w_out(1:ntot) = -500.d0
  w1_out(1:ntot) = -500.d0
  w2_out(1:ntot) = -500.d0
  TSTART = eta_out(ntot)
  YSTART(1) = winf
  YSTART(2) = 0.0d0
  TEND = eta_out(1)*0.9d0
  TOL = 5.0D-5
  THRES = 1.0D-10
  MESAGE = .FALSE.
  ERRASS = .FALSE.
  HSTART = 0.d0
  CALL SETUP_RK(NEQ, TSTART, YSTART, TEND, TOL, THRES, METHOD, 'Usual Task', &
                ERRASS, HSTART, WORK, LENWRK, MESAGE)
  w_out(ntot) = 0.d0
  w1_out(ntot) = 0.d0
  w2_out(ntot) = 0.d0
  nout = ntot
  DO L = ntot-1, 1, -1
    TWANT = eta_out(L)
    DO
      CALL UT_RK(wgc_rhs, TWANT, T, Y, YP, YMAX, WORK, UFLAG)
      IF (UFLAG /= 2 .AND. UFLAG /= 3) EXIT
    END DO
    IF (UFLAG == 1) THEN
      w_out(L) = Y(1)
      w1_out(L) = Y(2)
      w2_out(L) = YP(2)
    ELSE
      nout = -L
      RETURN
    END IF
  END DO
END FUNCTION ref_wgc_eta_table


SUBROUTINE ref_lnsrch(alp, f, g, task_c, c1, c2, xtol, stpmin, stpmax, points_out) BIND(C, name="ref_lnsrch")
  USE ISO_C_BINDING
  USE LnSrch_MOD, ONLY: LnSrch, points
  IMPLICIT NONE
  REAL(C_DOUBLE) :: alp
  REAL(C_DOUBLE), VALUE :: f, g, c1, c2, xtol, stpmin, stpmax
  CHARACTER(KIND=C_CHAR) :: task_c(70)
  REAL(C_DOUBLE) :: points_out(6)
  CHARACTER(LEN=70) :: task
  REAL(KIND=8) :: a1, a2
  INTEGER :: i
  DO i = 1, 70
    task(i:i) = task_c(i)
  END DO
  a1 = c1
  a2 = c2
  CALL LnSrch(alp, f, g, task, a1, a2, xtol, stpmin, stpmax)
  DO i = 1, 70
    task_c(i) = task(i:i)
  END DO
  points_out(1:6) = RESHAPE(points, [6])
END SUBROUTINE ref_lnsrch
