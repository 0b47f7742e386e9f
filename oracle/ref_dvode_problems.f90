! ORACLE / TEST INFRASTRUCTURE ONLY.  Never linked into the product library.
!
! bind(C) entry that runs the reference's own DVODE_F90 (src/fortran_src/dvode.f90, compiled in
! place by oracle/Makefile) with the options UCLCHEM passes it (METHOD_FLAG=22, internal
! finite-difference Jacobian, vector absolute tolerance, MXSTEP; chemistry.f90:250-253) on
! standard stiff test problems with published answers: linear decay, Robertson's kinetics
! and HIRES.  The tests compare the solver's BDF core with the reference integrator there.
MODULE ucref_dvode_problems
    USE iso_c_binding
    USE DVODE_F90_M, only: VODE_OPTS, SET_OPTS, DVODE_F90, GET_STATS
    IMPLICIT NONE
    INTEGER :: which = 0
CONTAINS

    SUBROUTINE fprob(NEQ, T, Y, YDOT)
        INTEGER, PARAMETER :: WP = KIND(1.0D0)
        INTEGER NEQ
        REAL(WP) T
        REAL(WP), DIMENSION(NEQ) :: Y, YDOT
        INTENT(IN) :: NEQ, T, Y
        INTENT(OUT) :: YDOT
        SELECT CASE (which)
        CASE (0)
            YDOT(1) = -1.0d0 * Y(1)
            YDOT(2) = -1.0d3 * Y(2)
            YDOT(3) = -1.0d6 * Y(3)
        CASE (1)
            YDOT(1) = -0.04d0 * Y(1) + 1.0d4 * Y(2) * Y(3)
            YDOT(3) = 3.0d7 * Y(2) * Y(2)
            YDOT(2) = -YDOT(1) - YDOT(3)
        CASE DEFAULT
            YDOT(1) = -1.71d0 * Y(1) + 0.43d0 * Y(2) + 8.32d0 * Y(3) + 0.0007d0
            YDOT(2) = 1.71d0 * Y(1) - 8.75d0 * Y(2)
            YDOT(3) = -10.03d0 * Y(3) + 0.43d0 * Y(4) + 0.035d0 * Y(5)
            YDOT(4) = 8.32d0 * Y(2) + 1.71d0 * Y(3) - 1.12d0 * Y(4)
            YDOT(5) = -1.745d0 * Y(5) + 0.43d0 * Y(6) + 0.43d0 * Y(7)
            YDOT(6) = -280.0d0 * Y(6) * Y(8) + 0.69d0 * Y(4) + 1.71d0 * Y(5) - 0.43d0 * Y(6) + 0.69d0 * Y(7)
            YDOT(7) = 280.0d0 * Y(6) * Y(8) - 1.81d0 * Y(7)
            YDOT(8) = -YDOT(7)
        END SELECT
    END SUBROUTINE fprob

    ! problem: 0 decay (n=3), 1 Robertson (n=3), 2 HIRES (n=8).  One cold-start call
    ! (ITASK=1, ISTATE=1) from t0 to tout; y in/out; istate = DVODE's ISTATE on return;
    ! stats = NST, NFE, NJE.
    SUBROUTINE ucref_dvode_problem(problem, t0, tout, rtol, atol, y, n, mxstep, istate, stats) &
            & bind(C, name="ucref_dvode_problem")
        INTEGER(c_int), VALUE :: problem, n, mxstep
        REAL(c_double), VALUE :: t0, tout, rtol
        REAL(c_double), INTENT(IN) :: atol(n)
        REAL(c_double), INTENT(INOUT) :: y(n)
        INTEGER(c_int), INTENT(OUT) :: istate
        INTEGER(c_int), INTENT(OUT) :: stats(3)
        TYPE(VODE_OPTS) :: OPTIONS
        DOUBLE PRECISION :: t, tt, abstol(n), yy(n), rstats(22)
        INTEGER :: itask, ist, istats(31)
        which = problem
        abstol = atol
        yy = y
        t = t0
        tt = tout
        itask = 1
        ist = 1
        OPTIONS = SET_OPTS(METHOD_FLAG=22, ABSERR_VECTOR=abstol, RELERR=rtol, USER_SUPPLIED_JACOBIAN=.FALSE., &
            & MXSTEP=mxstep)
        CALL DVODE_F90(fprob, n, yy, t, tt, itask, ist, OPTIONS)
        CALL GET_STATS(rstats, istats)
        y = yy
        istate = ist
        stats(1) = istats(11)
        stats(2) = istats(12)
        stats(3) = istats(13)
    END SUBROUTINE ucref_dvode_problem

END MODULE ucref_dvode_problems
