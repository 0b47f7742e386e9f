! ORACLE / TEST INFRASTRUCTURE ONLY.  Never linked into the product library.
!
! bind(C) driver over the UCLCHEM reference Fortran hot path, compiled in place
! from the reference sources by oracle/Makefile into oracle/_ref/libucref.so.
! It restates only the glue the reference keeps in files that do not compile
! with amdflang (wrap.f90, hotcore.f90):
!   * the solveAbundances time loop for points=1   (src/fortran_src/wrap.f90:598-690)
!   * hotcore updateTargetTime / updatePhysics      (src/fortran_src/hotcore.f90:56-90)
!   * the choice of physics module per point: cloud, hot core, or the reference's own
!     collapse module (collapse.f90, compiled in place) for model 2
!   * a per-point reset of every module global the parameter parser can touch,
!     so each grid point starts from the documented defaults
!     (src/fortran_src/defaultparameters.f90:18-122; wrap.f90:699-983 never resets).
! Everything numeric (rates, RHS, DVODE, chemistry driver, physics core, cloud
! schedule) is the reference's own compiled code.
MODULE ucref_driver
    USE iso_c_binding
    USE constants
    USE defaultparameters
    USE physicscore, only: points, dstep, cloudsize, h2crprate, &
        & currentTime, targetTime, timeinyears, currentTimeOld, density, gastemp, dusttemp, av, coldens, &
        & coreInitializePhysics, coreUpdatePhysics
    USE network
    USE f2py_constants, only: nspec, nreac
    USE chemistry
    USE rates, only: calculateReactionRates, lastTemp
    USE surfacereactions, only: safeMantle, safeBulk, bulkLayersReciprocal, NUM_SITES_PER_GRAIN, &
        & GAS_DUST_DENSITY_RATIO, bulkGainFromMantleBuildUp
    USE postprocess_mod, only: lusecoldens, usepostprocess
    USE cloud_mod, only: cloudInitializePhysics => initializePhysics, cloudUpdateTargetTime => updateTargetTime
    USE collapse_mod, only: collapse_mode, writePhysics, collapseInitializePhysics => initializePhysics, &
        & collapseUpdateTargetTime => updateTargetTime, collapseUpdatePhysics => updatePhysics
    USE DVODE_F90_M, only: GET_STATS, VODE_OPTS, SET_OPTS, DVODE_F90
    IMPLICIT NONE

    ! hotcore.f90:20-22 tables (restated: hotcore.f90 itself does not compile with amdflang)
    INTEGER, PARAMETER :: nMasses = 6
    REAL(dp), PARAMETER :: tempa(nMasses)=(/1.927d-1,4.8560d-2,7.8470d-3,9.6966d-4,1.706d-4,4.74d-7/)
    REAL(dp), PARAMETER :: tempb(nMasses)=(/0.5339,0.6255,0.8395,1.085,1.289,1.98/)
    INTEGER :: modelKind = 0, tempIndx = 1
    REAL(dp) :: maxTemp = 300.0d0
    ! given start abundances of the next model (wrap.f90:636-640, abundanceStart)
    LOGICAL :: have_start = .false.
    REAL(dp), ALLOCATABLE :: start_abund(:)
    ! rate coefficients per output row (returnRates, io.f90:90-93), while ucref_model_rates runs
    LOGICAL :: want_rates = .false.
    REAL(dp), ALLOCATABLE :: rates_buf(:,:)

    ! parameter vector layout; mirrored by oracle/oracle.py and uclchem_amd/params.py
    INTEGER, PARAMETER :: P_DENS=1, P_TEMP=2, P_ZETA=3, P_RADFIELD=4, P_BASEAV=5, P_ROUT=6, P_RIN=7, &
        & P_FINALTIME=8, P_FREEFALL=9, P_FINALDENS=10, P_RELTOL=11, P_ABSFAC=12, P_ABSMIN=13, &
        & P_MODEL=14, P_TEMPINDX=15, P_MAXTEMP=16, P_FREEZEFAC=17, P_ENDATFINALDENS=18, P_FREEFALLFAC=19, &
        & P_MXSTEP=20, P_MAXSTEPS=21, P_METALLICITY=22, P_ION=23, P_FH=24, P_FHE=25, P_FC=26, P_FO=27, &
        & P_FN=28, P_FS=29, P_FMG=30, P_FSI=31, P_FCL=32, P_FP=33, P_FFE=34, P_FF=35, P_FD=36, P_FLI=37, &
        & P_FNA=38, P_FPAH=39, P_F15N=40, P_F13C=41, P_F18O=42, P_EBMAXH2=43, P_EBMAXCR=44, P_EBMAXUVCR=45, &
        & P_EPSILON=46, P_UV_YIELD=47, P_PHI=48, P_UVCREFF=49, P_OMEGA=50, P_DESORB=51, P_H2DESORB=52, &
        & P_CRDESORB=53, P_UVDESORB=54, P_THERMDESORB=55, P_CRATTEN=56, P_IONMODEL=57, P_IMPROVEDH2CRP=58, &
        & P_ENFORCECHARGE=59, P_COLLAPSE=60, P_NPARAM=64
CONTAINS

    SUBROUTINE reset_globals()
        ! defaults exactly as declared in defaultparameters.f90:18-122
        initialTemp=10.0; initialDens=1.00d2; finalDens=1.00d5; currentTime=0.0; finalTime=5.0d6
        radfield=1.0; zeta=1.0; rout=0.05; rin=0.0; baseAv=2.0; points=1; bm0=1.0
        freezeFactor=1.0; endAtFinalDensity=.False.; freefall=.False.; freefallFactor=1.0
        desorb=.True.; h2desorb=.True.; crdesorb=.True.; uvdesorb=.True.; thermdesorb=.True.
        instantSublimation=.False.; cosmicRayAttenuation=.False.; ionModel='L'
        improvedH2CRPDissociation=.False.; enforceChargeConservation=.false.
        metallicity=1.0; ion=2; fh=0.5; fhe=0.1; fc=1.77d-04; fo=3.34d-04; fn=6.18d-05; fs=3.51d-6
        fmg=2.256d-06; fsi=1.78d-06; fcl=3.39d-08; fp=7.78d-08; ffe=2.01d-7; ff=3.6d-08; fd=0.0
        fli=0.0; fna=0.0; fpah=0.0; f15n=0.0; f13c=0.0; f18O=0.0
        reltol=1d-8; abstol_factor=1.0d-14; abstol_min=1.0d-25; MXSTEP=10000
        ebmaxh2=1.21d3; ebmaxcr=1.21d3; ebmaxuvcr=1.0d4; epsilon=0.01; uv_yield=0.03; phi=1.0d5
        uvcreff=1.0d-3; omega=0.5
        usepostprocess=.true.; lusecoldens=.false.
        modelKind=0; tempIndx=1; maxTemp=300.0d0
        collapse_mode=1; writePhysics=.false.
        h2CRPRate=0.0d0   ! module state the physics update sets; each point starts from zero
    END SUBROUTINE reset_globals

    SUBROUTINE apply_params(p)
        REAL(c_double), INTENT(IN) :: p(P_NPARAM)
        initialDens=p(P_DENS); initialTemp=p(P_TEMP); zeta=p(P_ZETA); radfield=p(P_RADFIELD)
        baseAv=p(P_BASEAV); rout=p(P_ROUT); rin=p(P_RIN); finalTime=p(P_FINALTIME)
        freefall=(p(P_FREEFALL) .ne. 0.0d0); finalDens=p(P_FINALDENS)
        reltol=p(P_RELTOL); abstol_factor=p(P_ABSFAC); abstol_min=p(P_ABSMIN)
        modelKind=NINT(p(P_MODEL)); tempIndx=NINT(p(P_TEMPINDX)); maxTemp=p(P_MAXTEMP)
        collapse_mode=NINT(p(P_COLLAPSE))
        freezeFactor=p(P_FREEZEFAC); endAtFinalDensity=(p(P_ENDATFINALDENS) .ne. 0.0d0)
        freefallFactor=p(P_FREEFALLFAC); MXSTEP=NINT(p(P_MXSTEP))
        metallicity=p(P_METALLICITY); ion=NINT(p(P_ION), dp); fh=p(P_FH); fhe=p(P_FHE); fc=p(P_FC); fo=p(P_FO)
        fn=p(P_FN); fs=p(P_FS); fmg=p(P_FMG); fsi=p(P_FSI); fcl=p(P_FCL); fp=p(P_FP); ffe=p(P_FFE); ff=p(P_FF)
        fd=p(P_FD); fli=p(P_FLI); fna=p(P_FNA); fpah=p(P_FPAH); f15n=p(P_F15N); f13c=p(P_F13C); f18O=p(P_F18O)
        ebmaxh2=p(P_EBMAXH2); ebmaxcr=p(P_EBMAXCR); ebmaxuvcr=p(P_EBMAXUVCR); epsilon=p(P_EPSILON)
        uv_yield=p(P_UV_YIELD); phi=p(P_PHI); uvcreff=p(P_UVCREFF); omega=p(P_OMEGA)
        desorb=(p(P_DESORB) .ne. 0.0d0); h2desorb=(p(P_H2DESORB) .ne. 0.0d0); crdesorb=(p(P_CRDESORB) .ne. 0.0d0)
        uvdesorb=(p(P_UVDESORB) .ne. 0.0d0); thermdesorb=(p(P_THERMDESORB) .ne. 0.0d0)
        cosmicRayAttenuation=(p(P_CRATTEN) .ne. 0.0d0)
        IF (NINT(p(P_IONMODEL)) .eq. 1) THEN
            ionModel='H'
        ELSE
            ionModel='L'
        END IF
        improvedH2CRPDissociation=(p(P_IMPROVEDH2CRP) .ne. 0.0d0)
        enforceChargeConservation=(p(P_ENFORCECHARGE) .ne. 0.0d0)
    END SUBROUTINE apply_params

    SUBROUTINE hotcoreUpdateTargetTime
        ! hotcore.f90:56-72
        IF (timeInYears .gt. 1.0d6) THEN
            targetTime=(timeInYears+1.0d5)*SECONDS_PER_YEAR
        ELSE  IF (timeInYears .gt. 1.0d5) THEN
            targetTime=(timeInYears+1.0d4)*SECONDS_PER_YEAR
        ELSE IF (timeInYears .gt. 1.0d4) THEN
            targetTime=(timeInYears+1000.0)*SECONDS_PER_YEAR
        ELSE IF (timeInYears .gt. 1000) THEN
            targetTime=(timeInYears+100.0)*SECONDS_PER_YEAR
        ELSE IF (timeInYears .gt. 100) THEN
            targetTime=(timeInYears+10.0)*SECONDS_PER_YEAR
        ELSE IF (timeInYears .gt. 0.0) THEN
            targetTime=(timeInYears*10.0)*SECONDS_PER_YEAR
        ELSE
            targetTime=SECONDS_PER_YEAR*1.0d-7
        ENDIF
    END SUBROUTINE hotcoreUpdateTargetTime

    SUBROUTINE hotcoreUpdatePhysics
        ! hotcore.f90:78-90
        IF (gasTemp(dstep) .lt. maxTemp) THEN
            gasTemp(dstep)=(cloudSize/(rout*pc))*(real(dstep)/real(points))
            gasTemp(dstep)=gasTemp(dstep)**(-0.5)
            gasTemp(dstep)=initialTemp + ((tempa(tempindx)*(currentTime/SECONDS_PER_YEAR)**tempb(tempindx))*gasTemp(dstep))
            if (gasTemp(dstep) .gt. maxTemp) gasTemp(dstep)=maxTemp
        END IF
        dustTemp=gasTemp
    END SUBROUTINE hotcoreUpdatePhysics

    SUBROUTINE setup_point(p, flag)
        REAL(c_double), INTENT(IN) :: p(P_NPARAM)
        INTEGER, INTENT(OUT) :: flag
        CALL reset_globals()
        CALL apply_params(p)
        dstep=1
        currentTime=0.0
        timeInYears=0.0
        flag=0
        CALL coreInitializePhysics(flag)
        IF (flag .lt. 0) RETURN
        IF (modelKind .eq. 2) THEN
            CALL collapseInitializePhysics(flag)   ! reads flag before setting it: 0 here, as in wrap.f90
        ELSE
            CALL cloudInitializePhysics(flag)   ! hotcore's initializePhysics does the same for three-phase networks
        END IF
        IF (flag .lt. 0) RETURN
        CALL initializeChemistry(.false.)
        dstep=1   ! coreInitializePhysics leaves its loop counter at points+1 (wrap.f90:643 resets it too)
    END SUBROUTINE setup_point

    ! updateChemistry (chemistry.f90:145-239) for one point (points=1, no post-processed
    ! column densities, no refractory species: the default network's refractoryList is
    ! (/-999/), network.f90:262) restated with one addition: before every integrator call
    ! the steps taken so far by the point (nsteps, all attempts of all intervals, from
    ! GET_STATS NST) are compared with a total-step bound, and the point stops with
    ! INT_TOO_MANY_FAILS_ERROR once it is reached.  The rates, the integrator call and its
    ! retry handling are the reference's own (calculateReactionRates, integrateODESystem).
    ! With usepostprocess (its default, postprocess.f90:14) the reference never raises its
    ! failure counters, so they are not restated.
    SUBROUTINE bounded_update_chemistry(maxsteps, nsteps, flag, attempts, work)
        REAL(dp), INTENT(IN) :: maxsteps
        REAL(dp), INTENT(INOUT) :: nsteps
        INTEGER, INTENT(OUT) :: flag
        INTEGER(c_int), INTENT(OUT) :: attempts
        INTEGER(c_int), INTENT(INOUT), OPTIONAL :: work(6)   ! += NFE NJE NLU NNI NCFN NETF of every attempt
        REAL(dp) :: originalTargetTime, surfCov, rstats(22)
        INTEGER :: loops, istats(31)
        loops = 0
        attempts = 0
        flag = 0
        originalTargetTime = targetTime
        DO WHILE ((currentTime .lt. targetTime) .and. (loops .lt. 10))
            IF (nsteps >= maxsteps) THEN
                flag = INT_TOO_MANY_FAILS_ERROR
                RETURN
            END IF
            IF (.not. freefall) abund(nspec+1,dstep)=density(dstep)
            h2ColToCell=0.0
            coColToCell=0.0
            cColToCell=0.0
            h2Col=h2ColToCell+0.5*abund(nh2,dstep)*density(dstep)*(cloudSize/real(points))
            coCol=coColToCell+0.5*abund(nco,dstep)*density(dstep)*(cloudSize/real(points))
            cCol=cColToCell+0.5*abund(nc,dstep)*density(dstep)*(cloudSize/real(points))
            abund(nBulk,dstep)=sum(abund(bulkList,dstep))
            abund(nSurface,dstep)=sum(abund(surfaceList,dstep))
            safeMantle=MAX(1d-30,abund(nSurface,dstep))
            safeBulk=MAX(1d-30,abund(nBulk,dstep))
            bulkLayersReciprocal=MIN(1.0,NUM_SITES_PER_GRAIN/(GAS_DUST_DENSITY_RATIO*safeBulk))
            surfCov=bulkGainFromMantleBuildUp()
            CALL calculateReactionRates(abund,safeMantle, h2col, cocol, ccol, rate)
            CALL integrateODESystem(flag)
            CALL GET_STATS(rstats, istats)
            nsteps = nsteps + real(istats(11), dp)
            IF (PRESENT(work)) work = work + (/istats(12), istats(13), istats(19), istats(20), istats(21), istats(22)/)
            IF (flag .lt. 0) RETURN
            WHERE(abund<MIN_ABUND) abund=MIN_ABUND
            density(dstep)=abund(NEQ,dstep)
            loops = loops + 1
            attempts = loops
            targetTime = originalTargetTime
        END DO
    END SUBROUTINE bounded_update_chemistry

    ! ---------------------------------------------------------------- C entry points

    INTEGER(c_int) FUNCTION ucref_sizes(neq, nr) BIND(C, name="ucref_sizes")
        INTEGER(c_int), INTENT(OUT) :: neq, nr
        neq = nspec+1
        nr = nreac
        ucref_sizes = 0
    END FUNCTION ucref_sizes

    ! Set up one point (parameters, physics, elemental abundances); returns y0.
    INTEGER(c_int) FUNCTION ucref_init(p, y0) BIND(C, name="ucref_init")
        REAL(c_double), INTENT(IN) :: p(P_NPARAM)
        REAL(c_double), INTENT(OUT) :: y0(nspec+1)
        INTEGER :: flag
        CALL setup_point(p, flag)
        y0 = abund(:,1)
        ucref_init = flag
    END FUNCTION ucref_init

    ! The pre-integration part of updateChemistry (chemistry.f90:169-203) on the given
    ! state, then calculateReactionRates.  Module state (phi, lastTemp) persists
    ! between calls exactly as in the reference.  y is updated with the BULK/SURFACE reset.
    INTEGER(c_int) FUNCTION ucref_rates(y, rate_out, scal) BIND(C, name="ucref_rates")
        REAL(c_double), INTENT(INOUT) :: y(nspec+1)
        REAL(c_double), INTENT(OUT) :: rate_out(nreac), scal(8)
        REAL(dp) :: sc
        abund(:,1) = y
        IF (.not. freefall) abund(nspec+1,dstep)=density(dstep)
        h2ColToCell=0.0; coColToCell=0.0; cColToCell=0.0
        h2Col=h2ColToCell+0.5*abund(nh2,dstep)*density(dstep)*(cloudSize/real(points))
        coCol=coColToCell+0.5*abund(nco,dstep)*density(dstep)*(cloudSize/real(points))
        cCol=cColToCell+0.5*abund(nc,dstep)*density(dstep)*(cloudSize/real(points))
        abund(nBulk,dstep)=sum(abund(bulkList,dstep))
        abund(nSurface,dstep)=sum(abund(surfaceList,dstep))
        safeMantle=MAX(1d-30,abund(nSurface,dstep))
        safeBulk=MAX(1d-30,abund(nBulk,dstep))
        bulkLayersReciprocal=MIN(1.0,NUM_SITES_PER_GRAIN/(GAS_DUST_DENSITY_RATIO*safeBulk))
        sc=bulkGainFromMantleBuildUp()
        CALL calculateReactionRates(abund,safeMantle, h2col, cocol, ccol, rate)
        rate_out = rate
        y = abund(:,1)
        scal(1)=h2col; scal(2)=cocol; scal(3)=ccol; scal(4)=safeMantle; scal(5)=safeBulk
        scal(6)=bulkLayersReciprocal; scal(7)=phi; scal(8)=av(1)
        ucref_rates = 0
    END FUNCTION ucref_rates

    ! RHS F(y) with the module's current rates (chemistry.f90:294-352).
    INTEGER(c_int) FUNCTION ucref_rhs(y, ydot) BIND(C, name="ucref_rhs")
        REAL(c_double), INTENT(IN) :: y(nspec+1)
        REAL(c_double), INTENT(OUT) :: ydot(nspec+1)
        REAL(dp) :: t
        t = currentTime
        CALL F(nspec+1, t, y, ydot)
        ucref_rhs = 0
    END FUNCTION ucref_rhs

    ! One DVODE_F90 call of the chemistry (integrateODESystem, chemistry.f90:241-292) on the
    ! given state from t0 to t1 seconds, with the module's current rates (ucref_rates) and
    ! tolerances: y in/out, istate_out = DVODE's ISTATE, stats = NST, NFE, NJE, NQU (order of
    ! the last step), NNI (Newton iterations), NCFN (convergence failures) of the call.
    INTEGER(c_int) FUNCTION ucref_integrate(y, t0, t1, istate_out, stats) BIND(C, name="ucref_integrate")
        REAL(c_double), INTENT(INOUT) :: y(nspec+1)
        REAL(c_double), VALUE :: t0, t1
        INTEGER(c_int), INTENT(OUT) :: istate_out
        INTEGER(c_int), INTENT(OUT) :: stats(6)
        INTEGER :: flag
        REAL(dp) :: rstats(22)
        INTEGER :: istats(31)
        dstep = 1
        abund(:,1) = y
        currentTime = t0
        targetTime = t1
        CALL integrateODESystem(flag)
        y = abund(:,1)
        istate_out = ISTATE
        CALL GET_STATS(rstats, istats)
        stats(1) = istats(11)
        stats(2) = istats(12)
        stats(3) = istats(13)
        stats(4) = istats(14)
        stats(5) = istats(20)
        stats(6) = istats(21)
        ucref_integrate = flag
    END FUNCTION ucref_integrate

    ! DVODE_F90 on the chemistry one step at a time (ITASK=2) from t0 until t1 is passed, with
    ! the tolerances integrateODESystem sets (chemistry.f90:248-253): per step the time
    ! reached, the step size used and its order (GET_STATS HU, NQU).  Returns the step count.
    INTEGER(c_int) FUNCTION ucref_step_trace(y, t0, t1, maxsteps, tt, hh, qq) BIND(C, name="ucref_step_trace")
        REAL(c_double), INTENT(INOUT) :: y(nspec+1)
        REAL(c_double), VALUE :: t0, t1
        INTEGER(c_int), VALUE :: maxsteps
        REAL(c_double), INTENT(OUT) :: tt(maxsteps), hh(maxsteps)
        INTEGER(c_int), INTENT(OUT) :: qq(maxsteps)
        TYPE(VODE_OPTS) :: OPTS
        REAL(dp) :: t, tout, rstats(22)
        INTEGER :: istats(31), itask, ist, k
        dstep = 1
        abund(:,1) = y
        abstol = abstol_factor*abund(:,1)
        WHERE(abstol < abstol_min) abstol = abstol_min
        OPTS = SET_OPTS(METHOD_FLAG=22, ABSERR_VECTOR=abstol, RELERR=reltol, USER_SUPPLIED_JACOBIAN=.FALSE., &
            & MXSTEP=MXSTEP)
        t = t0
        tout = t1
        itask = 2
        ist = 1
        k = 0
        DO WHILE (t < t1 .AND. k < maxsteps)
            CALL DVODE_F90(F, NEQ, abund(:,1), t, tout, itask, ist, OPTS)
            IF (ist < 0) EXIT
            CALL GET_STATS(rstats, istats)
            k = k + 1
            tt(k) = t
            hh(k) = rstats(11)
            qq(k) = istats(14)
        END DO
        y = abund(:,1)
        ucref_step_trace = k
    END FUNCTION ucref_step_trace

    ! The physics update the solveAbundances loop makes after an interval ends at t_s seconds
    ! (wrap.f90:672-676: coreUpdatePhysics, then the hot-core temperature for model 1).
    INTEGER(c_int) FUNCTION ucref_update_physics(t_s, temp_out) BIND(C, name="ucref_update_physics")
        REAL(c_double), VALUE :: t_s
        REAL(c_double), INTENT(OUT) :: temp_out
        dstep = 1
        currentTime = t_s
        timeInYears = t_s/SECONDS_PER_YEAR
        CALL coreUpdatePhysics
        IF (modelKind .eq. 1) CALL hotcoreUpdatePhysics
        temp_out = gasTemp(1)
        ucref_update_physics = 0
    END FUNCTION ucref_update_physics

    ! Overwrite the module rate array (so RHS tests can use arbitrary rates).
    INTEGER(c_int) FUNCTION ucref_set_rates(rate_in) BIND(C, name="ucref_set_rates")
        REAL(c_double), INTENT(IN) :: rate_in(nreac)
        rate = rate_in
        ucref_set_rates = 0
    END FUNCTION ucref_set_rates

    INTEGER(c_int) FUNCTION ucref_set_phi(v) BIND(C, name="ucref_set_phi")
        REAL(c_double), VALUE :: v
        phi = v
        ucref_set_phi = 0
    END FUNCTION ucref_set_phi

    ! Full single-point model (solveAbundances loop, wrap.f90:598-690, points=1).
    ! rec(:, k) = (time_yr, density, gasTemp, av, y(1:nspec)) for each output row k,
    ! row 1 being t=0.  stats = (NST, NFE, NJE summed over the last integrator call of each
    ! interval, n_intervals, steps of every call when p(P_MAXSTEPS) > 0, 0, 0, 0).
    ! p(P_MAXSTEPS) > 0 bounds the point's total steps (bounded_update_chemistry).
    INTEGER(c_int) FUNCTION ucref_model(p, maxrec, rec, nrec, stats) BIND(C, name="ucref_model")
        REAL(c_double), INTENT(IN) :: p(P_NPARAM)
        INTEGER(c_int), VALUE :: maxrec
        REAL(c_double), INTENT(OUT) :: rec(4+nspec, maxrec)
        INTEGER(c_int), INTENT(OUT) :: nrec
        INTEGER(c_int), INTENT(OUT) :: stats(8)
        ucref_model = run_model(p, HUGE(1.0d0), maxrec, rec, nrec, stats)
    END FUNCTION ucref_model

    ! ucref_model that also returns the rate array as each row is written: the reference's
    ! ratesarray(dtime, dstep, :) = rate(:nreac) (io.f90:90-93).  Row 1 holds whatever the
    ! module's rate array held before the first interval.
    INTEGER(c_int) FUNCTION ucref_model_rates(p, maxrec, rec, nrec, stats, rates) BIND(C, name="ucref_model_rates")
        REAL(c_double), INTENT(IN) :: p(P_NPARAM)
        INTEGER(c_int), VALUE :: maxrec
        REAL(c_double), INTENT(OUT) :: rec(4+nspec, maxrec), rates(nreac, maxrec)
        INTEGER(c_int), INTENT(OUT) :: nrec
        INTEGER(c_int), INTENT(OUT) :: stats(8)
        IF (ALLOCATED(rates_buf)) DEALLOCATE(rates_buf)
        ALLOCATE(rates_buf(nreac, maxrec))
        rates_buf = 0.0d0
        want_rates = .true.
        ucref_model_rates = run_model(p, HUGE(1.0d0), maxrec, rec, nrec, stats)
        want_rates = .false.
        rates = rates_buf
    END FUNCTION ucref_model_rates

    ! A reaction coefficient replaced as the alpha/beta/gamma dictionary keys do
    ! (wrap.f90:744-761, coefficientParser :985-1019): which 0 alpha, 1 beta, 2 gama; k 1-based.
    ! Like the reference's, the change stays in the module arrays until set back.
    INTEGER(c_int) FUNCTION ucref_set_coefficient(which, k, v) BIND(C, name="ucref_set_coefficient")
        INTEGER(c_int), VALUE :: which, k
        REAL(c_double), VALUE :: v
        ucref_set_coefficient = -1
        IF (k < 1 .or. k > nreac) RETURN
        SELECT CASE (which)
            CASE (0)
                alpha(k) = v
            CASE (1)
                beta(k) = v
            CASE (2)
                gama(k) = v
            CASE DEFAULT
                RETURN
        END SELECT
        ucref_set_coefficient = 0
    END FUNCTION ucref_set_coefficient

    ! The model loop from given start abundances y0(1:nspec+1) instead of the elemental start
    ! (the abundanceStart path of wrap.f90:636-640: set after initializeChemistry).
    INTEGER(c_int) FUNCTION ucref_model_from(p, y0, maxrec, rec, nrec, stats) BIND(C, name="ucref_model_from")
        REAL(c_double), INTENT(IN) :: p(P_NPARAM), y0(nspec+1)
        INTEGER(c_int), VALUE :: maxrec
        REAL(c_double), INTENT(OUT) :: rec(4+nspec, maxrec)
        INTEGER(c_int), INTENT(OUT) :: nrec
        INTEGER(c_int), INTENT(OUT) :: stats(8)
        IF (ALLOCATED(start_abund)) DEALLOCATE(start_abund)
        ALLOCATE(start_abund(nspec+1))
        start_abund = y0
        have_start = .true.
        ucref_model_from = run_model(p, HUGE(1.0d0), maxrec, rec, nrec, stats)
        have_start = .false.
    END FUNCTION ucref_model_from

    ! The same loop stopped after the first interval that ends at or beyond t_stop years;
    ! y_out and phys_out are the point's state there as the next interval starts from it
    ! (after the physics update): phys = (density, gasTemp, dustTemp, av, cloudSize, phi,
    ! abstol_factor, currentTime [s]).
    INTEGER(c_int) FUNCTION ucref_model_until(p, t_stop, maxrec, rec, nrec, stats, y_out, phys_out) &
            BIND(C, name="ucref_model_until")
        REAL(c_double), INTENT(IN) :: p(P_NPARAM)
        REAL(c_double), VALUE :: t_stop
        INTEGER(c_int), VALUE :: maxrec
        REAL(c_double), INTENT(OUT) :: rec(4+nspec, maxrec)
        INTEGER(c_int), INTENT(OUT) :: nrec
        INTEGER(c_int), INTENT(OUT) :: stats(8)
        REAL(c_double), INTENT(OUT) :: y_out(nspec+1), phys_out(8)
        ucref_model_until = run_model(p, t_stop, maxrec, rec, nrec, stats)
        CALL get_state(y_out, phys_out)
    END FUNCTION ucref_model_until

    ! The point's current state in the layout of ucref_model_until (e.g. after
    ! ucref_interval and ucref_update_physics, to continue the loop interval by interval).
    INTEGER(c_int) FUNCTION ucref_get_state(y_out, phys_out) BIND(C, name="ucref_get_state")
        REAL(c_double), INTENT(OUT) :: y_out(nspec+1), phys_out(8)
        CALL get_state(y_out, phys_out)
        ucref_get_state = 0
    END FUNCTION ucref_get_state

    SUBROUTINE get_state(y, phys)
        REAL(c_double), INTENT(OUT) :: y(nspec+1), phys(8)
        y = abund(:,1)
        phys(1) = density(1)
        phys(2) = gasTemp(1)
        phys(3) = dustTemp(1)
        phys(4) = av(1)
        phys(5) = cloudSize
        phys(6) = phi
        phys(7) = abstol_factor
        phys(8) = currentTime
    END SUBROUTINE get_state

    ! One output interval (updateChemistry, chemistry.f90:145-239, restated as
    ! bounded_update_chemistry with an unreachable bound unless maxsteps > 0) of the point
    ! with parameters p from the state (y, phys) given as ucref_model_until returns it, to
    ! t_target seconds; no physics update afterwards.  stats = (steps, attempts, and DVODE's
    ! NFE, NJE, NLU, NNI, NCFN, NETF, each summed over the attempts).
    INTEGER(c_int) FUNCTION ucref_interval(p, y, phys, t_target, maxsteps, y_out, phys_out, stats) &
            BIND(C, name="ucref_interval")
        REAL(c_double), INTENT(IN) :: p(P_NPARAM), y(nspec+1), phys(8)
        REAL(c_double), VALUE :: t_target, maxsteps
        REAL(c_double), INTENT(OUT) :: y_out(nspec+1), phys_out(8)
        INTEGER(c_int), INTENT(OUT) :: stats(8)
        INTEGER :: flag
        REAL(dp) :: nsteps, bound
        stats = 0
        CALL setup_point(p, flag)
        IF (flag .lt. 0) THEN
            ucref_interval = flag
            RETURN
        END IF
        abund(:,1) = y
        density(1) = phys(1)
        gasTemp(1) = phys(2)
        dustTemp(1) = phys(3)
        av(1) = phys(4)
        cloudSize = phys(5)
        phi = phys(6)
        abstol_factor = phys(7)
        currentTime = phys(8)
        timeInYears = currentTime/SECONDS_PER_YEAR
        targetTime = t_target
        dstep = 1
        nsteps = 0.0d0
        bound = HUGE(1.0d0)
        IF (maxsteps > 0.0d0) bound = maxsteps
        CALL bounded_update_chemistry(bound, nsteps, flag, stats(2), stats(3:8))
        stats(1) = INT(MIN(nsteps, 2.0d9))
        CALL get_state(y_out, phys_out)
        ucref_interval = flag
    END FUNCTION ucref_interval

    INTEGER FUNCTION run_model(p, t_stop, maxrec, rec, nrec, stats)
        REAL(c_double), INTENT(IN) :: p(P_NPARAM)
        REAL(c_double), INTENT(IN) :: t_stop
        INTEGER(c_int), VALUE :: maxrec
        REAL(c_double), INTENT(OUT) :: rec(4+nspec, maxrec)
        INTEGER(c_int), INTENT(OUT) :: nrec
        INTEGER(c_int), INTENT(OUT) :: stats(8)
        INTEGER :: flag, attempts
        REAL(dp) :: rstats(22), nsteps
        INTEGER :: istats(31)
        stats = 0
        nrec = 0
        nsteps = 0.0d0
        CALL setup_point(p, flag)
        IF (flag .lt. 0) THEN
            run_model = flag
            RETURN
        END IF
        IF (have_start) abund(:nspec+1,1) = start_abund(:nspec+1)
        CALL record_row(rec, maxrec, nrec)
        DO WHILE ((flag .eq. 0) .and. (timeInYears < t_stop) .and. &
            &(((endAtFinalDensity) .and. (density(1) < finalDens)) .or. &
            &((.not. endAtFinalDensity) .and. (timeInYears < finalTime))))
            currentTimeold=currentTime
            timeInYears=currentTime/SECONDS_PER_YEAR
            IF (modelKind .eq. 1) THEN
                CALL hotcoreUpdateTargetTime
            ELSE IF (modelKind .eq. 2) THEN
                CALL collapseUpdateTargetTime
            ELSE
                CALL cloudUpdateTargetTime
            END IF
            dstep=1
            currentTime=currentTimeold
            IF (p(P_MAXSTEPS) > 0.0d0) THEN
                CALL bounded_update_chemistry(p(P_MAXSTEPS), nsteps, flag, attempts)
            ELSE
                CALL updateChemistry(flag)
                attempts = loopCounter
            END IF
            IF (flag .lt. 0) EXIT
            CALL GET_STATS(rstats, istats)
            stats(1) = stats(1) + istats(11)
            stats(2) = stats(2) + istats(12)
            stats(3) = stats(3) + istats(13)
            stats(4) = stats(4) + 1
            ! integrator calls per interval: each further one recomputes the rates mid-interval
            stats(6) = MAX(stats(6), attempts)
            IF (attempts > 1) stats(7) = stats(7) + 1
            timeInYears= targetTime/SECONDS_PER_YEAR
            CALL coreUpdatePhysics
            IF (modelKind .eq. 1) CALL hotcoreUpdatePhysics
            IF (modelKind .eq. 2) CALL collapseUpdatePhysics
            CALL record_row(rec, maxrec, nrec)
            IF (nrec .ge. maxrec) THEN
                flag = NOT_ENOUGH_TIMEPOINTS_ERROR
                EXIT
            END IF
        END DO
        stats(5) = INT(MIN(nsteps, 2.0d9))
        run_model = flag
    END FUNCTION run_model

    SUBROUTINE record_row(rec, maxrec, nrec)
        INTEGER(c_int), VALUE :: maxrec
        REAL(c_double), INTENT(INOUT) :: rec(4+nspec, maxrec)
        INTEGER(c_int), INTENT(INOUT) :: nrec
        IF (nrec .ge. maxrec) RETURN
        nrec = nrec + 1
        rec(1, nrec) = timeInYears
        rec(2, nrec) = density(1)
        rec(3, nrec) = gasTemp(1)
        rec(4, nrec) = av(1)
        rec(5:, nrec) = abund(1:nspec, 1)
        IF (want_rates) rates_buf(:, nrec) = rate(1:nreac)
    END SUBROUTINE record_row

END MODULE ucref_driver
