! ISO_C_BINDING interface of libuclchem_gpu.so for the UCLCHEM Fortran core.
!
! The C prototypes are in include/uclchem_gpu.h.  Array layouts line up with the
! reference's own column-major arrays, so no copies are needed:
!   params(UCG_NPARAM, npoints)          <-> C params [n][UCG_NPARAM]
!   abund(NEQ, npoints)                  <-> C y_init / y_out [n][NEQ]
!                                            (chemistry.f90:40,55 abund(NEQ,points))
!   rec(4+nspec, max_rec, npoints)       <-> C rec [n][max_rec][4+nspec]
!                                            (one row per output time: time, density,
!                                             gasTemp, Av, abundances, as wrap.f90 writes
!                                             physicsarray/chemicalabunarray)
! Parameter slots are given 1-based here (UCG_P_* + 1 of the C header).
MODULE uclchem_gpu
    USE, INTRINSIC :: ISO_C_BINDING
    IMPLICIT NONE
    PRIVATE
    PUBLIC :: ucg_init, ucg_network_load_file_f, ucg_network_info, ucg_solve_batch_f, ucg_error_message, &
              ucg_default_params, ucg_cancel, ucg_queue_slot_bytes, ucg_queue_open, ucg_queue_submit, &
              ucg_queue_test, ucg_queue_result, ucg_queue_close, ucg_queue_open_shared, ucg_queue_submit_to, &
              ucg_queue_times
    INTEGER, PARAMETER, PUBLIC :: UCG_NPARAM = 64
    INTEGER, PARAMETER, PUBLIC :: UCG_E_ARG = -102   ! an array does not match the batch or network
    INTEGER, PARAMETER, PUBLIC :: UCGP_DENS = 1, UCGP_TEMP = 2, UCGP_ZETA = 3, UCGP_RADFIELD = 4, &
        UCGP_BASEAV = 5, UCGP_ROUT = 6, UCGP_RIN = 7, UCGP_FINALTIME = 8, UCGP_FREEFALL = 9, &
        UCGP_FINALDENS = 10, UCGP_RELTOL = 11, UCGP_ABSFAC = 12, UCGP_ABSMIN = 13, UCGP_MODEL = 14, &
        UCGP_TEMPINDX = 15, UCGP_MAXTEMP = 16, UCGP_FREEZEFAC = 17, UCGP_ENDATFINALDENS = 18, &
        UCGP_FREEFALLFAC = 19, UCGP_MXSTEP = 20, UCGP_MAXSTEPS = 21, UCGP_METALLICITY = 22, UCGP_ION = 23, &
        UCGP_FH = 24, UCGP_FHE = 25, UCGP_FC = 26, UCGP_FO = 27, UCGP_FN = 28, UCGP_FS = 29, UCGP_FMG = 30, &
        UCGP_FSI = 31, UCGP_FCL = 32, UCGP_FP = 33, UCGP_FFE = 34, UCGP_FF = 35, UCGP_FD = 36, UCGP_FLI = 37, &
        UCGP_FNA = 38, UCGP_FPAH = 39, UCGP_F15N = 40, UCGP_F13C = 41, UCGP_F18O = 42, UCGP_EBMAXH2 = 43, &
        UCGP_EBMAXCR = 44, UCGP_EBMAXUVCR = 45, UCGP_EPSILON = 46, UCGP_UV_YIELD = 47, UCGP_PHI = 48, &
        UCGP_UVCREFF = 49, UCGP_OMEGA = 50, UCGP_DESORB = 51, UCGP_H2DESORB = 52, UCGP_CRDESORB = 53, &
        UCGP_UVDESORB = 54, UCGP_THERMDESORB = 55, UCGP_CRATTENUATION = 56, UCGP_IONMODEL = 57, &
        UCGP_IMPROVEDH2CRP = 58, UCGP_ENFORCECHARGE = 59, UCGP_COLLAPSE = 60

    INTERFACE
        INTEGER(C_INT) FUNCTION ucg_init(device) BIND(C, name="ucg_init")
            IMPORT :: C_INT
            INTEGER(C_INT), VALUE :: device
        END FUNCTION ucg_init

        INTEGER(C_INT) FUNCTION c_network_load_file(path, handle) BIND(C, name="ucg_network_load_file")
            IMPORT :: C_INT, C_CHAR
            CHARACTER(KIND=C_CHAR), DIMENSION(*), INTENT(IN) :: path
            INTEGER(C_INT), INTENT(OUT) :: handle
        END FUNCTION c_network_load_file

        INTEGER(C_INT) FUNCTION ucg_network_info(handle, nspec, neq, nreac) BIND(C, name="ucg_network_info")
            IMPORT :: C_INT
            INTEGER(C_INT), VALUE :: handle
            INTEGER(C_INT), INTENT(OUT) :: nspec, neq, nreac
        END FUNCTION ucg_network_info

        INTEGER(C_INT) FUNCTION c_solve_batch(handle, n, params, y_init, y_out, max_rec, rec, n_rec, status, &
                                              stats, lu_bits) BIND(C, name="ucg_solve_batch")
            IMPORT :: C_INT, C_DOUBLE, C_PTR
            INTEGER(C_INT), VALUE :: handle, n, max_rec, lu_bits
            REAL(C_DOUBLE), INTENT(IN) :: params(*)
            TYPE(C_PTR), VALUE :: y_init
            REAL(C_DOUBLE), INTENT(OUT) :: y_out(*)
            TYPE(C_PTR), VALUE :: rec
            INTEGER(C_INT), INTENT(OUT) :: n_rec(*), status(*), stats(*)
        END FUNCTION c_solve_batch

        INTEGER(C_INT) FUNCTION c_solve_batch_rates(handle, n, params, y_init, y_out, max_rec, rec, rates, n_rec, &
                                                    status, stats, lu_bits) BIND(C, name="ucg_solve_batch_rates")
            IMPORT :: C_INT, C_DOUBLE, C_PTR
            INTEGER(C_INT), VALUE :: handle, n, max_rec, lu_bits
            REAL(C_DOUBLE), INTENT(IN) :: params(*)
            TYPE(C_PTR), VALUE :: y_init
            REAL(C_DOUBLE), INTENT(OUT) :: y_out(*)
            TYPE(C_PTR), VALUE :: rec, rates
            INTEGER(C_INT), INTENT(OUT) :: n_rec(*), status(*), stats(*)
        END FUNCTION c_solve_batch_rates

        ! stop running device solves on a handle (points not finished report cancelled)
        INTEGER(C_INT) FUNCTION ucg_cancel(handle) BIND(C, name="ucg_cancel")
            IMPORT :: C_INT
            INTEGER(C_INT), VALUE :: handle
        END FUNCTION ucg_cancel

        ! request queues: one persistent solve kernel per queue that takes single-point
        ! requests posted while it runs (continuous batching); params(UCG_NPARAM) per request,
        ! y(NEQ) and stats(8) per result
        INTEGER(C_INT) FUNCTION ucg_queue_slot_bytes(handle, lu_bits, bytes) BIND(C, name="ucg_queue_slot_bytes")
            IMPORT :: C_INT, C_SIZE_T
            INTEGER(C_INT), VALUE :: handle, lu_bits
            INTEGER(C_SIZE_T), INTENT(OUT) :: bytes
        END FUNCTION ucg_queue_slot_bytes

        INTEGER(C_INT) FUNCTION ucg_queue_open(handle, capacity, n_slots, lu_bits, idle_timeout_s, queue) &
                BIND(C, name="ucg_queue_open")
            IMPORT :: C_INT, C_DOUBLE
            INTEGER(C_INT), VALUE :: handle, capacity, n_slots, lu_bits
            REAL(C_DOUBLE), VALUE :: idle_timeout_s
            INTEGER(C_INT), INTENT(OUT) :: queue
        END FUNCTION ucg_queue_open

        INTEGER(C_INT) FUNCTION ucg_queue_submit(queue, params, request) BIND(C, name="ucg_queue_submit")
            IMPORT :: C_INT, C_DOUBLE
            INTEGER(C_INT), VALUE :: queue
            REAL(C_DOUBLE), INTENT(IN) :: params(*)
            INTEGER(C_INT), INTENT(OUT) :: request
        END FUNCTION ucg_queue_submit

        INTEGER(C_INT) FUNCTION ucg_queue_test(queue, request, done) BIND(C, name="ucg_queue_test")
            IMPORT :: C_INT
            INTEGER(C_INT), VALUE :: queue, request
            INTEGER(C_INT), INTENT(OUT) :: done
        END FUNCTION ucg_queue_test

        INTEGER(C_INT) FUNCTION ucg_queue_result(queue, request, y, status, stats) BIND(C, name="ucg_queue_result")
            IMPORT :: C_INT, C_DOUBLE
            INTEGER(C_INT), VALUE :: queue, request
            REAL(C_DOUBLE), INTENT(OUT) :: y(*)
            INTEGER(C_INT), INTENT(OUT) :: status, stats(*)
        END FUNCTION ucg_queue_result

        INTEGER(C_INT) FUNCTION ucg_queue_close(queue) BIND(C, name="ucg_queue_close")
            IMPORT :: C_INT
            INTEGER(C_INT), VALUE :: queue
        END FUNCTION ucg_queue_close

        ! one queue shared by several networks (handles(n_networks)); requests name the
        ! network by its 0-based position in handles
        INTEGER(C_INT) FUNCTION ucg_queue_open_shared(handles, n_networks, capacity, n_slots, lu_bits, &
                idle_timeout_s, queue) BIND(C, name="ucg_queue_open_shared")
            IMPORT :: C_INT, C_DOUBLE
            INTEGER(C_INT), INTENT(IN) :: handles(*)
            INTEGER(C_INT), VALUE :: n_networks, capacity, n_slots, lu_bits
            REAL(C_DOUBLE), VALUE :: idle_timeout_s
            INTEGER(C_INT), INTENT(OUT) :: queue
        END FUNCTION ucg_queue_open_shared

        INTEGER(C_INT) FUNCTION ucg_queue_submit_to(queue, network, params, request) BIND(C, name="ucg_queue_submit_to")
            IMPORT :: C_INT, C_DOUBLE
            INTEGER(C_INT), VALUE :: queue, network
            REAL(C_DOUBLE), INTENT(IN) :: params(*)
            INTEGER(C_INT), INTENT(OUT) :: request
        END FUNCTION ucg_queue_submit_to

        INTEGER(C_INT) FUNCTION ucg_queue_times(queue, request, t_taken, t_done) BIND(C, name="ucg_queue_times")
            IMPORT :: C_INT, C_DOUBLE
            INTEGER(C_INT), VALUE :: queue, request
            REAL(C_DOUBLE), INTENT(OUT) :: t_taken, t_done
        END FUNCTION ucg_queue_times

        TYPE(C_PTR) FUNCTION c_last_error() BIND(C, name="ucg_last_error")
            IMPORT :: C_PTR
        END FUNCTION c_last_error

        INTEGER(C_SIZE_T) FUNCTION c_strlen(s) BIND(C, name="strlen")
            IMPORT :: C_PTR, C_SIZE_T
            TYPE(C_PTR), VALUE :: s
        END FUNCTION c_strlen
    END INTERFACE

CONTAINS

    ! Load a network file written by MakeRates' HIP backend (<name>.ucgnet).
    INTEGER FUNCTION ucg_network_load_file_f(path, handle) RESULT(rc)
        CHARACTER(LEN=*), INTENT(IN) :: path
        INTEGER, INTENT(OUT) :: handle
        INTEGER(C_INT) :: h
        rc = c_network_load_file(TRIM(path)//C_NULL_CHAR, h)
        handle = h
    END FUNCTION ucg_network_load_file_f

    ! Solve npoints independent grid points (the solveAbundances loop of wrap.f90:549-697
    ! per point).  abund_init absent: UCLCHEM's elemental start.  rec optional.
    ! rates (optional, with rec): the rate coefficients of every output row, rates(nreac, max_rec, n)
    ! (the reference's ratesarray, io.f90:90-93; row 1 is written before any interval and is zero)
    INTEGER FUNCTION ucg_solve_batch_f(handle, params, abund_out, status, stats, abund_init, rec, n_rec, &
                                       lu_bits, rates) RESULT(rc)
        INTEGER, INTENT(IN) :: handle
        REAL(C_DOUBLE), INTENT(IN) :: params(:, :)
        REAL(C_DOUBLE), INTENT(OUT) :: abund_out(:, :)
        INTEGER(C_INT), INTENT(OUT) :: status(:), stats(:, :)
        REAL(C_DOUBLE), INTENT(IN), OPTIONAL, TARGET, CONTIGUOUS :: abund_init(:, :)
        REAL(C_DOUBLE), INTENT(OUT), OPTIONAL, TARGET, CONTIGUOUS :: rec(:, :, :)
        INTEGER(C_INT), INTENT(OUT), OPTIONAL :: n_rec(:)
        INTEGER, INTENT(IN), OPTIONAL :: lu_bits
        REAL(C_DOUBLE), INTENT(OUT), OPTIONAL, TARGET, CONTIGUOUS :: rates(:, :, :)
        INTEGER(C_INT) :: n, bits, max_rec, nspec, neq, nreac
        INTEGER(C_INT), ALLOCATABLE :: nrec_local(:)
        TYPE(C_PTR) :: p_init, p_rec
        n = SIZE(params, 2)
        ! every array must hold what the library reads or writes for n points
        rc = ucg_network_info(INT(handle, C_INT), nspec, neq, nreac)
        IF (rc /= 0) RETURN
        rc = UCG_E_ARG
        IF (SIZE(params, 1) /= UCG_NPARAM .OR. SIZE(abund_out, 1) /= neq .OR. SIZE(abund_out, 2) < n) RETURN
        IF (SIZE(status) < n .OR. SIZE(stats, 1) /= 8 .OR. SIZE(stats, 2) < n) RETURN
        IF (PRESENT(abund_init)) THEN
            IF (SIZE(abund_init, 1) /= neq .OR. SIZE(abund_init, 2) < n) RETURN
        END IF
        IF (PRESENT(rec)) THEN
            IF (SIZE(rec, 1) /= 4 + nspec .OR. SIZE(rec, 3) < n) RETURN
        END IF
        IF (PRESENT(n_rec)) THEN
            IF (SIZE(n_rec) < n) RETURN
        END IF
        IF (PRESENT(rates)) THEN
            IF (.NOT. PRESENT(rec)) RETURN
            IF (SIZE(rates, 1) /= nreac .OR. SIZE(rates, 2) /= SIZE(rec, 2) .OR. SIZE(rates, 3) < n) RETURN
        END IF
        bits = 64
        IF (PRESENT(lu_bits)) bits = lu_bits
        p_init = C_NULL_PTR
        IF (PRESENT(abund_init)) p_init = C_LOC(abund_init)
        p_rec = C_NULL_PTR
        max_rec = 0
        IF (PRESENT(rec)) THEN
            p_rec = C_LOC(rec)
            max_rec = SIZE(rec, 2)
        END IF
        ALLOCATE(nrec_local(n))
        IF (PRESENT(rates)) THEN
            rc = c_solve_batch_rates(INT(handle, C_INT), n, params, p_init, abund_out, max_rec, p_rec, C_LOC(rates), &
                                     nrec_local, status, stats, bits)
        ELSE
            rc = c_solve_batch(INT(handle, C_INT), n, params, p_init, abund_out, max_rec, p_rec, nrec_local, status, &
                               stats, bits)
        END IF
        IF (PRESENT(n_rec)) n_rec = nrec_local
    END FUNCTION ucg_solve_batch_f

    FUNCTION ucg_error_message() RESULT(msg)
        CHARACTER(LEN=:), ALLOCATABLE :: msg
        TYPE(C_PTR) :: p
        CHARACTER(KIND=C_CHAR), POINTER :: chars(:)
        INTEGER :: i, n
        p = c_last_error()
        n = INT(c_strlen(p))
        CALL C_F_POINTER(p, chars, [n])
        ALLOCATE(CHARACTER(LEN=n) :: msg)
        DO i = 1, n
            msg(i:i) = chars(i)
        END DO
    END FUNCTION ucg_error_message

    ! UCLCHEM's defaults (defaultparameters.f90:18-122), rout as the compiled reference
    ! holds it (a default-kind literal, float32-rounded).
    SUBROUTINE ucg_default_params(p)
        REAL(C_DOUBLE), INTENT(OUT) :: p(UCG_NPARAM)
        p = 0.0D0
        p(UCGP_DENS) = 1.0D2
        p(UCGP_TEMP) = 10.0D0
        p(UCGP_ZETA) = 1.0D0
        p(UCGP_RADFIELD) = 1.0D0
        p(UCGP_BASEAV) = 2.0D0
        p(UCGP_ROUT) = REAL(0.05, C_DOUBLE)
        p(UCGP_RIN) = 0.0D0
        p(UCGP_FINALTIME) = 5.0D6
        p(UCGP_FREEFALL) = 0.0D0
        p(UCGP_FINALDENS) = 1.0D5
        p(UCGP_RELTOL) = 1.0D-8
        p(UCGP_ABSFAC) = 1.0D-14
        p(UCGP_ABSMIN) = 1.0D-25
        p(UCGP_MODEL) = 0.0D0
        p(UCGP_TEMPINDX) = 1.0D0
        p(UCGP_MAXTEMP) = 300.0D0
        p(UCGP_FREEZEFAC) = 1.0D0
        p(UCGP_ENDATFINALDENS) = 0.0D0
        p(UCGP_FREEFALLFAC) = 1.0D0
        p(UCGP_MXSTEP) = 10000.0D0
        p(UCGP_MAXSTEPS) = 0.0D0
        p(UCGP_METALLICITY) = 1.0D0
        p(UCGP_ION) = 2.0D0
        p(UCGP_FH) = 0.5D0
        p(UCGP_FHE) = REAL(0.1, C_DOUBLE)
        p(UCGP_FC) = 1.77D-04
        p(UCGP_FO) = 3.34D-04
        p(UCGP_FN) = 6.18D-05
        p(UCGP_FS) = 3.51D-6
        p(UCGP_FMG) = 2.256D-06
        p(UCGP_FSI) = 1.78D-06
        p(UCGP_FCL) = 3.39D-08
        p(UCGP_FP) = 7.78D-08
        p(UCGP_FFE) = 2.01D-7
        p(UCGP_FF) = 3.6D-08
        p(UCGP_EBMAXH2) = 1.21D3
        p(UCGP_EBMAXCR) = 1.21D3
        p(UCGP_EBMAXUVCR) = 1.0D4
        p(UCGP_EPSILON) = REAL(0.01, C_DOUBLE)
        p(UCGP_UV_YIELD) = REAL(0.03, C_DOUBLE)
        p(UCGP_PHI) = 1.0D5
        p(UCGP_UVCREFF) = 1.0D-3
        p(UCGP_OMEGA) = 0.5D0
        p(UCGP_DESORB) = 1.0D0
        p(UCGP_H2DESORB) = 1.0D0
        p(UCGP_CRDESORB) = 1.0D0
        p(UCGP_UVDESORB) = 1.0D0
        p(UCGP_THERMDESORB) = 1.0D0
    END SUBROUTINE ucg_default_params

END MODULE uclchem_gpu
