! Example Fortran caller: a static-cloud grid solved through the C ABI, the way a
! batched entry in the reference's wrap.f90 (beside cloud, wrap.f90:15-80) would call it.
!
! usage: cloud_grid <network.ucgnet> <npoints> <lu_bits> <out.bin>
! Point q (1-based) has n_H = 10**(2 + 5*(q-1)/max(1,npoints-1)), T = 10 K, other
! parameters at UCLCHEM's defaults, t_end = 1e6 yr.  Writes the final abund(NEQ, npoints)
! as raw little-endian doubles and one status line per point on stdout.
PROGRAM cloud_grid
    USE, INTRINSIC :: ISO_C_BINDING
    USE uclchem_gpu
    IMPLICIT NONE
    CHARACTER(LEN=512) :: path, arg, outpath
    INTEGER :: npoints, bits, handle, rc, q, u
    INTEGER(C_INT) :: nspec, neq, nreac
    REAL(C_DOUBLE), ALLOCATABLE :: params(:, :), abund(:, :)
    INTEGER(C_INT), ALLOCATABLE :: status(:), stats(:, :)

    CALL GET_COMMAND_ARGUMENT(1, path)
    CALL GET_COMMAND_ARGUMENT(2, arg)
    READ (arg, *) npoints
    CALL GET_COMMAND_ARGUMENT(3, arg)
    READ (arg, *) bits
    CALL GET_COMMAND_ARGUMENT(4, outpath)

    rc = ucg_network_load_file_f(path, handle)
    IF (rc /= 0) THEN
        WRITE (*, '(A)') 'network load failed: '//ucg_error_message()
        STOP 1
    END IF
    rc = ucg_network_info(INT(handle, C_INT), nspec, neq, nreac)
    WRITE (*, '(A,I0,A,I0,A,I0)') 'network: nspec=', nspec, ' neq=', neq, ' nreac=', nreac

    ALLOCATE(params(UCG_NPARAM, npoints), abund(neq, npoints), status(npoints), stats(8, npoints))
    DO q = 1, npoints
        CALL ucg_default_params(params(:, q))
        params(UCGP_DENS, q) = 10.0D0**(2.0D0 + 5.0D0*DBLE(q - 1)/DBLE(MAX(1, npoints - 1)))
        params(UCGP_FINALTIME, q) = 1.0D6
    END DO
    rc = ucg_solve_batch_f(handle, params, abund, status, stats, lu_bits=bits)
    IF (rc /= 0) THEN
        WRITE (*, '(A)') 'solve failed: '//ucg_error_message()
        STOP 2
    END IF
    DO q = 1, npoints
        WRITE (*, '(A,I0,A,ES12.5,A,I0,A,I0)') 'point ', q, ' n_H=', params(UCGP_DENS, q), ' status=', status(q), &
            ' steps=', stats(1, q)
    END DO
    OPEN (NEWUNIT=u, FILE=TRIM(outpath), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
    WRITE (u) abund
    CLOSE (u)
END PROGRAM cloud_grid
