! pdac_drive -- a host program shaped like the reference's time loop
! (main_prog, prog.f:143-329) that advances a domain through the C-ABI of
! libpdacgpu from Fortran, through MODULE pdac_gpu.
!
!   pdac_drive.x --sizes          print C_SIZEOF of pdac_dims and pdac_params
!                                 (no device call)
!   pdac_drive.x IN OUT           read a domain from IN, step it, write OUT
!   pdac_drive.x IN OUT DIR       the same, and in DIR the output file of the
!                                 first step's boundary call (output.0000), the
!                                 one of the last step (output.0001) and the
!                                 restart file (pdac.res), written here with
!                                 the reference's unformatted records from
!                                 device snapshots (outp, taperd)
!
! IN is a little-endian stream file: the dims (8 int32); the params
! component by component (12 doubles, dk rl phis cps kap plim of
! PDAC_MAX_SOLID doubles each, atm_ygc of PDAC_MAX_GAS doubles, gas_type of
! PDAC_MAX_GAS int32, 16 int32, mus of PDAC_MAX_SOLID doubles, cmut); nsteps,
! rungekut, mode (int32); dt, the start time (double);
! immb (int32); dx(nx), dy(ny), dz(nz), p_atm(nz), t_atm(nz) (doubles);
! flag(ntot) (int32); when immb = 1 the immersed-boundary arrays of
! pdac_set_ib (nfp(3), fpi(7*sum nfp), fpd(3*sum nfp), coords(2nx+2ny+2nz),
! bmask(6*ntot), ivf(ntot)); nput, then per field id (int32), count
! (int64), values; nget, then the ids of the fields to write back.
! mode 0 calls the subroutines one by one in main_prog's order (boundary,
! fieldn, sgsg, allocate_fluxes, and per Runge-Kutta stage with dt/(rungekut+1-rk)
! tilde, iter, ygas, htilde, ftem; for time_integration = 1 boundary,
! initialize_fieldn, sgsg and per stage explicit_iter, ygas); mode 1 calls
! pdac_step for all steps at once.
! OUT: status, nit_last (int32), nit_total (int64), then per field written
! back its id (int32), count (int64) and values.
PROGRAM pdac_drive
  USE iso_c_binding
  USE pdac_gpu
  IMPLICIT NONE

  TYPE(pdac_dims) :: dims
  TYPE(pdac_params) :: prm
  TYPE(C_PTR) :: ctx
  CHARACTER(LEN=1024) :: arg_in, arg_out, arg_dir
  INTEGER(C_INT32_T) :: d8(8), pint(16), nsteps, rungekut, mode, immb, nput, nget, fid
  INTEGER(C_INT32_T) :: nfp(3), nit, conv, fail, nit_last
  INTEGER(C_INT64_T) :: nit_total, cnt, ntot
  REAL(C_DOUBLE) :: pdbl(12), dt, dts, cfl, time
  INTEGER(C_INT32_T) :: nfil
  REAL(C_DOUBLE), ALLOCATABLE :: dx(:), dy(:), dz(:), p_atm(:), t_atm(:), fpd(:), coords(:), &
       ivf(:), a(:)
  INTEGER(C_INT32_T), ALLOCATABLE :: flag(:), fpi(:), bmask(:), ids(:)
  INTEGER :: u, istep, rk, i, nfpt
  INTEGER(C_INT) :: rc

  arg_in = ''
  IF (COMMAND_ARGUMENT_COUNT() >= 1) CALL GET_COMMAND_ARGUMENT(1, arg_in)
  IF (COMMAND_ARGUMENT_COUNT() == 1 .AND. TRIM(arg_in) == '--sizes') THEN
    WRITE(*, '(A,I0,A,I0)') 'pdac_dims ', C_SIZEOF(dims), ' pdac_params ', C_SIZEOF(prm)
  ELSE IF (COMMAND_ARGUMENT_COUNT() == 2 .OR. COMMAND_ARGUMENT_COUNT() == 3) THEN
    CALL GET_COMMAND_ARGUMENT(2, arg_out)
    arg_dir = ''
    IF (COMMAND_ARGUMENT_COUNT() == 3) CALL GET_COMMAND_ARGUMENT(3, arg_dir)
    CALL run()
  ELSE
    WRITE(*, '(A)') 'usage: pdac_drive.x --sizes | pdac_drive.x IN OUT [DIR]'
    ERROR STOP 2
  END IF

CONTAINS

  ! read IN, set the domain up, step it and write OUT
  SUBROUTINE run()
    OPEN(NEWUNIT=u, FILE=TRIM(arg_in), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD', ACTION='READ')
    READ(u) d8
    dims%job_type = d8(1); dims%itc = d8(2); dims%nx = d8(3); dims%ny = d8(4); dims%nz = d8(5)
    dims%nsolid = d8(6); dims%ngas = d8(7); dims%precision = d8(8)
    READ(u) pdbl
    prm%alpha = pdbl(1); prm%alphagrav = pdbl(2); prm%gravx = pdbl(3); prm%gravy = pdbl(4)
    prm%gravz = pdbl(5); prm%beta = pdbl(6); prm%omega = pdbl(7); prm%delg = pdbl(8)
    prm%rlim = pdbl(9); prm%flim = pdbl(10); prm%tref = pdbl(11); prm%void_fraction = pdbl(12)
    READ(u) prm%dk, prm%rl, prm%phis, prm%cps, prm%kap, prm%plim, prm%atm_ygc, prm%gas_type
    READ(u) pint
    prm%muscl = pint(1); prm%lm = pint(2); prm%lv = pint(3); prm%inmax = pint(4)
    prm%maxout = pint(5); prm%pmodel = pint(6); prm%drag_model = pint(7); prm%icpc = pint(8)
    prm%ord = pint(9); prm%cstop = pint(10); prm%iter_order = pint(11)
    prm%time_integration = pint(12)
    prm%gas_viscosity = pint(13); prm%part_viscosity = pint(14); prm%repulsive_model = pint(15)
    prm%iturb = pint(16)
    READ(u) prm%mus, prm%cmut
    READ(u) nsteps, rungekut, mode, dt, time, immb
    ntot = INT(dims%nx, C_INT64_T) * dims%ny * dims%nz
    ALLOCATE(dx(dims%nx), dy(dims%ny), dz(dims%nz), p_atm(dims%nz), t_atm(dims%nz), flag(ntot))
    READ(u) dx, dy, dz, p_atm, t_atm, flag

    rc = pdac_create(ctx, dims, 0_C_INT)
    CALL check('pdac_create')
    rc = pdac_set_grid(ctx, dx, dy, dz, flag, p_atm, t_atm)
    CALL check('pdac_set_grid')
    rc = pdac_set_params(ctx, prm)
    CALL check('pdac_set_params')
    IF (immb == 1) THEN
      READ(u) nfp
      nfpt = nfp(1) + nfp(2) + nfp(3)
      ALLOCATE(fpi(7 * nfpt), fpd(3 * nfpt), coords(2 * (dims%nx + dims%ny + dims%nz)), &
           bmask(6 * ntot), ivf(ntot))
      READ(u) fpi, fpd, coords, bmask, ivf
      rc = pdac_set_ib(ctx, nfp, fpi, fpd, coords, bmask, ivf)
      CALL check('pdac_set_ib')
    END IF
    READ(u) nput
    DO i = 1, nput
      READ(u) fid, cnt
      ALLOCATE(a(cnt))
      READ(u) a
      rc = pdac_put(ctx, INT(fid, C_INT), a, cnt)
      CALL check('pdac_put')
      DEALLOCATE(a)
    END DO
    READ(u) nget
    ALLOCATE(ids(nget))
    READ(u) ids
    CLOSE(u)
    rc = pdac_cnvert(ctx)
    CALL check('pdac_cnvert')

    nit_last = 0
    nit_total = 0
    nfil = 0
    IF (mode == 1) THEN
      rc = pdac_step(ctx, dt, rungekut, nsteps, nit_last, nit_total)
      CALL check('pdac_step')
    ELSE
      DO istep = 1, nsteps
        rc = pdac_boundary(ctx)
        CALL check('boundary')
        ! the initial conditions' output file (prog.f:160-165)
        IF (istep == 1 .AND. LEN_TRIM(arg_dir) > 0) CALL write_output()
        IF (prm%time_integration == 1) THEN
          rc = pdac_initialize_fieldn(ctx, dt)
          CALL check('initialize_fieldn')
          rc = pdac_sgsg(ctx)
          CALL check('sgsg')
          DO rk = 1, rungekut
            dts = dt / DBLE(rungekut + 1 - rk)
            rc = pdac_explicit_iter(ctx, dts, fail, cfl)
            CALL check('explicit_iter')
            rc = pdac_ygas(ctx, dts)
            CALL check('ygas')
          END DO
        ELSE
          rc = pdac_fieldn(ctx)
          CALL check('fieldn')
          rc = pdac_sgsg(ctx)
          CALL check('sgsg')
          rc = pdac_begin_step(ctx)
          CALL check('allocate_fluxes')
          DO rk = 1, rungekut
            dts = dt / DBLE(rungekut + 1 - rk)
            rc = pdac_tilde(ctx, dts)
            CALL check('tilde')
            rc = pdac_iter(ctx, dts, nit, conv)
            CALL check('iter')
            nit_last = nit
            nit_total = nit_total + nit
            rc = pdac_ygas(ctx, dts)
            CALL check('ygas')
            rc = pdac_htilde(ctx, dts)
            CALL check('htilde')
            rc = pdac_ftem(ctx, dts, fail)
            CALL check('ftem')
          END DO
        END IF
        time = time + dt
      END DO
      IF (LEN_TRIM(arg_dir) > 0) THEN
        CALL write_output()
        CALL write_restart()
      END IF
    END IF

    OPEN(NEWUNIT=u, FILE=TRIM(arg_out), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE', &
         ACTION='WRITE')
    WRITE(u) 0_C_INT32_T, nit_last, nit_total
    DO i = 1, nget
      cnt = ntot * pdac_field_ncomp(ctx, INT(ids(i), C_INT))
      ALLOCATE(a(cnt))
      rc = pdac_get(ctx, INT(ids(i), C_INT), a, cnt)
      CALL check('pdac_get')
      WRITE(u) ids(i), cnt, a
      DEALLOCATE(a)
    END DO
    CLOSE(u)
    rc = pdac_destroy(ctx)
  END SUBROUTINE run

  ! outp (outp.f:194-384), unformatted: the time as a 4-byte real, then the
  ! arrays of a device snapshot one record each
  SUBROUTINE write_output()
    REAL(C_FLOAT), ALLOCATABLE, TARGET :: fb(:, :)
    INTEGER(C_INT64_T) :: nb
    CHARACTER(LEN=4) :: num
    INTEGER :: v, j
    rc = pdac_snapshot_async(ctx, PDAC_SNAP_OUTPUT)
    CALL check('pdac_snapshot_async')
    rc = pdac_snapshot_size(ctx, PDAC_SNAP_OUTPUT, nb)
    CALL check('pdac_snapshot_size')
    ALLOCATE(fb(ntot, nb / (4 * ntot)))
    rc = pdac_snapshot_fetch(ctx, PDAC_SNAP_OUTPUT, C_LOC(fb), nb)
    CALL check('pdac_snapshot_fetch')
    WRITE(num, '(I4.4)') nfil
    OPEN(NEWUNIT=v, FILE=TRIM(arg_dir)//'/output.'//num, FORM='UNFORMATTED', STATUS='REPLACE', &
         ACTION='WRITE')
    WRITE(v) REAL(time, 4)
    DO j = 1, SIZE(fb, 2)
      WRITE(v) fb(:, j)
    END DO
    CLOSE(v)
    DEALLOCATE(fb)
    nfil = nfil + 1
  END SUBROUTINE write_output

  ! taperd (io_restart.f:31-70): the header records, then the arrays of a
  ! device snapshot one record each
  SUBROUTINE write_restart()
    REAL(C_DOUBLE), ALLOCATABLE, TARGET :: db(:, :)
    INTEGER(C_INT64_T) :: nb
    INTEGER :: v, j
    rc = pdac_snapshot_async(ctx, PDAC_SNAP_RESTART)
    CALL check('pdac_snapshot_async')
    rc = pdac_snapshot_size(ctx, PDAC_SNAP_RESTART, nb)
    CALL check('pdac_snapshot_size')
    ALLOCATE(db(ntot, nb / (8 * ntot)))
    rc = pdac_snapshot_fetch(ctx, PDAC_SNAP_RESTART, C_LOC(db), nb)
    CALL check('pdac_snapshot_fetch')
    OPEN(NEWUNIT=v, FILE=TRIM(arg_dir)//'/pdac.res', FORM='UNFORMATTED', STATUS='REPLACE', &
         ACTION='WRITE')
    WRITE(v) time, dims%nx, dims%ny, dims%nz, dims%nsolid, dims%ngas, nfil
    WRITE(v) prm%gas_type(1:dims%ngas)
    DO j = 1, SIZE(db, 2)
      WRITE(v) db(:, j)
    END DO
    CLOSE(v)
    DEALLOCATE(db)
  END SUBROUTINE write_restart

  ! the reference's CALL error(...) on a failing call: message and stop
  SUBROUTINE check(where)
    CHARACTER(LEN=*), INTENT(IN) :: where
    INTEGER :: v
    IF (rc == 0) RETURN
    WRITE(*, '(A,A,I0,A,A)') where, ': status ', rc, ': ', pdac_error_text()
    OPEN(NEWUNIT=v, FILE=TRIM(arg_out), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE', &
         ACTION='WRITE')
    WRITE(v) INT(rc, C_INT32_T), 0_C_INT32_T, 0_C_INT64_T
    CLOSE(v)
    ERROR STOP 1
  END SUBROUTINE check

END PROGRAM pdac_drive
