! pdac_gpu -- Fortran (iso_c_binding) interface to libpdacgpu, the MI355X
! implementation of PDAC's per-time-step hot path declared in
! include/pdacgpu.h.
!
! The reference's time loop (main_prog, prog.f:143-329) calls argument-less
! module subroutines over module-global arrays (bdry.f:22, tilde.f:125,276,
! iter.f:38, ygas.f:27, htilde.f:25, ftem.f:19).  A host that keeps that
! loop replaces the bodies of those subroutines with the calls below; the
! arrays cross in the reference's own column-major (ntot, ncomp) layout.
! The derived types mirror the C structs field for field (BIND(C), same
! order and kinds), so C_SIZEOF of each equals sizeof() in C.
MODULE pdac_gpu
  USE iso_c_binding
  IMPLICIT NONE

  INTEGER, PARAMETER :: PDAC_MAX_SOLID = 4, PDAC_MAX_GAS = 7

  ! numerical status codes (> 0); < 0 is a HIP/runtime error
  INTEGER(C_INT), PARAMETER :: PDAC_ST_ITER_NOT_CONVERGED = 1, PDAC_ST_EOS_FAILURE = 2, &
       PDAC_ST_BAD_ARGUMENT = 3
  INTEGER(C_INT32_T), PARAMETER :: PDAC_PREC_FP64 = 0, PDAC_PREC_FP32FLUX = 1
  INTEGER(C_INT), PARAMETER :: PDAC_SNAP_OUTPUT = 0, PDAC_SNAP_RESTART = 1
  INTEGER(C_INT32_T), PARAMETER :: PDAC_ORDER_LEXICOGRAPHIC = 0, PDAC_ORDER_REDBLACK = 1, &
       PDAC_ORDER_STAGE = 2

  TYPE, BIND(C) :: pdac_dims
    INTEGER(C_INT32_T) :: job_type, itc, nx, ny, nz, nsolid, ngas, precision
  END TYPE pdac_dims

  TYPE, BIND(C) :: pdac_params
    REAL(C_DOUBLE) :: alpha, alphagrav, gravx, gravy, gravz, beta, omega, delg, &
         rlim, flim, tref, void_fraction
    REAL(C_DOUBLE) :: dk(PDAC_MAX_SOLID), rl(PDAC_MAX_SOLID), phis(PDAC_MAX_SOLID), &
         cps(PDAC_MAX_SOLID), kap(PDAC_MAX_SOLID), plim(PDAC_MAX_SOLID)
    REAL(C_DOUBLE) :: atm_ygc(PDAC_MAX_GAS)
    INTEGER(C_INT32_T) :: gas_type(PDAC_MAX_GAS)
    INTEGER(C_INT32_T) :: muscl, lm, lv, inmax, maxout, pmodel, drag_model, icpc, ord, &
         cstop, iter_order, time_integration
    INTEGER(C_INT32_T) :: gas_viscosity, part_viscosity, repulsive_model
    REAL(C_DOUBLE) :: mus(PDAC_MAX_SOLID)
    REAL(C_DOUBLE) :: cmut
    INTEGER(C_INT32_T) :: iturb
  END TYPE pdac_params

  ! field identifiers (enum pdac_field)
  INTEGER(C_INT), PARAMETER :: PDAC_P = 0, PDAC_EP = 1, PDAC_RGP = 2, PDAC_ROG = 3, &
       PDAC_TG = 4, PDAC_SIEG = 5, PDAC_UG = 6, PDAC_VG = 7, PDAC_WG = 8, PDAC_RLK = 9, &
       PDAC_TS = 10, PDAC_SIES = 11, PDAC_US = 12, PDAC_VS = 13, PDAC_WS = 14, &
       PDAC_YGC = 15, PDAC_XGC = 16

  INTERFACE
    INTEGER(C_INT) FUNCTION pdac_create(ctx, dims, device) BIND(C, name='pdac_create')
      IMPORT :: C_INT, C_PTR, pdac_dims
      TYPE(C_PTR), INTENT(OUT) :: ctx
      TYPE(pdac_dims), INTENT(IN) :: dims
      INTEGER(C_INT), VALUE :: device
    END FUNCTION pdac_create

    INTEGER(C_INT) FUNCTION pdac_destroy(ctx) BIND(C, name='pdac_destroy')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION pdac_destroy

    INTEGER(C_INT) FUNCTION pdac_set_grid(ctx, dx, dy, dz, flag, p_atm, t_atm) &
         BIND(C, name='pdac_set_grid')
      IMPORT :: C_INT, C_PTR, C_DOUBLE, C_INT32_T
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), INTENT(IN) :: dx(*), dy(*), dz(*), p_atm(*), t_atm(*)
      INTEGER(C_INT32_T), INTENT(IN) :: flag(*)
    END FUNCTION pdac_set_grid

    INTEGER(C_INT) FUNCTION pdac_set_params(ctx, prm) BIND(C, name='pdac_set_params')
      IMPORT :: C_INT, C_PTR, pdac_params
      TYPE(C_PTR), VALUE :: ctx
      TYPE(pdac_params), INTENT(IN) :: prm
    END FUNCTION pdac_set_params

    INTEGER(C_INT) FUNCTION pdac_set_ib(ctx, nfp, fpi, fpd, coords, bmask, ivf) &
         BIND(C, name='pdac_set_ib')
      IMPORT :: C_INT, C_PTR, C_DOUBLE, C_INT32_T
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT32_T), INTENT(IN) :: nfp(3), fpi(*), bmask(*)
      REAL(C_DOUBLE), INTENT(IN) :: fpd(*), coords(*), ivf(*)
    END FUNCTION pdac_set_ib

    INTEGER(C_INT) FUNCTION pdac_put(ctx, field, a, n) BIND(C, name='pdac_put')
      IMPORT :: C_INT, C_PTR, C_DOUBLE, C_INT64_T
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT), VALUE :: field
      REAL(C_DOUBLE), INTENT(IN) :: a(*)
      INTEGER(C_INT64_T), VALUE :: n
    END FUNCTION pdac_put

    INTEGER(C_INT) FUNCTION pdac_get(ctx, field, a, n) BIND(C, name='pdac_get')
      IMPORT :: C_INT, C_PTR, C_DOUBLE, C_INT64_T
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT), VALUE :: field
      REAL(C_DOUBLE), INTENT(OUT) :: a(*)
      INTEGER(C_INT64_T), VALUE :: n
    END FUNCTION pdac_get

    INTEGER(C_INT) FUNCTION pdac_field_ncomp(ctx, field) BIND(C, name='pdac_field_ncomp')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT), VALUE :: field
    END FUNCTION pdac_field_ncomp

    INTEGER(C_INT) FUNCTION pdac_cnvert(ctx) BIND(C, name='pdac_cnvert')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION pdac_cnvert

    INTEGER(C_INT) FUNCTION pdac_boundary(ctx) BIND(C, name='pdac_boundary')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION pdac_boundary

    INTEGER(C_INT) FUNCTION pdac_fieldn(ctx) BIND(C, name='pdac_fieldn')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION pdac_fieldn

    INTEGER(C_INT) FUNCTION pdac_sgsg(ctx) BIND(C, name='pdac_sgsg')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION pdac_sgsg

    INTEGER(C_INT) FUNCTION pdac_snapshot_size(ctx, kind, bytes) BIND(C, name='pdac_snapshot_size')
      IMPORT :: C_INT, C_PTR, C_INT64_T
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT), VALUE :: kind
      INTEGER(C_INT64_T), INTENT(OUT) :: bytes
    END FUNCTION pdac_snapshot_size

    INTEGER(C_INT) FUNCTION pdac_snapshot_async(ctx, kind) BIND(C, name='pdac_snapshot_async')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT), VALUE :: kind
    END FUNCTION pdac_snapshot_async

    INTEGER(C_INT) FUNCTION pdac_snapshot_fetch(ctx, kind, host, bytes) BIND(C, name='pdac_snapshot_fetch')
      IMPORT :: C_INT, C_PTR, C_INT64_T
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT), VALUE :: kind
      TYPE(C_PTR), VALUE :: host
      INTEGER(C_INT64_T), VALUE :: bytes
    END FUNCTION pdac_snapshot_fetch

    INTEGER(C_INT) FUNCTION pdac_ib_forcing(nx, ny, nz, axis, ord2d, cx, cy, cz, topo2d, ext_cells, n_ext, &
                                            cap, pts, nsl, count) BIND(C, name='pdac_ib_forcing')
      IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_DOUBLE
      INTEGER(C_INT32_T), VALUE :: nx, ny, nz, axis
      INTEGER(C_INT32_T), INTENT(IN) :: ord2d(*)
      REAL(C_DOUBLE), INTENT(IN) :: cx(*), cy(*), cz(*), topo2d(*)
      INTEGER(C_INT64_T), INTENT(IN) :: ext_cells(*)
      INTEGER(C_INT64_T), VALUE :: n_ext, cap
      INTEGER(C_INT32_T), INTENT(OUT) :: pts(*)
      REAL(C_DOUBLE), INTENT(OUT) :: nsl(*)
      INTEGER(C_INT64_T), INTENT(OUT) :: count
    END FUNCTION pdac_ib_forcing

    INTEGER(C_INT) FUNCTION pdac_ib_fill_cells(nx, ny, nz, tx, ty, tc, z, zb, flags, bd, vf) &
        BIND(C, name='pdac_ib_fill_cells')
      IMPORT :: C_INT, C_INT32_T, C_DOUBLE
      INTEGER(C_INT32_T), VALUE :: nx, ny, nz
      REAL(C_DOUBLE), INTENT(IN) :: tx(*), ty(*), tc(*), z(*), zb(*)
      INTEGER(C_INT32_T), INTENT(INOUT) :: flags(*)
      INTEGER(C_INT32_T), INTENT(OUT) :: bd(*)
      REAL(C_DOUBLE), INTENT(OUT) :: vf(*)
    END FUNCTION pdac_ib_fill_cells

    INTEGER(C_INT) FUNCTION pdac_begin_step(ctx) BIND(C, name='pdac_begin_step')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION pdac_begin_step

    INTEGER(C_INT) FUNCTION pdac_tilde(ctx, dt) BIND(C, name='pdac_tilde')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), VALUE :: dt
    END FUNCTION pdac_tilde

    INTEGER(C_INT) FUNCTION pdac_iter(ctx, dt, nit, converged) BIND(C, name='pdac_iter')
      IMPORT :: C_INT, C_PTR, C_DOUBLE, C_INT32_T
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), VALUE :: dt
      INTEGER(C_INT32_T), INTENT(OUT) :: nit, converged
    END FUNCTION pdac_iter

    INTEGER(C_INT) FUNCTION pdac_ygas(ctx, dt) BIND(C, name='pdac_ygas')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), VALUE :: dt
    END FUNCTION pdac_ygas

    INTEGER(C_INT) FUNCTION pdac_htilde(ctx, dt) BIND(C, name='pdac_htilde')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), VALUE :: dt
    END FUNCTION pdac_htilde

    INTEGER(C_INT) FUNCTION pdac_ftem(ctx, dt, eos_fail) BIND(C, name='pdac_ftem')
      IMPORT :: C_INT, C_PTR, C_DOUBLE, C_INT32_T
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), VALUE :: dt
      INTEGER(C_INT32_T), INTENT(OUT) :: eos_fail
    END FUNCTION pdac_ftem

    INTEGER(C_INT) FUNCTION pdac_initialize_fieldn(ctx, dt) BIND(C, name='pdac_initialize_fieldn')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), VALUE :: dt
    END FUNCTION pdac_initialize_fieldn

    INTEGER(C_INT) FUNCTION pdac_explicit_iter(ctx, dt, eos_fail, cfl) BIND(C, name='pdac_explicit_iter')
      IMPORT :: C_INT, C_PTR, C_DOUBLE, C_INT32_T
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), VALUE :: dt
      INTEGER(C_INT32_T), INTENT(OUT) :: eos_fail
      REAL(C_DOUBLE), INTENT(OUT) :: cfl
    END FUNCTION pdac_explicit_iter

    INTEGER(C_INT) FUNCTION pdac_step(ctx, dt, rungekut, nsteps, nit_last, nit_total) &
         BIND(C, name='pdac_step')
      IMPORT :: C_INT, C_PTR, C_DOUBLE, C_INT32_T, C_INT64_T
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), VALUE :: dt
      INTEGER(C_INT32_T), VALUE :: rungekut, nsteps
      INTEGER(C_INT32_T), INTENT(OUT) :: nit_last
      INTEGER(C_INT64_T), INTENT(OUT) :: nit_total
    END FUNCTION pdac_step

    INTEGER(C_INT) FUNCTION pdac_explicit_step_noalloc(ctx, dt, rungekut, nsteps, eos_fail, cfl) &
         BIND(C, name='pdac_explicit_step_noalloc')
      IMPORT :: C_INT, C_PTR, C_DOUBLE, C_INT32_T
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), VALUE :: dt
      INTEGER(C_INT32_T), VALUE :: rungekut, nsteps
      INTEGER(C_INT32_T), INTENT(OUT) :: eos_fail
      REAL(C_DOUBLE), INTENT(OUT) :: cfl
    END FUNCTION pdac_explicit_step_noalloc

    INTEGER(C_INT) FUNCTION pdac_sync(ctx) BIND(C, name='pdac_sync')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION pdac_sync

    INTEGER(C_INT) FUNCTION pdac_ens_create(ens, members, count) BIND(C, name='pdac_ens_create')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), INTENT(OUT) :: ens
      TYPE(C_PTR), INTENT(IN) :: members(*)
      INTEGER(C_INT), VALUE :: count
    END FUNCTION pdac_ens_create

    INTEGER(C_INT) FUNCTION pdac_ens_destroy(ens) BIND(C, name='pdac_ens_destroy')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ens
    END FUNCTION pdac_ens_destroy

    INTEGER(C_INT) FUNCTION pdac_ens_tune(ens, reps, tiles, us) BIND(C, name='pdac_ens_tune')
      IMPORT :: C_INT, C_PTR, C_INT32_T, C_DOUBLE
      TYPE(C_PTR), VALUE :: ens
      INTEGER(C_INT32_T), VALUE :: reps
      INTEGER(C_INT32_T), INTENT(OUT) :: tiles(8)
      REAL(C_DOUBLE), INTENT(OUT) :: us(32)
    END FUNCTION pdac_ens_tune

    INTEGER(C_INT) FUNCTION pdac_ens_step(ens, dt, rungekut, nsteps, eos_fail, cfl) &
         BIND(C, name='pdac_ens_step')
      IMPORT :: C_INT, C_PTR, C_DOUBLE, C_INT32_T
      TYPE(C_PTR), VALUE :: ens
      REAL(C_DOUBLE), INTENT(IN) :: dt(*)
      INTEGER(C_INT32_T), VALUE :: rungekut, nsteps
      INTEGER(C_INT32_T), INTENT(OUT) :: eos_fail(*)
      REAL(C_DOUBLE), INTENT(OUT) :: cfl(*)
    END FUNCTION pdac_ens_step

    INTEGER(C_INT) FUNCTION pdac_ens_set_tiles(ens, tiles) BIND(C, name='pdac_ens_set_tiles')
      IMPORT :: C_INT, C_PTR, C_INT32_T
      TYPE(C_PTR), VALUE :: ens
      INTEGER(C_INT32_T), INTENT(IN) :: tiles(8)
    END FUNCTION pdac_ens_set_tiles

    TYPE(C_PTR) FUNCTION pdac_last_error() BIND(C, name='pdac_last_error')
      IMPORT :: C_PTR
    END FUNCTION pdac_last_error

    ! diagnostics without a counterpart in the reference
    INTEGER(C_INT) FUNCTION pdac_get_momflux(ctx, comp, face, out) BIND(C, name='pdac_get_momflux')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT), VALUE :: comp, face
      REAL(C_DOUBLE), INTENT(OUT) :: out(*)
    END FUNCTION pdac_get_momflux

    INTEGER(C_INT) FUNCTION pdac_profile_enable(ctx, enable) BIND(C, name='pdac_profile_enable')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT), VALUE :: enable
    END FUNCTION pdac_profile_enable

    INTEGER(C_INT) FUNCTION pdac_profile_select(ctx, mask) BIND(C, name='pdac_profile_select')
      IMPORT :: C_INT, C_PTR, C_INT32_T
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT32_T), VALUE :: mask
    END FUNCTION pdac_profile_select

    INTEGER(C_INT) FUNCTION pdac_profile_reset(ctx) BIND(C, name='pdac_profile_reset')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION pdac_profile_reset

    INTEGER(C_INT) FUNCTION pdac_profile_read(ctx, kernel, launches, total_ms) &
         BIND(C, name='pdac_profile_read')
      IMPORT :: C_INT, C_PTR, C_INT64_T, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT), VALUE :: kernel
      INTEGER(C_INT64_T), INTENT(OUT) :: launches
      REAL(C_DOUBLE), INTENT(OUT) :: total_ms
    END FUNCTION pdac_profile_read

    INTEGER(C_INT) FUNCTION pdac_profile_visits(ctx, visits) BIND(C, name='pdac_profile_visits')
      IMPORT :: C_INT, C_PTR, C_INT64_T
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT64_T), INTENT(OUT) :: visits
    END FUNCTION pdac_profile_visits

    INTEGER(C_INT) FUNCTION pdac_profile_checks(ctx, checks) BIND(C, name='pdac_profile_checks')
      IMPORT :: C_INT, C_PTR, C_INT64_T
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT64_T), INTENT(OUT) :: checks
    END FUNCTION pdac_profile_checks
  END INTERFACE

CONTAINS

  ! The text of pdac_last_error() as a Fortran string (for CALL error(...)).
  FUNCTION pdac_error_text() RESULT(msg)
    CHARACTER(LEN=:), ALLOCATABLE :: msg
    TYPE(C_PTR) :: p
    CHARACTER(KIND=C_CHAR), POINTER :: chars(:)
    INTEGER :: n
    p = pdac_last_error()
    IF (.NOT. C_ASSOCIATED(p)) THEN
      msg = ''
      RETURN
    END IF
    CALL C_F_POINTER(p, chars, [4096])
    n = 0
    DO WHILE (n < 4096)
      IF (chars(n + 1) == C_NULL_CHAR) EXIT
      n = n + 1
    END DO
    ALLOCATE(CHARACTER(LEN=n) :: msg)
    msg = TRANSFER(chars(1:n), msg)
  END FUNCTION pdac_error_text

END MODULE pdac_gpu
