!-----------------------------------------------------------------------
! letkf_engine_mod.f90 -- iso_c_binding interface to libletkf_engine.so
! (include/letkf_engine.h) and a drop-in replacement of
!   SUBROUTINE da_main(fcst3d,fcst2d,anal3d,anal2d)
! (letkf/run/src/common_da.f90:22-124) for the sbPOM-LETKF driver.
!
! letkf_da_main_gpu reads the same module state that da_main reads
! (common_setting.f90:168-198 grid and observation arrays, and the
! PARAMETERs of common_setting.f90:20-131), passes it by address to the
! engine (no repacking: the engine takes the Fortran layouts as they are)
! and returns the analysis in anal3d/anal2d.  A failure is reported on unit
! 6 and the program stops, as the reference does.
!-----------------------------------------------------------------------
MODULE letkf_engine_mod
  USE iso_c_binding
  IMPLICIT NONE
  PRIVATE
  PUBLIC :: letkf_params_t, letkf_grid_t, letkf_obs_t, letkf_lpfgm_cfg_t
  PUBLIC :: letkf_params_default, letkf_create, letkf_destroy, letkf_set_params, letkf_da_main, letkf_da_main_lpfgm
  PUBLIC :: letkf_last_error_f
  PUBLIC :: LETKF_PREC_FP64, LETKF_PREC_FP32, LETKF_PREC_BF16X3, LETKF_PREC_FP32_PANEL, LETKF_PREC_FP32X3, &
            LETKF_PREC_FP32_DUAL64, LETKF_PREC_BF16X3_TRI32
  PUBLIC :: LETKF_FORM_PLAIN, LETKF_FORM_BATCHED, LETKF_FORM_BATCHED_QL

  INTEGER(c_int32_t), PARAMETER :: LETKF_PREC_FP64 = 0, LETKF_PREC_FP32 = 1, LETKF_PREC_BF16X3 = 2
  INTEGER(c_int32_t), PARAMETER :: LETKF_PREC_FP32_PANEL = 3, LETKF_PREC_FP32X3 = 4, LETKF_PREC_FP32_DUAL64 = 5
  INTEGER(c_int32_t), PARAMETER :: LETKF_PREC_BF16X3_TRI32 = 6
  INTEGER(c_int32_t), PARAMETER :: LETKF_FORM_PLAIN = 0, LETKF_FORM_BATCHED = 1, LETKF_FORM_BATCHED_QL = 2

  TYPE, BIND(C) :: letkf_params_t
     INTEGER(c_int32_t) :: nmem, nlon, nlat, nlev, nij1, nv3d, nv2d
     INTEGER(c_int32_t) :: iv3d_t, iv3d_s, id_z_obs, real_kind, roff, precision, form
     REAL(c_double) :: pi, re, cov_infl_mul, alpha_rtpp, alpha_rtps
     REAL(c_double) :: sigma_obs, sigma_obsv, dist_zero, dist_zerov
     REAL(c_double) :: tmin, tmax, smin, smax
     INTEGER(c_int32_t) :: stage_timing
     INTEGER(c_int64_t) :: workspace_bytes
  END TYPE letkf_params_t

  TYPE, BIND(C) :: letkf_grid_t
     TYPE(c_ptr) :: lon1, lat1, ri1, rj1, dlon1, dlat1, phi1, depth1
  END TYPE letkf_grid_t

  TYPE, BIND(C) :: letkf_obs_t
     INTEGER(c_int32_t) :: nobs
     TYPE(c_ptr) :: nobsgrd, elm, lon, lat, lev, err, dep, hdxf
  END TYPE letkf_obs_t

  ! iswitch_da, nmonte, RM, DP, cov_infl_gm (common_setting.f90:65,79-89)
  ! and the nmonte x nmem resampling draws
  TYPE, BIND(C) :: letkf_lpfgm_cfg_t
     INTEGER(c_int32_t) :: mode, nmonte, rm_sr, dp
     REAL(c_double) :: cov_infl_gm
     TYPE(c_ptr) :: rand
  END TYPE letkf_lpfgm_cfg_t

  INTERFACE
     INTEGER(c_int) FUNCTION letkf_params_default(p, real_kind) BIND(C, name="letkf_params_default")
       IMPORT :: c_int, c_int32_t, letkf_params_t
       TYPE(letkf_params_t), INTENT(INOUT) :: p
       INTEGER(c_int32_t), VALUE :: real_kind
     END FUNCTION letkf_params_default

     INTEGER(c_int) FUNCTION letkf_create(p, h) BIND(C, name="letkf_create")
       IMPORT :: c_int, c_ptr, letkf_params_t
       TYPE(letkf_params_t), INTENT(IN) :: p
       TYPE(c_ptr), INTENT(OUT) :: h
     END FUNCTION letkf_create

     INTEGER(c_int) FUNCTION letkf_destroy(h) BIND(C, name="letkf_destroy")
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: h
     END FUNCTION letkf_destroy

     INTEGER(c_int) FUNCTION letkf_set_params(h, p) BIND(C, name="letkf_set_params")
       IMPORT :: c_int, c_ptr, letkf_params_t
       TYPE(c_ptr), VALUE :: h
       TYPE(letkf_params_t), INTENT(IN) :: p
     END FUNCTION letkf_set_params

     INTEGER(c_int) FUNCTION letkf_da_main(h, grid, obs, fcst3d, fcst2d, anal3d, anal2d) &
          BIND(C, name="letkf_da_main")
       IMPORT :: c_int, c_ptr, letkf_grid_t, letkf_obs_t
       TYPE(c_ptr), VALUE :: h
       TYPE(letkf_grid_t), INTENT(IN) :: grid
       TYPE(letkf_obs_t), INTENT(IN) :: obs
       TYPE(c_ptr), VALUE :: fcst3d, fcst2d, anal3d, anal2d
     END FUNCTION letkf_da_main

     INTEGER(c_int) FUNCTION letkf_da_main_lpfgm(h, grid, obs, cfg, fcst3d, fcst2d, anal3d, anal2d) &
          BIND(C, name="letkf_da_main_lpfgm")
       IMPORT :: c_int, c_ptr, letkf_grid_t, letkf_obs_t, letkf_lpfgm_cfg_t
       TYPE(c_ptr), VALUE :: h
       TYPE(letkf_grid_t), INTENT(IN) :: grid
       TYPE(letkf_obs_t), INTENT(IN) :: obs
       TYPE(letkf_lpfgm_cfg_t), INTENT(IN) :: cfg
       TYPE(c_ptr), VALUE :: fcst3d, fcst2d, anal3d, anal2d
     END FUNCTION letkf_da_main_lpfgm

     TYPE(c_ptr) FUNCTION letkf_last_error() BIND(C, name="letkf_last_error")
       IMPORT :: c_ptr
     END FUNCTION letkf_last_error
  END INTERFACE

CONTAINS

  ! The engine's error text as a Fortran string.
  FUNCTION letkf_last_error_f() RESULT(msg)
    CHARACTER(LEN=512) :: msg
    CHARACTER(KIND=c_char), POINTER :: cs(:)
    INTEGER :: n
    msg = ' '
    CALL c_f_pointer(letkf_last_error(), cs, [512])
    DO n = 1, 512
       IF (cs(n) == c_null_char) EXIT
       msg(n:n) = cs(n)
    END DO
  END FUNCTION letkf_last_error_f

END MODULE letkf_engine_mod

!-----------------------------------------------------------------------
! Drop-in for common_da.f90:22-124.  Compiled together with the driver's
! common_setting module; main.f90:92 calls letkf_da_main_gpu instead of
! da_main with the same four arguments.
!-----------------------------------------------------------------------
MODULE letkf_da_gpu
  USE iso_c_binding
  USE letkf_engine_mod
  IMPLICIT NONE
  PRIVATE
  PUBLIC :: letkf_da_main_gpu

  TYPE(c_ptr), SAVE :: handle = c_null_ptr

CONTAINS

  ! rand (nmem, nmonte): the resampling draws when iswitch_da is 2 or 3 (the
  ! reference draws them with com_rand_seed(nmem, 0, .), common.f90:1066).
  SUBROUTINE letkf_da_main_gpu(fcst3d, fcst2d, anal3d, anal2d, precision, form, rand)
    USE common_setting
    REAL(r_size), INTENT(IN), TARGET :: fcst3d(nij1,nlev,nmem,nv3d), fcst2d(nij1,nmem,nv2d)
    REAL(r_size), INTENT(OUT), TARGET :: anal3d(nij1,nlev,nmem,nv3d), anal2d(nij1,nmem,nv2d)
    INTEGER, INTENT(IN), OPTIONAL :: precision, form
    REAL(c_double), INTENT(IN), TARGET, OPTIONAL :: rand(nmem, nmonte)
    TYPE(letkf_params_t) :: p
    TYPE(letkf_lpfgm_cfg_t) :: lc
    INTEGER(c_int) :: rc

    rc = letkf_params_default(p, INT(r_size, c_int32_t))
    p%nmem = nmem
    p%nlon = nlon
    p%nlat = nlat
    p%nlev = nlev
    p%nij1 = nij1
    p%nv3d = nv3d
    p%nv2d = nv2d
    p%iv3d_t = iv3d_t
    p%iv3d_s = iv3d_s
    p%id_z_obs = id_z_obs
    p%roff = MERGE(1, 0, ROFF)
    p%pi = REAL(pi, c_double)
    p%re = REAL(re, c_double)
    p%cov_infl_mul = REAL(cov_infl_mul, c_double)
    p%alpha_rtpp = REAL(ALPHA_RTPP, c_double)
    p%alpha_rtps = REAL(ALPHA_RTPS, c_double)
    p%sigma_obs = REAL(sigma_obs, c_double)
    p%sigma_obsv = REAL(sigma_obsv, c_double)
    p%dist_zero = REAL(dist_zero, c_double)
    p%dist_zerov = REAL(dist_zerov, c_double)
    p%tmin = REAL(tmin, c_double)
    p%tmax = REAL(tmax, c_double)
    p%smin = REAL(smin, c_double)
    p%smax = REAL(smax, c_double)
    IF (PRESENT(precision)) p%precision = precision
    IF (PRESENT(form)) p%form = form

    IF (.NOT. c_associated(handle)) THEN
       rc = letkf_create(p, handle)
       CALL check(rc)
    ELSE
       ! the handle outlives a call: this call's precision, form and
       ! scalars replace the previous ones (sizes are fixed per run)
       rc = letkf_set_params(handle, p)
       CALL check(rc)
    END IF

    IF (iswitch_da == 2 .OR. iswitch_da == 3) THEN
       IF (.NOT. PRESENT(rand)) THEN
          WRITE(6,'(A)') "***FATAL ERROR: iswitch_da 2/3 needs the resampling draws (rand)"
          STOP
       END IF
       lc%mode = iswitch_da
       lc%nmonte = nmonte
       lc%rm_sr = MERGE(1, 0, RM == "SR")
       lc%dp = MERGE(1, 0, DP)
       lc%cov_infl_gm = REAL(cov_infl_gm, c_double)
       lc%rand = c_loc(rand)
    END IF

    IF (nobs > 0) THEN
       CALL launch(p, nobs, obselm, obslon, obslat, obslev, obserr, obsdep, obshdxf)
    ELSE
       CALL launch(p, 0, [0], [0.0_r_size], [0.0_r_size], [0.0_r_size], [0.0_r_size], &
            & [0.0_r_size], RESHAPE([0.0_r_size], [1, 1]))
    END IF

  CONTAINS

    ! Explicit-shape TARGET dummies give the addresses of the module arrays
    ! (contiguous allocatables are passed without copies).
    SUBROUTINE launch(p, no, elm, olon, olat, olev, oerr, odep, ohdxf)
      TYPE(letkf_params_t), INTENT(IN) :: p
      INTEGER, INTENT(IN) :: no
      INTEGER, INTENT(IN), TARGET :: elm(*)
      REAL(r_size), INTENT(IN), TARGET :: olon(*), olat(*), olev(*), oerr(*), odep(*)
      REAL(r_size), INTENT(IN), TARGET :: ohdxf(*)
      TYPE(letkf_obs_t) :: o
      o%nobs = no
      o%elm = c_loc(elm)
      o%lon = c_loc(olon)
      o%lat = c_loc(olat)
      o%lev = c_loc(olev)
      o%err = c_loc(oerr)
      o%dep = c_loc(odep)
      o%hdxf = c_loc(ohdxf)
      CALL grid_and_run(o, nobsgrd, lon1, lat1, ri1, rj1, dlon1, dlat1, phi1, depth1)
    END SUBROUTINE launch

    SUBROUTINE grid_and_run(o, grd, lo, la, ri, rj, dlo, dla, ph, dep)
      TYPE(letkf_obs_t), INTENT(INOUT) :: o
      INTEGER, INTENT(IN), TARGET :: grd(*)
      REAL(r_size), INTENT(IN), TARGET :: lo(*), la(*), ri(*), rj(*), dlo(*), dla(*), ph(*), dep(*)
      TYPE(letkf_grid_t) :: g
      g%lon1 = c_loc(lo)
      g%lat1 = c_loc(la)
      g%ri1 = c_loc(ri)
      g%rj1 = c_loc(rj)
      g%dlon1 = c_loc(dlo)
      g%dlat1 = c_loc(dla)
      g%phi1 = c_loc(ph)
      g%depth1 = c_loc(dep)
      o%nobsgrd = c_loc(grd)
      IF (iswitch_da == 2 .OR. iswitch_da == 3) THEN
         rc = letkf_da_main_lpfgm(handle, g, o, lc, c_loc(fcst3d), c_loc(fcst2d), c_loc(anal3d), &
              & c_loc(anal2d))
      ELSE
         rc = letkf_da_main(handle, g, o, c_loc(fcst3d), c_loc(fcst2d), c_loc(anal3d), c_loc(anal2d))
      END IF
      CALL check(rc)
    END SUBROUTINE grid_and_run

  END SUBROUTINE letkf_da_main_gpu

  SUBROUTINE check(rc)
    INTEGER(c_int), INTENT(IN) :: rc
    IF (rc /= 0) THEN
       WRITE(6,'(A,I4,2A)') "***FATAL ERROR: LETKF engine status ", rc, ": ", TRIM(letkf_last_error_f())
       STOP
    END IF
  END SUBROUTINE check

END MODULE letkf_da_gpu
