! ref_driver.f90 -- build-owned bind(c) shim over the reference's own
! letkf_core (common_letkf.f90:61-210), mtx_eigen (common.f90:992-1050),
! distance_2p (common.f90:79-101), akima_spline (common.f90:276-385),
! Trans_XtoY (common.f90:391-500) and ensemble_mesp (common.f90:505-596),
! compiled unmodified from /root/reference/letkf/run/src.  The reference fixes nmem = 128 and
! r_size = kind(0.0e0) at common_setting.f90:20,36, so this library is the
! shipped FP32 build at K = 128.  TEST INFRASTRUCTURE ONLY.
module letkf_ref_shim
  use iso_c_binding
  use common_setting, only: r_size, r_dble, nmem, pi, re, nlon, nlat, nlev, nv3d, nv2d, &
       lon, lat, depth, fsm, undef, hnosea, id_u_obs, id_v_obs, id_t_obs, id_s_obs, id_z_obs, &
       iv3d_u, iv3d_v, iv3d_t, iv3d_s, iv2d_z, nij1
  use common, only: distance_2p, mtx_eigen, akima_spline, Trans_XtoY, ensemble_mesp
  use common_letkf, only: letkf_core
  implicit none
contains
  integer(c_int) function ref_nmem() bind(c, name="ref_nmem")
    ref_nmem = nmem
  end function ref_nmem

  integer(c_int) function ref_real_kind() bind(c, name="ref_real_kind")
    ref_real_kind = r_size
  end function ref_real_kind

  subroutine ref_letkf_core(nobsl, hdxf, rdiag, rloc, dep, wvec, wmat) &
       bind(c, name="ref_letkf_core")
    integer(c_int), value :: nobsl
    real(c_float), intent(in) :: hdxf(nobsl, nmem), rdiag(nobsl)
    real(c_float), intent(in) :: rloc(nobsl), dep(nobsl)
    real(c_float), intent(out) :: wvec(nmem), wmat(nmem, nmem)
    call letkf_core(nobsl, hdxf, rdiag, rloc, dep, wvec, wmat)
  end subroutine ref_letkf_core

  subroutine ref_mtx_eigen(n, a, eval, evec) bind(c, name="ref_mtx_eigen")
    integer(c_int), value :: n
    real(c_float), intent(in) :: a(n, n)
    real(c_double), intent(out) :: eval(n), evec(n, n)
    integer :: np
    call mtx_eigen(n, a, eval, evec, np)
  end subroutine ref_mtx_eigen

  subroutine ref_distance(lona, lata, lonb, latb, dist) &
       bind(c, name="ref_distance")
    real(c_float), value :: lona, lata, lonb, latb
    real(c_float), intent(out) :: dist
    call distance_2p(lona, lata, lonb, latb, dist)
  end subroutine ref_distance

  ! The module's grid arrays (common_setting.f90:155-161) are set from the
  ! arguments, then Trans_XtoY runs on the reference's fixed grid size.
  subroutine ref_trans_xtoy(n, elm, idx, idy, lono, lato, levo, v3d, v2d, lon_in, lat_in, &
       depth_in, fsm_in, hx) bind(c, name="ref_trans_xtoy")
    integer(c_int), value :: n
    integer(c_int), intent(in) :: elm(n), idx(n), idy(n)
    real(c_float), intent(in) :: lono(n), lato(n), levo(n)
    real(c_float), intent(in) :: v3d(nlon, nlat, nlev, nv3d), v2d(nlon, nlat, nv2d)
    real(c_float), intent(in) :: lon_in(nlon, nlat), lat_in(nlon, nlat)
    real(c_float), intent(in) :: depth_in(nlon, nlat, nlev), fsm_in(nlon, nlat)
    real(c_float), intent(out) :: hx(n)
    lon = lon_in
    lat = lat_in
    depth = depth_in
    fsm = fsm_in
    call Trans_XtoY(n, elm, idx, idy, lono, lato, levo, v3d, v2d, hx)
  end subroutine ref_trans_xtoy

  subroutine ref_akima(ndim, x, y, x5, y5) bind(c, name="ref_akima")
    integer(c_int), value :: ndim
    real(c_float), intent(in) :: x(ndim), y(ndim)
    real(c_float), value :: x5
    real(c_float), intent(out) :: y5
    real(c_float) :: xt(1), yt(1)
    xt(1) = x5
    call akima_spline(ndim, x, y, 1, xt, yt)
    y5 = yt(1)
  end subroutine ref_akima

  ! nlon nlat nlev nv3d nv2d, element ids u v t s z, slots u v t s z
  subroutine ref_xtoy_consts(dims, vals) bind(c, name="ref_xtoy_consts")
    integer(c_int), intent(out) :: dims(15)
    real(c_float), intent(out) :: vals(2)
    dims = [nlon, nlat, nlev, nv3d, nv2d, id_u_obs, id_v_obs, id_t_obs, id_s_obs, id_z_obs, &
         iv3d_u, iv3d_v, iv3d_t, iv3d_s, iv2d_z]
    vals = [hnosea, undef]
  end subroutine ref_xtoy_consts

  subroutine ref_constants(pi_out, re_out) bind(c, name="ref_constants")
    real(c_float), intent(out) :: pi_out, re_out
    pi_out = pi
    re_out = re
  end subroutine ref_constants
  ! ensemble_mesp on the reference's fixed nlev, nmem, nv3d and nv2d, with
  ! the module's nij1 (common_setting.f90:168) set from the argument.
  subroutine ref_ensemble_mesp(n, v3d, v2d, v3dm, v3ds, v2dm, v2ds) bind(c, name="ref_ensemble_mesp")
    integer(c_int), value :: n
    real(c_float), intent(in) :: v3d(n, nlev, nmem, nv3d), v2d(n, nmem, nv2d)
    real(c_float), intent(out) :: v3dm(n, nlev, nv3d), v3ds(n, nlev, nv3d)
    real(c_float), intent(out) :: v2dm(n, nv2d), v2ds(n, nv2d)
    nij1 = n
    call ensemble_mesp(v3d, v2d, v3dm, v3ds, v2dm, v2ds)
  end subroutine ref_ensemble_mesp

end module letkf_ref_shim
