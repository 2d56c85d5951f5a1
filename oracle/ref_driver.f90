! ref_driver.f90 -- build-owned bind(c) shim over the reference's own
! letkf_core (common_letkf.f90:61-210), mtx_eigen (common.f90:992-1050) and
! distance_2p (common.f90:79-101), compiled unmodified from
! /root/reference/letkf/run/src.  The reference fixes nmem = 128 and
! r_size = kind(0.0e0) at common_setting.f90:20,36, so this library is the
! shipped FP32 build at K = 128.  TEST INFRASTRUCTURE ONLY.
module letkf_ref_shim
  use iso_c_binding
  use common_setting, only: r_size, r_dble, nmem, pi, re
  use common, only: distance_2p, mtx_eigen
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

  subroutine ref_constants(pi_out, re_out) bind(c, name="ref_constants")
    real(c_float), intent(out) :: pi_out, re_out
    pi_out = pi
    re_out = re
  end subroutine ref_constants
end module letkf_ref_shim
