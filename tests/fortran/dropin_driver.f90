! dropin_driver.f90 -- a Fortran caller of the engine's drop-in for da_main,
! test infrastructure only.  It USEs the reference's own common_setting
! module (letkf/run/src/common_setting.f90, compiled unmodified: REAL(4)
! r_size, nmem 128, nlon x nlat 1442 x 562, nlev 75, nv3d 4, nv2d 3), fills
! its module state -- the local grid and the sorted observation tables that
! da_main reads (common_setting.f90:168-198) -- from a stream file written
! by the test, and calls letkf_da_main_gpu twice as main.f90:92 calls
! da_main: first with the defaults (FP64 Gram, plain form), then with the
! FP32 Gram on MFMA and the batched form on the same handle.
!
!   letkf_dropin <input> <output>
! input : int32 nij1, nobs; lon1 lat1 ri1 rj1 dlon1 dlat1 phi1 (nij1);
!         depth1 (nij1,nlev); int32 nobsgrd (nlon,nlat); int32 obselm (nobs);
!         obslon obslat obslev obserr obsdep (nobs); obshdxf (nobs,nmem);
!         fcst3d (nij1,nlev,nmem,nv3d); fcst2d (nij1,nmem,nv2d)
! output: anal3d anal2d of the first call, then of the second
program letkf_dropin
  use common_setting
  use letkf_da_gpu
  implicit none
  character(len=1024) :: fin, fout
  real(r_size), allocatable :: f3(:,:,:,:), f2(:,:,:), a3(:,:,:,:), a2(:,:,:), b3(:,:,:,:), b2(:,:,:)
  integer :: u

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old', action='read')
  read(u) nij1, nobs
  allocate(lon1(nij1), lat1(nij1), ri1(nij1), rj1(nij1), dlon1(nij1), dlat1(nij1), phi1(nij1))
  allocate(depth1(nij1, nlev))
  read(u) lon1, lat1, ri1, rj1, dlon1, dlat1, phi1
  read(u) depth1
  read(u) nobsgrd
  allocate(obselm(nobs), obslon(nobs), obslat(nobs), obslev(nobs), obserr(nobs), obsdep(nobs))
  allocate(obshdxf(nobs, nmem))
  read(u) obselm
  read(u) obslon, obslat, obslev, obserr, obsdep
  read(u) obshdxf
  allocate(f3(nij1, nlev, nmem, nv3d), f2(nij1, nmem, nv2d))
  allocate(a3(nij1, nlev, nmem, nv3d), a2(nij1, nmem, nv2d), b3(nij1, nlev, nmem, nv3d), b2(nij1, nmem, nv2d))
  read(u) f3
  read(u) f2
  close(u)

  call letkf_da_main_gpu(f3, f2, a3, a2)
  call letkf_da_main_gpu(f3, f2, b3, b2, precision=1, form=1)

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
  write(u) a3, a2, b3, b2
  close(u)
  write(6, '(A,I0,A,I0,A)') "letkf_dropin: analysed ", nij1 * nlev, " points with ", nobs, " obs twice"
end program letkf_dropin
