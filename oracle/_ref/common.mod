﻿!mod$ v1 sum:2dcd5c82c7cba8f9
module common
contains
subroutine timer_init()
end
subroutine timer(section_name)
character(*,1),intent(in)::section_name
end
subroutine timer_end()
end
subroutine distance_2p(lona,lata,lonb,latb,dist)
real(4),intent(in)::lona
real(4),intent(in)::lata
real(4),intent(in)::lonb
real(4),intent(in)::latb
real(4),intent(out)::dist
end
subroutine obs_id_uniform(no,lono,lato,idx,idy)
integer(4),intent(in)::no
real(4),intent(in)::lono(1_8:int(no,kind=8))
real(4),intent(in)::lato(1_8:int(no,kind=8))
integer(4),intent(out)::idx(1_8:int(no,kind=8))
integer(4),intent(out)::idy(1_8:int(no,kind=8))
end
subroutine obs_id_nonuniform(no,lono,lato,idx,idy)
integer(4),intent(in)::no
real(4),intent(in)::lono(1_8:int(no,kind=8))
real(4),intent(in)::lato(1_8:int(no,kind=8))
integer(4),intent(out)::idx(1_8:int(no,kind=8))
integer(4),intent(out)::idy(1_8:int(no,kind=8))
end
subroutine bilinear_interpolation(x0,x1,y0,y1,x,y,f00,f10,f01,f11,fxy)
real(4),intent(in)::x0
real(4),intent(in)::x1
real(4),intent(in)::y0
real(4),intent(in)::y1
real(4),intent(in)::x
real(4),intent(in)::y
real(4),intent(in)::f00
real(4),intent(in)::f10
real(4),intent(in)::f01
real(4),intent(in)::f11
real(4),intent(out)::fxy
end
subroutine akima_spline(ndim,x,y,n,x5,y5)
integer(4),intent(in)::ndim
real(4),intent(in)::x(1_8:int(ndim,kind=8))
real(4),intent(in)::y(1_8:int(ndim,kind=8))
integer(4),intent(in)::n
real(4),intent(in)::x5(1_8:int(n,kind=8))
real(4),intent(out)::y5(1_8:int(n,kind=8))
end
subroutine trans_xtoy(n,elm,idx,idy,lono,lato,levo,v3d,v2d,hx)
integer(4),intent(in)::n
integer(4),intent(in)::elm(1_8:int(n,kind=8))
integer(4),intent(in)::idx(1_8:int(n,kind=8))
integer(4),intent(in)::idy(1_8:int(n,kind=8))
real(4),intent(in)::lono(1_8:int(n,kind=8))
real(4),intent(in)::lato(1_8:int(n,kind=8))
real(4),intent(in)::levo(1_8:int(n,kind=8))
real(4),intent(in)::v3d(1_8:1442_8,1_8:562_8,1_8:75_8,1_8:4_8)
real(4),intent(in)::v2d(1_8:1442_8,1_8:562_8,1_8:3_8)
real(4),intent(out)::hx(1_8:int(n,kind=8))
end
subroutine ensemble_mesp(v3d,v2d,v3dm,v3ds,v2dm,v2ds)
use common_setting,only:nij1
real(4),intent(in)::v3d(1_8:int(nij1,kind=8),1_8:75_8,1_8:128_8,1_8:4_8)
real(4),intent(in)::v2d(1_8:int(nij1,kind=8),1_8:128_8,1_8:3_8)
real(4),intent(out)::v3dm(1_8:int(nij1,kind=8),1_8:75_8,1_8:4_8)
real(4),intent(out)::v3ds(1_8:int(nij1,kind=8),1_8:75_8,1_8:4_8)
real(4),intent(out)::v2dm(1_8:int(nij1,kind=8),1_8:3_8)
real(4),intent(out)::v2ds(1_8:int(nij1,kind=8),1_8:3_8)
end
subroutine monit_dep(no,elm,dep,qc)
integer(4),intent(in)::no
integer(4),intent(in)::elm(1_8:int(no,kind=8))
real(4),intent(in)::dep(1_8:int(no,kind=8))
integer(4),intent(in)::qc(1_8:int(no,kind=8))
end
subroutine monit_mean(name,hx)
use common_setting,only:nobs
character(4_4,1),intent(in)::name
real(4),intent(in)::hx(1_8:int(nobs,kind=8))
end
subroutine monit_sprd(name,hx)
use common_setting,only:nobs
character(4_4,1),intent(in)::name
real(4),intent(in)::hx(1_8:int(nobs,kind=8))
end
subroutine mtx_eigen(n,a,eval,evec,np)
integer(4),intent(in)::n
real(4),intent(in)::a(1_8:int(n,kind=8),1_8:int(n,kind=8))
real(8),intent(out)::eval(1_8:int(n,kind=8))
real(8),intent(out)::evec(1_8:int(n,kind=8),1_8:int(n,kind=8))
integer(4),intent(out)::np
end
subroutine com_rand_seed(ndim,iseed,rand)
integer(4),intent(in)::ndim
integer(4),intent(in)::iseed
real(8),intent(out)::rand(1_8:int(ndim,kind=8))
end
recursive subroutine quick_sort_asend(n,dat,idx,idx_s,idx_e)
integer(4),intent(in)::n
real(8),intent(inout)::dat(1_8:int(n,kind=8))
integer(4),intent(inout)::idx(1_8:int(n,kind=8))
integer(4),intent(in)::idx_s
integer(4),intent(in)::idx_e
end
end
