﻿!mod$ v1 sum:f584c75948b673c5
!need$ 0bde2ac47243ead2 i iso_c_binding
!need$ 2dcd5c82c7cba8f9 n common
!need$ 7d9124a989da7208 n common_letkf
!need$ 4ef8e355501b7eb4 n common_setting
module letkf_ref_shim
use common_setting,only:r_size
use common_setting,only:r_dble
use common_setting,only:nmem
use common_setting,only:pi
use common_setting,only:re
use common_setting,only:nlon
use common_setting,only:nlat
use common_setting,only:nlev
use common_setting,only:nv3d
use common_setting,only:nv2d
use common_setting,only:lon
use common_setting,only:lat
use common_setting,only:depth
use common_setting,only:fsm
use common_setting,only:undef
use common_setting,only:hnosea
use common_setting,only:id_u_obs
use common_setting,only:id_v_obs
use common_setting,only:id_t_obs
use common_setting,only:id_s_obs
use common_setting,only:id_z_obs
use common_setting,only:iv3d_u
use common_setting,only:iv3d_v
use common_setting,only:iv3d_t
use common_setting,only:iv3d_s
use common_setting,only:iv2d_z
use common_setting,only:nij1
use common,only:distance_2p
use common,only:mtx_eigen
use common,only:akima_spline
use common,only:trans_xtoy
use common,only:ensemble_mesp
use common_letkf,only:letkf_core
use,intrinsic::iso_c_binding,only:c_associated
use,intrinsic::iso_c_binding,only:c_funloc
use,intrinsic::iso_c_binding,only:c_funptr
use,intrinsic::iso_c_binding,only:c_f_pointer
use,intrinsic::iso_c_binding,only:c_loc
use,intrinsic::iso_c_binding,only:c_null_funptr
use,intrinsic::iso_c_binding,only:c_null_ptr
use,intrinsic::iso_c_binding,only:c_ptr
use,intrinsic::iso_c_binding,only:c_sizeof
use,intrinsic::iso_c_binding,only:operator(==)
use,intrinsic::iso_c_binding,only:operator(/=)
use,intrinsic::iso_c_binding,only:c_int8_t
use,intrinsic::iso_c_binding,only:c_int16_t
use,intrinsic::iso_c_binding,only:c_int32_t
use,intrinsic::iso_c_binding,only:c_int64_t
use,intrinsic::iso_c_binding,only:c_int128_t
use,intrinsic::iso_c_binding,only:c_int
use,intrinsic::iso_c_binding,only:c_short
use,intrinsic::iso_c_binding,only:c_long
use,intrinsic::iso_c_binding,only:c_long_long
use,intrinsic::iso_c_binding,only:c_signed_char
use,intrinsic::iso_c_binding,only:c_size_t
use,intrinsic::iso_c_binding,only:c_intmax_t
use,intrinsic::iso_c_binding,only:c_intptr_t
use,intrinsic::iso_c_binding,only:c_ptrdiff_t
use,intrinsic::iso_c_binding,only:c_int_least8_t
use,intrinsic::iso_c_binding,only:c_int_fast8_t
use,intrinsic::iso_c_binding,only:c_int_least16_t
use,intrinsic::iso_c_binding,only:c_int_fast16_t
use,intrinsic::iso_c_binding,only:c_int_least32_t
use,intrinsic::iso_c_binding,only:c_int_fast32_t
use,intrinsic::iso_c_binding,only:c_int_least64_t
use,intrinsic::iso_c_binding,only:c_int_fast64_t
use,intrinsic::iso_c_binding,only:c_int_least128_t
use,intrinsic::iso_c_binding,only:c_int_fast128_t
use,intrinsic::iso_c_binding,only:c_float
use,intrinsic::iso_c_binding,only:c_double
use,intrinsic::iso_c_binding,only:c_long_double
use,intrinsic::iso_c_binding,only:c_float_complex
use,intrinsic::iso_c_binding,only:c_double_complex
use,intrinsic::iso_c_binding,only:c_long_double_complex
use,intrinsic::iso_c_binding,only:c_bool
use,intrinsic::iso_c_binding,only:c_char
use,intrinsic::iso_c_binding,only:c_null_char
use,intrinsic::iso_c_binding,only:c_alert
use,intrinsic::iso_c_binding,only:c_backspace
use,intrinsic::iso_c_binding,only:c_form_feed
use,intrinsic::iso_c_binding,only:c_new_line
use,intrinsic::iso_c_binding,only:c_carriage_return
use,intrinsic::iso_c_binding,only:c_horizontal_tab
use,intrinsic::iso_c_binding,only:c_vertical_tab
use,intrinsic::iso_c_binding,only:c_float128
use,intrinsic::iso_c_binding,only:c_float128_complex
use,intrinsic::iso_c_binding,only:c_uint8_t
use,intrinsic::iso_c_binding,only:c_uint16_t
use,intrinsic::iso_c_binding,only:c_uint32_t
use,intrinsic::iso_c_binding,only:c_uint64_t
use,intrinsic::iso_c_binding,only:c_uint128_t
use,intrinsic::iso_c_binding,only:c_unsigned_char
use,intrinsic::iso_c_binding,only:c_unsigned_short
use,intrinsic::iso_c_binding,only:c_unsigned
use,intrinsic::iso_c_binding,only:c_unsigned_long
use,intrinsic::iso_c_binding,only:c_unsigned_long_long
use,intrinsic::iso_c_binding,only:c_uintmax_t
use,intrinsic::iso_c_binding,only:c_uint_fast8_t
use,intrinsic::iso_c_binding,only:c_uint_fast16_t
use,intrinsic::iso_c_binding,only:c_uint_fast32_t
use,intrinsic::iso_c_binding,only:c_uint_fast64_t
use,intrinsic::iso_c_binding,only:c_uint_fast128_t
use,intrinsic::iso_c_binding,only:c_uint_least8_t
use,intrinsic::iso_c_binding,only:c_uint_least16_t
use,intrinsic::iso_c_binding,only:c_uint_least32_t
use,intrinsic::iso_c_binding,only:c_uint_least64_t
use,intrinsic::iso_c_binding,only:c_uint_least128_t
use,intrinsic::iso_c_binding,only:c_f_procpointer
contains
function ref_nmem() bind(c,name="ref_nmem")
integer(4)::ref_nmem
end
function ref_real_kind() bind(c,name="ref_real_kind")
integer(4)::ref_real_kind
end
subroutine ref_letkf_core(nobsl,hdxf,rdiag,rloc,dep,wvec,wmat) bind(c,name="ref_letkf_core")
integer(4),value::nobsl
real(4),intent(in)::hdxf(1_8:int(nobsl,kind=8),1_8:128_8)
real(4),intent(in)::rdiag(1_8:int(nobsl,kind=8))
real(4),intent(in)::rloc(1_8:int(nobsl,kind=8))
real(4),intent(in)::dep(1_8:int(nobsl,kind=8))
real(4),intent(out)::wvec(1_8:128_8)
real(4),intent(out)::wmat(1_8:128_8,1_8:128_8)
end
subroutine ref_mtx_eigen(n,a,eval,evec) bind(c,name="ref_mtx_eigen")
integer(4),value::n
real(4),intent(in)::a(1_8:int(n,kind=8),1_8:int(n,kind=8))
real(8),intent(out)::eval(1_8:int(n,kind=8))
real(8),intent(out)::evec(1_8:int(n,kind=8),1_8:int(n,kind=8))
end
subroutine ref_distance(lona,lata,lonb,latb,dist) bind(c,name="ref_distance")
real(4),value::lona
real(4),value::lata
real(4),value::lonb
real(4),value::latb
real(4),intent(out)::dist
end
subroutine ref_trans_xtoy(n,elm,idx,idy,lono,lato,levo,v3d,v2d,lon_in,lat_in,depth_in,fsm_in,hx) bind(c,name="ref_trans_xtoy")
integer(4),value::n
integer(4),intent(in)::elm(1_8:int(n,kind=8))
integer(4),intent(in)::idx(1_8:int(n,kind=8))
integer(4),intent(in)::idy(1_8:int(n,kind=8))
real(4),intent(in)::lono(1_8:int(n,kind=8))
real(4),intent(in)::lato(1_8:int(n,kind=8))
real(4),intent(in)::levo(1_8:int(n,kind=8))
real(4),intent(in)::v3d(1_8:1442_8,1_8:562_8,1_8:75_8,1_8:4_8)
real(4),intent(in)::v2d(1_8:1442_8,1_8:562_8,1_8:3_8)
real(4),intent(in)::lon_in(1_8:1442_8,1_8:562_8)
real(4),intent(in)::lat_in(1_8:1442_8,1_8:562_8)
real(4),intent(in)::depth_in(1_8:1442_8,1_8:562_8,1_8:75_8)
real(4),intent(in)::fsm_in(1_8:1442_8,1_8:562_8)
real(4),intent(out)::hx(1_8:int(n,kind=8))
end
subroutine ref_akima(ndim,x,y,x5,y5) bind(c,name="ref_akima")
integer(4),value::ndim
real(4),intent(in)::x(1_8:int(ndim,kind=8))
real(4),intent(in)::y(1_8:int(ndim,kind=8))
real(4),value::x5
real(4),intent(out)::y5
end
subroutine ref_xtoy_consts(dims,vals) bind(c,name="ref_xtoy_consts")
integer(4),intent(out)::dims(1_8:15_8)
real(4),intent(out)::vals(1_8:2_8)
end
subroutine ref_constants(pi_out,re_out) bind(c,name="ref_constants")
real(4),intent(out)::pi_out
real(4),intent(out)::re_out
end
subroutine ref_ensemble_mesp(n,v3d,v2d,v3dm,v3ds,v2dm,v2ds) bind(c,name="ref_ensemble_mesp")
integer(4),value::n
real(4),intent(in)::v3d(1_8:int(n,kind=8),1_8:75_8,1_8:128_8,1_8:4_8)
real(4),intent(in)::v2d(1_8:int(n,kind=8),1_8:128_8,1_8:3_8)
real(4),intent(out)::v3dm(1_8:int(n,kind=8),1_8:75_8,1_8:4_8)
real(4),intent(out)::v3ds(1_8:int(n,kind=8),1_8:75_8,1_8:4_8)
real(4),intent(out)::v2dm(1_8:int(n,kind=8),1_8:3_8)
real(4),intent(out)::v2ds(1_8:int(n,kind=8),1_8:3_8)
end
end
