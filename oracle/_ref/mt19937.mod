﻿!mod$ v1 sum:66cd4fcb952ad95f
module mt19937
intrinsic::bit_size
private::bit_size
integer(4),parameter,private::intg=4_4
intrinsic::selected_int_kind
private::selected_int_kind
integer(4),parameter,private::long=8_4
integer(4),parameter,private::flot=4_4
intrinsic::selected_real_kind
private::selected_real_kind
integer(4),parameter,private::dobl=8_4
integer(4),parameter::wi=4_4
integer(4),parameter::wl=8_4
integer(4),parameter::wr=8_4
integer(4),parameter,private::n=624_4
integer(4),parameter,private::m=397_4
integer(4),parameter,private::hbs=16_4
integer(4),parameter,private::qbs=8_4
integer(4),parameter,private::tbs=24_4
integer(4),private::mt(1_8:624_8)
logical(4),private::mtinit
integer(4),private::mti
private::uiadd
private::uisub
private::uimlt
contains
elemental function uiadd(a,b) result(c)
integer(4),intent(in)::a
integer(4),intent(in)::b
integer(4)::c
end
elemental function uisub(a,b) result(c)
integer(4),intent(in)::a
integer(4),intent(in)::b
integer(4)::c
end
elemental function uimlt(a,b) result(c)
integer(4),intent(in)::a
integer(4),intent(in)::b
integer(4)::c
end
subroutine init_genrand(s)
integer(4),intent(in)::s
end
subroutine init_by_array(init_key)
integer(4),intent(in)::init_key(:)
end
function genrand_int32() result(y)
integer(4)::y
end
function genrand_int31() result(i)
integer(4)::i
end
function genrand_real1() result(r)
real(8)::r
end
function genrand_real2() result(r)
real(8)::r
end
function genrand_real3() result(r)
real(8)::r
end
function genrand_res53() result(r)
real(8)::r
end
end
