﻿!mod$ v1 sum:7d9124a989da7208
module common_letkf
contains
subroutine letkf_core_noobs(wvec,wmat)
real(4),intent(out)::wvec(1_8:128_8)
real(4),intent(out)::wmat(1_8:128_8,1_8:128_8)
end
subroutine letkf_core(nobsl,hdxf,rdiag,rloc,dep,wvec,wmat)
integer(4),intent(in)::nobsl
real(4),intent(in)::hdxf(1_8:int(nobsl,kind=8),1_8:128_8)
real(4),intent(in)::rdiag(1_8:int(nobsl,kind=8))
real(4),intent(in)::rloc(1_8:int(nobsl,kind=8))
real(4),intent(in)::dep(1_8:int(nobsl,kind=8))
real(4),intent(out)::wvec(1_8:128_8)
real(4),intent(out)::wmat(1_8:128_8,1_8:128_8)
end
end
