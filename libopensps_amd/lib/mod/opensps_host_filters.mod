﻿!mod$ v1 sum:b579c837613f74aa
!need$ f5118467a58a718d n opensps_host_types
module opensps_host_filters
use opensps_host_types,only:wp
use opensps_host_types,only:safe_floor
private::wp
private::safe_floor
integer(4),parameter::filter_ok=0_4
integer(4),parameter::filter_err_few_points=1_4
integer(4),parameter::filter_err_nonfinite=2_4
integer(4),parameter::filter_err_order=3_4
integer(4),parameter::filter_err_argument=4_4
integer(4),parameter::filter_err_coverage=5_4
integer(4),parameter::filter_err_dead=6_4
integer(4),parameter::flag_truncated=1_4
integer(4),parameter::flag_undersampled=2_4
integer(4),parameter::flag_dead=4_4
real(8),parameter,private::min_coverage=9.9899999999999999911182158029987476766109466552734375e-1_8
integer(4),parameter,private::min_fwhm_points=5_4
type::filter_curve_t
character(64_4,1)::name="                                                                "
real(8),allocatable::lam(:)
real(8),allocatable::resp(:)
integer(4)::detector=0_4
logical(4)::has_cal=.false._4
real(8)::cal_lref=0._8
real(8)::cal_expo=0._8
end type
type::filter_set_t
logical(4)::used=.false._4
integer(4)::n=0_4
type(filter_curve_t),allocatable::f(:)
end type
type(filter_set_t),allocatable,save::fsets(:)
private::curve_clean
private::curve_integral
private::interp_lin
private::interval_of
private::resample_point
private::cell_of
private::resample_hat
contains
function fset_valid(h)
integer(4),intent(in)::h
logical(4)::fset_valid
end
function fset_create() result(h)
integer(4)::h
end
subroutine fset_free(h)
integer(4),intent(in)::h
end
subroutine fset_add(h,name,lam,resp,detector,cal_lref,cal_expo,status)
integer(4),intent(in)::h
character(*,1),intent(in)::name
real(8),intent(in)::lam(:)
real(8),intent(in)::resp(:)
integer(4),intent(in)::detector
real(8),intent(in)::cal_lref
real(8),intent(in)::cal_expo
integer(4),intent(out)::status
end
subroutine curve_clean(lam,resp,lc,rc,nc,status)
real(8),intent(in)::lam(:)
real(8),intent(in)::resp(:)
real(8),allocatable,intent(out)::lc(:)
real(8),allocatable,intent(out)::rc(:)
integer(4),intent(out)::nc
integer(4),intent(out)::status
end
pure function curve_integral(lc,r,nc,lo,hi,p) result(s)
real(8),intent(in)::lc(:)
real(8),intent(in)::r(:)
integer(4),intent(in)::nc
real(8),intent(in)::lo
real(8),intent(in)::hi
integer(4),intent(in)::p
real(8)::s
end
pure function trapezoid(x,y) result(s)
real(8),intent(in)::x(:)
real(8),intent(in)::y(:)
real(8)::s
end
pure function interp_lin(x,y,n,xo) result(v)
real(8),intent(in)::x(:)
real(8),intent(in)::y(:)
integer(4),intent(in)::n
real(8),intent(in)::xo
real(8)::v
end
pure function interval_of(x,n,v) result(i)
real(8),intent(in)::x(:)
integer(4),intent(in)::n
real(8),intent(in)::v
integer(4)::i
end
subroutine resample_point(lc,r,nc,grid,band)
real(8),intent(in)::lc(:)
real(8),intent(in)::r(:)
integer(4),intent(in)::nc
real(8),intent(in)::grid(:)
real(8),intent(out)::band(:)
end
pure function cell_of(grid,x) result(j)
real(8),intent(in)::grid(:)
real(8),intent(in)::x
integer(4)::j
end
subroutine resample_hat(lc,r,nc,grid,band)
real(8),intent(in)::lc(:)
real(8),intent(in)::r(:)
integer(4),intent(in)::nc
real(8),intent(in)::grid(:)
real(8),intent(out)::band(:)
end
subroutine filters_build(h,grid,mode,policy,bands,flags,status,bad)
integer(4),intent(in)::h
real(8),intent(in)::grid(:)
integer(4),intent(in)::mode
integer(4),intent(in)::policy
real(8),intent(out)::bands(:,:)
integer(4),intent(out)::flags(:)
integer(4),intent(out)::status
integer(4),intent(out)::bad
end
function band_pivot(grid,band) result(leff)
real(8),intent(in)::grid(:)
real(8),intent(in)::band(:)
real(8)::leff
end
end
