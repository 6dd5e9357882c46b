﻿!mod$ v1 sum:9ef4185c7839806c
!need$ 0bde2ac47243ead2 i iso_c_binding
module opensps_hip_iface
use,intrinsic::__fortran_builtins,only:__builtin_c_ptr
use,intrinsic::iso_c_binding,only:c_int
use,intrinsic::iso_c_binding,only:c_double
use,intrinsic::iso_c_binding,only:c_ptr
use,intrinsic::iso_c_binding,only:c_size_t
use,intrinsic::iso_c_binding,only:c_char
use,intrinsic::iso_c_binding,only:c_null_ptr
private::__builtin_c_ptr
private::c_int
private::c_double
private::c_ptr
private::c_size_t
private::c_char
private::c_null_ptr
integer(4),parameter::hip_fp64=0_4
integer(4),parameter::hip_mixed=1_4
type,bind(c)::hip_batch_t
integer(4)::n=0_4
integer(4)::precision=0_4
integer(4)::split_parts=0_4
integer(4)::use_ops=0_4
type(c_ptr)::prm=__builtin_c_ptr(__address=0_8)
type(c_ptr)::ssp_weights=__builtin_c_ptr(__address=0_8)
type(c_ptr)::mags=__builtin_c_ptr(__address=0_8)
type(c_ptr)::spec=__builtin_c_ptr(__address=0_8)
type(c_ptr)::mass=__builtin_c_ptr(__address=0_8)
type(c_ptr)::lbol=__builtin_c_ptr(__address=0_8)
type(c_ptr)::fband=__builtin_c_ptr(__address=0_8)
type(c_ptr)::young=__builtin_c_ptr(__address=0_8)
type(c_ptr)::old=__builtin_c_ptr(__address=0_8)
type(c_ptr)::idx=__builtin_c_ptr(__address=0_8)
type(c_ptr)::weights=__builtin_c_ptr(__address=0_8)
type(c_ptr)::op_count=__builtin_c_ptr(__address=0_8)
type(c_ptr)::ssp_weights_z=__builtin_c_ptr(__address=0_8)
end type
type,bind(c)::hip_grad_t
type(hip_batch_t)::fwd
type(c_ptr)::mags_bar=__builtin_c_ptr(__address=0_8)
type(c_ptr)::spec_bar=__builtin_c_ptr(__address=0_8)
type(c_ptr)::mass_bar=__builtin_c_ptr(__address=0_8)
type(c_ptr)::lbol_bar=__builtin_c_ptr(__address=0_8)
type(c_ptr)::params_bar=__builtin_c_ptr(__address=0_8)
type(c_ptr)::weights_bar=__builtin_c_ptr(__address=0_8)
type(c_ptr)::weights_z_bar=__builtin_c_ptr(__address=0_8)
end type
interface
function opensps_hip_create(device,ctx) bind(c,name="opensps_hip_create")
import::c_ptr
integer(4),value::device
type(c_ptr),intent(out)::ctx
integer(4)::opensps_hip_create
end
end interface
interface
function opensps_hip_destroy(ctx) bind(c,name="opensps_hip_destroy")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::opensps_hip_destroy
end
end interface
interface
function opensps_hip_set_grid(ctx,nspec,lam,nt,log_age,nz,logz,spec,mass,lbol) bind(c,name="opensps_hip_set_grid")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::nspec
real(8),intent(in)::lam(1_8:*)
integer(4),value::nt
real(8),intent(in)::log_age(1_8:*)
integer(4),value::nz
real(8),intent(in)::logz(1_8:*)
real(8),intent(in)::spec(1_8:*)
real(8),intent(in)::mass(1_8:*)
real(8),intent(in)::lbol(1_8:*)
integer(4)::opensps_hip_set_grid
end
end interface
interface
function opensps_hip_set_bands(ctx,nbands,bands) bind(c,name="opensps_hip_set_bands")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::nbands
real(8),intent(in)::bands(1_8:*)
integer(4)::opensps_hip_set_bands
end
end interface
interface
function opensps_hip_set_mag_mask(ctx,nbands,mask) bind(c,name="opensps_hip_set_mag_mask")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::nbands
integer(4),intent(in)::mask(1_8:*)
integer(4)::opensps_hip_set_mag_mask
end
end interface
interface
function opensps_hip_set_cosmology(ctx,n,z,dl) bind(c,name="opensps_hip_set_cosmology")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::n
real(8),intent(in)::z(1_8:*)
real(8),intent(in)::dl(1_8:*)
integer(4)::opensps_hip_set_cosmology
end
end interface
interface
function opensps_hip_set_options(ctx,add_igm,smooth_fast) bind(c,name="opensps_hip_set_options")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::add_igm
integer(4),value::smooth_fast
integer(4)::opensps_hip_set_options
end
end interface
interface
function opensps_hip_csp_batch(ctx,args) bind(c,name="opensps_hip_csp_batch")
import::c_ptr
import::hip_batch_t
type(c_ptr),value::ctx
type(hip_batch_t),intent(in)::args
integer(4)::opensps_hip_csp_batch
end
end interface
interface
function opensps_hip_csp_batch_grad(ctx,args) bind(c,name="opensps_hip_csp_batch_grad")
import::c_ptr
import::hip_grad_t
type(c_ptr),value::ctx
type(hip_grad_t),intent(in)::args
integer(4)::opensps_hip_csp_batch_grad
end
end interface
interface
function opensps_hip_sync(ctx) bind(c,name="opensps_hip_sync")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::opensps_hip_sync
end
end interface
interface
function opensps_hip_malloc(ptr,bytes) bind(c,name="opensps_hip_malloc")
import::c_ptr
type(c_ptr),intent(out)::ptr
integer(8),value::bytes
integer(4)::opensps_hip_malloc
end
end interface
interface
function opensps_hip_free(ptr) bind(c,name="opensps_hip_free")
import::c_ptr
type(c_ptr),value::ptr
integer(4)::opensps_hip_free
end
end interface
interface
function opensps_hip_memcpy_h2d(dst,src,bytes) bind(c,name="opensps_hip_memcpy_h2d")
import::c_ptr
type(c_ptr),value::dst
type(c_ptr),value::src
integer(8),value::bytes
integer(4)::opensps_hip_memcpy_h2d
end
end interface
interface
function opensps_hip_memcpy_d2h(dst,src,bytes) bind(c,name="opensps_hip_memcpy_d2h")
import::c_ptr
type(c_ptr),value::dst
type(c_ptr),value::src
integer(8),value::bytes
integer(4)::opensps_hip_memcpy_d2h
end
end interface
interface
function opensps_hip_last_error(ctx,msg,len) bind(c,name="opensps_hip_last_error")
import::c_ptr
type(c_ptr),value::ctx
character(1_8,1),intent(out)::msg(1_8:*)
integer(4),value::len
integer(4)::opensps_hip_last_error
end
end interface
interface
function opensps_hip_set_profiling(ctx,on) bind(c,name="opensps_hip_set_profiling")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::on
integer(4)::opensps_hip_set_profiling
end
end interface
interface
function opensps_hip_kernel_times(ctx,ms,launches,reset) bind(c,name="opensps_hip_kernel_times")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(out)::ms(1_8:4_8)
integer(4),intent(out)::launches(1_8:4_8)
integer(4),value::reset
integer(4)::opensps_hip_kernel_times
end
end interface
interface
function opensps_hip_device_count(n) bind(c,name="opensps_hip_device_count")
integer(4),intent(out)::n
integer(4)::opensps_hip_device_count
end
end interface
end
