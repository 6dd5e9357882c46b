﻿!mod$ v1 sum:cd8a862e941f3165
!need$ 9ef4185c7839806c n opensps_hip_iface
!need$ 4fdd01fda3834dfe n opensps_host_cosmology
!need$ bb381bf46e508468 i __fortran_builtins
!need$ 0bde2ac47243ead2 i iso_c_binding
!need$ f5118467a58a718d n opensps_host_types
module opensps_host_context
use,intrinsic::__fortran_builtins,only:__builtin_c_ptr
use,intrinsic::iso_c_binding,only:c_ptr
use,intrinsic::iso_c_binding,only:c_null_ptr
use,intrinsic::iso_c_binding,only:c_associated
use,intrinsic::iso_c_binding,only:c_int
use,intrinsic::iso_c_binding,only:c_size_t
use,intrinsic::iso_c_binding,only:c_double
use,intrinsic::iso_c_binding,only:c_char
use,intrinsic::iso_c_binding,only:c_loc
use opensps_host_cosmology,only:cosmology_table
use opensps_host_types,only:wp
use opensps_host_types,only:safe_floor
use opensps_host_types,only:npar
use opensps_host_types,only:param_names
use opensps_host_types,only:param_defaults
use opensps_host_types,only:p_tage
use opensps_host_types,only:p_sfh
use opensps_host_types,only:p_logzsol
use opensps_host_types,only:p_dust1
use opensps_host_types,only:p_dust2
use opensps_host_types,only:p_dust3
use opensps_host_types,only:p_frac_obrun
use opensps_host_types,only:p_zred
use opensps_host_types,only:p_igm_factor
use opensps_host_types,only:p_sigma_smooth
use opensps_host_types,only:p_tau
use opensps_host_types,only:p_const
use opensps_host_types,only:p_fburst
use opensps_host_types,only:p_tburst
use opensps_host_types,only:p_sf_start
use opensps_host_types,only:p_sf_trunc
use opensps_host_types,only:p_imf1
use opensps_host_types,only:ssp_names
use opensps_host_types,only:nbgrad
use opensps_host_types,only:grad_names
use opensps_host_types,only:last_error
use opensps_host_types,only:last_status
use opensps_host_types,only:lower
use opensps_host_types,only:param_slot
use opensps_host_types,only:set_error
use opensps_hip_iface,only:hip_fp64
use opensps_hip_iface,only:hip_mixed
use opensps_hip_iface,only:hip_batch_t
use opensps_hip_iface,only:hip_grad_t
use opensps_hip_iface,only:opensps_hip_create
use opensps_hip_iface,only:opensps_hip_destroy
use opensps_hip_iface,only:opensps_hip_set_grid
use opensps_hip_iface,only:opensps_hip_set_bands
use opensps_hip_iface,only:opensps_hip_set_mag_mask
use opensps_hip_iface,only:opensps_hip_set_cosmology
use opensps_hip_iface,only:opensps_hip_set_options
use opensps_hip_iface,only:opensps_hip_csp_batch
use opensps_hip_iface,only:opensps_hip_csp_batch_grad
use opensps_hip_iface,only:opensps_hip_sync
use opensps_hip_iface,only:opensps_hip_malloc
use opensps_hip_iface,only:opensps_hip_free
use opensps_hip_iface,only:opensps_hip_memcpy_h2d
use opensps_hip_iface,only:opensps_hip_memcpy_d2h
use opensps_hip_iface,only:opensps_hip_last_error
use opensps_hip_iface,only:opensps_hip_set_profiling
use opensps_hip_iface,only:opensps_hip_kernel_times
use opensps_hip_iface,only:opensps_hip_device_count
use,intrinsic::__fortran_builtins,only:iso_c_binding$__fortran_builtins$c_associated_c_ptr=>c_associated_c_ptr
private::__builtin_c_ptr
private::c_ptr
private::c_null_ptr
private::c_associated
private::c_int
private::c_size_t
private::c_double
private::c_char
private::c_loc
private::cosmology_table
private::wp
private::safe_floor
private::npar
private::param_names
private::param_defaults
private::p_tage
private::p_sfh
private::p_logzsol
private::p_dust1
private::p_dust2
private::p_dust3
private::p_frac_obrun
private::p_zred
private::p_igm_factor
private::p_sigma_smooth
private::p_tau
private::p_const
private::p_fburst
private::p_tburst
private::p_sf_start
private::p_sf_trunc
private::p_imf1
private::ssp_names
private::nbgrad
private::grad_names
private::last_error
private::last_status
private::lower
private::param_slot
private::set_error
private::hip_fp64
private::hip_mixed
private::hip_batch_t
private::hip_grad_t
private::opensps_hip_create
private::opensps_hip_destroy
private::opensps_hip_set_grid
private::opensps_hip_set_bands
private::opensps_hip_set_mag_mask
private::opensps_hip_set_cosmology
private::opensps_hip_set_options
private::opensps_hip_csp_batch
private::opensps_hip_csp_batch_grad
private::opensps_hip_sync
private::opensps_hip_malloc
private::opensps_hip_free
private::opensps_hip_memcpy_h2d
private::opensps_hip_memcpy_d2h
private::opensps_hip_last_error
private::opensps_hip_set_profiling
private::opensps_hip_kernel_times
private::opensps_hip_device_count
private::iso_c_binding$__fortran_builtins$c_associated_c_ptr
type::dbuf_t
type(c_ptr)::p=__builtin_c_ptr(__address=0_8)
integer(8)::bytes=0_8
end type
type::host_ctx_t
logical(4)::used=.false._4
integer(4)::add_igm_absorption=0_4
integer(4)::add_dust_emission=1_4
integer(4)::add_neb_emission=0_4
integer(4)::add_agn_dust=1_4
integer(4)::smooth_velocity=1_4
integer(4)::smoothspec_fast=1_4
integer(4)::dust_type=0_4
integer(4)::compute_vega_mags=0_4
integer(4)::redshift_colors=0_4
integer(4)::keep_maggies=0_4
integer(4)::imf_type=2_4
integer(4)::mixed=0_4
integer(4)::profile=0_4
logical(4)::dev_prof=.false._4
real(8)::h0=7.2e1_8
real(8)::om0=2.70000000000000017763568394002504646778106689453125e-1_8
real(8)::ol0=7.29999999999999982236431605997495353221893310546875e-1_8
real(8)::pset(1_8:28_8)=[REAL(8)::0._8,1._8,0._8,0._8,1.1e1_8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,-6.999999999999999555910790149937383830547332763671875e-1_8,-1._8,7._8,0._8,0._8,0._8,1._8,0._8,1.e3_8,1.e4_8,1.3000000000000000444089209850062616169452667236328125_8,2.29999999999999982236431605997495353221893310546875_8,2.29999999999999982236431605997495353221893310546875_8,1.2e2_8,8.000000000000000166533453693773481063544750213623046875e-2_8]
logical(4)::has_basis=.false._4
integer(4)::nspec=0_4
integer(4)::nt=0_4
integer(4)::nz=0_4
real(8)::zsol=1.42000000000000008160139230994900572113692760467529296875e-2_8
real(8),allocatable::lam(:)
real(8),allocatable::log_age(:)
real(8),allocatable::zlegend(:)
real(8),allocatable::logz(:)
real(8),allocatable::spec(:,:,:)
real(8),allocatable::mass(:,:)
real(8),allocatable::lbol(:,:)
integer(4)::nbands=0_4
real(8),allocatable::bands(:,:)
real(8),allocatable::wave_eff(:)
integer(4),allocatable::mag_mask(:)
integer(4),allocatable::bflags(:)
integer(4)::resample=0_4
integer(4)::grid_policy=2_4
real(8)::cz(1_8:500_8)=[REAL(8)::0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8]
real(8)::cdl(1_8:500_8)=[REAL(8)::0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8]
logical(4)::cosmo_dirty=.true._4
type(c_ptr)::dctx=__builtin_c_ptr(__address=0_8)
logical(4)::dev_grid=.false._4
logical(4)::dev_bands=.false._4
logical(4)::dev_cosmo=.false._4
logical(4)::dev_mask=.false._4
type(dbuf_t)::d_prm
type(dbuf_t)::d_w
type(dbuf_t)::d_mags
type(dbuf_t)::d_spec
type(dbuf_t)::d_mass
type(dbuf_t)::d_lbol
type(dbuf_t)::d_fband
type(dbuf_t)::d_young
type(dbuf_t)::d_old
type(dbuf_t)::d_idx
integer(4)::last_n=0_4
integer(4)::last_bytes=8_4
real(8),allocatable::last_maggies(:,:)
end type
type(host_ctx_t),allocatable,save,target::ctxs(:)
type(host_ctx_t),private,save::blank_ctx
private::dbuf_free
contains
function ctx_valid(h)
integer(4),intent(in)::h
logical(4)::ctx_valid
end
function ctx_create() result(h)
integer(4)::h
end
subroutine dbuf_free(b)
type(dbuf_t),intent(inout)::b
end
subroutine ctx_destroy(h)
integer(4),intent(in)::h
end
function dbuf_reserve(b,bytes) result(st)
type(dbuf_t),intent(inout)::b
integer(8),intent(in)::bytes
integer(4)::st
end
subroutine ctx_set_int(h,key,v,status)
integer(4),intent(in)::h
character(*,1),intent(in)::key
integer(4),intent(in)::v
integer(4),intent(out)::status
end
subroutine ctx_set_float(h,key,v,status)
integer(4),intent(in)::h
character(*,1),intent(in)::key
real(8),intent(in)::v
integer(4),intent(out)::status
end
subroutine ctx_set_str(h,key,v,status)
integer(4),intent(in)::h
character(*,1),intent(in)::key
character(*,1),intent(in)::v
integer(4),intent(out)::status
end
subroutine ctx_set_basis(h,nspec,lam,nt,log_age,nz,zlegend,zsol,spec,mass,lbol,status)
integer(4),intent(in)::h
integer(4),intent(in)::nspec
real(8),intent(in)::lam(1_8:int(nspec,kind=8))
integer(4),intent(in)::nt
real(8),intent(in)::log_age(1_8:int(nt,kind=8))
integer(4),intent(in)::nz
real(8),intent(in)::zlegend(1_8:int(nz,kind=8))
real(8),intent(in)::zsol
real(8),intent(in)::spec(1_8:int(nspec,kind=8),1_8:int(nt,kind=8),1_8:int(nz,kind=8))
real(8),intent(in)::mass(1_8:int(nt,kind=8),1_8:int(nz,kind=8))
real(8),intent(in)::lbol(1_8:int(nt,kind=8),1_8:int(nz,kind=8))
integer(4),intent(out)::status
end
subroutine ctx_refresh_cosmology(h)
integer(4),intent(in)::h
end
function hip_message(dctx) result(msg)
type(c_ptr),intent(in)::dctx
character(512_4,1)::msg
end
function ctx_device(h) result(st)
integer(4),intent(in)::h
integer(4)::st
end
end
