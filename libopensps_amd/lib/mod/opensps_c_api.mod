﻿!mod$ v1 sum:7b340f3eda27cdf3
!need$ b579c837613f74aa n opensps_host_filters
!need$ 555e300ca16dfe42 n opensps_host_batch
!need$ 0bde2ac47243ead2 i iso_c_binding
!need$ f5118467a58a718d n opensps_host_types
!need$ cd8a862e941f3165 n opensps_host_context
!need$ bb381bf46e508468 i __fortran_builtins
module opensps_c_api
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
use opensps_host_context,only:dbuf_t
use opensps_host_context,only:host_ctx_t
use opensps_host_context,only:ctxs
use opensps_host_context,only:ctx_valid
use opensps_host_context,only:ctx_create
use opensps_host_context,only:ctx_destroy
use opensps_host_context,only:dbuf_reserve
use opensps_host_context,only:ctx_set_int
use opensps_host_context,only:ctx_set_float
use opensps_host_context,only:ctx_set_str
use opensps_host_context,only:ctx_set_basis
use opensps_host_context,only:ctx_refresh_cosmology
use opensps_host_context,only:hip_message
use opensps_host_context,only:ctx_device
use opensps_host_filters,only:filter_ok
use opensps_host_filters,only:filter_err_few_points
use opensps_host_filters,only:filter_err_nonfinite
use opensps_host_filters,only:filter_err_order
use opensps_host_filters,only:filter_err_argument
use opensps_host_filters,only:filter_err_coverage
use opensps_host_filters,only:filter_err_dead
use opensps_host_filters,only:flag_truncated
use opensps_host_filters,only:flag_undersampled
use opensps_host_filters,only:flag_dead
use opensps_host_filters,only:filter_curve_t
use opensps_host_filters,only:filter_set_t
use opensps_host_filters,only:fsets
use opensps_host_filters,only:fset_valid
use opensps_host_filters,only:fset_create
use opensps_host_filters,only:fset_free
use opensps_host_filters,only:fset_add
use opensps_host_filters,only:trapezoid
use opensps_host_filters,only:filters_build
use opensps_host_filters,only:band_pivot
use opensps_host_batch,only:cstring
use opensps_host_batch,only:model_checks
use opensps_host_batch,only:resolve_params
use opensps_host_batch,only:batch_run
use opensps_host_batch,only:batch_run_device
use,intrinsic::__fortran_builtins,only:iso_c_binding$__fortran_builtins$c_associated_c_ptr=>c_associated_c_ptr
private::c_associated
private::c_funloc
private::c_funptr
private::c_f_pointer
private::c_loc
private::c_null_funptr
private::c_null_ptr
private::c_ptr
private::c_sizeof
private::operator(==)
private::operator(/=)
private::c_int8_t
private::c_int16_t
private::c_int32_t
private::c_int64_t
private::c_int128_t
private::c_int
private::c_short
private::c_long
private::c_long_long
private::c_signed_char
private::c_size_t
private::c_intmax_t
private::c_intptr_t
private::c_ptrdiff_t
private::c_int_least8_t
private::c_int_fast8_t
private::c_int_least16_t
private::c_int_fast16_t
private::c_int_least32_t
private::c_int_fast32_t
private::c_int_least64_t
private::c_int_fast64_t
private::c_int_least128_t
private::c_int_fast128_t
private::c_float
private::c_double
private::c_long_double
private::c_float_complex
private::c_double_complex
private::c_long_double_complex
private::c_bool
private::c_char
private::c_null_char
private::c_alert
private::c_backspace
private::c_form_feed
private::c_new_line
private::c_carriage_return
private::c_horizontal_tab
private::c_vertical_tab
private::c_float128
private::c_float128_complex
private::c_uint8_t
private::c_uint16_t
private::c_uint32_t
private::c_uint64_t
private::c_uint128_t
private::c_unsigned_char
private::c_unsigned_short
private::c_unsigned
private::c_unsigned_long
private::c_unsigned_long_long
private::c_uintmax_t
private::c_uint_fast8_t
private::c_uint_fast16_t
private::c_uint_fast32_t
private::c_uint_fast64_t
private::c_uint_fast128_t
private::c_uint_least8_t
private::c_uint_least16_t
private::c_uint_least32_t
private::c_uint_least64_t
private::c_uint_least128_t
private::c_f_procpointer
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
private::dbuf_t
private::host_ctx_t
private::ctxs
private::ctx_valid
private::ctx_create
private::ctx_destroy
private::dbuf_reserve
private::ctx_set_int
private::ctx_set_float
private::ctx_set_str
private::ctx_set_basis
private::ctx_refresh_cosmology
private::hip_message
private::ctx_device
private::filter_ok
private::filter_err_few_points
private::filter_err_nonfinite
private::filter_err_order
private::filter_err_argument
private::filter_err_coverage
private::filter_err_dead
private::flag_truncated
private::flag_undersampled
private::flag_dead
private::filter_curve_t
private::filter_set_t
private::fsets
private::fset_valid
private::fset_create
private::fset_free
private::fset_add
private::trapezoid
private::filters_build
private::band_pivot
private::cstring
private::model_checks
private::resolve_params
private::batch_run
private::batch_run_device
private::iso_c_binding$__fortran_builtins$c_associated_c_ptr
character(32_4,1),private,save,target::version_string
character(8_4,1),private,save,target::build_id
private::fstr
private::to_cstr
private::bad_handle
private::opensps_set_debug
private::opensps_clear_error
private::opensps_get_last_error
private::opensps_get_driver_version
private::opensps_get_version_string
private::opensps_get_build_id
private::opensps_get_nbgrad
private::opensps_get_grad_param_name
private::opensps_context_create
private::opensps_context_destroy
private::opensps_context_setup
private::opensps_context_set_ssp_basis
private::opensps_context_set_int
private::opensps_context_set_float
private::opensps_context_set_str
private::opensps_context_set_fast_mode
private::opensps_context_get_dims
private::opensps_context_get_nspec
private::opensps_context_get_ntfull
private::opensps_context_get_nbands
private::opensps_context_get_nz
private::opensps_context_get_nindx
private::opensps_context_get_nemline
private::opensps_context_get_zsol
private::opensps_context_get_lambda
private::opensps_context_get_ages
private::opensps_context_get_zlegend
private::opensps_context_get_filter_data
private::opensps_filterset_create
private::opensps_filterset_add
private::opensps_filterset_count
private::opensps_filterset_free
private::opensps_context_set_filters
private::opensps_context_get_bands
private::opensps_context_set_mag_compute
private::opensps_context_compute_csp_batch
private::opensps_context_compute_csp_batch_parts
private::opensps_context_compute_csp_batch_grad
private::opensps_context_compute_csp_batch_grad_rows
private::opensps_context_compute_csp_batch_device
private::opensps_context_get_params
private::opensps_context_kernel_times
private::opensps_context_batch_spec_device
private::opensps_context_get_batch_maggies
contains
function fstr(a) result(s)
character(1_8,1),intent(in)::a(1_8:*)
character(:,1),allocatable::s
end
subroutine to_cstr(s,dst,max_len)
character(*,1),intent(in)::s
character(1_8,1),intent(out)::dst(1_8:*)
integer(4),intent(in)::max_len
end
function bad_handle(handle,status)
integer(4),intent(in)::handle
integer(4),intent(out)::status
logical(4)::bad_handle
end
subroutine opensps_set_debug(flag) bind(c,name="opensps_set_debug")
integer(4),value::flag
end
subroutine opensps_clear_error() bind(c,name="opensps_clear_error")
end
subroutine opensps_get_last_error(status,message,message_len) bind(c,name="opensps_get_last_error")
integer(4),intent(out)::status
character(1_8,1),intent(out)::message(1_8:*)
integer(4),value::message_len
end
subroutine opensps_get_driver_version(major,minor,patch) bind(c,name="opensps_get_driver_version")
integer(4),intent(out)::major
integer(4),intent(out)::minor
integer(4),intent(out)::patch
end
function opensps_get_version_string() bind(c,name="opensps_get_version_string")
type(c_ptr)::opensps_get_version_string
end
function opensps_get_build_id() bind(c,name="opensps_get_build_id")
type(c_ptr)::opensps_get_build_id
end
subroutine opensps_get_nbgrad(nbgrad_out) bind(c,name="opensps_get_nbgrad")
integer(4),intent(out)::nbgrad_out
end
subroutine opensps_get_grad_param_name(grad_idx,name,max_len,status) bind(c,name="opensps_get_grad_param_name")
integer(4),value::grad_idx
character(1_8,1),intent(out)::name(1_8:*)
integer(4),value::max_len
integer(4),intent(out)::status
end
subroutine opensps_context_create(handle,status) bind(c,name="opensps_context_create")
integer(4),intent(out)::handle
integer(4),intent(out)::status
end
subroutine opensps_context_destroy(handle,status) bind(c,name="opensps_context_destroy")
integer(4),value::handle
integer(4),intent(out)::status
end
subroutine opensps_context_setup(zin,isoc_type,spec_type,dust_type,handle,status) bind(c,name="opensps_context_setup")
integer(4),value::zin
character(1_8,1),intent(in)::isoc_type(1_8:*)
character(1_8,1),intent(in)::spec_type(1_8:*)
character(1_8,1),intent(in)::dust_type(1_8:*)
integer(4),value::handle
integer(4),intent(out)::status
end
subroutine opensps_context_set_ssp_basis(handle,nspec,lambda,ntfull,log_age,nz,zlegend,zsol,spec,mass,lbol,status) bind(c,name="opensps_context_set_ssp_basis")
integer(4),value::handle
integer(4),value::nspec
real(8),intent(in)::lambda(1_8:*)
integer(4),value::ntfull
real(8),intent(in)::log_age(1_8:*)
integer(4),value::nz
real(8),intent(in)::zlegend(1_8:*)
real(8),value::zsol
real(8),intent(in)::spec(1_8:*)
real(8),intent(in)::mass(1_8:*)
real(8),intent(in)::lbol(1_8:*)
integer(4),intent(out)::status
end
subroutine opensps_context_set_int(handle,key,value,status) bind(c,name="opensps_context_set_int")
integer(4),value::handle
character(1_8,1),intent(in)::key(1_8:*)
integer(4),value::value
integer(4),intent(out)::status
end
subroutine opensps_context_set_float(handle,key,value,status) bind(c,name="opensps_context_set_float")
integer(4),value::handle
character(1_8,1),intent(in)::key(1_8:*)
real(8),value::value
integer(4),intent(out)::status
end
subroutine opensps_context_set_str(handle,key,value,status) bind(c,name="opensps_context_set_str")
integer(4),value::handle
character(1_8,1),intent(in)::key(1_8:*)
character(1_8,1),intent(in)::value(1_8:*)
integer(4),intent(out)::status
end
subroutine opensps_context_set_fast_mode(handle,fast_mode,status) bind(c,name="opensps_context_set_fast_mode")
integer(4),value::handle
integer(4),value::fast_mode
integer(4),intent(out)::status
end
subroutine opensps_context_get_dims(handle,n_spec,n_time,status) bind(c,name="opensps_context_get_dims")
integer(4),value::handle
integer(4),intent(out)::n_spec
integer(4),intent(out)::n_time
integer(4),intent(out)::status
end
subroutine opensps_context_get_nspec(handle,n,status) bind(c,name="opensps_context_get_nspec")
integer(4),value::handle
integer(4),intent(out)::n
integer(4),intent(out)::status
end
subroutine opensps_context_get_ntfull(handle,n,status) bind(c,name="opensps_context_get_ntfull")
integer(4),value::handle
integer(4),intent(out)::n
integer(4),intent(out)::status
end
subroutine opensps_context_get_nbands(handle,n,status) bind(c,name="opensps_context_get_nbands")
integer(4),value::handle
integer(4),intent(out)::n
integer(4),intent(out)::status
end
subroutine opensps_context_get_nz(handle,n,status) bind(c,name="opensps_context_get_nz")
integer(4),value::handle
integer(4),intent(out)::n
integer(4),intent(out)::status
end
subroutine opensps_context_get_nindx(handle,n,status) bind(c,name="opensps_context_get_nindx")
integer(4),value::handle
integer(4),intent(out)::n
integer(4),intent(out)::status
end
subroutine opensps_context_get_nemline(handle,n,status) bind(c,name="opensps_context_get_nemline")
integer(4),value::handle
integer(4),intent(out)::n
integer(4),intent(out)::status
end
subroutine opensps_context_get_zsol(handle,z,status) bind(c,name="opensps_context_get_zsol")
integer(4),value::handle
real(8),intent(out)::z
integer(4),intent(out)::status
end
subroutine opensps_context_get_lambda(handle,lambda,status) bind(c,name="opensps_context_get_lambda")
integer(4),value::handle
real(8),intent(out)::lambda(1_8:*)
integer(4),intent(out)::status
end
subroutine opensps_context_get_ages(handle,log_age,status) bind(c,name="opensps_context_get_ages")
integer(4),value::handle
real(8),intent(out)::log_age(1_8:*)
integer(4),intent(out)::status
end
subroutine opensps_context_get_zlegend(handle,z,status) bind(c,name="opensps_context_get_zlegend")
integer(4),value::handle
real(8),intent(out)::z(1_8:*)
integer(4),intent(out)::status
end
subroutine opensps_context_get_filter_data(handle,wave_eff,mag_vega,mag_sun,status) bind(c,name="opensps_context_get_filter_data")
integer(4),value::handle
type(c_ptr),value::wave_eff
type(c_ptr),value::mag_vega
type(c_ptr),value::mag_sun
integer(4),intent(out)::status
end
subroutine opensps_filterset_create(fset,status) bind(c,name="opensps_filterset_create")
integer(4),intent(out)::fset
integer(4),intent(out)::status
end
subroutine opensps_filterset_add(fset,name,n,lambda_aa,response,detector,cal_lambda_ref,cal_exponent,status) bind(c,name="opensps_filterset_add")
integer(4),value::fset
character(1_8,1),intent(in)::name(1_8:*)
integer(4),value::n
real(8),intent(in)::lambda_aa(1_8:*)
real(8),intent(in)::response(1_8:*)
integer(4),value::detector
real(8),value::cal_lambda_ref
real(8),value::cal_exponent
integer(4),intent(out)::status
end
subroutine opensps_filterset_count(fset,n,status) bind(c,name="opensps_filterset_count")
integer(4),value::fset
integer(4),intent(out)::n
integer(4),intent(out)::status
end
subroutine opensps_filterset_free(fset,status) bind(c,name="opensps_filterset_free")
integer(4),value::fset
integer(4),intent(out)::status
end
subroutine opensps_context_set_filters(handle,fset,resample,grid_policy,vega_ref_band,status) bind(c,name="opensps_context_set_filters")
integer(4),value::handle
integer(4),value::fset
integer(4),value::resample
integer(4),value::grid_policy
integer(4),value::vega_ref_band
integer(4),intent(out)::status
end
subroutine opensps_context_get_bands(handle,bands,status) bind(c,name="opensps_context_get_bands")
integer(4),value::handle
real(8),intent(out)::bands(1_8:*)
integer(4),intent(out)::status
end
subroutine opensps_context_set_mag_compute(handle,n_bands,mask,status) bind(c,name="opensps_context_set_mag_compute")
integer(4),value::handle
integer(4),value::n_bands
type(c_ptr),value::mask
integer(4),intent(out)::status
end
subroutine opensps_context_compute_csp_batch(handle,n,n_params,c_names,c_values,c_weights,c_mags,c_spec,c_mass,c_lbol,c_mdust,c_sfr,c_emlines,c_indx,status) bind(c,name="opensps_context_compute_csp_batch")
integer(4),value::handle
integer(4),value::n
integer(4),value::n_params
type(c_ptr),value::c_names
type(c_ptr),value::c_values
type(c_ptr),value::c_weights
type(c_ptr),value::c_mags
type(c_ptr),value::c_spec
type(c_ptr),value::c_mass
type(c_ptr),value::c_lbol
type(c_ptr),value::c_mdust
type(c_ptr),value::c_sfr
type(c_ptr),value::c_emlines
type(c_ptr),value::c_indx
integer(4),intent(out)::status
end
subroutine opensps_context_compute_csp_batch_parts(handle,n,n_params,c_names,c_values,c_weights,c_mags,c_spec,c_mass,c_lbol,c_mdust,c_sfr,c_emlines,c_indx,c_young,c_old,c_mass_young,c_mass_old,c_lbol_young,c_lbol_old,status) bind(c,name="opensps_context_compute_csp_batch_parts")
integer(4),value::handle
integer(4),value::n
integer(4),value::n_params
type(c_ptr),value::c_names
type(c_ptr),value::c_values
type(c_ptr),value::c_weights
type(c_ptr),value::c_mags
type(c_ptr),value::c_spec
type(c_ptr),value::c_mass
type(c_ptr),value::c_lbol
type(c_ptr),value::c_mdust
type(c_ptr),value::c_sfr
type(c_ptr),value::c_emlines
type(c_ptr),value::c_indx
type(c_ptr),value::c_young
type(c_ptr),value::c_old
type(c_ptr),value::c_mass_young
type(c_ptr),value::c_mass_old
type(c_ptr),value::c_lbol_young
type(c_ptr),value::c_lbol_old
integer(4),intent(out)::status
end
subroutine opensps_context_compute_csp_batch_grad(handle,n,n_params,c_names,c_values,c_weights,c_weights_z,c_zred_mags,c_mags_bar,c_spec_bar,c_mass_bar,c_lbol_bar,c_mags,c_spec,c_mass,c_lbol,c_mdust,c_sfr,c_emlines,c_indx,c_params_bar,c_weights_bar,c_weights_z_bar,status) bind(c,name="opensps_context_compute_csp_batch_grad")
integer(4),value::handle
integer(4),value::n
integer(4),value::n_params
type(c_ptr),value::c_names
type(c_ptr),value::c_values
type(c_ptr),value::c_weights
type(c_ptr),value::c_weights_z
type(c_ptr),value::c_zred_mags
type(c_ptr),value::c_mags_bar
type(c_ptr),value::c_spec_bar
type(c_ptr),value::c_mass_bar
type(c_ptr),value::c_lbol_bar
type(c_ptr),value::c_mags
type(c_ptr),value::c_spec
type(c_ptr),value::c_mass
type(c_ptr),value::c_lbol
type(c_ptr),value::c_mdust
type(c_ptr),value::c_sfr
type(c_ptr),value::c_emlines
type(c_ptr),value::c_indx
type(c_ptr),value::c_params_bar
type(c_ptr),value::c_weights_bar
type(c_ptr),value::c_weights_z_bar
integer(4),intent(out)::status
end
subroutine opensps_context_compute_csp_batch_grad_rows(handle,n,n_params,c_names,c_values,c_weights,c_weights_z,c_zred_mags,c_mags_bar,c_spec_bar,c_mass_bar,c_lbol_bar,c_mags,c_spec,c_mass,c_lbol,c_mdust,c_sfr,c_emlines,c_indx,c_params_bar,params_bar_rows,c_weights_bar,c_weights_z_bar,status) bind(c,name="opensps_context_compute_csp_batch_grad_rows")
integer(4),value::handle
integer(4),value::n
integer(4),value::n_params
type(c_ptr),value::c_names
type(c_ptr),value::c_values
type(c_ptr),value::c_weights
type(c_ptr),value::c_weights_z
type(c_ptr),value::c_zred_mags
type(c_ptr),value::c_mags_bar
type(c_ptr),value::c_spec_bar
type(c_ptr),value::c_mass_bar
type(c_ptr),value::c_lbol_bar
type(c_ptr),value::c_mags
type(c_ptr),value::c_spec
type(c_ptr),value::c_mass
type(c_ptr),value::c_lbol
type(c_ptr),value::c_mdust
type(c_ptr),value::c_sfr
type(c_ptr),value::c_emlines
type(c_ptr),value::c_indx
type(c_ptr),value::c_params_bar
integer(4),value::params_bar_rows
type(c_ptr),value::c_weights_bar
type(c_ptr),value::c_weights_z_bar
integer(4),intent(out)::status
end
subroutine opensps_context_compute_csp_batch_device(handle,n,d_prm,d_weights,precision,split,d_mags,d_spec,d_mass,d_lbol,status) bind(c,name="opensps_context_compute_csp_batch_device")
integer(4),value::handle
integer(4),value::n
type(c_ptr),value::d_prm
type(c_ptr),value::d_weights
integer(4),value::precision
integer(4),value::split
type(c_ptr),value::d_mags
type(c_ptr),value::d_spec
type(c_ptr),value::d_mass
type(c_ptr),value::d_lbol
integer(4),intent(out)::status
end
subroutine opensps_context_get_params(handle,prm,status) bind(c,name="opensps_context_get_params")
integer(4),value::handle
real(8),intent(out)::prm(1_8:28_8)
integer(4),intent(out)::status
end
subroutine opensps_context_kernel_times(handle,ms,launches,reset,status) bind(c,name="opensps_context_kernel_times")
integer(4),value::handle
real(8),intent(out)::ms(1_8:4_8)
integer(4),intent(out)::launches(1_8:4_8)
integer(4),value::reset
integer(4),intent(out)::status
end
subroutine opensps_context_batch_spec_device(handle,ptr,nspec,n,bytes,status) bind(c,name="opensps_context_batch_spec_device")
integer(4),value::handle
type(c_ptr),intent(out)::ptr
integer(4),intent(out)::nspec
integer(4),intent(out)::n
integer(4),intent(out)::bytes
integer(4),intent(out)::status
end
subroutine opensps_context_get_batch_maggies(handle,maggies,n_bands,n,status) bind(c,name="opensps_context_get_batch_maggies")
integer(4),value::handle
type(c_ptr),value::maggies
integer(4),intent(out)::n_bands
integer(4),intent(out)::n
integer(4),intent(out)::status
end
end
