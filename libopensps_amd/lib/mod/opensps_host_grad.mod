﻿!mod$ v1 sum:1d09ef0846d40523
!need$ 555e300ca16dfe42 n opensps_host_batch
!need$ 0bde2ac47243ead2 i iso_c_binding
!need$ f5118467a58a718d n opensps_host_types
!need$ cd8a862e941f3165 n opensps_host_context
!need$ 9ef4185c7839806c n opensps_hip_iface
!need$ bb381bf46e508468 i __fortran_builtins
module opensps_host_grad
use opensps_host_batch,only:resolve_params
use opensps_host_batch,only:model_checks
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
private::resolve_params
private::model_checks
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
contains
subroutine batch_run_grad(h,n,n_params,c_names,c_values,c_weights,c_weights_z,c_zred_mags,c_mags_bar,c_spec_bar,c_mass_bar,c_lbol_bar,c_mags,c_spec,c_mass,c_lbol,c_mdust,c_sfr,c_emlines,c_indx,c_params_bar,prows,c_weights_bar,c_weights_z_bar,status)
integer(4),intent(in)::h
integer(4),intent(in)::n
integer(4),intent(in)::n_params
type(c_ptr),intent(in)::c_names
type(c_ptr),intent(in)::c_values
type(c_ptr),intent(in)::c_weights
type(c_ptr),intent(in)::c_weights_z
type(c_ptr),intent(in)::c_zred_mags
type(c_ptr),intent(in)::c_mags_bar
type(c_ptr),intent(in)::c_spec_bar
type(c_ptr),intent(in)::c_mass_bar
type(c_ptr),intent(in)::c_lbol_bar
type(c_ptr),intent(in)::c_mags
type(c_ptr),intent(in)::c_spec
type(c_ptr),intent(in)::c_mass
type(c_ptr),intent(in)::c_lbol
type(c_ptr),intent(in)::c_mdust
type(c_ptr),intent(in)::c_sfr
type(c_ptr),intent(in)::c_emlines
type(c_ptr),intent(in)::c_indx
type(c_ptr),intent(in)::c_params_bar
integer(4),intent(in)::prows
type(c_ptr),intent(in)::c_weights_bar
type(c_ptr),intent(in)::c_weights_z_bar
integer(4),intent(out)::status
end
end
