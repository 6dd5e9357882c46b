!> iso_c_binding interfaces of libopensps_hip.so (include/opensps_hip.h): the device layer the Fortran host
!> calls with device pointers, as the reference's batch driver calls its HIP launchers through bind(C)
!> interface blocks (src/physics/opensps_csp_batch.f90:260-295, 2341-2349, 3742-3757).
module opensps_hip_iface
    use, intrinsic :: iso_c_binding, only: c_int, c_double, c_ptr, c_size_t, c_char, c_null_ptr
    implicit none
    private

    integer(c_int), parameter, public :: HIP_FP64 = 0, HIP_MIXED = 1

    !> opensps_hip_batch_t
    type, bind(C), public :: hip_batch_t
        integer(c_int) :: n = 0, precision = 0, split_parts = 0, use_ops = 0
        type(c_ptr) :: prm = c_null_ptr, ssp_weights = c_null_ptr, mags = c_null_ptr, spec = c_null_ptr
        type(c_ptr) :: mass = c_null_ptr, lbol = c_null_ptr, fband = c_null_ptr, young = c_null_ptr, old = c_null_ptr
        type(c_ptr) :: idx = c_null_ptr, weights = c_null_ptr, op_count = c_null_ptr, ssp_weights_z = c_null_ptr
        type(c_ptr) :: zred_mags = c_null_ptr, mdust = c_null_ptr, emlines = c_null_ptr
    end type hip_batch_t

    !> opensps_hip_grad_t
    type, bind(C), public :: hip_grad_t
        type(hip_batch_t) :: fwd
        type(c_ptr) :: mags_bar = c_null_ptr, spec_bar = c_null_ptr, mass_bar = c_null_ptr, lbol_bar = c_null_ptr
        type(c_ptr) :: params_bar = c_null_ptr, weights_bar = c_null_ptr, weights_z_bar = c_null_ptr
    end type hip_grad_t

    !> opensps_hip_ssp_ops_t
    type, bind(C), public :: hip_ssp_ops_t
        integer(c_int) :: nz = 0, nt = 0, npts = 0, ntpl = 0, nent = 0
        type(c_ptr) :: age_ptr = c_null_ptr, mini = c_null_ptr, m1 = c_null_ptr, m2 = c_null_ptr, first = c_null_ptr
        type(c_ptr) :: cm = c_null_ptr, cl = c_null_ptr, q = c_null_ptr, age_mmax = c_null_ptr
        type(c_ptr) :: ent_tpl = c_null_ptr, ent_val = c_null_ptr, etpl = c_null_ptr
    end type hip_ssp_ops_t

    public :: opensps_hip_set_ssp_ops, opensps_hip_free_ssp_ops, opensps_hip_set_imf
    public :: opensps_hip_create, opensps_hip_destroy, opensps_hip_set_grid, opensps_hip_set_bands, &
              opensps_hip_set_mag_mask, opensps_hip_set_vega, opensps_hip_set_sfh_tab, opensps_hip_set_dust_curve, opensps_hip_set_cosmology, opensps_hip_set_options, opensps_hip_csp_batch, &
              opensps_hip_sync, opensps_hip_malloc, opensps_hip_free, opensps_hip_memcpy_h2d, opensps_hip_memcpy_d2h, &
              opensps_hip_last_error, opensps_hip_device_count, opensps_hip_set_profiling, opensps_hip_kernel_times, &
              opensps_hip_op_kernel_times, opensps_hip_set_nebular, opensps_hip_set_neb_options, &
              opensps_hip_csp_batch_grad, opensps_hip_set_emission, opensps_hip_set_dust_emission, opensps_hip_set_agn

    interface
        integer(c_int) function opensps_hip_create(device, ctx) bind(C, name="opensps_hip_create")
            import :: c_int, c_ptr
            integer(c_int), value :: device
            type(c_ptr), intent(out) :: ctx
        end function

        integer(c_int) function opensps_hip_destroy(ctx) bind(C, name="opensps_hip_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function

        integer(c_int) function opensps_hip_set_grid(ctx, nspec, lam, nt, log_age, nz, logz, spec, mass, lbol) &
            bind(C, name="opensps_hip_set_grid")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: nspec, nt, nz
            real(c_double), intent(in) :: lam(*), log_age(*), logz(*), spec(*), mass(*), lbol(*)
        end function

        integer(c_int) function opensps_hip_set_bands(ctx, nbands, bands) bind(C, name="opensps_hip_set_bands")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: nbands
            real(c_double), intent(in) :: bands(*)
        end function

        integer(c_int) function opensps_hip_set_mag_mask(ctx, nbands, mask) bind(C, name="opensps_hip_set_mag_mask")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: nbands
            integer(c_int), intent(in) :: mask(*)
        end function

        integer(c_int) function opensps_hip_set_dust_curve(ctx, nspec, curve) bind(C, name="opensps_hip_set_dust_curve")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, curve
            integer(c_int), value :: nspec
        end function

        integer(c_int) function opensps_hip_set_sfh_tab(ctx, ntab, age_yr, sfr) bind(C, name="opensps_hip_set_sfh_tab")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: ntab
            real(c_double), intent(in) :: age_yr(*), sfr(*)
        end function

        integer(c_int) function opensps_hip_set_vega(ctx, nbands, magvega, vref) bind(C, name="opensps_hip_set_vega")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: nbands, vref
            real(c_double), intent(in) :: magvega(*)
        end function

        integer(c_int) function opensps_hip_set_cosmology(ctx, n, z, dl) bind(C, name="opensps_hip_set_cosmology")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: n
            real(c_double), intent(in) :: z(*), dl(*)
        end function

        integer(c_int) function opensps_hip_set_options(ctx, add_igm, smooth_fast) bind(C, name="opensps_hip_set_options")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: add_igm, smooth_fast
        end function

        integer(c_int) function opensps_hip_set_emission(ctx, add_dust_emission, add_agn_dust) &
                bind(C, name="opensps_hip_set_emission")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: add_dust_emission, add_agn_dust
        end function

        integer(c_int) function opensps_hip_set_dust_emission(ctx, nqpah, qpah, numin, umin, table) &
                bind(C, name="opensps_hip_set_dust_emission")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: nqpah, numin
            type(c_ptr), value :: qpah, umin, table
        end function

        integer(c_int) function opensps_hip_set_agn(ctx, ntau, tau, table) bind(C, name="opensps_hip_set_agn")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: ntau
            type(c_ptr), value :: tau, table
        end function

        integer(c_int) function opensps_hip_csp_batch(ctx, args) bind(C, name="opensps_hip_csp_batch")
            import :: c_int, c_ptr, hip_batch_t
            type(c_ptr), value :: ctx
            type(hip_batch_t), intent(in) :: args
        end function

        integer(c_int) function opensps_hip_csp_batch_grad(ctx, args) bind(C, name="opensps_hip_csp_batch_grad")
            import :: c_int, c_ptr, hip_grad_t
            type(c_ptr), value :: ctx
            type(hip_grad_t), intent(in) :: args
        end function

        integer(c_int) function opensps_hip_sync(ctx) bind(C, name="opensps_hip_sync")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function

        integer(c_int) function opensps_hip_malloc(ptr, bytes) bind(C, name="opensps_hip_malloc")
            import :: c_int, c_ptr, c_size_t
            type(c_ptr), intent(out) :: ptr
            integer(c_size_t), value :: bytes
        end function

        integer(c_int) function opensps_hip_free(ptr) bind(C, name="opensps_hip_free")
            import :: c_int, c_ptr
            type(c_ptr), value :: ptr
        end function

        integer(c_int) function opensps_hip_memcpy_h2d(dst, src, bytes) bind(C, name="opensps_hip_memcpy_h2d")
            import :: c_int, c_ptr, c_size_t
            type(c_ptr), value :: dst, src
            integer(c_size_t), value :: bytes
        end function

        integer(c_int) function opensps_hip_memcpy_d2h(dst, src, bytes) bind(C, name="opensps_hip_memcpy_d2h")
            import :: c_int, c_ptr, c_size_t
            type(c_ptr), value :: dst, src
            integer(c_size_t), value :: bytes
        end function

        integer(c_int) function opensps_hip_last_error(ctx, msg, len) bind(C, name="opensps_hip_last_error")
            import :: c_int, c_ptr, c_char
            type(c_ptr), value :: ctx
            character(kind=c_char), intent(out) :: msg(*)
            integer(c_int), value :: len
        end function

        integer(c_int) function opensps_hip_set_profiling(ctx, on) bind(C, name="opensps_hip_set_profiling")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: on
        end function

        integer(c_int) function opensps_hip_kernel_times(ctx, ms, launches, reset) bind(C, name="opensps_hip_kernel_times")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: ms(4)
            integer(c_int), intent(out) :: launches(4)
            integer(c_int), value :: reset
        end function

        integer(c_int) function opensps_hip_op_kernel_times(ctx, ms, launches, reset) &
            bind(C, name="opensps_hip_op_kernel_times")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: ms(2)
            integer(c_int), intent(out) :: launches(2)
            integer(c_int), value :: reset
        end function

        integer(c_int) function opensps_hip_set_nebular(ctx, nnz, logz, nna, age, nnu, logu, nlam, lam, cont, nline, &
                                                        line_lam, lines) bind(C, name="opensps_hip_set_nebular")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: nnz, nna, nnu, nlam, nline
            real(c_double), intent(in) :: logz(*), age(*), logu(*), lam(*), cont(*), line_lam(*), lines(*)
        end function

        integer(c_int) function opensps_hip_set_neb_options(ctx, add_neb_emission, add_neb_continuum, nebemlineinspec) &
            bind(C, name="opensps_hip_set_neb_options")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: add_neb_emission, add_neb_continuum, nebemlineinspec
        end function

        integer(c_int) function opensps_hip_set_ssp_ops(ctx, ops) bind(C, name="opensps_hip_set_ssp_ops")
            import :: c_int, c_ptr, hip_ssp_ops_t
            type(c_ptr), value :: ctx
            type(hip_ssp_ops_t), intent(in) :: ops
        end function

        integer(c_int) function opensps_hip_free_ssp_ops(ctx) bind(C, name="opensps_hip_free_ssp_ops")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function

        integer(c_int) function opensps_hip_set_imf(ctx, imf_type, remnants, salp_ind, mdave, mlim_bh, mlim_ns) &
            bind(C, name="opensps_hip_set_imf")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: imf_type, remnants
            real(c_double), value :: salp_ind, mdave, mlim_bh, mlim_ns
        end function

        integer(c_int) function opensps_hip_device_count(n) bind(C, name="opensps_hip_device_count")
            import :: c_int
            integer(c_int), intent(out) :: n
        end function
    end interface
end module opensps_hip_iface
