!> The public C ABI of libopensps.so (include/opensps.h): the reference's context API for the batched
!> composite-population path (include/opensps.h:132-455; src/bindings/c/opensps_c_api.f90), with int handles,
!> column-major double arrays and status out-parameters. Two entry points have no reference counterpart and
!> are marked so: opensps_context_set_ssp_basis (a resident basis from the caller, instead of set-up from a
!> data release) and opensps_context_compute_csp_batch_device (a batch on device-resident arrays).
module opensps_c_api
    use, intrinsic :: iso_c_binding
    use opensps_host_types
    use opensps_host_context
    use opensps_host_filters
    use opensps_host_batch
    use opensps_hip_iface, only: hip_ssp_ops_t, opensps_hip_set_ssp_ops, opensps_hip_free_ssp_ops
    implicit none
    private

    character(kind=c_char, len=32), target, save :: version_string = '0.3.0-mi355x' // c_null_char
    character(kind=c_char, len=8), target, save :: build_id = c_null_char

contains

    function fstr(a) result(s)
        character(kind=c_char), intent(in) :: a(*)
        character(len=:), allocatable :: s
        integer :: n
        n = 0
        do while (a(n + 1) /= c_null_char)
            n = n + 1
            if (n >= 4096) exit
        end do
        allocate(character(len=n) :: s)
        s = transfer(a(1:n), s)
    end function fstr

    subroutine to_cstr(s, dst, max_len)
        character(len=*), intent(in) :: s
        character(kind=c_char), intent(out) :: dst(*)
        integer, intent(in) :: max_len
        integer :: i, n
        if (max_len <= 0) return
        n = min(len_trim(s), max_len - 1)
        do i = 1, n
            dst(i) = s(i:i)
        end do
        dst(n + 1) = c_null_char
    end subroutine to_cstr

    logical function bad_handle(handle, status)
        integer(c_int), intent(in) :: handle
        integer(c_int), intent(out) :: status
        bad_handle = .not. ctx_valid(int(handle))
        status = 0
        if (bad_handle) then
            status = 1
            call set_error(1, 'bad handle')
        end if
    end function bad_handle

    ! ---- errors, version ---------------------------------------------------------------------------------------
    subroutine opensps_set_debug(flag) bind(C, name="opensps_set_debug")
        integer(c_int), value :: flag
    end subroutine

    subroutine opensps_clear_error() bind(C, name="opensps_clear_error")
        call set_error(0, '')
    end subroutine

    subroutine opensps_get_last_error(status, message, message_len) bind(C, name="opensps_get_last_error")
        integer(c_int), intent(out) :: status
        character(kind=c_char), intent(out) :: message(*)
        integer(c_int), value :: message_len
        status = last_status
        call to_cstr(last_error, message, int(message_len))
    end subroutine

    subroutine opensps_get_driver_version(major, minor, patch) bind(C, name="opensps_get_driver_version")
        integer(c_int), intent(out) :: major, minor, patch
        major = 0; minor = 3; patch = 0
    end subroutine

    type(c_ptr) function opensps_get_version_string() bind(C, name="opensps_get_version_string")
        opensps_get_version_string = c_loc(version_string)
    end function

    type(c_ptr) function opensps_get_build_id() bind(C, name="opensps_get_build_id")
        opensps_get_build_id = c_loc(build_id)
    end function

    subroutine opensps_get_nbgrad(nbgrad_out) bind(C, name="opensps_get_nbgrad")
        integer(c_int), intent(out) :: nbgrad_out
        nbgrad_out = NBGRAD
    end subroutine

    subroutine opensps_get_grad_param_name(grad_idx, name, max_len, status) bind(C, name="opensps_get_grad_param_name")
        integer(c_int), value :: grad_idx, max_len
        character(kind=c_char), intent(out) :: name(*)
        integer(c_int), intent(out) :: status
        status = 0
        if (grad_idx < 0 .or. grad_idx >= NBGRAD) then
            status = 3
            return
        end if
        call to_cstr(GRAD_NAMES(grad_idx + 1), name, int(max_len))
    end subroutine

    ! ---- contexts -----------------------------------------------------------------------------------------------
    subroutine opensps_context_create(handle, status) bind(C, name="opensps_context_create")
        integer(c_int), intent(out) :: handle, status
        handle = ctx_create()
        status = 0
    end subroutine

    subroutine opensps_context_destroy(handle, status) bind(C, name="opensps_context_destroy")
        integer(c_int), value :: handle
        integer(c_int), intent(out) :: status
        if (bad_handle(handle, status)) return
        call ctx_destroy(int(handle))
    end subroutine

    !> Set-up from a data release (isochrones, spectral library, dust) is not part of this build: status 301
    subroutine opensps_context_setup(zin, isoc_type, spec_type, dust_type, handle, status) &
        bind(C, name="opensps_context_setup")
        integer(c_int), value :: zin, handle
        character(kind=c_char), intent(in) :: isoc_type(*), spec_type(*), dust_type(*)
        integer(c_int), intent(out) :: status
        if (bad_handle(handle, status)) return
        status = 301
        call set_error(301, 'setup: data-release set-up is not part of this build; install a basis with ' // &
                       'opensps_context_set_ssp_basis')
    end subroutine

    !> No reference counterpart (the reference's set-up builds the basis from a data release and installs it with
    !> opensps_context_update_ssp_basis, src/core/opensps_context.f90:757): installs a resident SSP basis given by
    !> the caller. lambda (nspec, Angstrom, increasing), log_age (ntfull, log10 yr, increasing), zlegend (nz,
    !> metallicity mass fractions, increasing), zsol, spec (nspec x ntfull x nz, L_sun/Hz per solar mass),
    !> mass (ntfull x nz), lbol (ntfull x nz, log10 L_sun). status 1 bad handle, 3 bad sizes or grids.
    subroutine opensps_context_set_ssp_basis(handle, nspec, lambda, ntfull, log_age, nz, zlegend, zsol, spec, mass, &
                                             lbol, status) bind(C, name="opensps_context_set_ssp_basis")
        integer(c_int), value :: handle, nspec, ntfull, nz
        real(c_double), value :: zsol
        real(c_double), intent(in) :: lambda(*), log_age(*), zlegend(*), spec(*), mass(*), lbol(*)
        integer(c_int), intent(out) :: status
        integer :: st
        if (bad_handle(handle, status)) return
        if (nspec < 2 .or. ntfull < 2 .or. nz < 1) then
            status = 3
            call set_error(3, 'set_ssp_basis: bad sizes')
            return
        end if
        call ctx_set_basis(int(handle), int(nspec), lambda(1:nspec), int(ntfull), log_age(1:ntfull), int(nz), &
                           zlegend(1:nz), zsol, reshape(spec(1:nspec * ntfull * nz), [nspec, ntfull, nz]), &
                           reshape(mass(1:ntfull * nz), [ntfull, nz]), reshape(lbol(1:ntfull * nz), [ntfull, nz]), st)
        status = st
        if (st /= 0) call set_error(st, 'set_ssp_basis: grids must be increasing and Z, zsol positive')
    end subroutine

    subroutine opensps_context_set_int(handle, key, value, status) bind(C, name="opensps_context_set_int")
        integer(c_int), value :: handle, value
        character(kind=c_char), intent(in) :: key(*)
        integer(c_int), intent(out) :: status
        integer :: st
        if (bad_handle(handle, status)) return
        call ctx_set_int(int(handle), fstr(key), int(value), st)
        status = st
        if (st /= 0) call set_error(st, 'set_int: unknown key ' // fstr(key))
    end subroutine

    subroutine opensps_context_set_float(handle, key, value, status) bind(C, name="opensps_context_set_float")
        integer(c_int), value :: handle
        character(kind=c_char), intent(in) :: key(*)
        real(c_double), value :: value
        integer(c_int), intent(out) :: status
        integer :: st
        if (bad_handle(handle, status)) return
        call ctx_set_float(int(handle), fstr(key), value, st)
        status = st
        if (st /= 0) call set_error(st, 'set_float: unknown key ' // fstr(key))
    end subroutine

    subroutine opensps_context_set_str(handle, key, value, status) bind(C, name="opensps_context_set_str")
        integer(c_int), value :: handle
        character(kind=c_char), intent(in) :: key(*), value(*)
        integer(c_int), intent(out) :: status
        integer :: st
        if (bad_handle(handle, status)) return
        call ctx_set_str(int(handle), fstr(key), fstr(value), st)
        status = st
        if (st /= 0) call set_error(st, 'set_str: unknown key or value ' // fstr(key))
    end subroutine

    subroutine opensps_context_set_fast_mode(handle, fast_mode, status) bind(C, name="opensps_context_set_fast_mode")
        integer(c_int), value :: handle, fast_mode
        integer(c_int), intent(out) :: status
        if (bad_handle(handle, status)) return
    end subroutine

    ! ---- dimensions and grids -----------------------------------------------------------------------------------
    subroutine opensps_context_get_dims(handle, n_spec, n_time, status) bind(C, name="opensps_context_get_dims")
        integer(c_int), value :: handle
        integer(c_int), intent(out) :: n_spec, n_time, status
        n_spec = 0; n_time = 0
        if (bad_handle(handle, status)) return
        n_spec = ctxs(handle)%nspec
        n_time = ctxs(handle)%nt
        if (.not. ctxs(handle)%has_basis) status = 301
    end subroutine

    subroutine opensps_context_get_nspec(handle, n, status) bind(C, name="opensps_context_get_nspec")
        integer(c_int), value :: handle
        integer(c_int), intent(out) :: n, status
        n = 0
        if (bad_handle(handle, status)) return
        n = ctxs(handle)%nspec
    end subroutine

    subroutine opensps_context_get_ntfull(handle, n, status) bind(C, name="opensps_context_get_ntfull")
        integer(c_int), value :: handle
        integer(c_int), intent(out) :: n, status
        n = 0
        if (bad_handle(handle, status)) return
        n = ctxs(handle)%nt
    end subroutine

    subroutine opensps_context_get_nbands(handle, n, status) bind(C, name="opensps_context_get_nbands")
        integer(c_int), value :: handle
        integer(c_int), intent(out) :: n, status
        n = 0
        if (bad_handle(handle, status)) return
        n = ctxs(handle)%nbands
    end subroutine

    subroutine opensps_context_get_nz(handle, n, status) bind(C, name="opensps_context_get_nz")
        integer(c_int), value :: handle
        integer(c_int), intent(out) :: n, status
        n = 0
        if (bad_handle(handle, status)) return
        n = ctxs(handle)%nz
    end subroutine

    subroutine opensps_context_get_nindx(handle, n, status) bind(C, name="opensps_context_get_nindx")
        integer(c_int), value :: handle
        integer(c_int), intent(out) :: n, status
        n = 0
        if (bad_handle(handle, status)) return
    end subroutine

    subroutine opensps_context_get_nemline(handle, n, status) bind(C, name="opensps_context_get_nemline")
        integer(c_int), value :: handle
        integer(c_int), intent(out) :: n, status
        n = 166
        if (bad_handle(handle, status)) return
        n = nemline(ctxs(handle))
    end subroutine

    subroutine opensps_context_get_zsol(handle, z, status) bind(C, name="opensps_context_get_zsol")
        integer(c_int), value :: handle
        real(c_double), intent(out) :: z
        integer(c_int), intent(out) :: status
        z = 0.0_wp
        if (bad_handle(handle, status)) return
        z = ctxs(handle)%zsol
    end subroutine

    subroutine opensps_context_get_lambda(handle, lambda, status) bind(C, name="opensps_context_get_lambda")
        integer(c_int), value :: handle
        real(c_double), intent(out) :: lambda(*)
        integer(c_int), intent(out) :: status
        if (bad_handle(handle, status)) return
        if (.not. ctxs(handle)%has_basis) then
            status = 301; return
        end if
        lambda(1:ctxs(handle)%nspec) = ctxs(handle)%lam
    end subroutine

    subroutine opensps_context_get_ages(handle, log_age, status) bind(C, name="opensps_context_get_ages")
        integer(c_int), value :: handle
        real(c_double), intent(out) :: log_age(*)
        integer(c_int), intent(out) :: status
        if (bad_handle(handle, status)) return
        if (.not. ctxs(handle)%has_basis) then
            status = 301; return
        end if
        log_age(1:ctxs(handle)%nt) = ctxs(handle)%log_age
    end subroutine

    subroutine opensps_context_get_zlegend(handle, z, status) bind(C, name="opensps_context_get_zlegend")
        integer(c_int), value :: handle
        real(c_double), intent(out) :: z(*)
        integer(c_int), intent(out) :: status
        if (bad_handle(handle, status)) return
        if (.not. ctxs(handle)%has_basis) then
            status = 301; return
        end if
        z(1:ctxs(handle)%nz) = ctxs(handle)%zlegend
    end subroutine

    !> Effective (pivot) wavelengths and the AB magnitudes of Vega and of the Sun in every band (the Sun's in the Vega
    !> system when compute_vega_mags was 1 at set_filters); 99 without reference spectra or without flux
    subroutine opensps_context_get_filter_data(handle, wave_eff, mag_vega, mag_sun, status) &
        bind(C, name="opensps_context_get_filter_data")
        integer(c_int), value :: handle
        type(c_ptr), value :: wave_eff, mag_vega, mag_sun
        integer(c_int), intent(out) :: status
        real(c_double), pointer :: o(:)
        integer :: nb
        if (bad_handle(handle, status)) return
        nb = ctxs(handle)%nbands
        if (nb == 0) return
        if (c_associated(wave_eff)) then
            call c_f_pointer(wave_eff, o, [nb])
            o = ctxs(handle)%wave_eff
        end if
        if (c_associated(mag_vega)) then
            call c_f_pointer(mag_vega, o, [nb])
            o = 99.0_wp
            if (allocated(ctxs(handle)%magvega)) o = ctxs(handle)%magvega
        end if
        if (c_associated(mag_sun)) then
            call c_f_pointer(mag_sun, o, [nb])
            o = 99.0_wp
            if (allocated(ctxs(handle)%magsun)) o = ctxs(handle)%magsun
        end if
    end subroutine

    !> No reference counterpart (the reference reads Vega and the Sun from its data release at set-up,
    !> load_standard_sed, src/io/opensps_io.f90:1516-1582): the reference spectra as tables, wavelength (Angstrom,
    !> increasing) and flux, F_lambda (flambda != 0, Vega's table) or f_nu (the Sun's, at 10 pc). They are placed on
    !> the model grid as the reference places them when filters are installed, and give the Vega and solar zero
    !> points (opensps_context_get_filter_data) and the Vega system (compute_vega_mags = 1). status 1 bad handle, 3
    !> fewer than 2 points, a non-increasing or non-positive wavelength, a negative or non-finite flux.
    subroutine opensps_context_set_reference_spectra(handle, n_vega, lam_vega, flux_vega, vega_flambda, n_sun, &
                                                     lam_sun, flux_sun, sun_flambda, status) &
        bind(C, name="opensps_context_set_reference_spectra")
        integer(c_int), value :: handle, n_vega, vega_flambda, n_sun, sun_flambda
        real(c_double), intent(in) :: lam_vega(*), flux_vega(*), lam_sun(*), flux_sun(*)
        integer(c_int), intent(out) :: status
        if (bad_handle(handle, status)) return
        if (.not. (table_ok(lam_vega(1:max(n_vega, 0)), flux_vega(1:max(n_vega, 0))) .and. &
                   table_ok(lam_sun(1:max(n_sun, 0)), flux_sun(1:max(n_sun, 0))))) then
            status = 3
            call set_error(3, 'set_reference_spectra: a table has fewer than 2 points, a non-increasing or '// &
                           'non-positive wavelength or a negative or non-finite flux')
            return
        end if
        associate(c => ctxs(handle))
            c%vega_lam = lam_vega(1:n_vega)
            c%vega_flux = flux_vega(1:n_vega)
            c%sun_lam = lam_sun(1:n_sun)
            c%sun_flux = flux_sun(1:n_sun)
            c%vega_flambda = vega_flambda /= 0
            c%sun_flambda = sun_flambda /= 0
            c%has_ref = .true.
        end associate
        if (ctxs(handle)%nbands > 0) call install_zero_points(handle)
    contains
        logical function table_ok(l, f)
            real(c_double), intent(in) :: l(:), f(:)
            table_ok = size(l) >= 2
            if (.not. table_ok) return
            table_ok = all(l == l) .and. all(f == f) .and. all(abs(l) <= huge(1.0_wp)) .and. &
                       all(abs(f) <= huge(1.0_wp)) .and. all(l > 0.0_wp) .and. all(f >= 0.0_wp)
            if (table_ok) table_ok = all(l(2:) > l(:size(l) - 1))
        end function table_ok
    end subroutine

    !> opensps_context_set_sfh_tab (include/opensps.h:229-232; src/bindings/c/opensps_c_api.f90:2062-2089): the SFH
    !> table of sfh = 2, 3, ntab forward ages (Gyr, stored in years), SFRs (M_sun / yr) and metallicities (met may be
    !> NULL; the batch uses one metallicity, as the reference's). status 1 bad handle, 3 ntab out of 1 .. 20000.
    subroutine opensps_context_set_sfh_tab(handle, ntab, age, sfr, met, status) bind(C, name="opensps_context_set_sfh_tab")
        integer(c_int), value :: handle, ntab
        type(c_ptr), value :: age, sfr, met
        integer(c_int), intent(out) :: status
        real(c_double), pointer :: f(:)
        if (bad_handle(handle, status)) return
        if (ntab < 1 .or. ntab > 20000 .or. .not. c_associated(age) .or. .not. c_associated(sfr)) then
            status = 3
            call set_error(3, 'opensps_context_set_sfh_tab: ntab out of range')
            return
        end if
        associate(c => ctxs(handle))
            c%ntab = ntab
            call c_f_pointer(age, f, [ntab])
            c%tab_age = f * 1.0e9_wp
            call c_f_pointer(sfr, f, [ntab])
            c%tab_sfr = f
            if (allocated(c%tab_met)) deallocate(c%tab_met)
            allocate(c%tab_met(ntab))
            c%tab_met = 0.0_wp
            if (c_associated(met)) then
                call c_f_pointer(met, f, [ntab])
                c%tab_met = f
            end if
            c%dev_tab = .false.
        end associate
    end subroutine

    !> The filters' zero points from the reference spectra (opensps_context_set_filters's filters_zero_points after
    !> the calibration, src/spectra/opensps_context_filters.f90:113-124, with sun_to_vega_system when
    !> compute_vega_mags is 1); without reference spectra none (99)
    subroutine install_zero_points(handle)
        integer(c_int), intent(in) :: handle
        real(WP), allocatable :: spec(:)
        associate(c => ctxs(handle))
            if (allocated(c%magvega)) deallocate(c%magvega)
            if (allocated(c%magsun)) deallocate(c%magsun)
            c%dev_vega = .false.
            if (.not. c%has_ref .or. c%nbands < 1) return
            allocate(c%magvega(c%nbands), c%magsun(c%nbands), spec(c%nspec))
            call ref_spectrum_on_grid(c%vega_lam, c%vega_flux, c%vega_flambda, c%lam, spec)
            call zero_points(c%lam, spec, c%bands, c%magvega)
            call ref_spectrum_on_grid(c%sun_lam, c%sun_flux, c%sun_flambda, c%lam, spec)
            call zero_points(c%lam, spec, c%bands, c%magsun)
            if (c%compute_vega_mags == 1) call sun_to_vega(c%magsun, c%magvega, c%vega_ref)
        end associate
    end subroutine install_zero_points

    ! ---- filters ------------------------------------------------------------------------------------------------
    subroutine opensps_filterset_create(fset, status) bind(C, name="opensps_filterset_create")
        integer(c_int), intent(out) :: fset, status
        fset = fset_create()
        status = 0
    end subroutine

    subroutine opensps_filterset_add(fset, name, n, lambda_aa, response, detector, cal_lambda_ref, cal_exponent, &
                                     status) bind(C, name="opensps_filterset_add")
        integer(c_int), value :: fset, n, detector
        character(kind=c_char), intent(in) :: name(*)
        real(c_double), intent(in) :: lambda_aa(*), response(*)
        real(c_double), value :: cal_lambda_ref, cal_exponent
        integer(c_int), intent(out) :: status
        integer :: st
        status = 0
        if (.not. fset_valid(int(fset))) then
            status = 1
            call set_error(1, 'filterset_add: bad filter set')
            return
        end if
        if (n < 1) then
            status = 3
            call set_error(3, 'filterset_add: n < 1')
            return
        end if
        call fset_add(int(fset), fstr(name), lambda_aa(1:n), response(1:n), int(detector), cal_lambda_ref, &
                      cal_exponent, st)
        if (st /= FILTER_OK) then
            status = 3
            call set_error(3, 'filterset_add: bad sizes, detector or calibration')
        end if
    end subroutine

    subroutine opensps_filterset_count(fset, n, status) bind(C, name="opensps_filterset_count")
        integer(c_int), value :: fset
        integer(c_int), intent(out) :: n, status
        n = 0
        status = 0
        if (.not. fset_valid(int(fset))) then
            status = 1; return
        end if
        n = fsets(fset)%n
    end subroutine

    subroutine opensps_filterset_free(fset, status) bind(C, name="opensps_filterset_free")
        integer(c_int), value :: fset
        integer(c_int), intent(out) :: status
        status = 0
        if (.not. fset_valid(int(fset))) then
            status = 1; return
        end if
        call fset_free(int(fset))
    end subroutine

    !> opensps_context_set_filters (include/opensps.h:209-219): status 301 no basis, 320 unusable curve, 321 a
    !> filter beyond the grid (grid_policy 0), 322 no throughput (grid_policy 0); the context keeps its filters on
    !> an error. vega_ref_band: the Vega system's reference band (0: Vega is magnitude 0 in every band; k > 0: band k
    !> stays AB and the others are shifted relative to it).
    subroutine opensps_context_set_filters(handle, fset, resample, grid_policy, vega_ref_band, status) &
        bind(C, name="opensps_context_set_filters")
        integer(c_int), value :: handle, fset, resample, grid_policy, vega_ref_band
        integer(c_int), intent(out) :: status
        real(WP), allocatable :: bands(:, :)
        integer, allocatable :: flags(:)
        integer :: st, bad, nb, b
        character(len=16) :: num
        if (bad_handle(handle, status)) return
        if (.not. ctxs(handle)%has_basis) then
            status = 301
            call set_error(301, 'set_filters: no basis installed')
            return
        end if
        if (.not. fset_valid(int(fset))) then
            status = 3
            call set_error(3, 'set_filters: bad filter set')
            return
        end if
        nb = fsets(fset)%n
        if (nb < 1 .or. (resample /= 0 .and. resample /= 1) .or. grid_policy < 0 .or. grid_policy > 2 .or. &
            vega_ref_band < 0 .or. vega_ref_band > nb) then
            status = 3
            call set_error(3, 'set_filters: empty set, or resample, grid policy or Vega band out of range')
            return
        end if
        allocate(bands(ctxs(handle)%nspec, nb), flags(nb))
        call filters_build(int(fset), ctxs(handle)%lam, int(resample), int(grid_policy), bands, flags, st, bad)
        if (st /= FILTER_OK) then
            write(num, '(I0)') bad
            select case (st)
            case (FILTER_ERR_COVERAGE)
                status = 321
            case (FILTER_ERR_DEAD)
                status = 322
            case default
                status = 320
            end select
            call set_error(int(status), 'set_filters: filter ' // trim(num) // ' (' // &
                           trim(fsets(fset)%f(bad)%name) // ') is unusable or outside the model grid')
            return
        end if
        associate(c => ctxs(handle))
            c%nbands = nb
            call move_alloc(bands, c%bands)
            call move_alloc(flags, c%bflags)
            if (allocated(c%mag_mask)) deallocate(c%mag_mask)
            allocate(c%mag_mask(nb), c%wave_eff(nb))
            c%mag_mask = 1
            do b = 1, nb
                c%wave_eff(b) = band_pivot(c%lam, c%bands(:, b))
            end do
            c%resample = resample
            c%grid_policy = grid_policy
            c%vega_ref = vega_ref_band
            c%dev_bands = .false.
        end associate
        call install_zero_points(handle)
    end subroutine

    !> No reference counterpart: the context's bands on the model grid (nspec x nbands, column-major, normalised as
    !> set_filters leaves them), for callers that integrate spectra themselves. status 1 bad handle, 301 no filters.
    subroutine opensps_context_get_bands(handle, bands, status) bind(C, name="opensps_context_get_bands")
        integer(c_int), value :: handle
        real(c_double), intent(out) :: bands(*)
        integer(c_int), intent(out) :: status
        integer :: nb, ns
        if (bad_handle(handle, status)) return
        nb = ctxs(handle)%nbands
        ns = ctxs(handle)%nspec
        if (nb == 0) then
            status = 301
            return
        end if
        bands(1:ns * nb) = reshape(ctxs(handle)%bands, [ns * nb])
    end subroutine

    subroutine opensps_context_set_mag_compute(handle, n_bands, mask, status) bind(C, name="opensps_context_set_mag_compute")
        integer(c_int), value :: handle, n_bands
        type(c_ptr), value :: mask
        integer(c_int), intent(out) :: status
        integer(c_int), pointer :: m(:)
        if (bad_handle(handle, status)) return
        associate(c => ctxs(handle))
            if (n_bands < 1 .or. n_bands > c%nbands) then
                status = 3
                call set_error(3, 'set_mag_compute: n_bands out of range')
                return
            end if
            if (.not. allocated(c%mag_mask)) allocate(c%mag_mask(c%nbands))
            c%mag_mask = 0
            if (c_associated(mask)) then
                call c_f_pointer(mask, m, [n_bands])
                c%mag_mask(1:n_bands) = merge(1, 0, m /= 0)
            else
                c%mag_mask = 1
            end if
            c%dev_mask = .false.
        end associate
    end subroutine

    ! ---- the batch ----------------------------------------------------------------------------------------------
    subroutine opensps_context_compute_csp_batch(handle, n, n_params, c_names, c_values, c_weights, c_mags, c_spec, &
                                                 c_mass, c_lbol, c_mdust, c_sfr, c_emlines, c_indx, status) &
        bind(C, name="opensps_context_compute_csp_batch")
        integer(c_int), value :: handle, n, n_params
        type(c_ptr), value :: c_names, c_values, c_weights, c_mags, c_spec, c_mass, c_lbol, c_mdust, c_sfr, &
                              c_emlines, c_indx
        integer(c_int), intent(out) :: status
        integer :: st
        call batch_run(int(handle), int(n), int(n_params), c_names, c_values, c_weights, c_mags, c_spec, c_mass, &
                       c_lbol, c_mdust, c_sfr, c_emlines, c_indx, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, &
                       c_null_ptr, c_null_ptr, st)
        status = st
    end subroutine

    subroutine opensps_context_compute_csp_batch_parts(handle, n, n_params, c_names, c_values, c_weights, c_mags, &
                                                       c_spec, c_mass, c_lbol, c_mdust, c_sfr, c_emlines, c_indx, &
                                                       c_young, c_old, c_mass_young, c_mass_old, c_lbol_young, &
                                                       c_lbol_old, status) &
        bind(C, name="opensps_context_compute_csp_batch_parts")
        integer(c_int), value :: handle, n, n_params
        type(c_ptr), value :: c_names, c_values, c_weights, c_mags, c_spec, c_mass, c_lbol, c_mdust, c_sfr, &
                              c_emlines, c_indx, c_young, c_old, c_mass_young, c_mass_old, c_lbol_young, c_lbol_old
        integer(c_int), intent(out) :: status
        integer :: st
        if (.not. c_associated(c_young) .or. .not. c_associated(c_old)) then
            status = 3
            call set_error(3, 'compute_csp_batch_parts: young and old are required')
            return
        end if
        call batch_run(int(handle), int(n), int(n_params), c_names, c_values, c_weights, c_mags, c_spec, c_mass, &
                       c_lbol, c_mdust, c_sfr, c_emlines, c_indx, c_young, c_old, c_mass_young, c_mass_old, &
                       c_lbol_young, c_lbol_old, st)
        status = st
    end subroutine

    subroutine opensps_context_compute_csp_batch_grad(handle, n, n_params, c_names, c_values, c_weights, &
                                                      c_weights_z, c_zred_mags, c_mags_bar, c_spec_bar, c_mass_bar, &
                                                      c_lbol_bar, c_mags, c_spec, c_mass, c_lbol, c_mdust, c_sfr, &
                                                      c_emlines, c_indx, c_params_bar, c_weights_bar, &
                                                      c_weights_z_bar, status) &
        bind(C, name="opensps_context_compute_csp_batch_grad")
        use opensps_host_grad, only: batch_run_grad
        integer(c_int), value :: handle, n, n_params
        type(c_ptr), value :: c_names, c_values, c_weights, c_weights_z, c_zred_mags, c_mags_bar, c_spec_bar, &
                              c_mass_bar, c_lbol_bar, c_mags, c_spec, c_mass, c_lbol, c_mdust, c_sfr, c_emlines, &
                              c_indx, c_params_bar, c_weights_bar, c_weights_z_bar
        integer(c_int), intent(out) :: status
        integer :: st
        call batch_run_grad(int(handle), int(n), int(n_params), c_names, c_values, c_weights, c_weights_z, &
                            c_zred_mags, c_mags_bar, c_spec_bar, c_mass_bar, c_lbol_bar, c_mags, c_spec, c_mass, &
                            c_lbol, c_mdust, c_sfr, c_emlines, c_indx, c_params_bar, NBGRAD, c_weights_bar, &
                            c_weights_z_bar, st)
        status = st
    end subroutine

    subroutine opensps_context_compute_csp_batch_grad_rows(handle, n, n_params, c_names, c_values, c_weights, &
                                                           c_weights_z, c_zred_mags, c_mags_bar, c_spec_bar, &
                                                           c_mass_bar, c_lbol_bar, c_mags, c_spec, c_mass, c_lbol, &
                                                           c_mdust, c_sfr, c_emlines, c_indx, c_params_bar, &
                                                           params_bar_rows, c_weights_bar, c_weights_z_bar, status) &
        bind(C, name="opensps_context_compute_csp_batch_grad_rows")
        use opensps_host_grad, only: batch_run_grad
        integer(c_int), value :: handle, n, n_params, params_bar_rows
        type(c_ptr), value :: c_names, c_values, c_weights, c_weights_z, c_zred_mags, c_mags_bar, c_spec_bar, &
                              c_mass_bar, c_lbol_bar, c_mags, c_spec, c_mass, c_lbol, c_mdust, c_sfr, c_emlines, &
                              c_indx, c_params_bar, c_weights_bar, c_weights_z_bar
        integer(c_int), intent(out) :: status
        integer :: st
        call batch_run_grad(int(handle), int(n), int(n_params), c_names, c_values, c_weights, c_weights_z, &
                            c_zred_mags, c_mags_bar, c_spec_bar, c_mass_bar, c_lbol_bar, c_mags, c_spec, c_mass, &
                            c_lbol, c_mdust, c_sfr, c_emlines, c_indx, c_params_bar, int(params_bar_rows), &
                            c_weights_bar, c_weights_z_bar, st)
        status = st
    end subroutine

    !> No reference counterpart: the batch on device-resident arrays, for callers that keep their data on the
    !> GPU (the reference's opensps_context_batch_spec_device serves the same callers on the output side,
    !> include/opensps.h:308-316). d_prm: device (OPS_NPAR x n) doubles, the per-model parameters in the slots of
    !> opensps_params.h; d_weights (ntfull x n) or NULL; precision -1 (the context's), 0 FP64, 1 mixed; split 1
    !> for the young and old sums apart (birth-cloud or old-star screens). Outputs are device arrays: d_mags
    !> (nbands x n, required), d_spec (nspec x n, float with mixed precision, double with FP64), d_mass, d_lbol
    !> (n); any but d_mags may be NULL. status as the batch; 5 without a device.
    subroutine opensps_context_compute_csp_batch_device(handle, n, d_prm, d_weights, precision, split, d_mags, d_spec, &
                                                        d_mass, d_lbol, status) &
        bind(C, name="opensps_context_compute_csp_batch_device")
        integer(c_int), value :: handle, n, precision, split
        type(c_ptr), value :: d_prm, d_weights, d_mags, d_spec, d_mass, d_lbol
        integer(c_int), intent(out) :: status
        integer :: st
        call batch_run_device(int(handle), int(n), d_prm, d_weights, int(precision), int(split), d_mags, d_spec, &
                              d_mass, d_lbol, st)
        status = st
    end subroutine

    !> No reference counterpart: the context's parameter values in the slot order of opensps_params.h (the
    !> defaults of every slot a device-resident batch does not vary).
    subroutine opensps_context_get_params(handle, prm, status) bind(C, name="opensps_context_get_params")
        integer(c_int), value :: handle
        real(c_double), intent(out) :: prm(NPAR)
        integer(c_int), intent(out) :: status
        if (bad_handle(handle, status)) return
        prm = ctxs(handle)%pset
    end subroutine

    !> No reference counterpart: the accumulated device time of the batch's kernels [prep, group, age-sum product,
    !> chain] in milliseconds and their launch counts, recorded with HIP events on the context's stream after
    !> set_int("profile_kernels", 1); reset != 0 clears them. status 1 bad handle, 5 before any device batch.
    subroutine opensps_context_kernel_times(handle, ms, launches, reset, status) &
        bind(C, name="opensps_context_kernel_times")
        use opensps_hip_iface, only: opensps_hip_kernel_times
        integer(c_int), value :: handle, reset
        real(c_double), intent(out) :: ms(4)
        integer(c_int), intent(out) :: launches(4), status
        integer(c_int) :: rc
        ms = 0.0_wp
        launches = 0
        if (bad_handle(handle, status)) return
        if (.not. c_associated(ctxs(handle)%dctx)) then
            status = 5
            return
        end if
        rc = opensps_hip_kernel_times(ctxs(handle)%dctx, ms, launches, reset)
        status = rc
    end subroutine

    !> No reference counterpart: with profile_kernels on, the SSP operator's phases of each batch whose operator
    !> launches take one round, [the IMF weights, the fold and mass sums], in accumulated milliseconds and batch
    !> counts; reset != 0 clears them. status 1 bad handle, 5 before any device batch.
    subroutine opensps_context_op_kernel_times(handle, ms, launches, reset, status) &
        bind(C, name="opensps_context_op_kernel_times")
        use opensps_hip_iface, only: opensps_hip_op_kernel_times
        integer(c_int), value :: handle, reset
        real(c_double), intent(out) :: ms(2)
        integer(c_int), intent(out) :: launches(2), status
        integer(c_int) :: rc
        ms = 0.0_wp
        launches = 0
        if (bad_handle(handle, status)) return
        if (.not. c_associated(ctxs(handle)%dctx)) then
            status = 5
            return
        end if
        rc = opensps_hip_op_kernel_times(ctxs(handle)%dctx, ms, launches, reset)
        status = rc
    end subroutine

    subroutine opensps_context_batch_spec_device(handle, ptr, nspec, n, bytes, status) &
        bind(C, name="opensps_context_batch_spec_device")
        integer(c_int), value :: handle
        type(c_ptr), intent(out) :: ptr
        integer(c_int), intent(out) :: nspec, n, bytes, status
        ptr = c_null_ptr
        nspec = 0; n = 0; bytes = 0
        if (bad_handle(handle, status)) return
        associate(c => ctxs(handle))
            nspec = c%nspec
            bytes = c%last_bytes
            n = c%last_n
            if (n > 0) ptr = c%d_spec%p
        end associate
    end subroutine

    subroutine opensps_context_get_batch_maggies(handle, maggies, n_bands, n, status) &
        bind(C, name="opensps_context_get_batch_maggies")
        integer(c_int), value :: handle
        type(c_ptr), value :: maggies
        integer(c_int), intent(out) :: n_bands, n, status
        real(c_double), pointer :: o(:, :)
        n_bands = 0; n = 0
        if (bad_handle(handle, status)) return
        associate(c => ctxs(handle))
            if (.not. allocated(c%last_maggies)) then
                status = 302
                call set_error(302, 'get_batch_maggies: no batch has run with keep_maggies = 1')
                return
            end if
            n_bands = size(c%last_maggies, 1)
            n = size(c%last_maggies, 2)
            if (c_associated(maggies)) then
                call c_f_pointer(maggies, o, [n_bands, n])
                o = c%last_maggies
            end if
        end associate
    end subroutine

    ! ---- the SSP operator (reference include/opensps.h:425-455) --------------------------------------------------
    !> opensps_context_build_ssp_ops (src/bindings/c/opensps_c_api.f90:1172-1223) builds the operator from the
    !> isochrones and spectral libraries of a data release set-up, which is not part of this build: status 301, as
    !> the reference returns for a context that is not set up. The operator is given by opensps_context_set_ssp_ops.
    subroutine opensps_context_build_ssp_ops(handle, all_metallicities, status) bind(C, name="opensps_context_build_ssp_ops")
        integer(c_int), value :: handle, all_metallicities
        integer(c_int), intent(out) :: status
        if (bad_handle(handle, status)) return
        status = 301
        call set_error(301, 'opensps_context_build_ssp_ops: context not set up from a data release; give the ' // &
                       'operator with opensps_context_set_ssp_ops')
    end subroutine

    !> opensps_context_build_ssp_ops_free_mods (src/bindings/c/opensps_c_api.f90:1229-1262): status 301 (as above)
    subroutine opensps_context_build_ssp_ops_free_mods(handle, all_metallicities, status) &
        bind(C, name="opensps_context_build_ssp_ops_free_mods")
        integer(c_int), value :: handle, all_metallicities
        integer(c_int), intent(out) :: status
        if (bad_handle(handle, status)) return
        status = 301
        call set_error(301, 'opensps_context_build_ssp_ops_free_mods: context not set up from a data release')
    end subroutine

    !> No reference counterpart (the reference reads its dust emission library at set-up from its data directory,
    !> load_dust_emission_table, src/io/opensps_io.f90:959-1069): the Draine & Li (2007) templates (or any library of
    !> that layout) as given: qpah (nqpah) and umin (numin) ascending, lam (nlam, A) ascending, spectra (nlam, 2, numin,
    !> nqpah) column-major = the data library's [qpah][umin][component][wavelength] (per (q_PAH, U_min) the spectrum
    !> of U_max = U_min, then of the power-law U distribution). The batch places them on the grid as the reference
    !> does and, with add_dust_emission = 1, adds each dusty model's emission by energy balance. status 1 bad handle,
    !> 3 bad sizes or order.
    subroutine opensps_context_set_dust_emission_templates(handle, nqpah, qpah, numin, umin, nlam, lam, spectra, &
                                                           status) bind(C, name="opensps_context_set_dust_emission_templates")
        integer(c_int), value :: handle, nqpah, numin, nlam
        type(c_ptr), value :: qpah, umin, lam, spectra
        integer(c_int), intent(out) :: status
        real(c_double), pointer :: q(:), u(:), l(:), sp(:, :, :)
        integer :: st
        if (bad_handle(handle, status)) return
        if (nqpah < 2 .or. numin < 2 .or. nlam < 2 .or. .not. c_associated(qpah) .or. .not. c_associated(umin) .or. &
            .not. c_associated(lam) .or. .not. c_associated(spectra)) then
            status = 3
            call set_error(3, 'set_dust_emission_templates: bad sizes or NULL arrays')
            return
        end if
        call c_f_pointer(qpah, q, [nqpah])
        call c_f_pointer(umin, u, [numin])
        call c_f_pointer(lam, l, [nlam])
        call c_f_pointer(spectra, sp, [int(nlam), 2 * int(numin), int(nqpah)])
        call ctx_set_dust_emission(int(handle), int(nqpah), q, int(numin), u, int(nlam), l, sp, st)
        status = st
    end subroutine

    !> No reference counterpart (the reference reads its CLOUDY library at set-up, read_nebular_standard,
    !> src/io/opensps_io.f90:914-937, and places it on the grid, load_nebular_grid, :749-823): the tables as given:
    !> logz (nnz), age (nna, yr), logu (nnu), lam (nlam, A) ascending; cont (nnz x nna x nnu x nlam, L_sun/Hz) and lines
    !> (nnz x nna x nnu x nline, L_sun), per unit ionizing photon rate, C order; line_lam (nline, A). With
    !> add_neb_emission = 1 every model adds the nebular continuum (add_neb_continuum = 1) and lines (into the
    !> spectrum with nebemlineinspec = 1) of its young stars' ionizing photons at its gas_logz, gas_logu, and emlines
    !> holds the lines' luminosities. status 1 bad handle, 3 bad sizes, NULL arrays or axes not ascending.
    subroutine opensps_context_set_nebular_tables(handle, nnz, logz, nna, age, nnu, logu, nlam, lam, cont, nline, &
                                                  line_lam, lines, status) bind(C, name="opensps_context_set_nebular_tables")
        integer(c_int), value :: handle, nnz, nna, nnu, nlam, nline
        type(c_ptr), value :: logz, age, logu, lam, cont, line_lam, lines
        integer(c_int), intent(out) :: status
        real(c_double), pointer :: z(:), a(:), u(:), l(:), co(:), ll(:), li(:)
        integer :: st
        if (bad_handle(handle, status)) return
        if (nnz < 2 .or. nna < 2 .or. nnu < 2 .or. nlam < 2 .or. nline < 1 .or. .not. c_associated(logz) .or. &
            .not. c_associated(age) .or. .not. c_associated(logu) .or. .not. c_associated(lam) .or. &
            .not. c_associated(cont) .or. .not. c_associated(line_lam) .or. .not. c_associated(lines)) then
            status = 3
            call set_error(3, 'set_nebular_tables: bad sizes or NULL arrays')
            return
        end if
        call c_f_pointer(logz, z, [nnz])
        call c_f_pointer(age, a, [nna])
        call c_f_pointer(logu, u, [nnu])
        call c_f_pointer(lam, l, [nlam])
        call c_f_pointer(cont, co, [int(nnz) * int(nna) * int(nnu) * int(nlam)])
        call c_f_pointer(line_lam, ll, [nline])
        call c_f_pointer(lines, li, [int(nnz) * int(nna) * int(nnu) * int(nline)])
        call ctx_set_nebular(int(handle), int(nnz), z, int(nna), a, int(nnu), u, int(nlam), l, co, int(nline), ll, &
                             li, st)
        status = st
    end subroutine

    !> No reference counterpart (the reference reads its AGN torus library at set-up, load_agn_dust_models,
    !> src/io/opensps_io.f90:1101-1155): the templates as given: tau (ntau) and lam (nlam, A) ascending, spectra
    !> (nlam, ntau) column-major = the data library's [tau][wavelength]. With add_agn_dust = 1 each model with
    !> fagn > 0 adds fagn L_bol times its agn_tau template, attenuated by the diffuse dust. status 1 bad handle,
    !> 3 bad sizes or order.
    subroutine opensps_context_set_agn_templates(handle, ntau, tau, nlam, lam, spectra, status) &
        bind(C, name="opensps_context_set_agn_templates")
        integer(c_int), value :: handle, ntau, nlam
        type(c_ptr), value :: tau, lam, spectra
        integer(c_int), intent(out) :: status
        real(c_double), pointer :: t(:), l(:), sp(:, :)
        integer :: st
        if (bad_handle(handle, status)) return
        if (ntau < 2 .or. nlam < 2 .or. .not. c_associated(tau) .or. .not. c_associated(lam) .or. &
            .not. c_associated(spectra)) then
            status = 3
            call set_error(3, 'set_agn_templates: bad sizes or NULL arrays')
            return
        end if
        call c_f_pointer(tau, t, [ntau])
        call c_f_pointer(lam, l, [nlam])
        call c_f_pointer(spectra, sp, [int(nlam), int(ntau)])
        call ctx_set_agn(int(handle), int(ntau), t, int(nlam), l, sp, st)
        status = st
    end subroutine

    !> No reference counterpart (the reference builds the operator set from isochrones, ssp_operator_build and
    !> ssp_opset_build, src/physics/opensps_ssp_operator.f90, opensps_ssp_batch.f90:99-147): installs an operator
    !> given by the caller on the GPU. With it every batch takes each model's IMF (imf1, imf2, imf3,
    !> imf_lower_limit, imf_upper_limit may vary per model). Arrays as opensps_hip_ssp_ops_t (include/opensps_hip.h).
    !> status 1 bad handle or no basis, 3 bad sizes or tables, 5 no HIP device, 10 device error.
    subroutine opensps_context_set_ssp_ops(handle, nz, ntfull, npts, ntpl, nent, age_ptr, mini, m1, m2, first, cm, &
                                           cl, q, age_mmax, ent_tpl, ent_val, etpl, status) &
        bind(C, name="opensps_context_set_ssp_ops")
        integer(c_int), value :: handle, nz, ntfull, npts, ntpl, nent
        type(c_ptr), value :: age_ptr, mini, m1, m2, first, cm, cl, q, age_mmax, ent_tpl, ent_val, etpl
        integer(c_int), intent(out) :: status
        type(hip_ssp_ops_t) :: o
        integer(c_int) :: rc
        integer :: st
        if (bad_handle(handle, status)) return
        associate(c => ctxs(handle))
            if (.not. c%has_basis) then
                status = 1
                call set_error(1, 'set_ssp_ops: no SSP basis installed')
                return
            end if
            if (nz /= c%nz .or. ntfull /= c%nt) then
                status = 3
                call set_error(3, 'set_ssp_ops: the operator''s metallicities and ages must be the basis''s')
                return
            end if
            st = ctx_device(int(handle))
            if (st /= 0) then
                status = st
                return
            end if
            o%nz = nz; o%nt = ntfull; o%npts = npts; o%ntpl = ntpl; o%nent = nent
            o%age_ptr = age_ptr; o%mini = mini; o%m1 = m1; o%m2 = m2; o%first = first
            o%cm = cm; o%cl = cl; o%q = q; o%age_mmax = age_mmax
            o%ent_tpl = ent_tpl; o%ent_val = ent_val; o%etpl = etpl
            rc = opensps_hip_set_ssp_ops(c%dctx, o)
            if (rc /= 0) then
                status = merge(int(rc), 10, rc == 1 .or. rc == 2 .or. rc == 3)
                call set_error(int(status), 'set_ssp_ops: ' // trim(hip_message(c%dctx)))
                c%has_ops = .false.
                return
            end if
            c%has_ops = .true.
            c%dev_imf = .false.
        end associate
    end subroutine

    !> opensps_context_free_ssp_ops (reference include/opensps.h:443; src/bindings/c/opensps_c_api.f90:1266-1278)
    subroutine opensps_context_free_ssp_ops(handle, status) bind(C, name="opensps_context_free_ssp_ops")
        integer(c_int), value :: handle
        integer(c_int), intent(out) :: status
        integer(c_int) :: rc
        if (bad_handle(handle, status)) return
        associate(c => ctxs(handle))
            if (c%has_ops .and. c_associated(c%dctx)) rc = opensps_hip_free_ssp_ops(c%dctx)
            c%has_ops = .false.
        end associate
    end subroutine

    !> opensps_context_has_ssp_ops (reference include/opensps.h:446; src/bindings/c/opensps_c_api.f90:1281-1293)
    subroutine opensps_context_has_ssp_ops(handle, has_ops, status) bind(C, name="opensps_context_has_ssp_ops")
        integer(c_int), value :: handle
        integer(c_int), intent(out) :: has_ops, status
        has_ops = 0
        if (bad_handle(handle, status)) return
        if (ctxs(handle)%has_ops) has_ops = 1
    end subroutine

    !> opensps_context_evaluate_ssp_op (reference include/opensps.h:451-455): the SSPs of every age for one IMF
    !> (opensps_host_batch's ssp_op_evaluate)
    subroutine opensps_context_evaluate_ssp_op(handle, imf_type, imf_params, mass_grid, lbol_grid, spec_grid, status) &
        bind(C, name="opensps_context_evaluate_ssp_op")
        integer(c_int), value :: handle, imf_type
        type(c_ptr), value :: imf_params, mass_grid, lbol_grid, spec_grid
        integer(c_int), intent(out) :: status
        integer :: st
        if (bad_handle(handle, status)) return
        call ssp_op_evaluate(int(handle), int(imf_type), imf_params, mass_grid, lbol_grid, spec_grid, st)
        status = st
    end subroutine

    !> No reference counterpart: the context's cosmology table (redshift, luminosity distance in pc; nrow = 500 rows,
    !> the table the batch's distance term interpolates, cosmospl, src/physics/opensps_cosmology.f90:288-300), for callers
    !> that drive the device layer (libopensps_hip) themselves. z, dl may be NULL (nrow only).
    subroutine opensps_context_get_cosmology(handle, nrow, z, dl, status) bind(C, name="opensps_context_get_cosmology")
        integer(c_int), value :: handle
        integer(c_int), intent(out) :: nrow, status
        type(c_ptr), value :: z, dl
        real(c_double), pointer :: oz(:), odl(:)
        nrow = 0
        if (bad_handle(handle, status)) return
        call ctx_refresh_cosmology(int(handle))
        associate(c => ctxs(handle))
            nrow = size(c%cz)
            if (c_associated(z)) then
                call c_f_pointer(z, oz, [nrow])
                oz = c%cz
            end if
            if (c_associated(dl)) then
                call c_f_pointer(dl, odl, [nrow])
                odl = c%cdl
            end if
        end associate
    end subroutine

end module opensps_c_api
