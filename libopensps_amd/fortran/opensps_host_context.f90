!> Contexts of the Fortran host: settings and parameters (the reference's opensps_context_t, its key/value
!> setters src/core/opensps_context.f90:294-666 and defaults :60-100), the resident SSP basis and grids, the
!> filters, the cosmology table, and the device side (a libopensps_hip context plus the device buffers of the
!> batch, kept between calls as the reference keeps batch_out, src/core/opensps_context_types.f90:62-85).
!> The device is touched on the first batch, so that a context can be set up and checked without a GPU.
module opensps_host_context
    use, intrinsic :: iso_c_binding, only: c_ptr, c_null_ptr, c_associated, c_int, c_size_t, c_double, c_char, c_loc
    use opensps_host_types
    use opensps_hip_iface
    use opensps_host_cosmology, only: cosmology_table
    implicit none
    private

    type, public :: dbuf_t
        type(c_ptr) :: p = c_null_ptr
        integer(c_size_t) :: bytes = 0
    end type dbuf_t

    type, public :: host_ctx_t
        logical :: used = .false.
        ! settings (integer keys)
        integer :: add_igm_absorption = 0, add_dust_emission = 1, add_neb_emission = 0, add_agn_dust = 1
        integer :: smooth_velocity = 1, smoothspec_fast = 1, dust_type = 0, compute_vega_mags = 0
        integer :: redshift_colors = 0, keep_maggies = 0, imf_type = 2, mixed = 0, profile = 0
        logical :: dev_prof = .false.
        real(WP) :: h0 = 72.0_wp, om0 = 0.27_wp, ol0 = 0.73_wp
        real(WP) :: pset(NPAR) = PARAM_DEFAULTS
        ! the members no stage here reads (INERT_NAMES), kept as set; the context's IMF settings beyond the slots
        real(WP) :: inert(NINERT) = INERT_DEFAULTS
        real(WP) :: mdave = 0.5_wp, vdmc = 0.08_wp
        integer :: add_stellar_remnants = 1, zmet = 1
        ! an SSP operator installed on the device (opensps_context_set_ssp_ops): the batch then takes each model's IMF
        logical :: has_ops = .false.
        ! the dust emission and AGN torus templates as given (opensps_context_set_dust_emission_templates,
        ! opensps_context_set_agn_templates): the library's wavelengths and spectra, placed on the grid at upload
        logical :: has_dustem = .false., has_agn = .false.
        real(WP), allocatable :: dl_qpah(:), dl_umin(:), dl_lam(:), dl_spec(:, :, :)   ! (nlam, 2 numin, nqpah)
        real(WP), allocatable :: agn_tau(:), agn_lam(:), agn_spec(:, :)                 ! (nlam, ntau)
        ! the CLOUDY nebular library as given (opensps_context_set_nebular_tables): its axes, the continuum's
        ! wavelengths, the continuum (nnz, nna, nnu, nlam) and the lines (nnz, nna, nnu, nline) flat in C order, the
        ! lines' wavelengths; placed on the grid by the device at upload. add_neb_continuum, nebemlineinspec: the
        ! reference's settings of the same names (defaults 1, 1)
        logical :: has_neb = .false.
        integer :: add_neb_continuum = 1, nebemlineinspec = 1, neb_nline = 0
        real(WP), allocatable :: nb_logz(:), nb_age(:), nb_logu(:), nb_lam(:), nb_cont(:), nb_llam(:), nb_lines(:)
        ! the resident basis
        logical :: has_basis = .false.
        integer :: nspec = 0, nt = 0, nz = 0
        real(WP) :: zsol = 0.0142_wp
        real(WP), allocatable :: lam(:), log_age(:), zlegend(:), logz(:), spec(:, :, :), mass(:, :), lbol(:, :)
        ! filters
        integer :: nbands = 0
        real(WP), allocatable :: bands(:, :), wave_eff(:)
        integer, allocatable :: mag_mask(:), bflags(:)
        integer :: resample = 0, grid_policy = 2
        ! the reference spectra as given (Vega, the Sun: wavelength, flux, F_lambda or f_nu) and the filters' zero
        ! points from them (magvega, magsun: 99 without them), the Vega system's reference band
        logical :: has_ref = .false., vega_flambda = .true., sun_flambda = .false.
        real(WP), allocatable :: vega_lam(:), vega_flux(:), sun_lam(:), sun_flux(:), magvega(:), magsun(:)
        integer :: vega_ref = 0
        ! the SFH table of sfh 2, 3 (opensps_context_set_sfh_tab): forward ages (yr), SFRs (M_sun / yr), metallicities
        integer :: ntab = 0
        real(WP), allocatable :: tab_age(:), tab_sfr(:), tab_met(:)
        ! cosmology
        real(WP) :: cz(500) = 0.0_wp, cdl(500) = 0.0_wp
        logical :: cosmo_dirty = .true.
        ! device side
        type(c_ptr) :: dctx = c_null_ptr
        logical :: dev_grid = .false., dev_bands = .false., dev_cosmo = .false., dev_mask = .false.
        logical :: dev_vega = .false., dev_tab = .false., dev_dust = .false., dev_imf = .false., dev_emis = .false.
        logical :: dev_neb = .false.
        type(dbuf_t) :: d_prm, d_w, d_mags, d_spec, d_mass, d_lbol, d_fband, d_young, d_old, d_idx, d_mdust, d_eml
        integer :: last_n = 0, last_bytes = 8
        real(WP), allocatable :: last_maggies(:, :)
    end type host_ctx_t

    type(host_ctx_t), allocatable, target, public, save :: ctxs(:)
    type(host_ctx_t), save :: blank_ctx   ! (default initialised: the state of a new context)

    public :: ctx_create, ctx_destroy, ctx_valid, ctx_set_int, ctx_set_float, ctx_set_str, ctx_set_basis, &
              ctx_device, dbuf_reserve, ctx_refresh_cosmology, hip_message, ctx_set_dust_emission, ctx_set_agn, &
              ctx_set_nebular, nemline

contains

    logical function ctx_valid(h)
        integer, intent(in) :: h
        ctx_valid = .false.
        if (.not. allocated(ctxs)) return
        if (h < 1 .or. h > size(ctxs)) return
        ctx_valid = ctxs(h)%used
    end function ctx_valid

    integer function ctx_create() result(h)
        type(host_ctx_t), allocatable :: tmp(:)
        integer :: i, n
        if (.not. allocated(ctxs)) allocate(ctxs(8))
        do i = 1, size(ctxs)
            if (.not. ctxs(i)%used) then
                ctxs(i) = blank_ctx
                ctxs(i)%used = .true.
                h = i
                return
            end if
        end do
        n = size(ctxs)
        allocate(tmp(2 * n))
        tmp(1:n) = ctxs
        call move_alloc(tmp, ctxs)
        ctxs(n + 1) = blank_ctx
        ctxs(n + 1)%used = .true.
        h = n + 1
    end function ctx_create

    subroutine dbuf_free(b)
        type(dbuf_t), intent(inout) :: b
        integer(c_int) :: st
        if (c_associated(b%p)) st = opensps_hip_free(b%p)
        b%p = c_null_ptr
        b%bytes = 0
    end subroutine dbuf_free

    subroutine ctx_destroy(h)
        integer, intent(in) :: h
        integer(c_int) :: st
        if (.not. ctx_valid(h)) return
        associate(c => ctxs(h))
            call dbuf_free(c%d_prm); call dbuf_free(c%d_w); call dbuf_free(c%d_mags); call dbuf_free(c%d_spec)
            call dbuf_free(c%d_mass); call dbuf_free(c%d_lbol); call dbuf_free(c%d_fband); call dbuf_free(c%d_young)
            call dbuf_free(c%d_old); call dbuf_free(c%d_idx); call dbuf_free(c%d_mdust); call dbuf_free(c%d_eml)
            if (c_associated(c%dctx)) st = opensps_hip_destroy(c%dctx)
        end associate
        ctxs(h) = blank_ctx
    end subroutine ctx_destroy

    !> A device buffer of at least `bytes` (grown, never shrunk); status 10 on a device error
    integer function dbuf_reserve(b, bytes) result(st)
        type(dbuf_t), intent(inout) :: b
        integer(c_size_t), intent(in) :: bytes
        st = 0
        if (b%bytes >= bytes .and. c_associated(b%p)) return
        call dbuf_free(b)
        st = opensps_hip_malloc(b%p, max(bytes, 256_c_size_t))
        if (st == 0) b%bytes = max(bytes, 256_c_size_t)
    end function dbuf_reserve

    !> An SSP parameter of the context changed: an installed SSP operator is dropped, as the reference's setters
    !> invalidate it (invalidate_ssps, src/bindings/c/opensps_c_api.f90:2196-2203, 470-473, 497-502)
    subroutine drop_ops(c)
        type(host_ctx_t), intent(inout) :: c
        integer(c_int) :: rc
        if (c%has_ops .and. c_associated(c%dctx)) rc = opensps_hip_free_ssp_ops(c%dctx)
        c%has_ops = .false.
    end subroutine drop_ops

    !> Integer settings (opensps_context_set_int). status 0, or 101 for an unknown key
    subroutine ctx_set_int(h, key, v, status)
        integer, intent(in) :: h, v
        character(len=*), intent(in) :: key
        integer, intent(out) :: status
        integer :: k
        status = 0
        associate(c => ctxs(h))
        select case (lower(trim(key)))
        case ('add_igm_absorption'); c%add_igm_absorption = v; c%dev_grid = .false.
        case ('add_dust_emission'); c%add_dust_emission = v; c%dev_emis = .false.
        case ('add_neb_emission'); c%add_neb_emission = v; c%dev_neb = .false.
        case ('add_neb_continuum'); c%add_neb_continuum = v; c%dev_neb = .false.
        case ('nebemlineinspec'); c%nebemlineinspec = v; c%dev_neb = .false.
        case ('add_agn_dust'); c%add_agn_dust = v; c%dev_emis = .false.
        case ('smooth_velocity'); c%smooth_velocity = v; c%dev_grid = .false.
        case ('smoothspec_fast'); c%smoothspec_fast = v; c%dev_grid = .false.
        case ('dust_type'); c%dust_type = v; c%dev_dust = .false.
        case ('compute_vega_mags'); c%compute_vega_mags = v; c%dev_vega = .false.
        case ('redshift_colors'); c%redshift_colors = v
        case ('keep_maggies'); c%keep_maggies = v
        case ('imf_type')
            if (v /= c%imf_type) call drop_ops(c)
            c%imf_type = v; c%dev_imf = .false.
        case ('add_stellar_remnants')
            if (v /= c%add_stellar_remnants) call drop_ops(c)
            c%add_stellar_remnants = v; c%dev_imf = .false.
        case ('zmet'); c%zmet = v
        case ('mixed_precision'); c%mixed = merge(1, 0, v /= 0)
        case ('profile_kernels'); c%profile = merge(1, 0, v /= 0); c%dev_prof = .false.
        case ('sfh'); c%pset(P_SFH) = real(v, WP)
        case default
            k = param_slot(key)
            if (k > 0) then
                c%pset(k) = real(v, WP)
            else
                status = 101
            end if
        end select
        end associate
    end subroutine ctx_set_int

    !> Float settings and parameters (opensps_context_set_float). status 0, or 102 for an unknown key
    subroutine ctx_set_float(h, key, v, status)
        integer, intent(in) :: h
        character(len=*), intent(in) :: key
        real(WP), intent(in) :: v
        integer, intent(out) :: status
        integer :: k
        status = 0
        associate(c => ctxs(h))
        select case (lower(trim(key)))
        case ('h0')
            if (v > 0.0_wp) then
                if (v /= c%h0) c%cosmo_dirty = .true.
                c%h0 = v
            end if
        case ('om0')
            if (v >= 0.0_wp .and. v + c%ol0 > 0.0_wp) then
                if (v /= c%om0) c%cosmo_dirty = .true.
                c%om0 = v
            end if
        case ('ol0')
            if (c%om0 + v > 0.0_wp) then
                if (v /= c%ol0) c%cosmo_dirty = .true.
                c%ol0 = v
            end if
        case ('mdave')
            if (v /= c%mdave) call drop_ops(c)
            c%mdave = v; c%dev_imf = .false.
        case ('vdmc')
            if (v /= c%vdmc) call drop_ops(c)
            c%vdmc = v
        case default
            k = param_slot(key)
            if (k > 0) then
                if (any(IMF_NAMES == lower(trim(key))) .and. v /= c%pset(k)) call drop_ops(c)
                c%pset(k) = v
            else if (inert_slot(key) > 0) then
                c%inert(inert_slot(key)) = v
            else
                status = 102
            end if
        end select
        end associate
    end subroutine ctx_set_float

    !> String settings. kernel_precision: "double" (FP64 age sums and chain, the reference's default build) or
    !> "mixed" / "single" (split-BF16 age sums, FP32 chain). status 0, or 103 for an unknown key or value
    subroutine ctx_set_str(h, key, v, status)
        integer, intent(in) :: h
        character(len=*), intent(in) :: key, v
        integer, intent(out) :: status
        status = 0
        select case (lower(trim(key)))
        case ('kernel_precision')
            select case (lower(trim(v)))
            case ('double', 'fp64'); ctxs(h)%mixed = 0
            case ('mixed', 'single', 'fp32'); ctxs(h)%mixed = 1
            case default; status = 103
            end select
        case default
            status = 103
        end select
    end subroutine ctx_set_str

    !> Installs a resident SSP basis (the reference's set-up plus opensps_context_update_ssp_basis,
    !> src/core/opensps_context.f90:757): lam (nspec), log_age (nt), zlegend (nz, mass fractions), spec (nspec,
    !> nt, nz), mass (nt, nz), lbol (nt, nz). status 3 for bad sizes or non-increasing grids.
    subroutine ctx_set_basis(h, nspec, lam, nt, log_age, nz, zlegend, zsol, spec, mass, lbol, status)
        integer, intent(in) :: h, nspec, nt, nz
        real(WP), intent(in) :: lam(nspec), log_age(nt), zlegend(nz), zsol, spec(nspec, nt, nz), mass(nt, nz), &
                                lbol(nt, nz)
        integer, intent(out) :: status
        integer :: i
        status = 0
        if (nspec < 2 .or. nt < 2 .or. nz < 1 .or. zsol <= 0.0_wp) then
            status = 3; return
        end if
        do i = 2, nspec
            if (.not. (lam(i) > lam(i - 1))) then
                status = 3; return
            end if
        end do
        do i = 2, nt
            if (.not. (log_age(i) > log_age(i - 1))) then
                status = 3; return
            end if
        end do
        if (any(zlegend <= 0.0_wp)) then
            status = 3; return
        end if
        do i = 2, nz
            if (.not. (zlegend(i) > zlegend(i - 1))) then
                status = 3; return
            end if
        end do
        associate(c => ctxs(h))
            c%nspec = nspec; c%nt = nt; c%nz = nz; c%zsol = zsol
            c%lam = lam; c%log_age = log_age; c%zlegend = zlegend
            c%logz = log10(zlegend / zsol)
            c%spec = spec; c%mass = mass; c%lbol = lbol
            c%has_basis = .true.
            c%dev_grid = .false.
            ! an SSP operator belongs to the grid it was given for
            if (c%has_ops .and. c_associated(c%dctx)) i = opensps_hip_free_ssp_ops(c%dctx)
            c%has_ops = .false.
            ! a new grid invalidates the filters (they are built on the grid)
            c%nbands = 0
            if (allocated(c%bands)) deallocate(c%bands)
            if (allocated(c%magvega)) deallocate(c%magvega)
            if (allocated(c%magsun)) deallocate(c%magsun)
            c%dev_bands = .false.
            c%dev_vega = .false.
        end associate
    end subroutine ctx_set_basis

    !> The fixed attenuation curves A_lambda / A_V at wavelength lam (Angstrom): dust_type 2, Calzetti et al. (2000)
    !> eq. 4 extended beyond 0.12-2.2 um as the reference extends it (clipped at zero), and 6, Reddy et al. (2015)
    !> (compute_attenuation_curve_point, src/physics/opensps_dust.f90:567-594, 630-662)
    pure real(WP) function dust_curve_point(lam, dt) result(curve)
        real(WP), intent(in) :: lam
        integer, intent(in) :: dt
        real(WP) :: x, k
        x = 1.0e4_wp / lam
        curve = 0.0_wp
        if (dt == 2) then
            if (lam > 6300.0_wp) then
                k = 2.659_wp * (-1.857_wp + 1.040_wp * x) + 4.05_wp
            else
                k = 4.05_wp + 2.659_wp * (-2.156_wp + x * (1.509_wp + x * (-0.198_wp + x * 0.011_wp)))
            end if
            curve = max(k, 0.0_wp) / 4.05_wp
        else if (dt == 6) then
            k = 0.0_wp
            if (lam < 6000.0_wp) then
                x = min(x, 1.0e4_wp / 1500.0_wp)
                k = -5.726_wp + x * (4.004_wp + x * (-0.525_wp + x * 0.029_wp)) + 2.505_wp
            else if (lam < 28500.0_wp) then
                k = -2.672_wp + x * (-0.010_wp + x * (1.532_wp + x * (-0.412_wp))) + 2.505_wp - 0.036221981_wp
            end if
            curve = k / 2.505_wp
        end if
    end function dust_curve_point

    subroutine ctx_refresh_cosmology(h)
        integer, intent(in) :: h
        associate(c => ctxs(h))
            if (c%cosmo_dirty) then
                call cosmology_table(c%h0, c%om0, c%ol0, c%cz, c%cdl)
                c%cosmo_dirty = .false.
                c%dev_cosmo = .false.
            end if
        end associate
    end subroutine ctx_refresh_cosmology

    function hip_message(dctx) result(msg)
        type(c_ptr), intent(in) :: dctx
        character(len=512) :: msg
        character(kind=c_char) :: buf(512)
        integer(c_int) :: st
        integer :: i
        st = opensps_hip_last_error(dctx, buf, 512_c_int)
        msg = ''
        do i = 1, 512
            if (buf(i) == char(0)) exit
            msg(i:i) = buf(i)
        end do
    end function hip_message

    !> The device side of a context, created and brought up to date: the grid, the bands and their mask,
    !> the cosmology table and the options. status 0, 5 without a HIP device, 10 on a device error.
    integer function ctx_device(h) result(st)
        integer, intent(in) :: h
        integer(c_int) :: ndev, rc
        integer, allocatable, target :: mask(:)
        real(WP), allocatable, target :: curve(:)
        integer :: j
        st = 0
        associate(c => ctxs(h))
        if (.not. c_associated(c%dctx)) then
            rc = opensps_hip_device_count(ndev)
            if (ndev < 1) then
                call set_error(5, 'batch: no HIP device (this build runs the batch on the GPU only)')
                st = 5
                return
            end if
            rc = opensps_hip_create(0_c_int, c%dctx)
            if (rc /= 0) then
                call set_error(10, 'batch: ' // trim(hip_message(c_null_ptr)))
                st = 10
                return
            end if
            c%dev_grid = .false.; c%dev_bands = .false.; c%dev_cosmo = .false.; c%dev_mask = .false.
            c%dev_prof = .false.; c%dev_vega = .false.; c%dev_tab = .false.; c%dev_dust = .false.
            c%dev_imf = .false.; c%dev_emis = .false.; c%dev_neb = .false.
        end if
        if (.not. c%dev_grid) then
            rc = opensps_hip_set_grid(c%dctx, int(c%nspec, c_int), c%lam, int(c%nt, c_int), c%log_age, &
                                      int(c%nz, c_int), c%logz, c%spec, c%mass, c%lbol)
            if (rc == 0) rc = opensps_hip_set_options(c%dctx, int(c%add_igm_absorption, c_int), &
                                                      int(merge(1, 0, c%smooth_velocity == 1 .and. &
                                                                c%smoothspec_fast == 1), c_int))
            if (rc /= 0) goto 900
            c%dev_grid = .true.
            c%dev_bands = .false.
            c%dev_dust = .false.
            c%dev_emis = .false.
            c%dev_neb = .false.
        end if
        if (.not. c%dev_bands) then
            if (c%nbands > 0) then
                rc = opensps_hip_set_bands(c%dctx, int(c%nbands, c_int), c%bands)
            else
                rc = opensps_hip_set_bands(c%dctx, 0_c_int, [0.0_wp])
            end if
            if (rc /= 0) goto 900
            c%dev_bands = .true.
            c%dev_mask = .false.
        end if
        if (.not. c%dev_mask .and. c%nbands > 0) then
            allocate(mask(c%nbands))
            mask = 1
            if (allocated(c%mag_mask)) mask = c%mag_mask(1:c%nbands)
            rc = opensps_hip_set_mag_mask(c%dctx, int(c%nbands, c_int), mask)
            if (rc /= 0) goto 900
            c%dev_mask = .true.
            c%dev_vega = .false.
        end if
        if (.not. c%dev_dust) then
            if (c%dust_type == 2 .or. c%dust_type == 6) then
                allocate(curve(c%nspec))
                do j = 1, c%nspec
                    curve(j) = dust_curve_point(c%lam(j), c%dust_type)
                end do
                rc = opensps_hip_set_dust_curve(c%dctx, int(c%nspec, c_int), c_loc(curve))
            else
                rc = opensps_hip_set_dust_curve(c%dctx, 0_c_int, c_null_ptr)
            end if
            if (rc /= 0) goto 900
            c%dev_dust = .true.
        end if
        if (.not. c%dev_emis) then
            st = emission_upload(c)
            if (st /= 0) return
            c%dev_emis = .true.
        end if
        if (.not. c%dev_neb) then
            if (c%has_neb) then
                rc = opensps_hip_set_nebular(c%dctx, int(size(c%nb_logz), c_int), c%nb_logz, &
                                             int(size(c%nb_age), c_int), c%nb_age, int(size(c%nb_logu), c_int), &
                                             c%nb_logu, int(size(c%nb_lam), c_int), c%nb_lam, c%nb_cont, &
                                             int(c%neb_nline, c_int), c%nb_llam, c%nb_lines)
            else
                rc = opensps_hip_set_nebular(c%dctx, 0_c_int, [0.0_wp], 0_c_int, [0.0_wp], 0_c_int, [0.0_wp], &
                                             0_c_int, [0.0_wp], [0.0_wp], 0_c_int, [0.0_wp], [0.0_wp])
            end if
            if (rc == 0) rc = opensps_hip_set_neb_options(c%dctx, int(merge(1, 0, c%add_neb_emission == 1), c_int), &
                                                          int(c%add_neb_continuum, c_int), &
                                                          int(c%nebemlineinspec, c_int))
            if (rc /= 0) goto 900
            c%dev_neb = .true.
        end if
        if (.not. c%dev_tab) then
            if (c%ntab > 0) then
                rc = opensps_hip_set_sfh_tab(c%dctx, int(c%ntab, c_int), c%tab_age, c%tab_sfr)
            else
                rc = opensps_hip_set_sfh_tab(c%dctx, 0_c_int, [0.0_wp], [0.0_wp])
            end if
            if (rc /= 0) goto 900
            c%dev_tab = .true.
        end if
        if (.not. c%dev_vega) then
            ! the Vega system (magnitudes_from_fluxes, src/spectra/opensps_batch_photometry.f90:185-232)
            if (c%compute_vega_mags == 1 .and. c%nbands > 0 .and. allocated(c%magvega)) then
                rc = opensps_hip_set_vega(c%dctx, int(c%nbands, c_int), c%magvega, int(c%vega_ref, c_int))
            else
                rc = opensps_hip_set_vega(c%dctx, 0_c_int, [0.0_wp], 0_c_int)
            end if
            if (rc /= 0) goto 900
            c%dev_vega = .true.
        end if
        if (.not. c%dev_imf .and. c%has_ops) then
            ! the IMF of the SSP operator: the context's type, remnants and Dave's mdave; Salpeter's slope 2.35 and the
            ! remnant progenitor limits 40 and 8.5 M_sun are the context state's defaults
            ! (src/core/opensps_context_types.f90:120-125)
            rc = opensps_hip_set_imf(c%dctx, int(c%imf_type, c_int), int(c%add_stellar_remnants, c_int), 2.35_wp, &
                                     c%mdave, 40.0_wp, 8.5_wp)
            if (rc /= 0) goto 900
            c%dev_imf = .true.
        end if
        if (.not. c%dev_prof) then
            rc = opensps_hip_set_profiling(c%dctx, int(c%profile, c_int))
            if (rc /= 0) goto 900
            c%dev_prof = .true.
        end if
        call ctx_refresh_cosmology(h)
        if (.not. c%dev_cosmo) then
            rc = opensps_hip_set_cosmology(c%dctx, 500_c_int, c%cz, c%cdl)
            if (rc /= 0) goto 900
            c%dev_cosmo = .true.
        end if
        return
900     continue
        st = merge(10, int(rc), rc == 10)
        call set_error(st, 'device: ' // trim(hip_message(c%dctx)))
        end associate
    end function ctx_device

    ! ---------------------------------------------------------------------------------------------
    ! Dust emission and AGN torus templates
    ! ---------------------------------------------------------------------------------------------

    !> The reference's find_interval (src/math/opensps_interpolation.f90:149-198) on an ascending table: i with
    !> a(i) <= v < a(i+1), 0 below, n above, 1 at a(1), n-1 at a(n)
    pure integer function find_interval_asc(a, v) result(lower)
        real(WP), intent(in) :: a(:), v
        integer :: upper, mid, n
        n = size(a)
        lower = 0
        upper = n + 1
        do while (upper - lower > 1)
            mid = (upper + lower) / 2
            if (v >= a(mid)) then
                lower = mid
            else
                upper = mid
            end if
        end do
        if (v == a(1)) lower = 1
        if (v == a(n)) lower = n - 1
    end function find_interval_asc

    !> interpolate_linear_scalar (src/math/opensps_interpolation.f90:45-72): the bracketing interval clamped to the
    !> table, so that the ends extrapolate linearly
    pure real(WP) function interp_lin(x, y, xo) result(v)
        real(WP), intent(in) :: x(:), y(:), xo
        integer :: i
        i = max(1, min(find_interval_asc(x, xo), size(x) - 1))
        v = y(i) + (y(i + 1) - y(i)) / (x(i + 1) - x(i)) * (xo - x(i))
    end function interp_lin

    logical function ascending(a)
        real(WP), intent(in) :: a(:)
        integer :: i
        ascending = size(a) >= 2
        do i = 2, size(a)
            if (.not. (a(i) > a(i - 1))) ascending = .false.
        end do
    end function ascending

    !> The Draine & Li (or any two-component) dust emission library as given: q_PAH (nqpah), U_min (numin) and
    !> wavelengths (nlam, A) ascending; spectra (nlam, 2, numin, nqpah): per (q_PAH, U_min) the spectrum of U_max =
    !> U_min, then of the power-law U distribution (the data library's layout, load_dust_emission_table,
    !> src/io/opensps_io.f90:959-1069). Status 3 for bad sizes or tables not ascending.
    subroutine ctx_set_dust_emission(h, nqpah, qpah, numin, umin, nlam, lam, spectra, status)
        integer, intent(in) :: h, nqpah, numin, nlam
        real(WP), intent(in) :: qpah(nqpah), umin(numin), lam(nlam), spectra(nlam, 2 * numin, nqpah)
        integer, intent(out) :: status
        status = 0
        if (nqpah < 2 .or. numin < 2 .or. nlam < 2) then
            status = 3
            call set_error(3, 'set_dust_emission_templates: at least 2 q_PAH, 2 U_min and 2 wavelengths are required')
            return
        end if
        if (.not. (ascending(qpah) .and. ascending(umin) .and. ascending(lam))) then
            status = 3
            call set_error(3, 'set_dust_emission_templates: q_PAH, U_min and wavelengths must ascend')
            return
        end if
        associate(c => ctxs(h))
            c%dl_qpah = qpah
            c%dl_umin = umin
            c%dl_lam = lam
            c%dl_spec = spectra
            c%has_dustem = .true.
            c%dev_emis = .false.
        end associate
    end subroutine ctx_set_dust_emission

    !> The AGN torus library as given: optical depths tau (ntau) and wavelengths (nlam, A) ascending, spectra
    !> (nlam, ntau) (load_agn_dust_models, src/io/opensps_io.f90:1101-1155). Status 3 for bad sizes or order.
    subroutine ctx_set_agn(h, ntau, tau, nlam, lam, spectra, status)
        integer, intent(in) :: h, ntau, nlam
        real(WP), intent(in) :: tau(ntau), lam(nlam), spectra(nlam, ntau)
        integer, intent(out) :: status
        status = 0
        if (ntau < 2 .or. nlam < 2) then
            status = 3
            call set_error(3, 'set_agn_templates: at least 2 optical depths and 2 wavelengths are required')
            return
        end if
        if (.not. (ascending(tau) .and. ascending(lam))) then
            status = 3
            call set_error(3, 'set_agn_templates: optical depths and wavelengths must ascend')
            return
        end if
        associate(c => ctxs(h))
            c%agn_tau = tau
            c%agn_lam = lam
            c%agn_spec = spectra
            c%has_agn = .true.
            c%dev_emis = .false.
        end associate
    end subroutine ctx_set_agn

    !> The CLOUDY nebular library as read_nebular_standard reads it (src/io/opensps_io.f90:914-937): axes logz (nnz),
    !> age (nna, yr), logu (nnu) and the continuum's wavelengths lam (nlam, A), all ascending; cont (nnz, nna, nnu,
    !> nlam) in L_sun/Hz and lines (nnz, nna, nnu, nline) in L_sun, per unit ionizing photon rate, flat in C order;
    !> the lines' wavelengths line_lam (nline, A). Status 3 for bad sizes or axes not ascending.
    subroutine ctx_set_nebular(h, nnz, logz, nna, age, nnu, logu, nlam, lam, cont, nline, line_lam, lines, status)
        integer, intent(in) :: h, nnz, nna, nnu, nlam, nline
        real(WP), intent(in) :: logz(nnz), age(nna), logu(nnu), lam(nlam), cont(nnz * nna * nnu * nlam)
        real(WP), intent(in) :: line_lam(nline), lines(nnz * nna * nnu * nline)
        integer, intent(out) :: status
        status = 0
        if (nnz < 2 .or. nna < 2 .or. nnu < 2 .or. nlam < 2 .or. nline < 1) then
            status = 3
            call set_error(3, 'set_nebular_tables: at least 2 gas metallicities, ages, ionization parameters and ' // &
                           'wavelengths and 1 line are required')
            return
        end if
        if (.not. (ascending(logz) .and. ascending(age) .and. ascending(logu) .and. ascending(lam))) then
            status = 3
            call set_error(3, 'set_nebular_tables: the axes and wavelengths must ascend')
            return
        end if
        if (any(age <= 0.0_wp)) then
            status = 3
            call set_error(3, 'set_nebular_tables: ages must be positive (yr)')
            return
        end if
        associate(c => ctxs(h))
            c%nb_logz = logz
            c%nb_age = age
            c%nb_logu = logu
            c%nb_lam = lam
            c%nb_cont = cont
            c%nb_llam = line_lam
            c%nb_lines = lines
            c%neb_nline = nline
            c%has_neb = .true.
            c%dev_neb = .false.
        end associate
    end subroutine ctx_set_nebular

    !> The rows of emlines: the nebular library's lines once it is set, else the reference's NEMLINE (166,
    !> src/core/opensps_constants.f90:173)
    integer function nemline(c)
        type(host_ctx_t), intent(in) :: c
        nemline = merge(c%neb_nline, 166, c%has_neb)
    end function nemline

    !> The templates on the context's grid and to the device, with the two settings. Dust emission as
    !> load_dust_emission_table places it: every column linear in wavelength from the pixel of 1 micron
    !> (find_interval) on, extrapolating linearly past the library's ends, zero below; the AGN as load_agn_dust_models:
    !> log-linear (log10 of flux + tiny) over the pixels from the one of the library's first wavelength to the one of
    !> its last, zero elsewhere.
    integer function emission_upload(c) result(st)
        type(host_ctx_t), intent(inout), target :: c
        real(WP), allocatable, target :: tb(:, :, :), ta(:, :)
        real(WP), allocatable :: llam(:), lflux(:)
        integer :: i, k, j, i0, i1, i2, nq, nu, nta
        integer(c_int) :: rc
        st = 0
        rc = opensps_hip_set_emission(c%dctx, int(c%add_dust_emission, c_int), int(c%add_agn_dust, c_int))
        if (rc /= 0) goto 900
        if (c%has_dustem) then
            nq = size(c%dl_qpah)
            nu = size(c%dl_umin)
            allocate(tb(c%nspec, nq, 2 * nu))
            tb = 0.0_wp
            i0 = max(find_interval_asc(c%lam, 1.0e4_wp), 1)
            do k = 1, nq
                do j = 1, 2 * nu
                    do i = i0, c%nspec
                        tb(i, k, j) = interp_lin(c%dl_lam, c%dl_spec(:, j, k), c%lam(i))
                    end do
                end do
            end do
            rc = opensps_hip_set_dust_emission(c%dctx, int(nq, c_int), c_loc(c%dl_qpah), int(nu, c_int), &
                                               c_loc(c%dl_umin), c_loc(tb))
        else
            rc = opensps_hip_set_dust_emission(c%dctx, 0_c_int, c_null_ptr, 0_c_int, c_null_ptr, c_null_ptr)
        end if
        if (rc /= 0) goto 900
        if (c%has_agn) then
            nta = size(c%agn_tau)
            allocate(ta(c%nspec, nta), llam(size(c%agn_lam)), lflux(size(c%agn_lam)))
            ta = 0.0_wp
            i1 = max(find_interval_asc(c%lam, c%agn_lam(1)), 1)
            i2 = max(find_interval_asc(c%lam, c%agn_lam(size(c%agn_lam))), 1)
            llam = log10(c%agn_lam)
            do k = 1, nta
                lflux = log10(c%agn_spec(:, k) + SAFE_FLOOR)
                do i = i1, i2
                    ta(i, k) = 10.0_wp**interp_lin(llam, lflux, log10(c%lam(i))) - SAFE_FLOOR
                end do
            end do
            rc = opensps_hip_set_agn(c%dctx, int(nta, c_int), c_loc(c%agn_tau), c_loc(ta))
        else
            rc = opensps_hip_set_agn(c%dctx, 0_c_int, c_null_ptr, c_null_ptr)
        end if
        if (rc /= 0) goto 900
        return
900     st = 10
        call set_error(10, 'batch: ' // trim(hip_message(c%dctx)))
    end function emission_upload

end module opensps_host_context
