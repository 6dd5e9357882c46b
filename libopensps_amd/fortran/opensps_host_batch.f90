!> The batch driver of the Fortran host: the C API's names / values become one FP64 array of per-model
!> parameters (a slot per parameter, as the reference's batch_handle resolves each name to one word of
!> PARAMS_CORE, src/bindings/c/opensps_c_api.f90:663-961), the checks of compute_csp_batch's stage 0 run on the
!> host (src/physics/opensps_csp_batch.f90:608-787), and stages 1-8 run on the GPU through libopensps_hip.
module opensps_host_batch
    use, intrinsic :: iso_c_binding
    use opensps_host_types
    use opensps_host_context
    use opensps_hip_iface
    implicit none
    private
    public :: batch_run, batch_run_device, cstring, resolve_params, model_checks, ssp_op_evaluate

contains

    !> A NUL-terminated C string as a Fortran string
    function cstring(p) result(s)
        type(c_ptr), intent(in) :: p
        character(len=:), allocatable :: s
        character(kind=c_char), pointer :: a(:)
        integer :: n
        s = ''
        if (.not. c_associated(p)) return
        call c_f_pointer(p, a, [4096])
        n = 0
        do while (n < 4096)
            if (a(n + 1) == c_null_char) exit
            n = n + 1
        end do
        allocate(character(len=n) :: s)
        s = transfer(a(1:n), s)
    end function cstring

    pure real(WP) function expm1_ref(x) result(v)   ! opensps_expm1 (src/math/opensps_special_functions.f90:219-233)
        real(WP), intent(in) :: x
        if (abs(x) < 1.0e-5_wp) then
            v = x * (1.0_wp + x * (0.5_wp + x * (1.0_wp / 6.0_wp)))
        else
            v = exp(x) - 1.0_wp
        end if
    end function expm1_ref

    pure real(WP) function gammainc12(p, x) result(r)   ! gammainc for p = 1, 2 (opensps_special_functions.f90:142-211)
        integer, intent(in) :: p
        real(WP), intent(in) :: x
        if (x < 0.0_wp) then
            r = 0.0_wp
        else if (p == 1) then
            r = -expm1_ref(-x)
        else
            r = -expm1_ref(-x) - x * exp(-x)
        end if
    end function gammainc12

    !> The SFR of a tabular model (sfh 2, 3) at its age (Gyr): the table interpolated linearly in time, at least 0, in
    !> M_sun / yr (get_sfh_properties_at_age_pc, src/physics/opensps_sfh.f90:460-481; its mass fraction stays 1)
    real(WP) function sfr_tab(c, age_gyr) result(sfr)
        type(host_ctx_t), intent(in) :: c
        real(WP), intent(in) :: age_gyr
        real(WP) :: t, slope, a1, a2, s1, s2
        integer :: lo, hi, mid, i, n
        n = c%ntab
        sfr = 0.0_wp
        if (n < 1) return
        t = age_gyr * 1.0e9_wp
        ! find_interval_tab_time (:1041-1073) on the stored table, clamped to a bin; rows beyond the table read 0
        i = 1
        if (n >= 2) then
            lo = 0
            hi = n + 1
            do while (hi - lo > 1)
                mid = (hi + lo) / 2
                if ((c%tab_age(n) >= c%tab_age(1)) .eqv. (t >= c%tab_age(mid))) then
                    lo = mid
                else
                    hi = mid
                end if
            end do
            i = lo
            if (t == c%tab_age(1)) i = 1
            if (t == c%tab_age(n)) i = n - 1
        end if
        i = max(min(i, n - 1), 1)
        a1 = c%tab_age(i); s1 = c%tab_sfr(i)
        a2 = 0.0_wp; s2 = 0.0_wp
        if (i + 1 <= n) then
            a2 = c%tab_age(i + 1); s2 = c%tab_sfr(i + 1)
        end if
        slope = (s2 - s1) / (a2 - a1)
        sfr = max(s1 + slope * (t - a1), 0.0_wp)
    end function sfr_tab

    !> The SFR of a model at its age per unit mass formed: get_sfh_properties_at_age (src/physics/opensps_sfh.f90:282-486)
    !> for sfh 0, 1, 4, 5, divided by the mass fraction formed as batch_model_indices does (opensps_csp_batch.f90:1556-1560)
    real(WP) function sfr_at_age(p, tmax_log) result(sfr)
        real(WP), intent(in) :: p(NPAR), tmax_log
        real(WP) :: t_max, t_prime, t_trunc, tau_g, tr, total_tau, mass_tau, sfr_tau, mfrac_b, sfr_c, mass_c, norm, mfrac
        real(WP) :: slope_m, t_zero, t_hi, sfr_trunc, total_lin, mass_lin
        integer :: pw, sfh
        sfh = nint(p(P_SFH))
        sfr = 0.0_wp
        if (sfh == 5) then
            ! Simha: delayed tau up to the truncation, then linear (opensps_sfh.f90:401-455)
            t_max = 10.0_wp**(tmax_log - 9.0_wp) - p(P_SF_START)
            t_prime = p(P_TAGE) - p(P_SF_START)
            if (p(P_SF_TRUNC) < SAFE_FLOOR .or. p(P_SF_TRUNC) < p(P_SF_START)) then
                t_trunc = t_max
            else
                t_trunc = p(P_SF_TRUNC) - p(P_SF_START)
            end if
            tau_g = max(p(P_TAU), SAFE_FLOOR)
            total_tau = tau_g * gammainc12(2, min(t_max, t_trunc) / tau_g)
            if (t_prime < 0.0_wp) then
                mass_tau = 0.0_wp
                sfr_tau = 0.0_wp
            else
                tr = min(t_prime, t_trunc) / tau_g
                mass_tau = tau_g * gammainc12(2, tr)
                sfr_tau = tr * exp(-tr)
            end if
            slope_m = -p(P_SF_SLOPE)
            if (slope_m > 0.0_wp) then
                t_zero = t_trunc + 1.0_wp / slope_m
            else
                t_zero = t_max
            end if
            mass_lin = 0.0_wp
            total_lin = 0.0_wp
            if (t_trunc <= 0.0_wp .or. t_trunc > t_max) then
                sfr = sfr_tau / total_tau
            else
                t_hi = min(t_zero, t_max)
                sfr_trunc = (t_trunc / p(P_TAU)) * exp(-t_trunc / p(P_TAU))
                total_lin = sfr_trunc * ((t_hi - t_trunc) - (slope_m / 2.0_wp) * (t_hi**2 + t_trunc**2) + &
                                         (slope_m * t_hi * t_trunc))
                if (t_prime <= t_trunc) then
                    sfr = sfr_tau / (total_tau + total_lin)
                else
                    t_hi = min(t_zero, t_prime)
                    mass_lin = sfr_trunc * ((t_hi - t_trunc) - (slope_m / 2.0_wp) * (t_hi**2 + t_trunc**2) + &
                                            (slope_m * t_hi * t_trunc))
                    sfr = max(sfr_trunc * (1.0_wp - slope_m * (t_prime - t_trunc)), 0.0_wp) / (total_tau + total_lin)
                end if
            end if
            if (t_prime >= 0.0_wp) then
                mfrac = (mass_tau + mass_lin) / (total_tau + total_lin)
            else
                mfrac = 0.0_wp
                sfr = 0.0_wp
            end if
            sfr = sfr / 1.0e9_wp
            if (mfrac > SAFE_FLOOR) sfr = sfr / mfrac
            return
        end if
        if (sfh /= 1 .and. sfh /= 4) return
        t_max = 10.0_wp**(tmax_log - 9.0_wp) - p(P_SF_START)
        t_prime = p(P_TAGE) - p(P_SF_START)
        if (p(P_SF_TRUNC) < SAFE_FLOOR .or. p(P_SF_TRUNC) < p(P_SF_START)) then
            t_trunc = t_max
        else
            t_trunc = p(P_SF_TRUNC) - p(P_SF_START)
        end if
        tau_g = max(p(P_TAU), SAFE_FLOOR)
        pw = merge(1, 2, sfh == 1)
        total_tau = tau_g * gammainc12(pw, min(t_max, t_trunc) / tau_g)
        if (t_prime < 0.0_wp) then
            mass_tau = 0.0_wp
            sfr_tau = 0.0_wp
        else
            tr = min(t_prime, t_trunc) / tau_g
            mass_tau = tau_g * gammainc12(pw, tr)
            sfr_tau = (tr**(pw - 1)) * exp(-tr)
        end if
        mfrac_b = merge(1.0_wp, 0.0_wp, t_prime > (p(P_TBURST) - p(P_SF_START)))
        if (t_prime > 0.0_wp) then
            sfr_c = 1.0_wp / min(t_max, t_trunc)
            mass_c = max(min(t_prime, t_trunc), 0.0_wp) * sfr_c
        else
            sfr_c = 0.0_wp
            mass_c = 0.0_wp
        end if
        norm = 1.0_wp - p(P_CONST) - p(P_FBURST)
        mfrac = (norm * mass_tau / total_tau) + (p(P_CONST) * mass_c) + (p(P_FBURST) * mfrac_b)
        if (t_prime > t_trunc) then
            sfr = 0.0_wp
        else
            sfr = (norm * sfr_tau / total_tau) + (p(P_CONST) * sfr_c)
        end if
        sfr = sfr / 1.0e9_wp
        if (mfrac > SAFE_FLOOR) sfr = sfr / mfrac
    end function sfr_at_age

    !> Per-model checks of stage 0 that this build adds to the reference's: options whose stages it does
    !> not compute. status 2 with a message, else 0.
    !> have_zm: the call gives each model's magnitude redshift (zred_mags), which redshifted colours need; grad: the
    !> batch with gradients (whose reverse pass of the dust emission and the AGN torus is not part of this build)
    subroutine model_checks(c, prm, n, st, have_zm, grad)
        type(host_ctx_t), intent(in) :: c
        real(WP), intent(in) :: prm(n, NPAR)
        integer, intent(in) :: n
        integer, intent(out) :: st
        logical, intent(in), optional :: have_zm, grad
        logical :: dusty(n), agn(n), gr
        logical :: zm
        zm = .false.
        if (present(have_zm)) zm = have_zm
        gr = .false.
        if (present(grad)) gr = grad
        st = 0
        ! who emits dust (batch_model_indices, opensps_csp_batch.f90:1569-1570) and who has an AGN (agn_indices :1591-1606)
        dusty = c%add_dust_emission == 1 .and. (prm(:, P_DUST1) > SAFE_FLOOR .or. prm(:, P_DUST2) > SAFE_FLOOR)
        agn = c%add_agn_dust == 1 .and. prm(:, P_FAGN) > SAFE_FLOOR
        if (any(agn) .and. .not. c%has_agn) then
            st = 2
            call set_error(2, 'batch: the AGN torus (fagn > 0 with add_agn_dust = 1) needs its templates ' // &
                           '(opensps_context_set_agn_templates)')
        else if (any(dusty) .and. .not. c%has_dustem) then
            st = 2
            call set_error(2, 'batch: dust emission (add_dust_emission = 1 with dust) needs its templates ' // &
                           '(opensps_context_set_dust_emission_templates)')
        else if (gr .and. (any(dusty) .or. any(agn))) then
            st = 2
            call set_error(2, 'batch_grad: the reverse pass of the dust emission and the AGN torus is not part of ' // &
                           'this build')
        else if (c%has_ops .and. c%imf_type /= 0 .and. c%imf_type /= 2 .and. c%imf_type /= 4) then
            st = 2
            call set_error(2, 'batch: the SSP operator integrates imf_type 0 (Salpeter), 2 (Kroupa) and 4 (Dave)')
        else if (c%add_neb_emission == 2) then
            st = 2
            call set_error(2, 'batch: nebular emission in the SSPs (add_neb_emission = 2) is not part of this build')
        else if (c%add_neb_emission == 1 .and. .not. c%has_neb) then
            st = 2
            call set_error(2, 'batch: nebular emission (add_neb_emission = 1) needs its tables ' // &
                           '(opensps_context_set_nebular_tables)')
        else if (gr .and. c%add_neb_emission == 1) then
            st = 2
            call set_error(2, 'batch_grad: the reverse pass of the nebular emission is not part of this build')
        else if (c%compute_vega_mags == 1 .and. .not. allocated(c%magvega)) then
            st = 2
            call set_error(2, 'batch: Vega magnitudes need the Vega spectrum (opensps_context_set_reference_spectra)')
        else if (c%redshift_colors /= 0 .and. .not. zm) then
            st = 2
            call set_error(2, 'batch: redshifted colours need zred_mags')
        else if (c%dust_type /= 0 .and. c%dust_type /= 2 .and. c%dust_type /= 6) then
            st = 2
            call set_error(2, 'batch: dust_type 0 (power law), 2 (Calzetti) and 6 (Reddy) are part of this build')
        end if
    end subroutine model_checks

    !> names[] / values[] -> prm(n, NPAR): the context's values, overwritten per model by the named parameters
    !> (batch_handle's name resolution, src/bindings/c/opensps_c_api.f90:797-853). Status 2 for a parameter that
    !> changes the SSPs, and for the IMF's without an SSP operator on the context (:710-714, 770-780); 3 for an
    !> unknown name or NULL values. Members no stage here reads are accepted.
    subroutine resolve_params(c, n, n_params, c_names, c_values, prm, status)
        type(host_ctx_t), intent(in) :: c
        integer, intent(in) :: n, n_params
        type(c_ptr), intent(in) :: c_names, c_values
        real(WP), allocatable, intent(out) :: prm(:, :)
        integer, intent(out) :: status
        type(c_ptr), pointer :: np(:), vp(:)
        real(c_double), pointer :: v(:)
        character(len=:), allocatable :: nm
        integer :: i, k, ki
        status = 0
        allocate(prm(n, NPAR))
        do k = 1, NPAR
            prm(:, k) = c%pset(k)
        end do
        if (n_params <= 0) return
        if (.not. c_associated(c_names) .or. .not. c_associated(c_values)) then
            status = 3
            call set_error(3, 'batch: names / values are NULL')
            return
        end if
        call c_f_pointer(c_names, np, [n_params])
        call c_f_pointer(c_values, vp, [n_params])
        do i = 1, n_params
            nm = lower(cstring(np(i)))
            if (any(SSP_NAMES == nm)) then
                status = 2
                call set_error(2, 'batch: ' // nm // ' changes the SSPs and cannot vary in a batch')
                return
            end if
            if (any(IMF_NAMES == nm)) then
                if (.not. c%has_ops) then
                    status = 2
                    call set_error(2, 'batch: ' // nm // ' changes the SSPs; the IMF varies per model only with an ' // &
                                   'SSP operator on the context (opensps_context_set_ssp_ops)')
                    return
                end if
                if (nm == 'vdmc' .or. nm == 'mdave') then
                    status = 2
                    call set_error(2, 'batch: ' // nm // ' per model is not part of this build (imf1, imf2, imf3 ' // &
                                   'and the IMF limits are)')
                    return
                end if
            end if
            if (nm == 'afe') then
                status = 2
                call set_error(2, 'batch: afe needs the [alpha/Fe] bases of opensps_context_build_afe_basis')
                return
            end if
            k = param_slot(nm)
            ki = inert_slot(nm)
            if (k == 0 .and. ki == 0) then
                status = 3
                call set_error(3, 'batch: unknown parameter ' // nm)
                return
            end if
            if (.not. c_associated(vp(i))) then
                status = 3
                call set_error(3, 'batch: NULL values for ' // nm)
                return
            end if
            call c_f_pointer(vp(i), v, [n])
            if (k > 0) prm(:, k) = v(1:n)
        end do
    end subroutine resolve_params

    !> compute_csp_batch through the C API (batch_handle without the gradients)
    subroutine batch_run(h, n, n_params, c_names, c_values, c_weights, c_mags, c_spec, c_mass, c_lbol, c_mdust, c_sfr, &
                         c_emlines, c_indx, c_young, c_old, c_mass_young, c_mass_old, c_lbol_young, c_lbol_old, status)
        integer, intent(in) :: h, n, n_params
        type(c_ptr), intent(in) :: c_names, c_values, c_weights, c_mags, c_spec, c_mass, c_lbol, c_mdust, c_sfr, &
                                   c_emlines, c_indx, c_young, c_old, c_mass_young, c_mass_old, c_lbol_young, c_lbol_old
        integer, intent(out) :: status
        type(c_ptr), pointer :: np(:), vp(:)
        real(c_double), pointer :: v(:), outv(:), outm(:, :), w_in(:, :)
        real(WP), allocatable, target :: prm(:, :), wts(:, :), hbuf(:, :), mblend(:)
        real(c_float), allocatable, target :: fbuf(:, :)
        integer(c_int), allocatable, target :: idx(:, :)
        integer :: i, k, m, st, nb, ns, nt, bytes, it
        logical :: parts, split, want_w, neb
        character(len=:), allocatable :: nm
        type(hip_batch_t) :: a
        integer(c_int) :: rc
        real(WP) :: mk, lk, dz, wk
        real(WP), allocatable :: msy(:), mso(:), lsy(:), lso(:)

        status = 0
        if (.not. ctx_valid(h)) then
            status = 1
            call set_error(1, 'batch: bad handle')
            return
        end if
        associate(c => ctxs(h))
        if (.not. c%has_basis) then
            status = 1
            call set_error(1, 'batch: no SSP basis installed')
            return
        end if
        if (n < 0 .or. n_params < 0) then
            status = 3
            call set_error(3, 'batch: n or n_params < 0')
            return
        end if
        if (n == 0) return
        if (.not. c_associated(c_mags)) then
            status = 3
            call set_error(3, 'batch: mags is required')
            return
        end if
        if (c_associated(c_indx)) then
            status = 2
            call set_error(2, 'batch: spectral indices are not part of this build')
            return
        end if
        nb = c%nbands
        ns = c%nspec
        nt = c%nt
        ! names -> slots; the values of the named parameters per model over the context's
        call resolve_params(c, n, n_params, c_names, c_values, prm, status)
        if (status /= 0) return
        call model_checks(c, prm, n, st)
        if (st /= 0) then
            status = st
            return
        end if
        parts = c_associated(c_young) .and. c_associated(c_old)
        split = parts .or. any((prm(:, P_DUST1) > SAFE_FLOOR .and. 1.0_wp - prm(:, P_FRAC_OBRUN) > SAFE_FLOOR) .or. &
                               abs(prm(:, P_DUST3)) > SAFE_FLOOR)
        want_w = c_associated(c_mass_young) .or. c_associated(c_mass_old) .or. c_associated(c_lbol_young) .or. &
                 c_associated(c_lbol_old)
        if (want_w .and. c%has_ops) then
            status = 2
            call set_error(2, 'batch: the young and old masses and L_bol with the SSP operator are not part of this build')
            return
        end if
        st = ctx_device(h)
        if (st /= 0) then
            status = st
            return
        end if
        bytes = merge(4, 8, c%mixed == 1)
        st = 0
        st = st + dbuf_reserve(c%d_prm, int(8, c_size_t) * NPAR * n)
        st = st + dbuf_reserve(c%d_mags, int(8, c_size_t) * max(nb, 1) * n)
        st = st + dbuf_reserve(c%d_spec, int(bytes, c_size_t) * ns * n)
        st = st + dbuf_reserve(c%d_mass, int(8, c_size_t) * n)
        st = st + dbuf_reserve(c%d_lbol, int(8, c_size_t) * n)
        st = st + dbuf_reserve(c%d_idx, int(4, c_size_t) * 3 * n)
        if (c%keep_maggies == 1) st = st + dbuf_reserve(c%d_fband, int(8, c_size_t) * max(nb, 1) * n)
        if (c_associated(c_weights) .or. want_w) st = st + dbuf_reserve(c%d_w, int(8, c_size_t) * nt * n)
        if (parts) then
            st = st + dbuf_reserve(c%d_young, int(bytes, c_size_t) * ns * n)
            st = st + dbuf_reserve(c%d_old, int(bytes, c_size_t) * ns * n)
        end if
        if (c_associated(c_mdust)) st = st + dbuf_reserve(c%d_mdust, int(8, c_size_t) * n)
        neb = c%add_neb_emission == 1 .and. c_associated(c_emlines)
        if (neb) st = st + dbuf_reserve(c%d_eml, int(8, c_size_t) * c%neb_nline * n)
        if (st /= 0) then
            status = 10
            call set_error(10, 'batch: device allocation failed')
            return
        end if
        rc = opensps_hip_memcpy_h2d(c%d_prm%p, c_loc(prm), int(8, c_size_t) * NPAR * n)
        if (c_associated(c_weights)) then
            call c_f_pointer(c_weights, w_in, [nt, n])
            allocate(wts(nt, n))
            wts = w_in
            rc = opensps_hip_memcpy_h2d(c%d_w%p, c_loc(wts), int(8, c_size_t) * nt * n)
        end if
        a%n = n
        a%precision = merge(HIP_MIXED, HIP_FP64, c%mixed == 1)
        a%split_parts = merge(1, 0, split)
        a%use_ops = merge(1, 0, c%has_ops)
        a%prm = c%d_prm%p
        a%ssp_weights = c_null_ptr
        if (c_associated(c_weights)) a%ssp_weights = c%d_w%p
        a%mags = c%d_mags%p
        a%spec = c%d_spec%p
        a%mass = c%d_mass%p
        a%lbol = c%d_lbol%p
        a%fband = c_null_ptr
        if (c%keep_maggies == 1) a%fband = c%d_fband%p
        a%young = c_null_ptr
        a%old = c_null_ptr
        if (parts) then
            a%young = c%d_young%p
            a%old = c%d_old%p
        end if
        a%idx = c%d_idx%p
        a%weights = c_null_ptr
        if (want_w .and. .not. c_associated(c_weights)) a%weights = c%d_w%p
        a%mdust = c_null_ptr
        if (c_associated(c_mdust)) a%mdust = c%d_mdust%p
        a%emlines = c_null_ptr
        if (neb) a%emlines = c%d_eml%p
        rc = opensps_hip_csp_batch(c%dctx, a)
        if (rc == 0) rc = opensps_hip_sync(c%dctx)
        if (rc /= 0) then
            status = merge(int(rc), 10, rc == 2 .or. rc == 3)
            call set_error(status, 'batch: ' // trim(hip_message(c%dctx)))
            return
        end if
        c%last_n = n
        c%last_bytes = bytes
        ! outputs to the caller
        if (nb > 0) rc = opensps_hip_memcpy_d2h(c_mags, c%d_mags%p, int(8, c_size_t) * nb * n)
        if (c_associated(c_mass)) rc = opensps_hip_memcpy_d2h(c_mass, c%d_mass%p, int(8, c_size_t) * n)
        if (c_associated(c_lbol)) rc = opensps_hip_memcpy_d2h(c_lbol, c%d_lbol%p, int(8, c_size_t) * n)
        call copy_spectra(c_spec, c%d_spec%p)
        if (parts) then
            call copy_spectra(c_young, c%d_young%p)
            call copy_spectra(c_old, c%d_old%p)
        end if
        if (c_associated(c_mdust)) rc = opensps_hip_memcpy_d2h(c_mdust, c%d_mdust%p, int(8, c_size_t) * n)
        if (neb) then
            rc = opensps_hip_memcpy_d2h(c_emlines, c%d_eml%p, int(8, c_size_t) * c%neb_nline * n)
        else if (c_associated(c_emlines)) then
            call c_f_pointer(c_emlines, outm, [nemline(c), n])
            outm = 0.0_wp
        end if
        if (c_associated(c_sfr)) then
            call c_f_pointer(c_sfr, outv, [n])
            do m = 1, n
                if (c_associated(c_weights)) then
                    outv(m) = ieee_nan()
                else
                    if (nint(prm(m, P_SFH)) == 2 .or. nint(prm(m, P_SFH)) == 3) then
                        outv(m) = sfr_tab(c, prm(m, P_TAGE))
                    else
                        outv(m) = sfr_at_age(prm(m, :), c%log_age(nt))
                    end if
                end if
            end do
        end if
        if (c%keep_maggies == 1 .and. nb > 0) then
            if (allocated(c%last_maggies)) deallocate(c%last_maggies)
            allocate(c%last_maggies(nb, n))
            call maggies_from_device(c, n, prm)
        end if
        if (want_w) then
            ! the masses and L_bol of the young and old parts from the weights (model_sums, :2129-2176)
            allocate(hbuf(nt, n), idx(3, n), msy(n), mso(n), lsy(n), lso(n))
            rc = opensps_hip_memcpy_d2h(c_loc(hbuf), c%d_w%p, int(8, c_size_t) * nt * n)
            rc = opensps_hip_memcpy_d2h(c_loc(idx), c%d_idx%p, int(4, c_size_t) * 3 * n)
            msy = 0.0_wp; mso = 0.0_wp; lsy = 0.0_wp; lso = 0.0_wp
            do m = 1, n
                it = idx(1, m)
                dz = 0.0_wp
                if (c%nz > 1) dz = (prm(m, P_LOGZSOL) - c%logz(idx(2, m))) / (c%logz(idx(2, m) + 1) - c%logz(idx(2, m)))
                do k = 1, nt
                    wk = hbuf(k, m)
                    if (.not. (wk > SAFE_FLOOR)) cycle
                    if (c%nz > 1) then
                        mk = wk * ((1.0_wp - dz) * c%mass(k, idx(2, m)) + dz * c%mass(k, idx(2, m) + 1))
                        lk = wk * 10.0_wp**((1.0_wp - dz) * c%lbol(k, idx(2, m)) + dz * c%lbol(k, idx(2, m) + 1))
                    else
                        mk = wk * c%mass(k, 1)
                        lk = wk * 10.0_wp**c%lbol(k, 1)
                    end if
                    if (k <= it) then
                        msy(m) = msy(m) + mk; lsy(m) = lsy(m) + lk
                    else
                        mso(m) = mso(m) + mk; lso(m) = lso(m) + lk
                    end if
                end do
            end do
            call put(c_mass_young, msy)
            call put(c_mass_old, mso)
            call put(c_lbol_young, log10(max(lsy, tiny(1.0_wp))))
            call put(c_lbol_old, log10(max(lso, tiny(1.0_wp))))
        end if
        end associate

    contains

        subroutine put(p, x)
            type(c_ptr), intent(in) :: p
            real(WP), intent(in) :: x(:)
            real(c_double), pointer :: o(:)
            if (.not. c_associated(p)) return
            call c_f_pointer(p, o, [size(x)])
            o = x
        end subroutine put

        !> spectra (nspec, n) from the device to the caller's doubles (converted from FP32 in mixed precision)
        subroutine copy_spectra(dst, src)
            type(c_ptr), intent(in) :: dst, src
            real(c_double), pointer :: o(:, :)
            integer(c_int) :: r
            if (.not. c_associated(dst)) return
            if (bytes == 8) then
                r = opensps_hip_memcpy_d2h(dst, src, int(8, c_size_t) * ns * n)
            else
                if (.not. allocated(fbuf)) allocate(fbuf(ns, n))
                r = opensps_hip_memcpy_d2h(c_loc(fbuf), src, int(4, c_size_t) * ns * n)
                call c_f_pointer(dst, o, [ns, n])
                o = real(fbuf, WP)
            end if
        end subroutine copy_spectra

    end subroutine batch_run

    !> Linear AB fluxes of the last batch (compute_magnitudes_columns' maggies_out,
    !> src/spectra/opensps_batch_photometry.f90:166-179): band flux x 10^(0.4 (48.60 + 2.5 ABS_MAG_ZEROPOINT_LOG -
    !> distance term)), NaN for bands not computed.
    subroutine maggies_from_device(c, n, prm)
        type(host_ctx_t), intent(inout) :: c
        integer, intent(in) :: n
        real(WP), intent(in) :: prm(n, NPAR)
        real(WP), allocatable, target :: fb(:, :)
        integer(c_int) :: rc
        integer :: m, b, i
        real(WP) :: z, dm, dist, scale, amzl, pi, pc, slope
        allocate(fb(c%nbands, n))
        rc = opensps_hip_memcpy_d2h(c_loc(fb), c%d_fband%p, int(8, c_size_t) * c%nbands * n)
        pi = acos(-1.0_wp)
        pc = (648000.0_wp / pi) * 1.495978707e13_wp
        amzl = log10(3.828e33_wp / (4.0_wp * pi * (10.0_wp * pc) * (10.0_wp * pc)))
        do m = 1, n
            z = prm(m, P_ZRED)
            dist = 0.0_wp
            if (abs(z) > SAFE_FLOOR) then
                i = 1
                do while (i < 499 .and. c%cz(i + 1) > z)
                    i = i + 1
                end do
                slope = (c%cdl(i + 1) - c%cdl(i)) / (c%cz(i + 1) - c%cz(i))
                dm = c%cdl(i) + slope * (z - c%cz(i))
                if (dm == dm .and. dm > SAFE_FLOOR) dist = 5.0_wp * log10(dm / 10.0_wp) - 2.5_wp * log10(1.0_wp + z)
            end if
            scale = 10.0_wp**(0.4_wp * (48.60_wp + 2.5_wp * amzl - dist))
            do b = 1, c%nbands
                if (allocated(c%mag_mask) .and. .not. (c%compute_vega_mags == 1 .and. b == c%vega_ref)) then
                    if (c%mag_mask(b) == 0) then
                        c%last_maggies(b, m) = ieee_nan()
                        cycle
                    end if
                end if
                c%last_maggies(b, m) = fb(b, m) * scale
            end do
        end do
    end subroutine maggies_from_device

    real(WP) function ieee_nan()
        use, intrinsic :: ieee_arithmetic, only: ieee_value, ieee_quiet_nan
        ieee_nan = ieee_value(1.0_wp, ieee_quiet_nan)
    end function ieee_nan

    !> A batch on device-resident data: d_prm (NPAR, n) per-model parameters in the slots of
    !> include/opensps_params.h, d_weights (ntfull, n) or NULL, device outputs (mags required). precision
    !> -1: the context's; split 1: young and old apart. No transfer to or from the host.
    subroutine batch_run_device(h, n, d_prm, d_weights, precision, split, d_mags, d_spec, d_mass, d_lbol, status)
        integer, intent(in) :: h, n, precision, split
        type(c_ptr), intent(in) :: d_prm, d_weights, d_mags, d_spec, d_mass, d_lbol
        integer, intent(out) :: status
        type(hip_batch_t) :: a
        integer(c_int) :: rc
        status = 0
        if (.not. ctx_valid(h)) then
            status = 1
            call set_error(1, 'batch_device: bad handle')
            return
        end if
        if (.not. ctxs(h)%has_basis) then
            status = 1
            call set_error(1, 'batch_device: no SSP basis installed')
            return
        end if
        if (n < 0 .or. .not. c_associated(d_prm) .or. (.not. c_associated(d_mags) .and. ctxs(h)%nbands > 0)) then
            status = 3
            call set_error(3, 'batch_device: bad arguments')
            return
        end if
        if (n == 0) return
        status = ctx_device(h)
        if (status /= 0) return
        a%n = n
        a%precision = merge(HIP_MIXED, HIP_FP64, (precision < 0 .and. ctxs(h)%mixed == 1) .or. precision == 1)
        a%split_parts = merge(1, 0, split /= 0)
        a%use_ops = merge(1, 0, ctxs(h)%has_ops)
        a%prm = d_prm
        a%ssp_weights = d_weights
        a%mags = d_mags
        a%spec = d_spec
        a%mass = d_mass
        a%lbol = d_lbol
        rc = opensps_hip_csp_batch(ctxs(h)%dctx, a)
        if (rc == 0) rc = opensps_hip_sync(ctxs(h)%dctx)
        if (rc /= 0) then
            status = merge(int(rc), 10, rc == 2 .or. rc == 3)
            call set_error(status, 'batch_device: ' // trim(hip_message(ctxs(h)%dctx)))
        end if
    end subroutine batch_run_device

    !> opensps_context_evaluate_ssp_op (src/bindings/c/opensps_c_api.f90:1299-1358; ssp_operator_evaluate,
    !> src/physics/opensps_ssp_operator.f90:1346-1450): the SSP of every age for one IMF from the context's operator,
    !> mass_grid (ntfull) surviving mass, lbol_grid (ntfull) log10 L_bol, spec_grid (nspec x ntfull). One batch of
    !> ntfull models on the GPU, model t holding unit weight at age t (its CSP is the SSP of that age) with the given
    !> IMF, no dust, no redshift, no smoothing, at the metallicity zmet of the context, in FP64 (the reference's
    !> working precision). imf_type < 0: the context's; imf_params (imf1, imf2, imf3, vdmc, mdave) or NULL: the
    !> context's values. Status 303 without an operator, 3 for NULL outputs, 2 for an IMF type the operator does not
    !> integrate or, with Dave's IMF, an mdave other than the context's.
    subroutine ssp_op_evaluate(h, imf_type, c_imf_params, c_mass, c_lbol, c_spec, status)
        integer, intent(in) :: h, imf_type
        type(c_ptr), intent(in) :: c_imf_params, c_mass, c_lbol, c_spec
        integer, intent(out) :: status
        integer, parameter :: NV = 9
        character(kind=c_char, len=20), target :: nm(NV)
        type(c_ptr), target :: np(NV), vp(NV)
        real(WP), allocatable, target :: vals(:, :), w(:, :), mags(:, :)
        real(c_double), pointer :: ip(:)
        integer :: nt, k, zidx, it_save, mixed_save
        logical :: imf_dirty
        status = 0
        if (.not. ctx_valid(h)) then
            status = 1
            call set_error(1, 'evaluate_ssp_op: bad handle')
            return
        end if
        associate(c => ctxs(h))
        if (.not. c%has_ops) then
            status = 303
            call set_error(303, 'opensps_context_evaluate_ssp_op: no SSP operator on the context')
            return
        end if
        if (.not. c_associated(c_mass) .or. .not. c_associated(c_lbol) .or. .not. c_associated(c_spec)) then
            status = 3
            call set_error(3, 'opensps_context_evaluate_ssp_op: output pointers must not be NULL')
            return
        end if
        nt = c%nt
        allocate(vals(nt, NV), w(nt, nt), mags(max(c%nbands, 1), nt))
        w = 0.0_wp
        do k = 1, nt
            w(k, k) = 1.0_wp
        end do
        zidx = max(1, min(c%nz, c%zmet))
        nm(1) = 'dust1'; vals(:, 1) = 0.0_wp
        nm(2) = 'dust2'; vals(:, 2) = 0.0_wp
        nm(3) = 'dust3'; vals(:, 3) = 0.0_wp
        nm(4) = 'zred'; vals(:, 4) = 0.0_wp
        nm(5) = 'sigma_smooth'; vals(:, 5) = 0.0_wp
        nm(6) = 'logzsol'; vals(:, 6) = c%logz(zidx)
        nm(7) = 'imf1'; vals(:, 7) = c%pset(P_IMF1)
        nm(8) = 'imf2'; vals(:, 8) = c%pset(P_IMF1 + 1)
        nm(9) = 'imf3'; vals(:, 9) = c%pset(P_IMF1 + 2)
        if (c_associated(c_imf_params)) then
            call c_f_pointer(c_imf_params, ip, [5])
            vals(:, 7) = ip(1); vals(:, 8) = ip(2); vals(:, 9) = ip(3)
            if (merge(imf_type, c%imf_type, imf_type >= 0) == 4 .and. ip(5) /= c%mdave) then
                status = 2
                call set_error(2, 'opensps_context_evaluate_ssp_op: mdave other than the context''s (set_float "mdave")')
                return
            end if
        end if
        do k = 1, NV
            nm(k) = trim(nm(k)) // c_null_char
            np(k) = c_loc(nm(k))
            vp(k) = c_loc(vals(1, k))
        end do
        it_save = c%imf_type
        mixed_save = c%mixed
        imf_dirty = imf_type >= 0 .and. imf_type /= c%imf_type
        if (imf_dirty) then
            c%imf_type = imf_type
            c%dev_imf = .false.
        end if
        c%mixed = 0
        end associate
        call batch_run(h, nt, NV, c_loc(np), c_loc(vp), c_loc(w), c_loc(mags), c_spec, c_mass, c_lbol, c_null_ptr, &
                       c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, &
                       c_null_ptr, c_null_ptr, status)
        ctxs(h)%mixed = mixed_save
        if (imf_dirty) then
            ctxs(h)%imf_type = it_save
            ctxs(h)%dev_imf = .false.
        end if
    end subroutine ssp_op_evaluate

end module opensps_host_batch
