!> The batch with gradients through the C API (compute_csp_batch with its cotangents and params_bar /
!> weights_bar / weights_z_bar; batch_handle's gradient path, src/bindings/c/opensps_c_api.f90:663-961, and
!> batch_backward, src/physics/opensps_csp_batch.f90:4713-5136): the host resolves the parameters and checks the
!> options, the device runs the forward batch and its reverse pass (opensps_hip_csp_batch_grad); with SSP operators
!> on the context the IMF per model and its rows of params_bar (ssp_batch_age_sums_adjoint,
!> src/physics/opensps_ssp_batch.f90:991-1248).
module opensps_host_grad
    use, intrinsic :: iso_c_binding
    use opensps_host_types
    use opensps_host_context
    use opensps_hip_iface
    use opensps_host_batch, only: resolve_params, model_checks
    implicit none
    private
    public :: batch_run_grad

contains

    !> prows: the caller's leading dimension of params_bar (rows beyond NBGRAD are left as they are); fewer than
    !> NBGRAD rows is status 3 with nothing written (opensps_c_api.f90:729-735).
    subroutine batch_run_grad(h, n, n_params, c_names, c_values, c_weights, c_weights_z, c_zred_mags, c_mags_bar, &
                              c_spec_bar, c_mass_bar, c_lbol_bar, c_mags, c_spec, c_mass, c_lbol, c_mdust, c_sfr, &
                              c_emlines, c_indx, c_params_bar, prows, c_weights_bar, c_weights_z_bar, status)
        integer, intent(in) :: h, n, n_params, prows
        type(c_ptr), intent(in) :: c_names, c_values, c_weights, c_weights_z, c_zred_mags, c_mags_bar, c_spec_bar, &
                                   c_mass_bar, c_lbol_bar, c_mags, c_spec, c_mass, c_lbol, c_mdust, c_sfr, c_emlines, &
                                   c_indx, c_params_bar, c_weights_bar, c_weights_z_bar
        integer, intent(out) :: status
        real(WP), allocatable, target :: prm(:, :), pb(:, :), hspec(:, :)
        real(c_float), allocatable, target :: fspec(:, :)
        real(c_double), pointer :: o2(:, :), o1(:)
        type(hip_grad_t) :: g
        type(c_ptr) :: d(16)
        integer(c_size_t) :: sz(16)
        integer :: st, nb, ns, nt, nz, k, bytes, i
        integer(c_int) :: rc

        status = 0
        if (.not. ctx_valid(h)) then
            status = 1
            call set_error(1, 'batch_grad: bad handle')
            return
        end if
        associate(c => ctxs(h))
        if (.not. c%has_basis) then
            status = 1
            call set_error(1, 'batch_grad: no SSP basis installed')
            return
        end if
        if (n < 0 .or. n_params < 0) then
            status = 3
            call set_error(3, 'batch_grad: n or n_params < 0')
            return
        end if
        if (c_associated(c_params_bar) .and. prows < NBGRAD) then
            status = 3
            call set_error(3, 'compute_csp_batch_grad: params_bar has fewer rows than opensps_get_nbgrad returns')
            return
        end if
        if (n == 0) return
        if (.not. c_associated(c_mags)) then
            status = 3
            call set_error(3, 'batch_grad: mags is required')
            return
        end if
        if (c_associated(c_indx)) then
            status = 2
            call set_error(2, 'batch_grad: spectral indices are not part of this build')
            return
        end if
        if (c_associated(c_weights) .and. c_associated(c_weights_z)) then
            status = 3
            call set_error(3, 'batch_grad: ssp_weights and ssp_weights_z both given')
            return
        end if
        if (c%has_ops .and. c_associated(c_weights_z)) then
            ! (the forward refuses weights over metallicities with the operator, opensps_csp_batch.f90:616-618)
            status = 2
            call set_error(2, 'batch_grad: ssp_weights_z with the SSP operator')
            return
        end if
        call resolve_params(c, n, n_params, c_names, c_values, prm, status)
        if (status /= 0) return
        call model_checks(c, prm, n, st, c_associated(c_zred_mags), .true.)
        if (st /= 0) then
            status = st
            return
        end if
        st = ctx_device(h)
        if (st /= 0) then
            status = st
            return
        end if
        nb = c%nbands
        ns = c%nspec
        nt = c%nt
        nz = c%nz
        bytes = merge(4, 8, c%mixed == 1)
        ! device buffers: 1 prm, 2 weights, 3 weights_z, 4-7 cotangents, 8 mags, 9 spec, 10 mass, 11 lbol,
        ! 12 params_bar, 13 weights_bar, 14 weights_z_bar, 15 zred_mags
        d = c_null_ptr
        sz = 0
        sz(1) = 8_c_size_t * NPAR * n
        if (c_associated(c_weights)) sz(2) = 8_c_size_t * nt * n
        if (c_associated(c_weights_z)) sz(3) = 8_c_size_t * nt * nz * n
        if (c_associated(c_mags_bar)) sz(4) = 8_c_size_t * max(nb, 1) * n
        if (c_associated(c_spec_bar)) sz(5) = 8_c_size_t * ns * n
        if (c_associated(c_mass_bar)) sz(6) = 8_c_size_t * n
        if (c_associated(c_lbol_bar)) sz(7) = 8_c_size_t * n
        sz(8) = 8_c_size_t * max(nb, 1) * n
        if (c_associated(c_spec)) sz(9) = int(bytes, c_size_t) * ns * n
        if (c_associated(c_mass)) sz(10) = 8_c_size_t * n
        if (c_associated(c_lbol)) sz(11) = 8_c_size_t * n
        if (c_associated(c_params_bar)) sz(12) = 8_c_size_t * NBGRAD * n
        if (c_associated(c_weights_bar)) sz(13) = 8_c_size_t * nt * n
        if (c_associated(c_weights_z_bar) .and. c_associated(c_weights_z)) sz(14) = 8_c_size_t * nt * nz * n
        if (c_associated(c_zred_mags)) sz(15) = 8_c_size_t * n
        do k = 1, 15
            if (sz(k) > 0) then
                rc = opensps_hip_malloc(d(k), sz(k))
                if (rc /= 0) then
                    status = 10
                    call set_error(10, 'batch_grad: device allocation failed')
                    call release()
                    return
                end if
            end if
        end do
        rc = opensps_hip_memcpy_h2d(d(1), c_loc(prm), sz(1))
        if (sz(2) > 0) rc = opensps_hip_memcpy_h2d(d(2), c_weights, sz(2))
        if (sz(3) > 0) rc = opensps_hip_memcpy_h2d(d(3), c_weights_z, sz(3))
        if (sz(4) > 0) rc = opensps_hip_memcpy_h2d(d(4), c_mags_bar, sz(4))
        if (sz(5) > 0) rc = opensps_hip_memcpy_h2d(d(5), c_spec_bar, sz(5))
        if (sz(6) > 0) rc = opensps_hip_memcpy_h2d(d(6), c_mass_bar, sz(6))
        if (sz(7) > 0) rc = opensps_hip_memcpy_h2d(d(7), c_lbol_bar, sz(7))
        if (sz(15) > 0) rc = opensps_hip_memcpy_h2d(d(15), c_zred_mags, sz(15))
        g%fwd%n = n
        g%fwd%precision = merge(HIP_MIXED, HIP_FP64, c%mixed == 1)
        g%fwd%split_parts = 1
        g%fwd%use_ops = merge(1, 0, c%has_ops)
        g%fwd%prm = d(1)
        g%fwd%ssp_weights = d(2)
        g%fwd%ssp_weights_z = d(3)
        g%fwd%zred_mags = d(15)
        g%fwd%mags = d(8)
        g%fwd%spec = d(9)
        g%fwd%mass = d(10)
        g%fwd%lbol = d(11)
        g%mags_bar = d(4)
        g%spec_bar = d(5)
        g%mass_bar = d(6)
        g%lbol_bar = d(7)
        g%params_bar = d(12)
        g%weights_bar = d(13)
        g%weights_z_bar = d(14)
        rc = opensps_hip_csp_batch_grad(c%dctx, g)
        if (rc == 0) rc = opensps_hip_sync(c%dctx)
        if (rc /= 0) then
            status = merge(int(rc), 10, rc == 2 .or. rc == 3)
            call set_error(status, 'batch_grad: ' // trim(hip_message(c%dctx)))
            call release()
            return
        end if
        ! outputs
        if (nb > 0) rc = opensps_hip_memcpy_d2h(c_mags, d(8), 8_c_size_t * nb * n)
        if (sz(10) > 0) rc = opensps_hip_memcpy_d2h(c_mass, d(10), sz(10))
        if (sz(11) > 0) rc = opensps_hip_memcpy_d2h(c_lbol, d(11), sz(11))
        if (sz(9) > 0) then
            if (bytes == 8) then
                rc = opensps_hip_memcpy_d2h(c_spec, d(9), sz(9))
            else
                allocate(fspec(ns, n))
                rc = opensps_hip_memcpy_d2h(c_loc(fspec), d(9), sz(9))
                call c_f_pointer(c_spec, o2, [ns, n])
                o2 = real(fspec, WP)
            end if
        end if
        if (sz(12) > 0) then
            allocate(pb(NBGRAD, n))
            rc = opensps_hip_memcpy_d2h(c_loc(pb), d(12), sz(12))
            call c_f_pointer(c_params_bar, o2, [prows, n])
            o2(1:NBGRAD, 1:n) = pb
        end if
        if (sz(13) > 0) rc = opensps_hip_memcpy_d2h(c_weights_bar, d(13), sz(13))
        if (sz(14) > 0) rc = opensps_hip_memcpy_d2h(c_weights_z_bar, d(14), sz(14))
        if (c_associated(c_weights_z_bar) .and. .not. c_associated(c_weights_z)) then
            call c_f_pointer(c_weights_z_bar, o1, [nt * nz * n])
            o1 = 0.0_wp
        end if
        if (c_associated(c_mdust)) then
            call c_f_pointer(c_mdust, o1, [n])
            o1 = SAFE_FLOOR
        end if
        if (c_associated(c_emlines)) then
            call c_f_pointer(c_emlines, o2, [nemline(c), n])
            o2 = 0.0_wp
        end if
        if (c_associated(c_sfr)) then
            call c_f_pointer(c_sfr, o1, [n])
            do i = 1, n
                o1(i) = ieee_nan_v()
            end do
        end if
        call release()
        end associate

    contains

        subroutine release()
            integer :: q
            integer(c_int) :: r
            do q = 1, 16
                if (c_associated(d(q))) r = opensps_hip_free(d(q))
                d(q) = c_null_ptr
            end do
        end subroutine release

        real(WP) function ieee_nan_v()
            use, intrinsic :: ieee_arithmetic, only: ieee_value, ieee_quiet_nan
            ieee_nan_v = ieee_value(1.0_wp, ieee_quiet_nan)
        end function ieee_nan_v

    end subroutine batch_run_grad

end module opensps_host_grad
