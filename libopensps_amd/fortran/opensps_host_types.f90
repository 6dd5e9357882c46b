!> Constants, parameter slots and the error state of the Fortran host.
!>
!> The per-model parameter slots are those of include/opensps_params.h (the members of the reference's
!> PARAMS_CORE that the batch reads, src/core/opensps_types.f90:33-45), 1-based here: slot(name) = the
!> header's OPS_P_<NAME> + 1.
module opensps_host_types
    use, intrinsic :: iso_c_binding, only: c_double, c_int
    implicit none
    private

    integer, parameter, public :: WP = c_double
    real(WP), parameter, public :: SAFE_FLOOR = tiny(1.0_wp)

    integer, parameter, public :: NPAR = 35
    character(len=16), parameter, public :: PARAM_NAMES(NPAR) = [character(len=16) :: &
        'tage', 'tau', 'const', 'fburst', 'tburst', 'sf_start', 'sf_trunc', 'sf_slope', 'sfh', 'logzsol', &
        'dust1', 'dust2', 'dust3', 'dust_index', 'dust1_index', 'dust_tesc', 'frac_obrun', 'frac_nodust', &
        'zred', 'igm_factor', 'sigma_smooth', 'min_wave_smooth', 'max_wave_smooth', 'imf1', 'imf2', 'imf3', &
        'imf_upper_limit', 'imf_lower_limit', 'duste_umin', 'duste_qpah', 'duste_gamma', 'fagn', 'agn_tau', &
        'gas_logz', 'gas_logu']
    real(WP), parameter, public :: PARAM_DEFAULTS(NPAR) = [ &
        0.0_wp, 1.0_wp, 0.0_wp, 0.0_wp, 11.0_wp, 0.0_wp, 0.0_wp, 0.0_wp, 0.0_wp, 0.0_wp, &
        0.0_wp, 0.0_wp, 0.0_wp, -0.7_wp, -1.0_wp, 7.0_wp, 0.0_wp, 0.0_wp, &
        0.0_wp, 1.0_wp, 0.0_wp, 1.0e3_wp, 1.0e4_wp, 1.3_wp, 2.3_wp, 2.3_wp, &
        120.0_wp, 0.08_wp, 1.0_wp, 3.5_wp, 0.01_wp, 0.0_wp, 10.0_wp, 0.0_wp, -2.0_wp]
    integer, parameter, public :: P_TAGE = 1, P_SFH = 9, P_LOGZSOL = 10, P_DUST1 = 11, P_DUST2 = 12, P_DUST3 = 13, &
                                  P_FRAC_OBRUN = 17, P_ZRED = 19, P_IGM_FACTOR = 20, P_SIGMA_SMOOTH = 21, &
                                  P_TAU = 2, P_CONST = 3, P_FBURST = 4, P_TBURST = 5, P_SF_START = 6, P_SF_TRUNC = 7, &
                                  P_SF_SLOPE = 8, &
                                  P_IMF1 = 24, P_DUSTE_UMIN = 29, P_DUSTE_QPAH = 30, P_DUSTE_GAMMA = 31, &
                                  P_FAGN = 32, P_AGN_TAU = 33, P_GAS_LOGZ = 34, P_GAS_LOGU = 35

    !> Parameters the batch refuses to vary per model: they change the SSPs (is_ssp_sensitive_float_param,
    !> src/core/opensps_context.f90:958-969, and the integer ones), so the reference's batch_handle returns status 2
    !> for them (src/bindings/c/opensps_c_api.f90:761-788)
    character(len=20), parameter, public :: SSP_NAMES(18) = [character(len=20) :: &
        'imf_type', 'zmet', 'pagb', 'dell', 'delt', 'fbhb', 'sbss', 'redgb', 'agb', 'masscut', 'evtype', 'fcstar', &
        'tiny_logt', 'logt_wmb_hot', 'nebular_smooth_init', 'frac_xrb', 'agb_dust', 'imf_filename']

    !> The IMF's parameters: per model only with an SSP operator on the context (the reference's IMF_PARAMS,
    !> src/bindings/c/opensps_c_api.f90:713-714, 772-776); status 2 without one
    character(len=16), parameter, public :: IMF_NAMES(7) = [character(len=16) :: &
        'imf1', 'imf2', 'imf3', 'vdmc', 'mdave', 'imf_lower_limit', 'imf_upper_limit']

    !> Members of PARAMS_CORE (src/core/opensps_types.f90:33-45) that no stage this build runs reads: the dust
    !> curves of dust_type 1, 3, 4, 5 (mwr, uvb), the clumpy screen and pmetals (not read by the batch) and cue's gas
    !> parameters (gas_logn, gas_logno, gas_logco: nebular_model = 1, not built). They are accepted per model and at
    !> the context, as the reference accepts them
    integer, parameter, public :: NINERT = 7
    character(len=16), parameter, public :: INERT_NAMES(NINERT) = [character(len=16) :: &
        'pmetals', 'dust_clumps', 'uvb', 'mwr', 'gas_logn', 'gas_logno', 'gas_logco']
    real(WP), parameter, public :: INERT_DEFAULTS(NINERT) = [ &
        0.02_wp, -99.0_wp, 1.0_wp, 3.1_wp, 2.0_wp, 0.0_wp, 0.0_wp]

    !> Rows of params_bar (include/opensps.h:12-59), as opensps_get_nbgrad reports them
    integer, parameter, public :: NBGRAD = 44
    character(len=16), parameter, public :: GRAD_NAMES(NBGRAD) = [character(len=16) :: &
        'fagn', 'dust1', 'dust2', 'dust3', 'dust_index', 'dust1_index', 'frac_obrun', 'frac_nodust', 'mwr', 'uvb', &
        'duste_qpah', 'duste_umin', 'duste_gamma', 'gas_logz', 'gas_logu', 'logzsol', 'tage', 'tau', 'const', &
        'fburst', 'tburst', 'sf_start', 'zred', 'sigma_smooth', 'sf_trunc', 'sf_slope', 'imf1', 'imf2', 'imf3', &
        'vdmc', 'mdave', 'agn_tau', 'igm_factor', 'gas_logn', 'gas_logno', 'gas_logco', 'afe', 'fbhb', 'sbss', &
        'agb', 'pagb', 'redgb', 'delt', 'dell']

    character(len=512), public :: last_error = ''
    integer, public :: last_status = 0

    public :: param_slot, inert_slot, set_error, lower

contains

    pure function lower(s) result(r)
        character(len=*), intent(in) :: s
        character(len=len(s)) :: r
        integer :: i, c
        r = s
        do i = 1, len(s)
            c = iachar(s(i:i))
            if (c >= 65 .and. c <= 90) r(i:i) = achar(c + 32)
        end do
    end function lower

    !> The slot of a parameter name (1-based), 0 if it is not a per-model parameter of the batch
    pure integer function param_slot(name) result(k)
        character(len=*), intent(in) :: name
        integer :: i
        k = 0
        do i = 1, NPAR
            if (trim(PARAM_NAMES(i)) == lower(trim(name))) then
                k = i
                return
            end if
        end do
    end function param_slot

    !> The index of a parameter in INERT_NAMES, 0 if it is not one
    pure integer function inert_slot(name) result(k)
        character(len=*), intent(in) :: name
        integer :: i
        k = 0
        do i = 1, NINERT
            if (trim(INERT_NAMES(i)) == lower(trim(name))) then
                k = i
                return
            end if
        end do
    end function inert_slot

    subroutine set_error(status, msg)
        integer, intent(in) :: status
        character(len=*), intent(in) :: msg
        last_status = status
        last_error = msg
    end subroutine set_error

end module opensps_host_types
