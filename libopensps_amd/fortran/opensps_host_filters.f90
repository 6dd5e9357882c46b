!> Caller-supplied filter sets and their bands on the model grid: what the reference's filters_build computes
!> (src/spectra/opensps_filters.f90:186-503). A curve becomes a piecewise-linear function of wavelength (the
!> caller's samples, sanitised); its band on the grid is either its value at the grid's nodes or its projection on
!> the grid's hat functions (a merge over curve knots and grid nodes with the exact integral of two linear
!> functions); energy counters are divided by lambda; negatives are clipped and the band is normalised so that the
!> trapezoid rule gives integral T / lambda d lambda = 1, then optionally calibrated by integral (lambda /
!> lambda_ref)^expo T / lambda d lambda. The zero points and the Vega anchoring follow
!> src/spectra/opensps_context_filters.f90:113-198.
module opensps_host_filters
    use opensps_host_types, only: WP, SAFE_FLOOR
    implicit none
    private

    integer, parameter, public :: FILTER_OK = 0, FILTER_ERR_FEW_POINTS = 1, FILTER_ERR_NONFINITE = 2, &
                                  FILTER_ERR_ORDER = 3, FILTER_ERR_ARGUMENT = 4, FILTER_ERR_COVERAGE = 5, &
                                  FILTER_ERR_DEAD = 6
    integer, parameter, public :: FLAG_TRUNCATED = 1, FLAG_UNDERSAMPLED = 2, FLAG_DEAD = 4
    real(WP), parameter :: MIN_COVERAGE = 0.999_wp
    integer, parameter :: MIN_FWHM_POINTS = 5

    type, public :: filter_curve_t
        character(len=64) :: name = ''
        real(WP), allocatable :: lam(:), resp(:)
        integer :: detector = 0
        logical :: has_cal = .false.
        real(WP) :: cal_lref = 0.0_wp, cal_expo = 0.0_wp
    end type filter_curve_t

    type, public :: filter_set_t
        logical :: used = .false.
        integer :: n = 0
        type(filter_curve_t), allocatable :: f(:)
    end type filter_set_t

    type(filter_set_t), allocatable, public, save :: fsets(:)

    !> A piecewise-linear function: knots x (strictly increasing) and its values y there; beyond the end knots it
    !> continues the end segments
    type :: plin_t
        real(WP), allocatable :: x(:), y(:)
    end type plin_t

    public :: fset_create, fset_add, fset_free, fset_valid, filters_build, trapezoid, band_pivot, ref_spectrum_on_grid, &
              zero_points, sun_to_vega

contains

    logical function fset_valid(h)
        integer, intent(in) :: h
        fset_valid = .false.
        if (.not. allocated(fsets)) return
        if (h < 1 .or. h > size(fsets)) return
        fset_valid = fsets(h)%used
    end function fset_valid

    integer function fset_create() result(h)
        type(filter_set_t), allocatable :: tmp(:)
        integer :: i
        if (.not. allocated(fsets)) allocate(fsets(8))
        do i = 1, size(fsets)
            if (.not. fsets(i)%used) then
                h = i
                fsets(i)%used = .true.
                fsets(i)%n = 0
                return
            end if
        end do
        allocate(tmp(2 * size(fsets)))
        tmp(1:size(fsets)) = fsets
        call move_alloc(tmp, fsets)
        h = fset_create()
    end function fset_create

    subroutine fset_free(h)
        integer, intent(in) :: h
        if (.not. fset_valid(h)) return
        if (allocated(fsets(h)%f)) deallocate(fsets(h)%f)
        fsets(h)%n = 0
        fsets(h)%used = .false.
    end subroutine fset_free

    !> Appends a curve (status FILTER_ERR_ARGUMENT for n < 1, a bad detector or calibration)
    subroutine fset_add(h, name, lam, resp, detector, cal_lref, cal_expo, status)
        integer, intent(in) :: h, detector
        character(len=*), intent(in) :: name
        real(WP), intent(in) :: lam(:), resp(:), cal_lref, cal_expo
        integer, intent(out) :: status
        type(filter_curve_t), allocatable :: tmp(:)
        integer :: k
        status = FILTER_OK
        if (size(lam) < 1 .or. size(lam) /= size(resp) .or. (detector /= 0 .and. detector /= 1) .or. &
            cal_expo /= cal_expo) then
            status = FILTER_ERR_ARGUMENT
            return
        end if
        if (.not. allocated(fsets(h)%f)) allocate(fsets(h)%f(4))
        if (fsets(h)%n == size(fsets(h)%f)) then
            allocate(tmp(2 * size(fsets(h)%f)))
            tmp(1:fsets(h)%n) = fsets(h)%f(1:fsets(h)%n)
            call move_alloc(tmp, fsets(h)%f)
        end if
        k = fsets(h)%n + 1
        fsets(h)%f(k)%name = name
        fsets(h)%f(k)%lam = lam
        fsets(h)%f(k)%resp = resp
        fsets(h)%f(k)%detector = detector
        fsets(h)%f(k)%has_cal = cal_lref > 0.0_wp
        fsets(h)%f(k)%cal_lref = cal_lref
        fsets(h)%f(k)%cal_expo = cal_expo
        fsets(h)%n = k
    end subroutine fset_add

    ! ---- piecewise-linear functions ----------------------------------------------------------------------------

    !> The number of entries of the increasing array x at or below t (bisection)
    pure integer function n_at_or_below(x, t) result(k)
        real(WP), intent(in) :: x(:), t
        integer :: lo, hi, mid
        lo = 0                 ! x(1:lo) <= t
        hi = size(x) + 1       ! x(hi:) > t
        do while (hi - lo > 1)
            mid = lo + (hi - lo) / 2
            if (x(mid) <= t) then
                lo = mid
            else
                hi = mid
            end if
        end do
        k = lo
    end function n_at_or_below

    !> The segment of f that serves t: the one whose left knot is the last at or below t, kept within 1 .. n - 1
    pure integer function segment_of(f, t) result(s)
        type(plin_t), intent(in) :: f
        real(WP), intent(in) :: t
        s = min(max(n_at_or_below(f%x, t), 1), size(f%x) - 1)
    end function segment_of

    !> f on segment s at t (the segment's line, also beyond its knots): the left value plus the fraction of the
    !> segment's width times its rise
    pure real(WP) function on_segment(f, s, t) result(v)
        type(plin_t), intent(in) :: f
        integer, intent(in) :: s
        real(WP), intent(in) :: t
        real(WP) :: frac
        frac = (t - f%x(s)) / (f%x(s + 1) - f%x(s))
        v = f%y(s) + frac * (f%y(s + 1) - f%y(s))
    end function on_segment

    pure real(WP) function plin_at(f, t) result(v)
        type(plin_t), intent(in) :: f
        real(WP), intent(in) :: t
        v = on_segment(f, segment_of(f, t), t)
    end function plin_at

    !> integral of f(lambda) lambda^(-p) d lambda over [lo, hi] within f's knots, p = 1 or 2, exact: on a piece [a, b]
    !> where f runs linearly from fa to fb, f = (fa b - fb a) / (b - a) + lambda (fb - fa) / (b - a)
    pure real(WP) function plin_moment(f, lo, hi, p) result(acc)
        type(plin_t), intent(in) :: f
        real(WP), intent(in) :: lo, hi
        integer, intent(in) :: p
        real(WP) :: a, b, fa, fb, c0, lg
        integer :: s
        acc = 0.0_wp
        do s = 1, size(f%x) - 1
            a = max(f%x(s), lo)
            b = min(f%x(s + 1), hi)
            if (.not. (b > a)) cycle
            fa = on_segment(f, s, a)
            fb = on_segment(f, s, b)
            c0 = (fa * b - fb * a) / (b - a)
            lg = log(b / a)
            if (p == 1) then
                acc = acc + c0 * lg + (fb - fa)
            else
                acc = acc + c0 * (b - a) / (a * b) + (fb - fa) / (b - a) * lg
            end if
        end do
    end function plin_moment

    ! ---- a caller's curve ------------------------------------------------------------------------------------------

    !> The caller's samples as a piecewise-linear function, by the rules of the reference's strict filter_clean
    !> (opensps_filters.f90:186-230): every value finite and every wavelength positive (else FILTER_ERR_NONFINITE);
    !> samples given in decreasing wavelength are taken in reverse; a sample at the wavelength of the one before it
    !> is dropped; responses below 0 count as 0; what remains must increase strictly (FILTER_ERR_ORDER) and hold
    !> two samples at least (FILTER_ERR_FEW_POINTS)
    subroutine curve_from_samples(lam, resp, f, status)
        real(WP), intent(in) :: lam(:), resp(:)
        type(plin_t), intent(out) :: f
        integer, intent(out) :: status
        integer, allocatable :: ord(:)
        logical, allocatable :: keep(:)
        integer :: n, i
        status = FILTER_OK
        n = size(lam)
        if (.not. (all(abs(lam) <= huge(1.0_wp)) .and. all(abs(resp) <= huge(1.0_wp)) .and. all(lam > 0.0_wp))) then
            status = FILTER_ERR_NONFINITE
            return
        end if
        ord = [(i, i = 1, n)]
        if (lam(n) < lam(1)) ord = ord(n:1:-1)
        allocate(keep(n))
        keep(1) = .true.
        do i = 2, n
            keep(i) = abs(lam(ord(i)) - lam(ord(i - 1))) >= SAFE_FLOOR
        end do
        f%x = pack(lam(ord), keep)
        f%y = max(pack(resp(ord), keep), 0.0_wp)
        if (size(f%x) >= 2) then
            if (any(f%x(2:) <= f%x(:size(f%x) - 1))) status = FILTER_ERR_ORDER
        else
            status = FILTER_ERR_FEW_POINTS
        end if
    end subroutine curve_from_samples

    ! ---- bands on the grid -----------------------------------------------------------------------------------------

    !> f's values at the grid's nodes (the reference's resample_point, opensps_filters.f90:359-381): from the last node
    !> at or below the first knot (the first node when the curve starts left of the grid; the one before the last when
    !> it starts on the last node) to the last node at or below the last knot, where these span two nodes at least;
    !> 0 elsewhere
    subroutine sample_points(f, grid, band)
        type(plin_t), intent(in) :: f
        real(WP), intent(in) :: grid(:)
        real(WP), intent(out) :: band(:)
        integer :: i0, i1, ng, i
        ng = size(grid)
        band = 0.0_wp
        i0 = max(n_at_or_below(grid, f%x(1)), 1)
        if (f%x(1) == grid(ng)) i0 = ng - 1
        i1 = n_at_or_below(grid, f%x(size(f%x)))
        if (i1 <= i0) return
        do i = i0, i1
            band(i) = plin_at(f, grid(i))
        end do
    end subroutine sample_points

    !> f projected on the grid's hat functions and divided by each hat's area (the reference's resample_hat,
    !> opensps_filters.f90:319-357): one pass over the pieces between consecutive break points (curve knots and grid
    !> nodes merged) of the overlap of f and the grid; on a piece [a, b] both f and the cell's two hats are linear,
    !> and the integral of the product of linear u, v is (b - a) (2 ua va + ua vb + ub va + 2 ub vb) / 6
    subroutine project_hats(f, grid, band)
        type(plin_t), intent(in) :: f
        real(WP), intent(in) :: grid(:)
        real(WP), intent(out) :: band(:)
        integer :: ng, nk, s, j
        real(WP) :: lo, hi, a, b, fa, fb, w, ra, rb, area
        ng = size(grid)
        nk = size(f%x)
        band = 0.0_wp
        lo = max(f%x(1), grid(1))
        hi = min(f%x(nk), grid(ng))
        if (.not. (hi > lo)) return
        s = segment_of(f, lo)
        j = min(max(n_at_or_below(grid, lo), 1), ng - 1)
        a = lo
        do while (a < hi)
            b = min(f%x(s + 1), grid(j + 1), hi)
            if (b > a) then
                fa = on_segment(f, s, a)
                fb = on_segment(f, s, b)
                w = grid(j + 1) - grid(j)
                area = (b - a) / 6.0_wp
                ! node j's hat falls as (grid(j + 1) - lambda) / w over the cell, node j + 1's rises as
                ! (lambda - grid(j)) / w: each from its own node, so that neither loses digits near the other
                ra = (grid(j + 1) - a) / w
                rb = (grid(j + 1) - b) / w
                band(j) = band(j) + area * (2.0_wp * fa * ra + fa * rb + fb * ra + 2.0_wp * fb * rb)
                ra = (a - grid(j)) / w
                rb = (b - grid(j)) / w
                band(j + 1) = band(j + 1) + area * (2.0_wp * fa * ra + fa * rb + fb * ra + 2.0_wp * fb * rb)
            end if
            if (b >= f%x(s + 1) .and. s < nk - 1) s = s + 1
            if (b >= grid(j + 1) .and. j < ng - 1) j = j + 1
            a = b
        end do
        band(1) = band(1) / (0.5_wp * (grid(2) - grid(1)))
        band(2:ng - 1) = band(2:ng - 1) / (0.5_wp * (grid(3:ng) - grid(1:ng - 2)))
        band(ng) = band(ng) / (0.5_wp * (grid(ng) - grid(ng - 1)))
        band = max(band, 0.0_wp)
    end subroutine project_hats

    !> The trapezoid rule of y over the nodes x
    pure real(WP) function trapezoid(x, y) result(s)
        real(WP), intent(in) :: x(:), y(:)
        integer :: n
        n = size(x)
        s = 0.0_wp
        if (n < 2) return
        s = 0.5_wp * sum((x(2:n) - x(1:n - 1)) * (y(2:n) + y(1:n - 1)))
    end function trapezoid

    !> The bands of filter set h on the grid, normalised (filters_build, opensps_filters.f90:395-503, for a caller's
    !> set; mode 0 point values, 1 hat projection). flags: FLAG_TRUNCATED when less than MIN_COVERAGE of the curve's
    !> integral T / lambda^p (p = 2 for energy counters) lies on the grid, FLAG_UNDERSAMPLED for fewer than
    !> MIN_FWHM_POINTS nodes at or above half the band's peak, FLAG_DEAD for no throughput on the grid. status:
    !> FILTER_OK or the failing filter's code, that filter's index in bad; policy 0 makes a truncated or dead filter an
    !> error
    subroutine filters_build(h, grid, mode, policy, bands, flags, status, bad)
        integer, intent(in) :: h, mode, policy
        real(WP), intent(in) :: grid(:)
        real(WP), intent(out) :: bands(:, :)
        integer, intent(out) :: flags(:), status, bad
        type(plin_t) :: f
        real(WP) :: total, on_grid, area, cal
        integer :: i, st, p, ng
        status = FILTER_OK
        bad = 0
        ng = size(grid)
        bands = 0.0_wp
        flags = 0
        do i = 1, fsets(h)%n
            associate(fc => fsets(h)%f(i), band => bands(:, i))
                call curve_from_samples(fc%lam, fc%resp, f, st)
                if (st /= FILTER_OK) then
                    status = st
                    bad = i
                    return
                end if
                p = merge(2, 1, fc%detector == 1)
                total = plin_moment(f, f%x(1), f%x(size(f%x)), p)
                on_grid = 0.0_wp
                if (total > 0.0_wp) on_grid = plin_moment(f, grid(1), grid(ng), p) / total
                if (on_grid < MIN_COVERAGE) flags(i) = ior(flags(i), FLAG_TRUNCATED)
                if (mode == 1) then
                    call project_hats(f, grid, band)
                else
                    call sample_points(f, grid, band)
                end if
                if (fc%detector == 1) band = band / grid
                band = max(band, 0.0_wp)
                area = trapezoid(grid, band / grid)
                if (.not. (area > SAFE_FLOOR)) then
                    band = 0.0_wp
                    flags(i) = ior(flags(i), FLAG_DEAD)
                    cycle
                end if
                band = band / area
                if (count(band >= 0.5_wp * maxval(band)) < MIN_FWHM_POINTS) flags(i) = ior(flags(i), FLAG_UNDERSAMPLED)
                if (fc%has_cal) then
                    cal = trapezoid(grid, (grid / fc%cal_lref)**fc%cal_expo * band / grid)
                    if (cal > SAFE_FLOOR) band = band / cal
                end if
            end associate
        end do
        if (policy /= 0) return
        do i = 1, fsets(h)%n
            if (iand(flags(i), FLAG_DEAD) /= 0) then
                status = FILTER_ERR_DEAD
            else if (iand(flags(i), FLAG_TRUNCATED) /= 0) then
                status = FILTER_ERR_COVERAGE
            else
                cycle
            end if
            bad = i
            return
        end do
    end subroutine filters_build

    !> The pivot wavelength of a band: the square root of its first moment over its minus-first (the reference's
    !> filter_pivot, opensps_filters.f90:521-532); 0 without throughput
    real(WP) function band_pivot(grid, band) result(leff)
        real(WP), intent(in) :: grid(:), band(:)
        real(WP) :: up, down
        up = trapezoid(grid, grid * band)
        down = trapezoid(grid, band / grid)
        leff = 0.0_wp
        if (up == up .and. down == down .and. down > SAFE_FLOOR) leff = sqrt(up / down)
    end function band_pivot

    ! ---- reference spectra and zero points -------------------------------------------------------------------------

    !> A reference spectrum (Vega, the Sun) on the model grid as the reference's load_standard_sed places it
    !> (src/io/opensps_io.f90:1516-1582): straight lines in (log10 lambda, log10(flux + tiny)), continued beyond the
    !> table's ends, F_lambda tables times lambda^2 (f_nu up to a constant), tiny redward of the table
    subroutine ref_spectrum_on_grid(lt, ft, flambda, grid, spec)
        real(WP), intent(in) :: lt(:), ft(:), grid(:)
        logical, intent(in) :: flambda
        real(WP), intent(out) :: spec(:)
        type(plin_t) :: loglog
        integer :: j
        loglog%x = log10(lt)
        loglog%y = log10(ft + SAFE_FLOOR)
        do j = 1, size(grid)
            if (grid(j) > maxval(lt)) then
                spec(j) = SAFE_FLOOR
            else
                spec(j) = 10.0_wp**plin_at(loglog, log10(grid(j)))
                if (flambda) spec(j) = spec(j) * grid(j)**2
            end if
        end do
    end subroutine ref_spectrum_on_grid

    !> The AB magnitude of f_nu (on the grid) in every band, -2.5 log10(integral f_nu T / lambda d lambda) - 48.60,
    !> 99 where it has no flux (the reference's filters_zero_points, opensps_filters.f90:536-556)
    subroutine zero_points(grid, spec, bands, mag)
        real(WP), intent(in) :: grid(:), spec(:), bands(:, :)
        real(WP), intent(out) :: mag(:)
        real(WP) :: flux
        integer :: b
        do b = 1, size(bands, 2)
            flux = trapezoid(grid, spec * bands(:, b) / grid)
            mag(b) = merge(-2.5_wp * log10(max(flux, SAFE_FLOOR)) - 48.60_wp, 99.0_wp, flux > SAFE_FLOOR)
        end do
    end subroutine zero_points

    !> The Sun's magnitudes in the Vega system (the reference's sun_to_vega_system,
    !> src/spectra/opensps_context_filters.f90:185-198): m_sun - m_vega in every band, plus Vega's AB magnitude in the
    !> reference band vref when there is one (that band stays AB); bands without flux (99) stay 99
    subroutine sun_to_vega(magsun, magvega, vref)
        real(WP), intent(inout) :: magsun(:)
        real(WP), intent(in) :: magvega(:)
        integer, intent(in) :: vref
        real(WP) :: anchor
        anchor = 0.0_wp
        if (vref > 0) anchor = magvega(vref)
        where (magsun /= 99.0_wp .and. magvega /= 99.0_wp) magsun = magsun - magvega + anchor
    end subroutine sun_to_vega

end module opensps_host_filters
