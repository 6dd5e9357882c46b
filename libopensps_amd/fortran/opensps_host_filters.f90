!> Caller-supplied filter sets and their bands on the model grid (the reference's opensps_filters,
!> src/spectra/opensps_filters.f90:186-503): the curves are cleaned, resampled onto the grid (point
!> interpolation or the conserving hat projection), counted as photons or energy, clipped at zero and
!> normalised to integral T / lambda d lambda = 1 by the trapezoid rule; an optional calibration divides
!> by integral (lambda / lambda_ref)^expo T / lambda d lambda.
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

    !> filter_clean, strict (opensps_filters.f90:186-230): finite, positive wavelengths, reversed if decreasing,
    !> repeated wavelengths dropped, negative responses clipped, then strictly increasing with >= 2 points
    subroutine curve_clean(lam, resp, lc, rc, nc, status)
        real(WP), intent(in) :: lam(:), resp(:)
        real(WP), allocatable, intent(out) :: lc(:), rc(:)
        integer, intent(out) :: nc, status
        real(WP), allocatable :: l(:), r(:)
        integer :: i, n
        status = FILTER_OK
        n = size(lam)
        allocate(lc(n), rc(n))
        nc = 0
        if (any(lam /= lam) .or. any(resp /= resp) .or. any(abs(lam) > huge(1.0_wp)) .or. &
            any(abs(resp) > huge(1.0_wp)) .or. any(lam <= 0.0_wp)) then
            status = FILTER_ERR_NONFINITE
            return
        end if
        l = lam
        r = resp
        if (n > 1) then
            if (l(n) < l(1)) then
                l = l(n:1:-1)
                r = r(n:1:-1)
            end if
        end if
        do i = 1, n
            if (nc > 0) then
                if (abs(l(i) - lc(nc)) < SAFE_FLOOR) cycle
            end if
            nc = nc + 1
            lc(nc) = l(i)
            rc(nc) = max(r(i), 0.0_wp)
        end do
        if (nc >= 2) then
            if (any(lc(2:nc) <= lc(1:nc - 1))) then
                status = FILTER_ERR_ORDER
                return
            end if
        end if
        if (nc < 2) status = FILTER_ERR_FEW_POINTS
    end subroutine curve_clean

    !> Exact integral of r / lambda^p over [lo, hi] of the piecewise-linear curve (opensps_filters.f90:264-291)
    pure real(WP) function curve_integral(lc, r, nc, lo, hi, p) result(s)
        real(WP), intent(in) :: lc(:), r(:), lo, hi
        integer, intent(in) :: nc, p
        integer :: k
        real(WP) :: a, b, ya, yb, t, beta, alpha
        s = 0.0_wp
        do k = 1, nc - 1
            a = max(lc(k), lo)
            b = min(lc(k + 1), hi)
            if (b <= a) cycle
            t = (a - lc(k)) / (lc(k + 1) - lc(k))
            ya = r(k) + t * (r(k + 1) - r(k))
            t = (b - lc(k)) / (lc(k + 1) - lc(k))
            yb = r(k) + t * (r(k + 1) - r(k))
            beta = (yb - ya) / (b - a)
            alpha = ya - beta * a
            if (p == 2) then
                s = s + alpha * (1.0_wp / a - 1.0_wp / b) + beta * log(b / a)
            else
                s = s + alpha * log(b / a) + beta * (b - a)
            end if
        end do
    end function curve_integral

    pure real(WP) function trapezoid(x, y) result(s)
        real(WP), intent(in) :: x(:), y(:)
        integer :: j
        s = 0.0_wp
        do j = 1, size(x) - 1
            s = s + 0.5_wp * (x(j + 1) - x(j)) * (y(j) + y(j + 1))
        end do
    end function trapezoid

    !> Linear interpolation with the reference's interval search (interpolate_linear_scalar)
    pure real(WP) function interp_lin(x, y, n, xo) result(v)
        real(WP), intent(in) :: x(:), y(:), xo
        integer, intent(in) :: n
        integer :: lo, hi, mid, i
        lo = 0
        hi = n + 1
        do while (hi - lo > 1)
            mid = (hi + lo) / 2
            if (xo >= x(mid)) then
                lo = mid
            else
                hi = mid
            end if
        end do
        i = lo
        if (xo == x(1)) i = 1
        if (xo == x(n)) i = n - 1
        i = max(1, min(i, n - 1))
        v = y(i) + (y(i + 1) - y(i)) / (x(i + 1) - x(i)) * (xo - x(i))
    end function interp_lin

    pure integer function interval_of(x, n, v) result(i)
        real(WP), intent(in) :: x(:), v
        integer, intent(in) :: n
        integer :: lo, hi, mid
        lo = 0
        hi = n + 1
        do while (hi - lo > 1)
            mid = (hi + lo) / 2
            if (v >= x(mid)) then
                lo = mid
            else
                hi = mid
            end if
        end do
        i = lo
        if (v == x(1)) i = 1
        if (v == x(n)) i = n - 1
    end function interval_of

    !> resample_point (opensps_filters.f90:359-381)
    subroutine resample_point(lc, r, nc, grid, band)
        real(WP), intent(in) :: lc(:), r(:), grid(:)
        integer, intent(in) :: nc
        real(WP), intent(out) :: band(:)
        integer :: si, ei, i, ng
        ng = size(grid)
        band = 0.0_wp
        if (nc <= 1) return
        if (lc(1) <= grid(1)) then
            si = 1
        else
            si = max(interval_of(grid, ng, lc(1)), 1)
        end if
        if (lc(nc) >= grid(ng)) then
            ei = ng
        else
            ei = min(interval_of(grid, ng, lc(nc)), ng)
        end if
        if (ei > si) then
            do i = si, ei
                band(i) = interp_lin(lc, r, nc, grid(i))
            end do
        end if
    end subroutine resample_point

    pure integer function cell_of(grid, x) result(j)
        real(WP), intent(in) :: grid(:), x
        integer :: lo, hi, mid
        lo = 1
        hi = size(grid)
        if (x <= grid(1)) then
            j = 1
            return
        end if
        if (x >= grid(hi)) then
            j = hi - 1
            return
        end if
        do while (hi - lo > 1)
            mid = (lo + hi) / 2
            if (grid(mid) <= x) then
                lo = mid
            else
                hi = mid
            end if
        end do
        j = lo
    end function cell_of

    !> resample_hat (opensps_filters.f90:319-357): hat-function projection, exact by Simpson's rule per piece
    subroutine resample_hat(lc, r, nc, grid, band)
        real(WP), intent(in) :: lc(:), r(:), grid(:)
        integer, intent(in) :: nc
        real(WP), intent(out) :: band(:)
        integer :: k, j, ng
        real(WP) :: x0, x1, a, b, ya, yb, dx, ym, ha, hb, hm, t, m
        real(WP), allocatable :: acc(:)
        ng = size(grid)
        allocate(acc(ng))
        acc = 0.0_wp
        do k = 1, nc - 1
            x0 = max(lc(k), grid(1))
            x1 = min(lc(k + 1), grid(ng))
            if (x1 <= x0) cycle
            j = cell_of(grid, x0)
            do while (j <= ng - 1)
                if (grid(j) >= x1) exit
                a = max(x0, grid(j))
                b = min(x1, grid(j + 1))
                if (b > a) then
                    dx = grid(j + 1) - grid(j)
                    t = (a - lc(k)) / (lc(k + 1) - lc(k)); ya = r(k) + t * (r(k + 1) - r(k))
                    t = (b - lc(k)) / (lc(k + 1) - lc(k)); yb = r(k) + t * (r(k + 1) - r(k))
                    m = 0.5_wp * (a + b); ym = 0.5_wp * (ya + yb)
                    ha = (grid(j + 1) - a) / dx; hb = (grid(j + 1) - b) / dx; hm = (grid(j + 1) - m) / dx
                    acc(j) = acc(j) + (b - a) / 6.0_wp * (ya * ha + 4.0_wp * ym * hm + yb * hb)
                    ha = (a - grid(j)) / dx; hb = (b - grid(j)) / dx; hm = (m - grid(j)) / dx
                    acc(j + 1) = acc(j + 1) + (b - a) / 6.0_wp * (ya * ha + 4.0_wp * ym * hm + yb * hb)
                end if
                j = j + 1
            end do
        end do
        band(1) = acc(1) / (0.5_wp * (grid(2) - grid(1)))
        do j = 2, ng - 1
            band(j) = acc(j) / (0.5_wp * (grid(j + 1) - grid(j - 1)))
        end do
        band(ng) = acc(ng) / (0.5_wp * (grid(ng) - grid(ng - 1)))
        where (band < 0.0_wp) band = 0.0_wp
    end subroutine resample_hat

    !> filters_build (opensps_filters.f90:395-503) for a caller set: bands(:, i) normalised; status FILTER_OK or the
    !> failing filter's code in `bad`; policy 0 makes a truncated or dead filter an error
    subroutine filters_build(h, grid, mode, policy, bands, flags, status, bad)
        integer, intent(in) :: h, mode, policy
        real(WP), intent(in) :: grid(:)
        real(WP), intent(out) :: bands(:, :)
        integer, intent(out) :: flags(:), status, bad
        real(WP), allocatable :: lc(:), rc(:), tmp(:)
        real(WP) :: norm, cov, d
        integer :: i, nc, st, ng, p, k
        status = FILTER_OK
        bad = 0
        ng = size(grid)
        bands = 0.0_wp
        flags = 0
        allocate(tmp(ng))
        do i = 1, fsets(h)%n
            associate(f => fsets(h)%f(i))
            call curve_clean(f%lam, f%resp, lc, rc, nc, st)
            if (st /= FILTER_OK) then
                status = st
                bad = i
                return
            end if
            p = merge(2, 1, f%detector == 1)
            norm = curve_integral(lc, rc, nc, lc(1), lc(nc), p)
            cov = 0.0_wp
            if (norm > 0.0_wp) cov = curve_integral(lc, rc, nc, grid(1), grid(ng), p) / norm
            if (mode == 1) then
                call resample_hat(lc, rc, nc, grid, bands(:, i))
            else
                call resample_point(lc, rc, nc, grid, bands(:, i))
            end if
            if (f%detector == 1) bands(:, i) = bands(:, i) / grid
            where (bands(:, i) < 0.0_wp) bands(:, i) = 0.0_wp
            tmp = bands(:, i) / (grid + SAFE_FLOOR)
            norm = trapezoid(grid, tmp)
            if (norm > SAFE_FLOOR) then
                bands(:, i) = bands(:, i) / norm
                k = count(bands(:, i) >= 0.5_wp * maxval(bands(:, i)))
                if (k < MIN_FWHM_POINTS) flags(i) = ior(flags(i), FLAG_UNDERSAMPLED)
                if (f%has_cal) then
                    tmp = (grid / f%cal_lref)**f%cal_expo * bands(:, i) / (grid + SAFE_FLOOR)
                    d = trapezoid(grid, tmp)
                    if (d > SAFE_FLOOR) bands(:, i) = bands(:, i) / d
                end if
            else
                bands(:, i) = 0.0_wp
                flags(i) = ior(flags(i), FLAG_DEAD)
            end if
            if (cov < MIN_COVERAGE) flags(i) = ior(flags(i), FLAG_TRUNCATED)
            end associate
        end do
        if (policy == 0) then
            do i = 1, fsets(h)%n
                if (iand(flags(i), FLAG_DEAD) /= 0) then
                    status = FILTER_ERR_DEAD
                    bad = i
                    return
                else if (iand(flags(i), FLAG_TRUNCATED) /= 0) then
                    status = FILTER_ERR_COVERAGE
                    bad = i
                    return
                end if
            end do
        end if
    end subroutine filters_build

    !> Pivot wavelength sqrt(integral lambda T / integral T / lambda) (filter_pivot, opensps_filters.f90:521-532)
    real(WP) function band_pivot(grid, band) result(leff)
        real(WP), intent(in) :: grid(:), band(:)
        real(WP) :: d
        leff = trapezoid(grid, grid * band)
        d = trapezoid(grid, band / grid)
        if (leff /= leff .or. d /= d .or. d <= SAFE_FLOOR) then
            leff = 0.0_wp
        else
            leff = sqrt(leff / d)
        end if
    end function band_pivot

    !> A reference spectrum (Vega, the Sun) on the model grid as load_standard_sed forms it (src/io/opensps_io.f90:
    !> 1516-1582): log10(flux + tiny) interpolated linearly in log10(lambda), the end intervals' slopes carried beyond
    !> the table, then exponentiated; an F_lambda table is multiplied by lambda^2 (f_nu up to a constant); tiny
    !> redward of the table's last wavelength.
    subroutine ref_spectrum_on_grid(lt, ft, flambda, grid, spec)
        real(WP), intent(in) :: lt(:), ft(:), grid(:)
        logical, intent(in) :: flambda
        real(WP), intent(out) :: spec(:)
        real(WP), allocatable :: x(:), y(:)
        integer :: j, n
        n = size(lt)
        x = log10(lt)
        y = log10(ft + SAFE_FLOOR)
        do j = 1, size(grid)
            spec(j) = 10.0_wp**interp_lin(x, y, n, log10(grid(j)))
        end do
        if (flambda) spec = spec * grid**2
        where (grid > maxval(lt)) spec = SAFE_FLOOR
    end subroutine ref_spectrum_on_grid

    !> filters_zero_points (opensps_filters.f90:536-556) for one spectrum: the AB magnitude -2.5 log10(integral
    !> f_nu T / lambda d lambda) - 48.60 of f_nu on the grid in every band, 99 where it has no flux
    subroutine zero_points(grid, spec, bands, mag)
        real(WP), intent(in) :: grid(:), spec(:), bands(:, :)
        real(WP), intent(out) :: mag(:)
        integer :: i
        do i = 1, size(bands, 2)
            mag(i) = trapezoid(grid, spec * bands(:, i) / grid)
            if (mag(i) > SAFE_FLOOR) then
                mag(i) = -2.5_wp * log10(mag(i)) - 48.60_wp
            else
                mag(i) = 99.0_wp
            end if
        end do
    end subroutine zero_points

    !> sun_to_vega_system (src/spectra/opensps_context_filters.f90:185-198): the Sun's magnitudes in the Vega system,
    !> anchored on band vref (0: Vega is magnitude 0 in every band); bands without flux (99) stay 99
    subroutine sun_to_vega(magsun, magvega, vref)
        real(WP), intent(inout) :: magsun(:)
        real(WP), intent(in) :: magvega(:)
        integer, intent(in) :: vref
        real(WP) :: s0
        integer :: i
        s0 = 0.0_wp
        if (vref > 0) s0 = magsun(vref)
        do i = 1, size(magsun)
            if (magsun(i) == 99.0_wp .or. magvega(i) == 99.0_wp) cycle
            if (vref > 0) then
                magsun(i) = (magsun(i) - s0) - (magvega(i) - magvega(vref)) + s0
            else
                magsun(i) = magsun(i) - magvega(i)
            end if
        end do
    end subroutine sun_to_vega

end module opensps_host_filters
