!> The cosmology table of a context (the reference's update_cosmology_table and get_luminosity_distance,
!> src/physics/opensps_cosmology.f90:249-300): 500 redshifts on a grid of scale factors and their luminosity
!> distances in parsecs, by the trapezoid rule over 10 000 steps in z. The distance modulus of every magnitude
!> at a redshift reads it (distance_term, src/spectra/opensps_batch_photometry.f90:784-794).
module opensps_host_cosmology
    use opensps_host_types, only: WP
    implicit none
    private
    integer, parameter :: N_COSMO_GRID = 10000
    real(WP), parameter :: C_LIGHT = 2.99792458e18_wp
    public :: cosmology_table, luminosity_distance

contains

    pure real(WP) function luminosity_distance(h0, om0, ol0, z) result(dl)
        real(WP), intent(in) :: h0, om0, ol0, z
        real(WP) :: hd, dz, s, prev, now, zc
        integer :: i
        dl = 0.0_wp
        if (z <= 0.0_wp) return
        hd = ((C_LIGHT * 1.0e-13_wp) / h0) * 1.0e6_wp
        dz = z / real(N_COSMO_GRID - 1, WP)
        prev = 1.0_wp / sqrt(om0 + ol0)
        s = 0.0_wp
        do i = 2, N_COSMO_GRID
            zc = real(i - 1, WP) * dz
            now = 1.0_wp / sqrt(om0 * (1.0_wp + zc)**3 + ol0)
            s = s + 0.5_wp * dz * (prev + now)
            prev = now
        end do
        dl = (1.0_wp + z) * hd * s
    end function luminosity_distance

    subroutine cosmology_table(h0, om0, ol0, cz, cdl)
        real(WP), intent(in) :: h0, om0, ol0
        real(WP), intent(out) :: cz(500), cdl(500)
        integer :: i
        real(WP) :: a
        do i = 1, 500
            a = real(i - 1, WP) / 499.0_wp * (1.0_wp - 1.0_wp / 1001.0_wp) + 1.0_wp / 1001.0_wp
            cz(i) = (1.0_wp / a) - 1.0_wp
            cdl(i) = luminosity_distance(h0, om0, ol0, cz(i))
        end do
    end subroutine cosmology_table

end module opensps_host_cosmology
