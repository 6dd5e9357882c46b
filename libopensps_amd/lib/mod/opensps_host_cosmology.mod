﻿!mod$ v1 sum:4fdd01fda3834dfe
!need$ f5118467a58a718d n opensps_host_types
module opensps_host_cosmology
use opensps_host_types,only:wp
private::wp
integer(4),parameter,private::n_cosmo_grid=10000_4
real(8),parameter,private::c_light=2.99792458e18_8
contains
pure function luminosity_distance(h0,om0,ol0,z) result(dl)
real(8),intent(in)::h0
real(8),intent(in)::om0
real(8),intent(in)::ol0
real(8),intent(in)::z
real(8)::dl
end
subroutine cosmology_table(h0,om0,ol0,cz,cdl)
real(8),intent(in)::h0
real(8),intent(in)::om0
real(8),intent(in)::ol0
real(8),intent(out)::cz(1_8:500_8)
real(8),intent(out)::cdl(1_8:500_8)
end
end
