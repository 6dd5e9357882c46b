﻿!mod$ v1 sum:f5118467a58a718d
!need$ 0bde2ac47243ead2 i iso_c_binding
module opensps_host_types
use,intrinsic::iso_c_binding,only:c_double
use,intrinsic::iso_c_binding,only:c_int
private::c_double
private::c_int
integer(4),parameter::wp=8_4
real(8),parameter::safe_floor=2.225073858507201383090232717332404064219215980462331830553327416887204434813918195854283159012511020564067339731035811005152434161553460108856012385377718821130777993532002330479610147442583636071921565046942503734208375250806650616658158948720491179968591639648500635908770118304874799780887753749949451580451605050915399856582470818645113537935804992115981085766051992433352114352390148795699609591288891602992641511063466313393663477586513029371762047325631781485664350872122828637642044846811407613911477062801689853244110024161447421618567166150540154285084716752901903161322778896729707373123334086988983175067838846926092773977972858659654941091369095406136467568702398678315290680984617210924625396728515625e-308_8
intrinsic::tiny
private::tiny
integer(4),parameter::npar=28_4
character(16_4,1),parameter::param_names(1_8:28_8)=[CHARACTER(KIND=1,LEN=16)::"tage            ","tau             ","const           ","fburst          ","tburst          ","sf_start        ","sf_trunc        ","sf_slope        ","sfh             ","logzsol         ","dust1           ","dust2           ","dust3           ","dust_index      ","dust1_index     ","dust_tesc       ","frac_obrun      ","frac_nodust     ","zred            ","igm_factor      ","sigma_smooth    ","min_wave_smooth ","max_wave_smooth ","imf1            ","imf2            ","imf3            ","imf_upper_limit ","imf_lower_limit "]
real(8),parameter::param_defaults(1_8:28_8)=[REAL(8)::0._8,1._8,0._8,0._8,1.1e1_8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,-6.999999999999999555910790149937383830547332763671875e-1_8,-1._8,7._8,0._8,0._8,0._8,1._8,0._8,1.e3_8,1.e4_8,1.3000000000000000444089209850062616169452667236328125_8,2.29999999999999982236431605997495353221893310546875_8,2.29999999999999982236431605997495353221893310546875_8,1.2e2_8,8.000000000000000166533453693773481063544750213623046875e-2_8]
integer(4),parameter::p_tage=1_4
integer(4),parameter::p_sfh=9_4
integer(4),parameter::p_logzsol=10_4
integer(4),parameter::p_dust1=11_4
integer(4),parameter::p_dust2=12_4
integer(4),parameter::p_dust3=13_4
integer(4),parameter::p_frac_obrun=17_4
integer(4),parameter::p_zred=19_4
integer(4),parameter::p_igm_factor=20_4
integer(4),parameter::p_sigma_smooth=21_4
integer(4),parameter::p_tau=2_4
integer(4),parameter::p_const=3_4
integer(4),parameter::p_fburst=4_4
integer(4),parameter::p_tburst=5_4
integer(4),parameter::p_sf_start=6_4
integer(4),parameter::p_sf_trunc=7_4
integer(4),parameter::p_imf1=24_4
character(16_4,1),parameter::ssp_names(1_8:14_8)=[CHARACTER(KIND=1,LEN=16)::"imf_type        ","zmet            ","pagb            ","dell            ","delt            ","fbhb            ","sbss            ","redgb           ","agb             ","masscut         ","vdmc            ","mdave           ","evtype          ","fcstar          "]
integer(4),parameter::nbgrad=44_4
character(16_4,1),parameter::grad_names(1_8:44_8)=[CHARACTER(KIND=1,LEN=16)::"fagn            ","dust1           ","dust2           ","dust3           ","dust_index      ","dust1_index     ","frac_obrun      ","frac_nodust     ","mwr             ","uvb             ","duste_qpah      ","duste_umin      ","duste_gamma     ","gas_logz        ","gas_logu        ","logzsol         ","tage            ","tau             ","const           ","fburst          ","tburst          ","sf_start        ","zred            ","sigma_smooth    ","sf_trunc        ","sf_slope        ","imf1            ","imf2            ","imf3            ","vdmc            ","mdave           ","agn_tau         ","igm_factor      ","gas_logn        ","gas_logno       ","gas_logco       ","afe             ","fbhb            ","sbss            ","agb             ","pagb            ","redgb           ","delt            ","dell            "]
character(512_4,1)::last_error
integer(4)::last_status
contains
pure function lower(s) result(r)
character(*,1),intent(in)::s
character(int(int(s%len,kind=8),kind=4),1)::r
end
pure function param_slot(name) result(k)
character(*,1),intent(in)::name
integer(4)::k
end
subroutine set_error(status,msg)
integer(4),intent(in)::status
character(*,1),intent(in)::msg
end
end
