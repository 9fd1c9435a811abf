! Drop-in for calc_h_gradient_ENO2.F90:1-2018 (an external subroutine of the
! reference, called by mg_driver.F90:140-150): the x and y components of the
! ENO2 gradient of h, stencils, blanking and vertices as in
! calc_phi_gradient_ENO2, on the GPU through liblsm_hip.
subroutine calc_h_gradient_ENO2( il,iu          ,&
                                 jl,ju          ,&
                                 kl,ku          ,&
                                 igp, jgp, kgp  ,&
                                 dc, de, dz     ,&
                                 csi            ,&
                                 eta            ,&
                                 zet            ,&
                                 h              ,&
                                 h_gradient      &
                                )
   use precision
   use global_mpi
   use global_app, only : nblk, nblke, blktype, li_blk_ia, li_blk_ib, li_blk_ja, li_blk_jb, &
                          li_blk_ka, li_blk_kb
   use global_lsm, only : OrderReinitialisationBoundaries , ENOBCReinitialisation
   use lsm_hip_f
   implicit none
   integer, intent(in) :: il,iu,jl,ju,kl,ku
   integer, intent(in) :: igp, jgp, kgp
   real (kind = rdf), intent(in) :: dc,de,dz
   real (kind = rdf), dimension(1:3,il:iu,jl:ju,kl:ku), intent(in) :: csi , eta, zet
   real (kind = rdf), dimension(il:iu,jl:ju,kl:ku) , intent(in) :: h
   real (kind = rdf), dimension(1:2,il:iu,jl:ju,kl:ku) , intent(inout) :: h_gradient
   integer(c_int) :: eno
   if ( nblk /= 0 .or. nblke /= 0 ) then
      call lsm_f_bind_blanking( nblk, nblke, li_blk_ia, li_blk_ib, li_blk_ja, li_blk_jb, &
                                li_blk_ka, li_blk_kb, blktype(:, :, myzone), consumers = .true. )
   else
      call lsm_f_check( lsm_host_consumer_blanking( 0_c_int32_t, [0_c_int32_t] ), 'lsm_host_consumer_blanking' )
   end if
   eno = 0
   if ( ENOBCReinitialisation ) eno = 1
   call lsm_f_check( lsm_host_h_gradient_eno2(                                                  &
                        int([il, iu, jl, ju, kl, ku, igp, jgp, kgp], c_int32_t),               &
                        lsm_f_faces([myback, myfront, myleft, myright, mydown, myup], mpi_proc_null), &
                        dc, de, dz, csi, eta, zet, h,                                           &
                        int(OrderReinitialisationBoundaries, c_int), eno, h_gradient ),         &
                     'calc_h_gradient_ENO2' )
end subroutine calc_h_gradient_ENO2
