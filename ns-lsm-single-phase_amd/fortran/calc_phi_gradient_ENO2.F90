! Drop-in for calc_phi_gradient_ENO2.F90:1-2025 (an external subroutine of
! the reference, called by mg_driver.F90:106-116): the ENO2 gradient of phi
! on interior, face and edge nodes (the vertices keep their values), boundary
! order OrderReinitialisationBoundaries, ENO boundary stencils when
! ENOBCReinitialisation, with the blanking regions' biased stencils and zero
! derivatives across their faces, on the GPU through liblsm_hip.
subroutine calc_phi_gradient_ENO2( il,iu          ,&
                                   jl,ju          ,&
                                   kl,ku          ,&
                                   igp, jgp, kgp  ,&
                                   dc, de, dz     ,&
                                   csi            ,&
                                   eta            ,&
                                   zet            ,&
                                   phi            ,&
                                   phi_gradient    &
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
   real (kind = rdf), dimension(il:iu,jl:ju,kl:ku) , intent(in) :: phi
   real (kind = rdf), dimension(1:3,il:iu,jl:ju,kl:ku) , intent(inout) :: phi_gradient
   integer(c_int) :: eno
   if ( nblk /= 0 .or. nblke /= 0 ) then
      call lsm_f_bind_blanking( nblk, nblke, li_blk_ia, li_blk_ib, li_blk_ja, li_blk_jb, &
                                li_blk_ka, li_blk_kb, blktype(:, :, myzone), consumers = .true. )
   else
      call lsm_f_check( lsm_host_consumer_blanking( 0_c_int32_t, [0_c_int32_t] ), 'lsm_host_consumer_blanking' )
   end if
   eno = 0
   if ( ENOBCReinitialisation ) eno = 1
   call lsm_f_check( lsm_host_phi_gradient_eno2(                                                &
                        int([il, iu, jl, ju, kl, ku, igp, jgp, kgp], c_int32_t),               &
                        lsm_f_faces([myback, myfront, myleft, myright, mydown, myup], mpi_proc_null), &
                        dc, de, dz, csi, eta, zet, phi,                                         &
                        int(OrderReinitialisationBoundaries, c_int), eno, phi_gradient ),       &
                     'calc_phi_gradient_ENO2' )
end subroutine calc_phi_gradient_ENO2
