! Drop-in for calc_h.F90:1-394 with find_z_free_surface.F90:1-62 (an
! external subroutine of the reference, called by mg_driver.F90:127-135 and
! init_LSM.F90:458): h = z_fs - z in every column of the block, z_fs at the
! first sign change of phi from the bottom, h = 0 (and the reference's
! message) in a column without one, then h extrapolated onto the faces of
! the blanking regions (calc_h.F90:195-345), on the GPU through liblsm_hip.
subroutine calc_h( il,iu          ,&
                   jl,ju          ,&
                   kl,ku          ,&
                   igp, jgp, kgp  ,&
                   z              ,&
                   phi            ,&
                   h               &
                  )
   use precision
   use global_mpi
   use global_app, only : nblk, nblke, blktype, li_blk_ia, li_blk_ib, li_blk_ja, li_blk_jb, &
                          li_blk_ka, li_blk_kb
   use lsm_hip_f
   implicit none
   integer, intent(in) :: il,iu,jl,ju,kl,ku
   integer, intent(in) :: igp, jgp, kgp
   real (kind = rdf), dimension(il:iu,jl:ju,kl:ku) , intent(in) :: z
   real (kind = rdf), dimension(il:iu,jl:ju,kl:ku) , intent(in) :: phi
   real (kind = rdf), dimension(il:iu,jl:ju,kl:ku) , intent(inout) :: h
   integer(c_int32_t) :: missing
   integer :: m
   if ( nblk /= 0 .or. nblke /= 0 ) then
      call lsm_f_bind_blanking( nblk, nblke, li_blk_ia, li_blk_ib, li_blk_ja, li_blk_jb, &
                                li_blk_ka, li_blk_kb, blktype(:, :, myzone), consumers = .true. )
   else
      call lsm_f_check( lsm_host_consumer_blanking( 0_c_int32_t, [0_c_int32_t] ), 'lsm_host_consumer_blanking' )
   end if
   missing = 0
   call lsm_f_check( lsm_host_calc_h( int([il, iu, jl, ju, kl, ku, igp, jgp, kgp], c_int32_t),   &
                        lsm_f_faces([myback, myfront, myleft, myright, mydown, myup], mpi_proc_null), &
                        z, phi, h, missing ), 'calc_h' )
   do m = 1, missing
      print *, 'free surface not found'
   end do
end subroutine calc_h
