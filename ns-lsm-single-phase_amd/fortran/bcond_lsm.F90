! Drop-in for bcond_lsm.F90:2-568: boundary extrapolation of the level-set
! function on the physical faces (btype_lsm, BigPhi switch), on the GPU
! through liblsm_hip.
subroutine bcond_lsm (phi)
   use lsm_hip_f
   implicit none
   real (kind = rdf), dimension(il:iu,jl:ju,kl:ku), intent(inout) :: phi
   include 'lsm_shim_bind.inc'
   call lsm_f_check( lsm_host_bcond( phi ), 'bcond_lsm' )
end subroutine bcond_lsm
