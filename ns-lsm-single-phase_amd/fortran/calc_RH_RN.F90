! Drop-in for calc_RH_RN.F90:1-244: right-hand side S(phi0)(|grad d| - 1)
! of the Godunov reinitialisation, on the GPU through liblsm_hip.
subroutine calc_RH_RN(sgndf_RN,signo_RN,rightH_RN)
   use lsm_hip_f
   implicit none
   real (kind = rdf), dimension(il:iu,jl:ju,kl:ku), intent(in)  :: sgndf_RN
   real (kind = rdf), dimension(il:iu,jl:ju,kl:ku), intent(in)  :: signo_RN
   real (kind = rdf), dimension(il:iu,jl:ju,kl:ku), intent(out) :: rightH_RN
   include 'lsm_shim_bind.inc'
   call lsm_f_check( lsm_host_rh_rn( sgndf_RN, signo_RN, rightH_RN ), 'calc_RH_RN' )
end subroutine calc_RH_RN
