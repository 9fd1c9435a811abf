! Drop-in for calc_NormGrad_d.F90:1-2448: ENO2 Godunov norm of the gradient
! of phi_RN (boundary order OrderReinitialisationBoundaries, ENO boundary
! stencils when ENOBCReinitialisation), on the GPU through liblsm_hip.
subroutine calc_NormGrad_d( phi_RN , phizero , NormGrad_RN )
   use lsm_hip_f
   use global_lsm, only : OrderReinitialisationBoundaries, ENOBCReinitialisation
   implicit none
   real (kind = rdf), dimension(il:iu,jl:ju,kl:ku), intent(in)  :: phi_RN
   real (kind = rdf), dimension(il:iu,jl:ju,kl:ku), intent(in)  :: phizero
   real (kind = rdf), dimension(il:iu,jl:ju,kl:ku), intent(out) :: NormGrad_RN
   integer(c_int) :: eno
   include 'lsm_shim_bind.inc'
   eno = 0
   if ( ENOBCReinitialisation ) eno = 1
   call lsm_f_check( lsm_host_normgrad( phi_RN, phizero, int(OrderReinitialisationBoundaries, c_int), &
                                        eno, NormGrad_RN ), 'calc_NormGrad_d' )
end subroutine calc_NormGrad_d
