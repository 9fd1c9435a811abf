! Drop-in for ad_RKTVD3.F90:1-88: one RK-TVD3 step of the level-set advection
! with the WENO3 right-hand side (orderAD = 1) or the first-order upwind one
! (orderAD = 0, ad_RKTVD3.F90:57-58) and bcond_lsm after every stage, on the
! GPU through liblsm_hip.  Uses the contravariant velocity of the preceding
! rhs_contra_j call.  For WENO3, lsm_advect_mode (module lsm_hip_f) selects
! the reference's arithmetic (default, bit-identical) or the one-division
! weights.
subroutine ad_RKTVD3( phiAD_n1, phiAD_n , AdvectionNodes )
   use lsm_hip_f
   implicit none
   real (kind = rdf), dimension(il:iu,jl:ju,kl:ku), intent(out) :: phiAD_n1
   real (kind = rdf), dimension(il:iu,jl:ju,kl:ku), intent(in)  :: phiAD_n
   integer, dimension(il:iu,jl:ju,kl:ku), intent(in)            :: AdvectionNodes
   integer(c_int) :: mode
   include 'lsm_shim_bind.inc'
   if ( orderAD /= 0 .and. orderAD /= 1 ) call lsm_f_unsupported( 'ad_RKTVD3', 'orderAD not 0 or 1' )
   mode = lsm_advect_mode
   if ( orderAD == 0 ) mode = LSM_ADVECT_UPWIND1
   call lsm_f_check( lsm_host_rk3( phiAD_n1, phiAD_n, AdvectionNodes, delti_lsm, mode ), &
                     'ad_RKTVD3' )
end subroutine ad_RKTVD3
