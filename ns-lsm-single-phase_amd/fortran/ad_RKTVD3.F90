! Drop-in for ad_RKTVD3.F90:1-88: one RK-TVD3 step of the level-set advection
! with the WENO3 right-hand side (orderAD = 1) and bcond_lsm after every
! stage, on the GPU through liblsm_hip.  Uses the contravariant velocity of
! the preceding rhs_contra_j call.  lsm_advect_mode (module lsm_hip_f)
! selects the reference's arithmetic (default, bit-identical) or the
! one-division WENO3 weights.
subroutine ad_RKTVD3( phiAD_n1, phiAD_n , AdvectionNodes )
   use lsm_hip_f
   implicit none
   real (kind = rdf), dimension(il:iu,jl:ju,kl:ku), intent(out) :: phiAD_n1
   real (kind = rdf), dimension(il:iu,jl:ju,kl:ku), intent(in)  :: phiAD_n
   integer, dimension(il:iu,jl:ju,kl:ku), intent(in)            :: AdvectionNodes
   include 'lsm_shim_bind.inc'
   if ( orderAD /= 1 ) call lsm_f_unsupported( 'ad_RKTVD3', 'orderAD /= 1 (only the WENO3 right-hand side)' )
   call lsm_f_check( lsm_host_rk3( phiAD_n1, phiAD_n, AdvectionNodes, delti_lsm, lsm_advect_mode ), &
                     'ad_RKTVD3' )
end subroutine ad_RKTVD3
