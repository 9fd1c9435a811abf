! Drop-in for calc_RH_AD_WENO3.F90:1-407: WENO3 right-hand side of the
! level-set advection in contravariant form at the interior nodes with
! AdvectionNodes > 0, on the GPU through liblsm_hip.  Other entries of
! rightH_AD keep their values.  Uses the contravariant velocity of the
! preceding rhs_contra_j call.
subroutine calc_RH_AD_WENO3( phi_AD , rightH_AD , AdvectionNodes )
   use lsm_hip_f
   implicit none
   real (kind = rdf), dimension(il:iu,jl:ju,kl:ku), intent(in)  :: phi_AD
   real (kind = rdf), dimension(il:iu,jl:ju,kl:ku), intent(out) :: rightH_AD
   integer, dimension(il:iu,jl:ju,kl:ku), intent(in)            :: AdvectionNodes
   include 'lsm_shim_bind.inc'
   call lsm_f_check( lsm_host_weno3_rhs( phi_AD, rightH_AD, AdvectionNodes, lsm_advect_mode ), &
                     'calc_RH_AD_WENO3' )
end subroutine calc_RH_AD_WENO3
