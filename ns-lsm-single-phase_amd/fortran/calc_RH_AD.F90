! Drop-in for calc_RH_AD.F90:1-273 (orderAD = 0): first-order upwind
! right-hand side of the level-set advection at every node with
! AdvectionNodes > 0 (one-sided differences on the physical faces), on the
! GPU through liblsm_hip.  Other entries of rightH_AD keep their values.  Uses
! the contravariant velocity of the preceding rhs_contra_j call.
subroutine calc_RH_AD( phi_AD , rightH_AD , AdvectionNodes )
   use lsm_hip_f
   implicit none
   real (kind = rdf), dimension(il:iu,jl:ju,kl:ku), intent(in)  :: phi_AD
   real (kind = rdf), dimension(il:iu,jl:ju,kl:ku), intent(out) :: rightH_AD
   integer, dimension(il:iu,jl:ju,kl:ku), intent(in)            :: AdvectionNodes
   include 'lsm_shim_bind.inc'
   call lsm_f_check( lsm_host_upwind_rhs( phi_AD, rightH_AD, AdvectionNodes ), 'calc_RH_AD' )
end subroutine calc_RH_AD
