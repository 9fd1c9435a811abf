! Drop-in for reinitialisation_test2.F90:1-218: Sussman pseudo-time
! reinitialisation with ENO2 gradients, SSP-RK3 pseudo-time stepping
! (reinitialisation_test2.F90:137-184) and the
! local mass correction, ended by the |grad d| convergence test or after
! RNitermax iterations; on the GPU through liblsm_hip.  The hybrid masking
! by InterfaceNodesID (reinitialisation_test2.F90:189-208) is done here.
subroutine reinitialisation_test2 ( PhiZero, PhiReinitialised, InterfaceNodesID )
   use lsm_hip_f
   use global_lsm, only : epslsm, epsReinitialisation, RNitermax, RNtimestep,         &
                          OrderReinitialisationBoundaries, ENOBCReinitialisation,    &
                          hybrid_reinitialisation
   implicit none
   real (kind = rdf), dimension(il:iu,jl:ju,kl:ku), intent(in)    :: PhiZero
   real (kind = rdf), dimension(il:iu,jl:ju,kl:ku), intent(inout) :: PhiReinitialised
   integer, dimension(il:iu,jl:ju,kl:ku), intent(in)              :: InterfaceNodesID
   real (kind = rdf), dimension(:,:,:), allocatable :: dn1
   type(lsm_reinit_params) :: p
   integer(c_int32_t) :: iters
   include 'lsm_shim_bind.inc'
   p%epslsm      = epslsm
   p%eps_reinit  = epsReinitialisation
   p%rn_timestep = RNtimestep
   p%rn_itermax  = RNitermax
   p%order_bc    = OrderReinitialisationBoundaries
   p%eno_bc      = 0
   if ( ENOBCReinitialisation ) p%eno_bc = 1
   p%m_sphi0     = 1
   p%m_ti        = 1
   allocate( dn1(il:iu,jl:ju,kl:ku) )
   call lsm_f_check( lsm_host_reinit_eno2( PhiZero, dn1, p, iters ), 'reinitialisation_test2' )
   if ( hybrid_reinitialisation ) then
      where( InterfaceNodesID == 0 ) PhiReinitialised = dn1
   else
      PhiReinitialised = dn1
   end if
   deallocate( dn1 )
end subroutine reinitialisation_test2
