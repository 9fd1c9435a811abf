! Drop-in for reinitialization_benchmark.F90:1-228: Godunov pseudo-time
! reinitialisation (first-order gradients, M_Grad = 0) with RK3 (M_TI = 1)
! or Euler stepping and the Sussman-Fatemi correction, RNitermax
! iterations, on the GPU through liblsm_hip.  The hybrid masking by
! InterfaceNodesID (reinitialization_benchmark.F90:192-220) is done here.
! M_eps and M_lambda are accepted and unused, as in the reference.
subroutine reinitialization_benchmark(  phizero, phicorregido,                  &
                                        M_Sphi0, M_TI, M_Grad, M_eps, M_lambda, &
                                        InterfaceNodesID                          )
   use lsm_hip_f
   use global_lsm, only : epslsm, RNitermax, RNtimestep, sussman_correction_method, &
                          hybrid_reinitialisation
   use global_obstacle, only : is_obstacle
   implicit none
   real (kind = rdf), dimension(il:iu,jl:ju,kl:ku), intent(in)    :: phizero
   real (kind = rdf), dimension(il:iu,jl:ju,kl:ku), intent(inout) :: phicorregido
   integer, intent(in) :: M_Sphi0, M_TI, M_Grad, M_eps, M_lambda
   integer, dimension(il:iu,jl:ju,kl:ku), intent(in)              :: InterfaceNodesID
   real (kind = rdf), dimension(:,:,:), allocatable :: sgndf
   type(lsm_reinit_params) :: p
   include 'lsm_shim_bind.inc'
   if ( M_Grad /= 0 ) call lsm_f_unsupported( 'reinitialization_benchmark', 'M_Grad /= 0' )
   if ( is_obstacle /= 0 ) call lsm_f_unsupported( 'reinitialization_benchmark', 'obstacles' )
   if ( sussman_correction_method /= 0 .and. sussman_correction_method /= 1 ) &
      call lsm_f_unsupported( 'reinitialization_benchmark', 'sussman_correction_method not 0 or 1' )
   p%epslsm      = epslsm
   p%eps_reinit  = 0.0_rdf
   p%rn_timestep = RNtimestep
   p%rn_itermax  = RNitermax
   p%order_bc    = 1
   p%eno_bc      = 0
   p%m_sphi0     = M_Sphi0
   p%m_ti        = M_TI
   p%lambda_method = sussman_correction_method
   allocate( sgndf(il:iu,jl:ju,kl:ku) )
   call lsm_f_check( lsm_host_reinit_godunov( phizero, sgndf, p ), 'reinitialization_benchmark' )
   if ( hybrid_reinitialisation ) then
      where( InterfaceNodesID == 0 ) phicorregido = sgndf
   else
      phicorregido = sgndf
   end if
   deallocate( sgndf )
end subroutine reinitialization_benchmark
