! Drop-in for geometric_reinitialisation.F90:1-94 with GetTetrahedraList,
! GetIsosurfaceAndDistances and NeighboursPhiCorrection (the hybrid
! reinitialisation of levelsetmethod.F90:299-309): the signed distance to the
! marching-tetrahedra isosurface of phizero at the interface nodes and
! min(|phicorrected|, distance) within 4 nodes of them, on the GPU through
! liblsm_hip.  InterfaceNodesID receives 1 / -1 / 0 as GetTetrahedraList and
! GetIsosurfaceAndDistances set it.  The global mass correction
! (PiecewiseConstantFunction, OrthogonalProjection) and triangulations other
! than case 1 are not implemented by the library.
subroutine geometric_reinitialisation( phizero , phicorrected, TotalWaterVolume, InterfaceNodesID , TriangulationCase )
   use lsm_hip_f
   implicit none
   real (kind = rdf), target, dimension(il:iu,jl:ju,kl:ku), intent(in) :: phizero
   real (kind = rdf), dimension(il:iu,jl:ju,kl:ku), intent(inout)      :: phicorrected
   real (kind = rdf), intent(in)                                       :: TotalWaterVolume
   integer, dimension(il:iu,jl:ju,kl:ku), intent(inout)                :: InterfaceNodesID
   integer, intent(in), optional                                       :: TriangulationCase
   include 'lsm_shim_bind.inc'
   if ( GlobalMassCorrection ) &
      call lsm_f_unsupported( 'geometric_reinitialisation', 'GlobalMassCorrection = .true.' )
   if ( present( TriangulationCase ) ) then
      if ( TriangulationCase /= 1 ) &
         call lsm_f_unsupported( 'geometric_reinitialisation', 'TriangulationCase /= 1' )
   end if
   call lsm_f_check( lsm_host_reinit_geometric( phizero, phicorrected, InterfaceNodesID ), &
                     'geometric_reinitialisation' )
end subroutine geometric_reinitialisation
