! Drop-in for TotalWaterVolume.F90:1-297: water volume of the block by
! marching tetrahedra (cut-cell tetrahedra, second-pass tetrahedra, bulk
! cells) and, when present, the AdvectionNodes mask (the 3x3x3 node
! neighbourhood of every cut cell), on the GPU through liblsm_hip.
! TriangulationCase 1 (the value levelsetmethod passes) is implemented;
! WriteGlobalVolume is accepted and unused, as in the reference.
subroutine TotalWaterVolume( phizero           , &
                             VolumeKTetrahedra , VolumeSecondTetrahedra , VolumeBulkCells , &
                             AdvectionNodes    , &
                             WriteGlobalVolume , &
                             TriangulationCase                                                 )
   use lsm_hip_f
   implicit none
   real(kind = rdf), target, dimension(il:iu,jl:ju,kl:ku), intent(in) :: phizero
   logical, optional :: WriteGlobalVolume
   integer, intent(in), optional :: TriangulationCase
   real(kind = rdf), intent(out) :: VolumeKTetrahedra, VolumeSecondTetrahedra, VolumeBulkCells
   integer, dimension(il:iu,jl:ju,kl:ku), intent(out), optional :: AdvectionNodes
   real(c_double) :: vol(3)
   include 'lsm_shim_bind.inc'
   if ( present( TriangulationCase ) ) then
      if ( TriangulationCase /= 1 ) call lsm_f_unsupported( 'TotalWaterVolume', 'TriangulationCase /= 1' )
   end if
   call lsm_f_check( lsm_host_water_volume( phizero, vol, AdvectionNodes ), 'TotalWaterVolume' )
   VolumeKTetrahedra      = vol(1)
   VolumeSecondTetrahedra = vol(2)
   VolumeBulkCells        = vol(3)
end subroutine TotalWaterVolume
