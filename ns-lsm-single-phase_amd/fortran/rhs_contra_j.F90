! Drop-in for rhs_contra_j.F90:2-50 (contravariant velocity over Jacobian,
! ucn_j = (1/aj) grad(xi_d) . u) on the GPU through liblsm_hip.  The result
! is kept on the device for the advection calls that follow and is also
! written to the host's ucn_j, as the reference routine does.
subroutine rhs_contra_j
   use lsm_hip_f
   implicit none
   include 'lsm_shim_bind.inc'
   call lsm_f_check( lsm_host_contra_velocity( q, ucn_j ), 'rhs_contra_j' )
end subroutine rhs_contra_j
