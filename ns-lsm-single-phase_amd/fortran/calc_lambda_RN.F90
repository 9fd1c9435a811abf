! Drop-in for calc_lambda_RN.F90:1-122: the Sussman-Fatemi mass-correction
! multiplier lambda, pointwise (sussman_correction_method = 0,
! calc_lambda_RN.F90:42-64) or with the numerator and denominator filtered by
! testfilter (sussman_correction_method = 1, calc_lambda_RN.F90:65-120), on the
! GPU through liblsm_hip.
subroutine calc_lambda_RN(sgndf_zero,normagrad_zero,sgndf_RN,lambda_RN,timeStep)
   use lsm_hip_f
   use global_lsm, only : sussman_correction_method
   implicit none
   real (kind = rdf), dimension(il:iu,jl:ju,kl:ku), intent(in)  :: sgndf_zero, normagrad_zero
   real (kind = rdf), dimension(il:iu,jl:ju,kl:ku), intent(in)  :: sgndf_RN
   real (kind = rdf), intent(in) :: timeStep
   real (kind = rdf), dimension(il:iu,jl:ju,kl:ku), intent(out) :: lambda_RN
   include 'lsm_shim_bind.inc'
   if ( sussman_correction_method == 0 ) then
      call lsm_f_check( lsm_host_lambda_rn( sgndf_zero, normagrad_zero, sgndf_RN, timeStep, epslsm, lambda_RN ), &
                        'calc_lambda_RN' )
   else if ( sussman_correction_method == 1 ) then
      call lsm_f_check( lsm_host_lambda_rn_simpson( sgndf_zero, normagrad_zero, sgndf_RN, timeStep, epslsm, &
                                                    lambda_RN ), 'calc_lambda_RN' )
   else
      call lsm_f_unsupported( 'calc_lambda_RN', 'sussman_correction_method not 0 or 1' )
   end if
end subroutine calc_lambda_RN
