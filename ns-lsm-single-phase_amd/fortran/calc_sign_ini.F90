! Drop-in for calc_sign_ini.F90:1-286: smoothed sign S(phi0) (M_Sphi0
! selects the form) and the Godunov norm of grad phi0, on the GPU through
! liblsm_hip.
subroutine calc_sign_ini(sgndf_RN,signo_RN,normagrad_RN,epslsm,M_Sphi0)
   use lsm_hip_f
   implicit none
   real (kind = rdf), dimension(il:iu,jl:ju,kl:ku), intent(in)  :: sgndf_RN
   real (kind = rdf), dimension(il:iu,jl:ju,kl:ku), intent(out) :: signo_RN, normagrad_RN
   real (kind = rdf), intent(in) :: epslsm
   integer, intent(in) :: M_Sphi0
   include 'lsm_shim_bind.inc'
   call lsm_f_check( lsm_host_sign_ini( sgndf_RN, epslsm, int(M_Sphi0, c_int), signo_RN, normagrad_RN ), &
                     'calc_sign_ini' )
end subroutine calc_sign_ini
