! Drop-in for levelsetmethod.F90:1-372 as a whole (the argument list of
! levelsetmethod.F90:1-13, called by mg_driver.F90:563-581): the water volume
! before advection, rhs_contra_j, numiter_lsm ad_RKTVD3 steps (WENO3, or the
! upwind right-hand side for orderAD = 0), every RNfreq-th step the Sussman or
! (hybrid_reinitialisation) the geometric reinitialisation, and bcond_lsm, all
! on the GPU through liblsm_hip with phi resident on the device for the whole
! call: phi, phi_n and q go up once and phi comes back once (the callee
! drop-ins move phi across PCIe at every callee).  The volume is reduced over
! the ranks and written to GlobalVolume.txt as levelsetmethod.F90:187-196
! does.  One block per rank with every face physical, no blanking, no global
! mass correction of the hybrid path.
subroutine levelsetmethod( il,iu                   , &
                           jl,ju                   , &
                           kl,ku                   , &
                           igp, jgp, kgp           , &
                           dc, de, dz              , &
                           q                       , &
                           csi                     , &
                           eta                     , &
                           zet                     , &
                           aj                      , &
                           phi, phi_n              , &
                           x, y, z                 , &
                           iteraciontiempo   )
   use global_param
   use global_app
   use global_mpi
   use global_lsm, only : RNitermax, RNtimestep, RNfreq, btype_lsm, numiter_lsm, delti_lsm, orderAD,      &
                          call_levelsetmethod, call_reinitialisation, hybrid_reinitialisation,          &
                          OrderReinitialisationBoundaries, ENOBCReinitialisation, GlobalMassCorrection, &
                          BigPhi, epslsm, epsReinitialisation
   use lsm_hip_f
   implicit none
   integer, intent(in) :: il,iu,jl,ju,kl,ku
   integer, intent(in) :: igp,jgp,kgp
   real (kind = rdf), intent(in) :: dc,de,dz
   real (kind = rdf), dimension(1:4,il:iu,jl:ju,kl:ku), intent(inout) :: q
   real (kind = rdf), dimension(1:3,il:iu,jl:ju,kl:ku), intent(in) :: csi , eta, zet
   real (kind = rdf), dimension(il:iu,jl:ju,kl:ku), intent(in) :: aj
   real (kind = rdf), dimension(il:iu,jl:ju,kl:ku), intent(in) :: x, y, z
   real (kind = rdf), dimension(il:iu,jl:ju,kl:ku), intent(in)    :: phi_n
   real (kind = rdf), dimension(il:iu,jl:ju,kl:ku), intent(inout) :: phi
   integer :: iteraciontiempo
   type(lsm_reinit_params) :: p
   real(c_double) :: vol(3)
   real (kind = rdf) :: TotVol_PreAdv_Local, TotVol_PreAdv
   integer(c_int) :: mode, cl, cr
   include 'lsm_shim_bind.inc'
   if ( orderAD /= 0 .and. orderAD /= 1 ) call lsm_f_unsupported( 'levelsetmethod', 'orderAD not 0 or 1' )
   if ( call_reinitialisation .and. hybrid_reinitialisation .and. GlobalMassCorrection ) &
      call lsm_f_unsupported( 'levelsetmethod', 'GlobalMassCorrection = .true. (hybrid path)' )
   mode = lsm_advect_mode
   if ( orderAD == 0 ) mode = LSM_ADVECT_UPWIND1
   p%epslsm      = epslsm
   p%eps_reinit  = epsReinitialisation
   p%rn_timestep = RNtimestep
   p%rn_itermax  = RNitermax
   p%order_bc    = OrderReinitialisationBoundaries
   p%eno_bc      = 0
   if ( ENOBCReinitialisation ) p%eno_bc = 1
   p%m_sphi0     = 1
   p%m_ti        = 1
   p%hybrid      = 0
   if ( hybrid_reinitialisation ) p%hybrid = 1
   cl = 0
   if ( call_levelsetmethod ) cl = 1
   cr = 0
   if ( call_reinitialisation ) cr = 1
   call lsm_f_check( lsm_host_levelset_step( phi, phi_n, q, p, int(numiter_lsm, c_int), delti_lsm, cl, cr, &
                                             int(RNfreq, c_int), int(ntime, c_int), mode, vol ),           &
                     'levelsetmethod' )
   TotVol_PreAdv_Local = vol(1) + vol(2) + vol(3)
   call mpi_allreduce( TotVol_PreAdv_Local, TotVol_PreAdv , &
                       1 , MPI_REAL_TYPE , mpi_sum , mpi_comm_world, ierr )
   if ( myid == root ) call WriteTotalWaterVolume( TotVol_PreAdv )

   contains

   include 'WriteTotalWaterVolume.F90'

end subroutine levelsetmethod
