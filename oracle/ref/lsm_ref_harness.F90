! ---------------------------------------------------------------------------
! Test-infrastructure harness around the reference Fortran level-set routines.
!
! The reference keeps its hot-path routines as internal procedures of
! `levelsetmethod` (levelsetmethod.F90:331-370), so they have no external
! linkage.  This file provides a host procedure, `ref_driver`, that declares
! the same host-associated names the routines read (extents, metrics, loop
! bounds, RK counter, ucn_j, ...) and pulls the UNMODIFIED reference sources in
! with `include`, exactly the way levelsetmethod.F90 does.  The sources are
! found through the compiler's -I path (the reference tree), nothing from the
! reference is copied here.  Mode 12 calls the compiled `levelsetmethod`
! itself.
!
! The harness is only used to write golden vectors and to time the reference
! on the host; it is never linked into or called by the product library.
! Started on several MPI ranks, every rank runs the same single-block case
! independently (all faces are physical boundaries on every rank).
!
! Mode 13 calls the compiled mg_metrics (metrics of the grid x, y, z).
! Modes 14-16 call the compiled caller-side consumers of phi, external
! subroutines of the reference (mg_driver.F90:106-150): calc_phi_gradient_ENO2,
! calc_h and calc_h_gradient_ENO2.
!
! I/O (stream, little endian):
!   input : int32 ip(64), real64 rp(32), then ni*nj*nk-sized arrays
!           x, y, z, csi(3,:), eta(3,:), zet(3,:), aj, q(4,:), phi, phi0,
!           aux (real64) and mask (int32), all i-fastest.
!   output: int32 nout, then per output: int32 kind (0 real array,
!           1 int array), int32 n, data.
! ---------------------------------------------------------------------------

program lsm_ref_harness

   use precision
   use global,          only : le_idx, le_idx_a, le_idx_b,                 &
                               li_ia, li_ib, li_ja, li_jb, li_ka, li_kb
   use global_mpi
   use global_lsm
   use global_app,      only : nblk, nblke, blktype,                               &
                               li_blk_ia, li_blk_ib, li_blk_ja, li_blk_jb, li_blk_ka, li_blk_kb, &
                               le_blk_ia, le_blk_ib, le_blk_ja, le_blk_jb, le_blk_ka, le_blk_kb
   use global_param,    only : ntime, delti
   use global_obstacle, only : is_obstacle, act_obstacle_ad, act_obstacle_rn
   use global_debug,    only : OUTPUTdir

   implicit none

   integer(4)        :: ip(64)
   real(kind=rdf)    :: rp(32)
   integer           :: ni, nj, nk, ii, jj, kk, l, nargs
   character(len=512):: fin, fout
   real(kind=rdf), allocatable :: x(:,:,:), y(:,:,:), z(:,:,:), aj(:,:,:)
   real(kind=rdf), allocatable :: csi(:,:,:,:), eta(:,:,:,:), zet(:,:,:,:)
   real(kind=rdf), allocatable :: q(:,:,:,:)
   real(kind=rdf), allocatable :: fphi(:,:,:), fphi0(:,:,:), faux(:,:,:)
   integer(4),     allocatable :: mask(:,:,:)
   real(kind=rdf) :: t0, t1

   call mpi_init(ierr)
   call mpi_comm_rank(mpi_comm_world, myid, ierr)
   call mpi_comm_size(mpi_comm_world, nproc, ierr)

   nargs = command_argument_count()
   if (nargs < 2) then
      print *, 'usage: lsm_ref_harness <input.bin> <output.bin>'
      call mpi_abort(mpi_comm_world, 2, ierr)
   end if
   call get_command_argument(1, fin)
   call get_command_argument(2, fout)
   ! several ranks (the host-core baseline) each run the same block; ranks
   ! other than 0 write their outputs next to rank 0's
   if (myid /= 0) write(fout, '(a,a,i0)') trim(fout), '.', myid

   open(21, file=trim(fin), access='stream', form='unformatted', status='old')
   read(21) ip
   read(21) rp
   ni = ip(2); nj = ip(3); nk = ip(4)

   allocate(x(ni,nj,nk), y(ni,nj,nk), z(ni,nj,nk), aj(ni,nj,nk))
   allocate(csi(3,ni,nj,nk), eta(3,ni,nj,nk), zet(3,ni,nj,nk), q(4,ni,nj,nk))
   allocate(fphi(ni,nj,nk), fphi0(ni,nj,nk), faux(ni,nj,nk), mask(ni,nj,nk))
   read(21) x
   read(21) y
   read(21) z
   read(21) csi
   read(21) eta
   read(21) zet
   read(21) aj
   read(21) q
   read(21) fphi
   read(21) fphi0
   read(21) faux
   read(21) mask
   close(21)

   ! one rank, one block: every face is a physical boundary
   myback  = mpi_proc_null ; myfront = mpi_proc_null
   myleft  = mpi_proc_null ; myright = mpi_proc_null
   mydown  = mpi_proc_null ; myup    = mpi_proc_null
   myzone  = 1
   MPI_REAL_TYPE = MPI_DOUBLE_PRECISION
   nblk  = 0
   nblke = 0
   ! a blanking region (init.F90:1216-1567 with one rank: the interior and the
   ! extended boxes coincide; nba_max = 1, global.F90:21): ip(31) = 1, the box
   ! li_blk ia, ib, ja, jb, ka, kb in ip(32:37), blktype(1:6) in ip(38:43)
   allocate(blktype(6,1,1))
   blktype = 0
   if (ip(31) /= 0) then
      nblk  = 1
      nblke = 1
      allocate(li_blk_ia(1,1), li_blk_ib(1,1), li_blk_ja(1,1), li_blk_jb(1,1), li_blk_ka(1,1), li_blk_kb(1,1))
      allocate(le_blk_ia(1,1), le_blk_ib(1,1), le_blk_ja(1,1), le_blk_jb(1,1), le_blk_ka(1,1), le_blk_kb(1,1))
      li_blk_ia(1,1) = ip(32) ; li_blk_ib(1,1) = ip(33)
      li_blk_ja(1,1) = ip(34) ; li_blk_jb(1,1) = ip(35)
      li_blk_ka(1,1) = ip(36) ; li_blk_kb(1,1) = ip(37)
      le_blk_ia = li_blk_ia ; le_blk_ib = li_blk_ib
      le_blk_ja = li_blk_ja ; le_blk_jb = li_blk_jb
      le_blk_ka = li_blk_ka ; le_blk_kb = li_blk_kb
      blktype(1:6,1,1) = ip(38:43)
   end if
   is_obstacle = 0
   ! the 1-D line communicators of one rank (calc_h reads comm1d_csi(1)%myrank)
   allocate(comm1d_csi(1), comm1d_eta(1), comm1d_zet(1))
   comm1d_csi(1) = cart_comm_1d(0, 1, mpi_proc_null, mpi_proc_null, mpi_comm_self)
   comm1d_eta(1) = comm1d_csi(1)
   comm1d_zet(1) = comm1d_csi(1)

   ! LSM parameters (init_LSM.F90:136-181 semantics)
   allocate(btype_lsm(6,1))
   btype_lsm(1:6,1)                = ip(6:11)
   orderAD                         = ip(12)
   numiter_lsm                     = ip(13)
   RNitermax                       = ip(14)
   OrderReinitialisationBoundaries = ip(15)
   ENOBCReinitialisation           = (ip(16) /= 0)
   hybrid_reinitialisation         = (ip(19) /= 0)
   call_reinitialisation           = (ip(20) /= 0)
   RNfreq                          = ip(21)
   sussman_correction_method       = ip(23)
   call_levelsetmethod             = (ip(24) /= 0)
   GlobalMassCorrection            = .false.
   TotalVolumeComputation          = .false.
   bc_extrapolation_order          = 1

   delti_lsm           = rp(4)
   epslsm              = rp(5)
   RNtimestep          = rp(6)
   epsReinitialisation = rp(7)
   BigPhi              = rp(8)
   delti               = rp(4) * real(numiter_lsm, kind=rdf)

   ! No obstacles: the reference allocates these masks on the block extents
   ! and sets them to 1 (init_obstacle.F90:28-35).  reinitialization_benchmark
   ! indexes act_obstacle_ad with the host's i,j,k after a loop has run past
   ! its end (reinitialization_benchmark.F90:119), so the mask is padded by one
   ! node on every side and filled with 1, the value every in-range entry has.
   allocate(act_obstacle_ad(0:ni+1,0:nj+1,0:nk+1), act_obstacle_rn(0:ni+1,0:nj+1,0:nk+1))
   act_obstacle_ad = 1
   act_obstacle_rn = 1

   ! index tables read by outputD1_real (global_debug.F90:15-58)
   allocate(le_idx(ni,nj,nk,1), le_idx_a(1), le_idx_b(1))
   allocate(li_ia(1), li_ib(1), li_ja(1), li_jb(1), li_ka(1), li_kb(1))
   l = 0
   do kk = 1, nk
   do jj = 1, nj
   do ii = 1, ni
      l = l + 1
      le_idx(ii,jj,kk,1) = l
   end do
   end do
   end do
   le_idx_a(1) = 1 ; le_idx_b(1) = l
   li_ia(1) = 1 ; li_ib(1) = ni
   li_ja(1) = 1 ; li_jb(1) = nj
   li_ka(1) = 1 ; li_kb(1) = nk
   OUTPUTdir = './'

   t0 = mpi_wtime()
   if (ip(1) == 13) then
      ! mg_metrics itself on the grid x, y, z (mg_metrics.F90:53-262)
      call metrics_driver( ni, nj, nk, x, y, z, rp(1), rp(2), rp(3), trim(fout) )
   else if (ip(1) >= 14 .and. ip(1) <= 16) then
      call consumer_driver( ni, nj, nk, rp(1), rp(2), rp(3), csi, eta, zet, z, &
                            fphi, faux, ip(1), trim(fout) )
   else if (ip(1) == 12) then
      ! levelsetmethod itself, called as mg_driver.F90:563-581 does
      do kk = 1, max(ip(5), 1)
         ntime = ip(22) + kk - 1
         fphi0 = fphi
         call levelsetmethod( 1, ni, 1, nj, 1, nk, 0, 0, 0, rp(1), rp(2), rp(3), &
                              q, csi, eta, zet, aj, fphi, fphi0, x, y, z, ntime )
      end do
      open(22, file=trim(fout), access='stream', form='unformatted', status='replace')
      write(22) int(1,4)
      write(22) int(0,4), int(ni*nj*nk,4)
      write(22) fphi
      close(22)
   else
      call ref_driver( 1, ni, 1, nj, 1, nk, 0, 0, 0, rp(1), rp(2), rp(3), &
                       q, csi, eta, zet, aj, x, y, z,                       &
                       fphi, fphi0, faux, mask, ip, rp, trim(fout) )
   end if
   t1 = mpi_wtime()
   if (ip(25) /= 0) print '(a,es16.8)', 'REF_SECONDS ', t1 - t0

   call mpi_finalize(ierr)

end program lsm_ref_harness


subroutine metrics_driver( ni, nj, nk, xin, yin, zin, dcin, dein, dzin, fout )

   ! Fills the module arrays mg_metrics reads (global.F90:46-54,150-155; one
   ! block, no ghost layers, every face physical) and calls it.
   use precision
   use global, only : x, y, z, csi, eta, zet, aj, xc, xe, xz, yc, ye, yz, zc, ze, zz, &
                      dc, de, dz, le_ia, le_ib, le_ja, le_jb, le_ka, le_kb, ns
   implicit none
   integer, intent(in) :: ni, nj, nk
   real (kind = rdf), intent(in) :: xin(ni*nj*nk), yin(ni*nj*nk), zin(ni*nj*nk)
   real (kind = rdf), intent(in) :: dcin, dein, dzin
   character(len=*), intent(in) :: fout
   integer :: n, c

   n = ni * nj * nk
   allocate( x(n), y(n), z(n), aj(n), csi(3,n), eta(3,n), zet(3,n) )
   allocate( xc(n), xe(n), xz(n), yc(n), ye(n), yz(n), zc(n), ze(n), zz(n) )
   allocate( dc(1), de(1), dz(1) )
   allocate( le_ia(1), le_ib(1), le_ja(1), le_jb(1), le_ka(1), le_kb(1) )
   x = xin ; y = yin ; z = zin
   dc(1) = dcin ; de(1) = dein ; dz(1) = dzin
   le_ia(1) = 1 ; le_ib(1) = ni
   le_ja(1) = 1 ; le_jb(1) = nj
   le_ka(1) = 1 ; le_kb(1) = nk
   ns = 1
   csi = 0.0_rdf ; eta = 0.0_rdf ; zet = 0.0_rdf ; aj = 0.0_rdf

   call mg_metrics

   open(22, file=fout, access='stream', form='unformatted', status='replace')
   write(22) int(10,4)
   do c = 1, 3
      write(22) int(0,4), int(n,4)
      write(22) csi(c,:)
   end do
   do c = 1, 3
      write(22) int(0,4), int(n,4)
      write(22) eta(c,:)
   end do
   do c = 1, 3
      write(22) int(0,4), int(n,4)
      write(22) zet(c,:)
   end do
   write(22) int(0,4), int(n,4)
   write(22) aj
   close(22)

end subroutine metrics_driver


subroutine consumer_driver( ni, nj, nk, dc, de, dz, csi, eta, zet, z, phi, aux, mode, fout )

   ! One block, no ghost layers, every face physical (the module state set by
   ! the main program).  The gradient outputs are intent(inout) in the
   ! reference and not written at the eight vertices, so they start as
   ! aux + c (component c); calc_h's output starts as aux.
   use precision
   implicit none
   integer, intent(in) :: ni, nj, nk, mode
   real (kind = rdf), intent(in) :: dc, de, dz
   real (kind = rdf), intent(in) :: csi(3,ni,nj,nk), eta(3,ni,nj,nk), zet(3,ni,nj,nk)
   real (kind = rdf), intent(in) :: z(ni,nj,nk), phi(ni,nj,nk), aux(ni,nj,nk)
   character(len=*), intent(in) :: fout
   real (kind = rdf), allocatable :: g(:,:,:,:), h(:,:,:)
   integer :: c, nc, n

   n = ni * nj * nk
   open(22, file=fout, access='stream', form='unformatted', status='replace')
   select case (mode)
   case (14, 16)   ! calc_phi_gradient_ENO2( phi ) / calc_h_gradient_ENO2( h = phi )
      nc = 3
      if (mode == 16) nc = 2
      allocate(g(nc,ni,nj,nk))
      do c = 1, nc
         g(c,:,:,:) = aux + real(c, rdf)
      end do
      if (mode == 14) then
         call calc_phi_gradient_ENO2( 1, ni, 1, nj, 1, nk, 0, 0, 0, dc, de, dz, &
                                      csi, eta, zet, phi, g )
      else
         call calc_h_gradient_ENO2( 1, ni, 1, nj, 1, nk, 0, 0, 0, dc, de, dz, &
                                    csi, eta, zet, phi, g )
      end if
      write(22) int(nc,4)
      do c = 1, nc
         write(22) int(0,4), int(n,4)
         write(22) g(c,:,:,:)
      end do
   case (15)       ! calc_h( z, phi ) -> h
      allocate(h(ni,nj,nk))
      h = aux
      call calc_h( 1, ni, 1, nj, 1, nk, 0, 0, 0, z, phi, h )
      write(22) int(1,4)
      write(22) int(0,4), int(n,4)
      write(22) h
   end select
   close(22)

end subroutine consumer_driver


subroutine ref_driver( il, iu, jl, ju, kl, ku, igp, jgp, kgp, dc, de, dz, &
                       q, csi, eta, zet, aj, x, y, z,                     &
                       phi, phi0, aux, mask, ip, rp, fout )

   use global_param
   use global_app
   use global_mpi
   use global_debug
   use global_lsm, only : RNitermax, RNtimestep, RNfreq, heaviside, deltaf,  &
                          btype_lsm, numiter_lsm, delti_lsm, orderAD,        &
                          bc_extrapolation_order, call_levelsetmethod,       &
                          call_reinitialisation, hybrid_reinitialisation,    &
                          OrderReinitialisationBoundaries,                   &
                          ENOBCReinitialisation, BigPhi, epslsm
   use global_obstacle, only : obstacle_lsm_ad, boundary_obstacle,           &
                               is_obstacle, act_obstacle_ad, act_obstacle_rn, &
                               li_obs_ia, li_obs_ib, li_obs_ja, li_obs_jb,   &
                               li_obs_ka, li_obs_kb, i_la_o
   use AdvectionMethods

   implicit none

   integer, intent(in) :: il, iu, jl, ju, kl, ku, igp, jgp, kgp
   real (kind = rdf), intent(in) :: dc, de, dz
   real (kind = rdf), dimension(1:4,il:iu,jl:ju,kl:ku), intent(inout) :: q
   real (kind = rdf), dimension(1:3,il:iu,jl:ju,kl:ku), intent(in) :: csi, eta, zet
   real (kind = rdf), dimension(il:iu,jl:ju,kl:ku), intent(in) :: aj, x, y, z
   real (kind = rdf), dimension(il:iu,jl:ju,kl:ku), intent(inout) :: phi, phi0, aux
   integer(4), dimension(il:iu,jl:ju,kl:ku), intent(inout) :: mask
   integer(4), intent(in) :: ip(64)
   real (kind = rdf), intent(in) :: rp(32)
   character(len=*), intent(in) :: fout

   ! host-associated names read by the included reference routines
   real (kind = rdf), dimension(:,:,:,:), allocatable :: ucn_j
   integer :: RK
   integer :: i, j, k
   integer :: i_mysta, j_mysta, k_mysta, i_myend, j_myend, k_myend
   integer :: iteraciontiempo
   character(len = 256) :: debugname

   ! harness locals
   real (kind = rdf), dimension(:,:,:), allocatable :: wa, wb, wc
   integer, dimension(:,:,:), allocatable :: imask
   real (kind = rdf) :: v1, v2, v3
   integer :: mode, nsteps, step
   integer :: n

   mode   = ip(1)
   nsteps = max(ip(5), 1)
   n      = (iu-il+1)*(ju-jl+1)*(ku-kl+1)
   iteraciontiempo = 1

   i_mysta = il + igp ; j_mysta = jl + jgp ; k_mysta = kl + kgp
   i_myend = iu - igp ; j_myend = ju - jgp ; k_myend = ku - kgp

   allocate(wa(il:iu,jl:ju,kl:ku), wb(il:iu,jl:ju,kl:ku), wc(il:iu,jl:ju,kl:ku))
   allocate(imask(il:iu,jl:ju,kl:ku))
   allocate(ucn_j(1:3,il:iu,jl:ju,kl:ku))
   wa = 0.0_rdf ; wb = 0.0_rdf ; wc = 0.0_rdf
   imask = mask
   ucn_j = 0.0_rdf

   open(22, file=fout, access='stream', form='unformatted', status='replace')

   select case (mode)

   case (1)   ! ad_RKTVD3, nsteps physical steps of numiter_lsm sub-steps
      call rhs_contra_j()
      do step = 1, nsteps
         wa = phi                        ! phi_n of this step
         call advect_one_step()
         if (step == 1) wb = phi
      end do
      call put_reals_header(2)
      call put_real(wb)
      call put_real(phi)

   case (2)   ! calc_RH_AD_WENO3 on interior nodes; rh zeroed first
      call rhs_contra_j()
      wa = 0.0_rdf
      call calc_RH_AD_WENO3( phi, wa, imask )
      call put_reals_header(1)
      call put_real(wa)

   case (3)   ! bcond_lsm
      call bcond_lsm( phi )
      call put_reals_header(1)
      call put_real(phi)

   case (4)   ! rhs_contra_j
      call rhs_contra_j()
      call put_reals_header(3)
      wa = ucn_j(1,:,:,:) ; call put_real(wa)
      wa = ucn_j(2,:,:,:) ; call put_real(wa)
      wa = ucn_j(3,:,:,:) ; call put_real(wa)

   case (5)   ! calc_NormGrad_d( d = phi, phizero = phi0 )
      call calc_NormGrad_d( phi, phi0, wa )
      call put_reals_header(1)
      call put_real(wa)

   case (6)   ! reinitialisation_test2( PhiZero = phi0, PhiReinitialised )
      wb = phi
      imask = 0
      call reinitialisation_test2( phi0, wb, imask )
      call put_reals_header(1)
      call put_real(wb)

   case (7)   ! calc_sign_ini( phi0 ) -> sign, |grad phi0|
      call calc_sign_ini( phi0, wa, wb, epslsm, ip(17) )
      call put_reals_header(2)
      call put_real(wa)
      call put_real(wb)

   case (8)   ! calc_RH_RN( d = phi, sign = aux )
      wa = 0.0_rdf
      call calc_RH_RN( phi, aux, wa )
      call put_reals_header(1)
      call put_real(wa)

   case (9)   ! calc_lambda_RN( phi0, |grad phi0| = aux, d = phi, dt = rp(9) )
      wa = 0.0_rdf
      call calc_lambda_RN( phi0, aux, phi, wa, rp(9) )
      call put_reals_header(1)
      call put_real(wa)

   case (10)  ! reinitialization_benchmark (Godunov path, M_Grad = 0)
      wb = phi
      imask = 0
      call reinitialization_benchmark( phi0, wb, ip(17), ip(18), 0, 0, 0, imask )
      call put_reals_header(1)
      call put_real(wb)

   case (17)  ! calc_RH_AD (orderAD = 0 upwind right-hand side); rh zeroed first
      call rhs_contra_j()
      wa = 0.0_rdf
      call calc_RH_AD( phi, wa, imask )
      call put_reals_header(1)
      call put_real(wa)

   case (11)  ! TotalWaterVolume with the AdvectionNodes mask
      call TotalWaterVolume( phi, v1, v2, v3, imask )
      write(22) int(2,4)
      write(22) int(0,4), int(3,4), v1, v2, v3
      write(22) int(1,4), int(n,4), int(imask,4)
      close(22)
      return

   end select

   close(22)

contains

   subroutine advect_one_step()
      ! levelsetmethod.F90:227-239 for one physical step
      integer :: it
      wc = wa
      do it = 1, numiter_lsm
         call ad_RKTVD3( phi, wc, imask )
         wc = phi
      end do
   end subroutine advect_one_step

   subroutine put_reals_header(m)
      integer, intent(in) :: m
      write(22) int(m,4)
   end subroutine put_reals_header

   subroutine put_real(a)
      real (kind = rdf), dimension(il:iu,jl:ju,kl:ku), intent(in) :: a
      write(22) int(0,4), int(n,4)
      write(22) a
   end subroutine put_real

   include 'bcond_lsm.F90'
   include 'bcond_lsm_obstacle.F90'
   include 'rhs_contra_j.F90'
   include 'ad_RKTVD3.F90'
   include 'calc_RH_AD_WENO3.F90'
   include 'calc_RH_AD.F90'
   include 'rhs_exchng3_3d.F90'
   include 'reinitialization_benchmark.F90'
   include 'reinitialisation_test2.F90'
   include 'calc_NormGrad_d.F90'
   include 'calc_RH_RN.F90'
   include 'calc_RH_RN_2.F90'
   include 'calc_RH_RN_obstacle.F90'
   include 'calc_sign_ini.F90'
   include 'calc_sign_ini_2.F90'
   include 'calc_sign_ini_obstacle.F90'
   include 'calc_lambda_RN.F90'
   include 'testfilter.F90'
   include 'TotalWaterVolume.F90'

end subroutine ref_driver
