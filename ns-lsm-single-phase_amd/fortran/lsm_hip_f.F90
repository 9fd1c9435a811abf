! ---------------------------------------------------------------------------
! Fortran interfaces (ISO_C_BINDING) to the host-memory adapter of
! liblsm_hip.so (include/lsm_hip.h, "host-memory adapter").  Used by the
! drop-in shim files in this directory, which replace the reference's
! level-set routines when levelsetmethod.F90 is compiled with this directory
! ahead of the reference sources on the include path.
! ---------------------------------------------------------------------------
module lsm_hip_f

   use iso_c_binding
   implicit none

   integer(c_int), parameter :: LSM_ADVECT_FAST  = 0
   integer(c_int), parameter :: LSM_ADVECT_EXACT = 1
   integer(c_int), parameter :: LSM_ADVECT_UPWIND1 = 2

   ! arithmetic of the advection calls: the reference's (bit-identical results)
   ! unless the caller opts into the one-division WENO3 weights
   integer(c_int) :: lsm_advect_mode = LSM_ADVECT_EXACT

   type, bind(C) :: lsm_reinit_params
      real(c_double)     :: epslsm, eps_reinit, rn_timestep
      integer(c_int32_t) :: rn_itermax, order_bc, eno_bc, m_sphi0, m_ti
      integer(c_int32_t) :: hybrid = 0
      integer(c_int32_t) :: lambda_method = 0
   end type lsm_reinit_params

   interface

      integer(c_int) function lsm_host_setup(ext, faces, btype, dc, de, dz, big_phi, &
                                             csi, eta, zet, aj, x, y, z) bind(C, name="lsm_host_setup")
         import :: c_int, c_int32_t, c_double
         integer(c_int32_t), intent(in) :: ext(9)
         integer(c_int32_t), value      :: faces
         integer(c_int32_t), intent(in) :: btype(6)
         real(c_double), value          :: dc, de, dz, big_phi
         real(c_double), intent(in)     :: csi(*), eta(*), zet(*), aj(*), x(*), y(*), z(*)
      end function lsm_host_setup

      ! blanking regions of the bound block: nblk rows of (li_blk_ia, ib, ja, jb, ka, kb, blktype(1:4))
      integer(c_int) function lsm_host_blanking(nblk, boxes) bind(C, name="lsm_host_blanking")
         import :: c_int, c_int32_t
         integer(c_int32_t), value      :: nblk
         integer(c_int32_t), intent(in) :: boxes(*)
      end function lsm_host_blanking

      ! blanking regions of the consumer calls (lsm_host_phi/h_gradient_eno2, lsm_host_calc_h)
      integer(c_int) function lsm_host_consumer_blanking(nblk, boxes) bind(C, name="lsm_host_consumer_blanking")
         import :: c_int, c_int32_t
         integer(c_int32_t), value      :: nblk
         integer(c_int32_t), intent(in) :: boxes(*)
      end function lsm_host_consumer_blanking

      ! the next lsm_host_setup uploads the grid again (after in-place edits of grid arrays)
      integer(c_int) function lsm_host_invalidate() bind(C, name="lsm_host_invalidate")
         import :: c_int
      end function lsm_host_invalidate

      integer(c_int) function lsm_host_contra_velocity(q, ucn) bind(C, name="lsm_host_contra_velocity")
         import :: c_int, c_double
         real(c_double), intent(in)    :: q(*)
         real(c_double), intent(inout) :: ucn(*)
      end function lsm_host_contra_velocity

      integer(c_int) function lsm_host_weno3_rhs(phi, rhs, mask, mode) bind(C, name="lsm_host_weno3_rhs")
         import :: c_int, c_int32_t, c_double
         real(c_double), intent(in)     :: phi(*)
         real(c_double), intent(inout)  :: rhs(*)
         integer(c_int32_t), intent(in) :: mask(*)
         integer(c_int), value          :: mode
      end function lsm_host_weno3_rhs

      integer(c_int) function lsm_host_upwind_rhs(phi, rhs, mask) bind(C, name="lsm_host_upwind_rhs")
         import :: c_int, c_int32_t, c_double
         real(c_double), intent(in)     :: phi(*)
         real(c_double), intent(inout)  :: rhs(*)
         integer(c_int32_t), intent(in) :: mask(*)
      end function lsm_host_upwind_rhs

      integer(c_int) function lsm_host_rk3(phi_np1, phi_n, mask, dt, mode) bind(C, name="lsm_host_rk3")
         import :: c_int, c_int32_t, c_double
         real(c_double), intent(inout)  :: phi_np1(*)
         real(c_double), intent(in)     :: phi_n(*)
         integer(c_int32_t), intent(in) :: mask(*)
         real(c_double), value          :: dt
         integer(c_int), value          :: mode
      end function lsm_host_rk3

      integer(c_int) function lsm_host_bcond(phi) bind(C, name="lsm_host_bcond")
         import :: c_int, c_double
         real(c_double), intent(inout) :: phi(*)
      end function lsm_host_bcond

      integer(c_int) function lsm_host_normgrad(d, phi0, order_bc, eno_bc, ng) bind(C, name="lsm_host_normgrad")
         import :: c_int, c_double
         real(c_double), intent(in)    :: d(*), phi0(*)
         integer(c_int), value         :: order_bc, eno_bc
         real(c_double), intent(inout) :: ng(*)
      end function lsm_host_normgrad

      integer(c_int) function lsm_host_reinit_eno2(phi0, phi_out, p, iters) bind(C, name="lsm_host_reinit_eno2")
         import :: c_int, c_int32_t, c_double, lsm_reinit_params
         real(c_double), intent(in)          :: phi0(*)
         real(c_double), intent(inout)       :: phi_out(*)
         type(lsm_reinit_params), intent(in) :: p
         integer(c_int32_t), intent(inout)   :: iters
      end function lsm_host_reinit_eno2

      ! levelsetmethod as a whole, phi resident on the device for the call
      integer(c_int) function lsm_host_levelset_step(phi, phi_n, q, p, numiter_lsm, delti_lsm,      &
                                                     call_levelset, call_reinit, rn_freq, ntime, mode, &
                                                     vol) bind(C, name="lsm_host_levelset_step")
         import :: c_int, c_double, lsm_reinit_params
         real(c_double), intent(inout)       :: phi(*)
         real(c_double), intent(in)          :: phi_n(*), q(*)
         type(lsm_reinit_params), intent(in) :: p
         integer(c_int), value               :: numiter_lsm, call_levelset, call_reinit, rn_freq, ntime, mode
         real(c_double), value               :: delti_lsm
         real(c_double), intent(inout)       :: vol(3)
      end function lsm_host_levelset_step

      integer(c_int) function lsm_host_sign_ini(phi0, epslsm, m_sphi0, sgn, ng) bind(C, name="lsm_host_sign_ini")
         import :: c_int, c_double
         real(c_double), intent(in)    :: phi0(*)
         real(c_double), value         :: epslsm
         integer(c_int), value         :: m_sphi0
         real(c_double), intent(inout) :: sgn(*), ng(*)
      end function lsm_host_sign_ini

      integer(c_int) function lsm_host_rh_rn(d, sgn, rh) bind(C, name="lsm_host_rh_rn")
         import :: c_int, c_double
         real(c_double), intent(in)    :: d(*), sgn(*)
         real(c_double), intent(inout) :: rh(*)
      end function lsm_host_rh_rn

      integer(c_int) function lsm_host_lambda_rn(phi0, ng0, d, dt, eps, lam) bind(C, name="lsm_host_lambda_rn")
         import :: c_int, c_double
         real(c_double), intent(in)    :: phi0(*), ng0(*), d(*)
         real(c_double), value         :: dt, eps
         real(c_double), intent(inout) :: lam(*)
      end function lsm_host_lambda_rn

      integer(c_int) function lsm_host_lambda_rn_simpson(phi0, ng0, d, dt, eps, lam) &
                              bind(C, name="lsm_host_lambda_rn_simpson")
         import :: c_int, c_double
         real(c_double), intent(in)    :: phi0(*), ng0(*), d(*)
         real(c_double), value         :: dt, eps
         real(c_double), intent(inout) :: lam(*)
      end function lsm_host_lambda_rn_simpson

      integer(c_int) function lsm_host_reinit_godunov(phi0, phi_out, p) bind(C, name="lsm_host_reinit_godunov")
         import :: c_int, c_double, lsm_reinit_params
         real(c_double), intent(in)          :: phi0(*)
         real(c_double), intent(inout)       :: phi_out(*)
         type(lsm_reinit_params), intent(in) :: p
      end function lsm_host_reinit_godunov

      integer(c_int) function lsm_host_reinit_geometric(phizero, phicorrected, ids) &
            bind(C, name="lsm_host_reinit_geometric")
         import :: c_int, c_int32_t, c_double
         real(c_double), intent(in)        :: phizero(*)
         real(c_double), intent(inout)     :: phicorrected(*)
         integer(c_int32_t), intent(inout) :: ids(*)
      end function lsm_host_reinit_geometric

      ! an absent mask is passed as NULL (no AdvectionNodes output)
      integer(c_int) function lsm_host_water_volume(phi, vol, mask) bind(C, name="lsm_host_water_volume")
         import :: c_int, c_int32_t, c_double
         real(c_double), intent(in)                  :: phi(*)
         real(c_double), intent(inout)               :: vol(3)
         integer(c_int32_t), intent(inout), optional :: mask(*)
      end function lsm_host_water_volume

      ! consumers of phi outside levelsetmethod (external subroutines)
      integer(c_int) function lsm_host_phi_gradient_eno2(ext, faces, dc, de, dz, csi, eta, zet, phi, &
                                                         order_bc, eno_bc, grad)                   &
                                                         bind(C, name="lsm_host_phi_gradient_eno2")
         import :: c_int, c_int32_t, c_double
         integer(c_int32_t), intent(in) :: ext(9)
         integer(c_int32_t), value      :: faces
         real(c_double), value          :: dc, de, dz
         real(c_double), intent(in)     :: csi(*), eta(*), zet(*), phi(*)
         integer(c_int), value          :: order_bc, eno_bc
         real(c_double), intent(inout)  :: grad(*)
      end function lsm_host_phi_gradient_eno2

      integer(c_int) function lsm_host_h_gradient_eno2(ext, faces, dc, de, dz, csi, eta, zet, h, &
                                                       order_bc, eno_bc, grad)                 &
                                                       bind(C, name="lsm_host_h_gradient_eno2")
         import :: c_int, c_int32_t, c_double
         integer(c_int32_t), intent(in) :: ext(9)
         integer(c_int32_t), value      :: faces
         real(c_double), value          :: dc, de, dz
         real(c_double), intent(in)     :: csi(*), eta(*), zet(*), h(*)
         integer(c_int), value          :: order_bc, eno_bc
         real(c_double), intent(inout)  :: grad(*)
      end function lsm_host_h_gradient_eno2

      integer(c_int) function lsm_host_calc_h(ext, faces, z, phi, h, missing) bind(C, name="lsm_host_calc_h")
         import :: c_int, c_int32_t, c_double
         integer(c_int32_t), intent(in)    :: ext(9)
         integer(c_int32_t), value         :: faces
         real(c_double), intent(in)        :: z(*), phi(*)
         real(c_double), intent(inout)     :: h(*)
         integer(c_int32_t), intent(inout) :: missing
      end function lsm_host_calc_h

      type(c_ptr) function lsm_last_error() bind(C, name="lsm_last_error")
         import :: c_ptr
      end function lsm_last_error

   end interface

contains

   ! stops the program with the library's message when a call failed
   subroutine lsm_f_check(rc, where)
      integer(c_int), intent(in)   :: rc
      character(len=*), intent(in) :: where
      if (rc /= 0) then
         write(*, '(a,a,a,i0,a,a)') 'lsm_hip: ', where, ' failed (', rc, '): ', trim(lsm_error_text())
         error stop 1
      end if
   end subroutine lsm_f_check

   ! stops the program for a configuration the library does not implement
   subroutine lsm_f_unsupported(where, what)
      character(len=*), intent(in) :: where, what
      write(*, '(a,a,a,a)') 'lsm_hip: ', where, ': unsupported configuration: ', what
      error stop 1
   end subroutine lsm_f_unsupported

   ! Binds the calling block (extents, neighbours, boundary types, metrics)
   ! to the library.  The grid is uploaded when the extents or the array
   ! addresses change and reused otherwise.  nbr holds the neighbour ranks in
   ! the order back, front, left, right, down, up; a face is a physical
   ! boundary when its neighbour is proc_null.
   subroutine lsm_f_bind(ext, nbr, proc_null, btype, dc, de, dz, big_phi, &
                         csi, eta, zet, aj, x, y, z)
      integer, intent(in)        :: ext(9), nbr(6), proc_null, btype(6)
      real(c_double), intent(in) :: dc, de, dz, big_phi
      real(c_double), intent(in) :: csi(*), eta(*), zet(*), aj(*), x(*), y(*), z(*)
      integer(c_int32_t) :: e(9), bt(6), faces
      integer :: f
      faces = 0
      do f = 1, 6
         if (nbr(f) == proc_null) faces = ior(faces, int(ishft(1, f - 1), c_int32_t))
      end do
      e  = int(ext, c_int32_t)
      bt = int(btype, c_int32_t)
      call lsm_f_check(lsm_host_setup(e, faces, bt, dc, de, dz, big_phi, &
                                      csi, eta, zet, aj, x, y, z), 'lsm_host_setup')
   end subroutine lsm_f_bind

   ! Passes the blanking regions of global_app (one block per rank: the
   ! interior boxes li_blk_* and the exterior boxes le_blk_* coincide) to the
   ! block bound by lsm_f_bind, or with consumers = .true. to the consumer
   ! calls of phi.
   subroutine lsm_f_bind_blanking(nblk, nblke, ia, ib, ja, jb, ka, kb, blktype, consumers)
      integer, intent(in) :: nblk, nblke
      integer, intent(in) :: ia(:,:), ib(:,:), ja(:,:), jb(:,:), ka(:,:), kb(:,:), blktype(:,:)
      logical, intent(in), optional :: consumers
      integer(c_int32_t) :: boxes(10, 4)
      integer :: nb
      if (nblk /= nblke) call lsm_f_unsupported('level-set drop-in', &
                                                'blanking regions across rank boundaries (nblk /= nblke)')
      if (nblk > 4) call lsm_f_unsupported('level-set drop-in', 'more than 4 blanking regions')
      boxes = 0
      do nb = 1, nblk
         boxes(1:6, nb) = int([ia(1, nb), ib(1, nb), ja(1, nb), jb(1, nb), ka(1, nb), kb(1, nb)], c_int32_t)
         boxes(7:10, nb) = int(blktype(1:4, nb), c_int32_t)
      end do
      if (present(consumers)) then
         if (consumers) then
            call lsm_f_check(lsm_host_consumer_blanking(int(nblk, c_int32_t), boxes), 'lsm_host_consumer_blanking')
            return
         end if
      end if
      call lsm_f_check(lsm_host_blanking(int(nblk, c_int32_t), boxes), 'lsm_host_blanking')
   end subroutine lsm_f_bind_blanking

   ! physical-face mask from the neighbour ranks (back, front, left, right,
   ! down, up; a face is physical when its neighbour is proc_null)
   integer(c_int32_t) function lsm_f_faces(nbr, proc_null) result(faces)
      integer, intent(in) :: nbr(6), proc_null
      integer :: f
      faces = 0
      do f = 1, 6
         if (nbr(f) == proc_null) faces = ior(faces, int(ishft(1, f - 1), c_int32_t))
      end do
   end function lsm_f_faces

   ! text of the library's last error message
   function lsm_error_text() result(msg)
      character(len=512) :: msg
      type(c_ptr) :: p
      character(kind=c_char), pointer :: s(:)
      integer :: n
      msg = ' '
      p = lsm_last_error()
      if (.not. c_associated(p)) return
      call c_f_pointer(p, s, [512])
      do n = 1, 512
         if (s(n) == c_null_char) exit
         msg(n:n) = s(n)
      end do
   end function lsm_error_text

end module lsm_hip_f
