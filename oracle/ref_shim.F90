! bind(c) driver around the reference's own advection kernels.
!
! TEST INFRASTRUCTURE ONLY: this file is compiled together with the
! reference sources src/operators/advection_schemes/{advection_base,
! advection_<scheme>}.F90 (which include advection.F90.template and
! limiters.h) by oracle/Makefile, into oracle/_ref/libgetm_ref.so.  It only
! forwards plain arrays to the reference's type-bound u2d/v2d/w3d procedures
! the way python/src/wrapper.F90:213-273 does (one call per layer for u/v,
! one call for w), so the reference arithmetic runs unmodified.
!
! Index convention: every array is a padded (nx_, ny_[, nz]) block with
! Fortran lower bound imin-halox = 1-halox, i.e. the interior is 1..nx.
module getm_ref_shim
   use iso_c_binding
   use iso_fortran_env, only: real64
   use advection_base,      only: type_advection_base
   use advection_hsimt,     only: type_advection_hsimt
   use advection_muscl,     only: type_advection_muscl
   use advection_p2_pdm,    only: type_advection_p2_pdm
   use advection_splmax13,  only: type_advection_splmax13
   use advection_superbee,  only: type_advection_superbee
   use advection_upstream,  only: type_advection_upstream
   implicit none

contains

   ! Scheme codes are pygetm.operators.AdvectionScheme (operators.py:11-18).
   subroutine make_op(scheme, op, ok)
      integer, intent(in) :: scheme
      class(type_advection_base), allocatable, intent(out) :: op
      logical, intent(out) :: ok
      ok = .true.
      select case (scheme)
      case (1); allocate(type_advection_hsimt :: op)
      case (2); allocate(type_advection_muscl :: op)
      case (3); allocate(type_advection_p2_pdm :: op)
      case (4); allocate(type_advection_splmax13 :: op)
      case (5); allocate(type_advection_superbee :: op)
      case (6); allocate(type_advection_upstream :: op)
      case default; ok = .false.
      end select
   end subroutine

   ! direction 1 = x (u2d), 2 = y (v2d); one reference call per layer.
   function ref_advection_uv(direction, scheme, nx, ny, nk, halox, haloy, &
         uvmask, metric1, metric2, huv, uv, tmask, iarea, has_ah, ah, dt, h, var) result(status) bind(c)
      integer(c_int), value, intent(in) :: direction, scheme, nx, ny, nk, halox, haloy, has_ah
      integer(c_int), intent(in) :: uvmask(1-halox:nx+halox, 1-haloy:ny+haloy)
      integer(c_int), intent(in) :: tmask(1-halox:nx+halox, 1-haloy:ny+haloy)
      real(c_double), intent(in) :: metric1(1-halox:nx+halox, 1-haloy:ny+haloy)
      real(c_double), intent(in) :: metric2(1-halox:nx+halox, 1-haloy:ny+haloy)
      real(c_double), intent(in) :: iarea(1-halox:nx+halox, 1-haloy:ny+haloy)
      real(c_double), intent(in) :: ah(1-halox:nx+halox, 1-haloy:ny+haloy)
      real(c_double), intent(in) :: huv(1-halox:nx+halox, 1-haloy:ny+haloy, nk)
      real(c_double), intent(in) :: uv(1-halox:nx+halox, 1-haloy:ny+haloy, nk)
      real(c_double), value, intent(in) :: dt
      real(c_double), intent(inout) :: h(1-halox:nx+halox, 1-haloy:ny+haloy, nk)
      real(c_double), intent(inout) :: var(1-halox:nx+halox, 1-haloy:ny+haloy, nk)
      integer(c_int) :: status

      class(type_advection_base), allocatable :: op
      logical :: ok, apply_diffusion
      integer :: k, halo(2)

      status = 1
      call make_op(scheme, op, ok)
      if (.not. ok) return
      halo = (/halox, haloy/)
      apply_diffusion = has_ah /= 0
      do k = 1, nk
         if (direction == 1) then
            call op%u2d(1, nx, 1, ny, halo, uvmask, metric1, metric2, huv(:,:,k), uv(:,:,k), &
                        tmask, iarea, apply_diffusion, ah, dt, h(:,:,k), var(:,:,k))
         else
            call op%v2d(1, nx, 1, ny, halo, uvmask, metric1, metric2, huv(:,:,k), uv(:,:,k), &
                        tmask, iarea, apply_diffusion, ah, dt, h(:,:,k), var(:,:,k))
         end if
      end do
      status = 0
   end function

   function ref_advection_w(scheme, nx, ny, nz, halox, haloy, w, w_var, tmask, dt, h, var) &
         result(status) bind(c)
      integer(c_int), value, intent(in) :: scheme, nx, ny, nz, halox, haloy
      real(c_double), intent(in) :: w(1-halox:nx+halox, 1-haloy:ny+haloy, 0:nz)
      real(c_double), intent(in) :: w_var(1-halox:nx+halox, 1-haloy:ny+haloy, 0:nz)
      integer(c_int), intent(in) :: tmask(1-halox:nx+halox, 1-haloy:ny+haloy)
      real(c_double), value, intent(in) :: dt
      real(c_double), intent(inout) :: h(1-halox:nx+halox, 1-haloy:ny+haloy, nz)
      real(c_double), intent(inout) :: var(1-halox:nx+halox, 1-haloy:ny+haloy, nz)
      integer(c_int) :: status

      class(type_advection_base), allocatable :: op
      logical :: ok

      status = 1
      call make_op(scheme, op, ok)
      if (.not. ok) return
      call op%w3d(1, nx, 1, ny, nz, (/halox, haloy/), w, w_var, tmask, dt, h, var)
      status = 0
   end function

end module
