!*******************************************************************************
! example_refine.f90 -- a Fortran host driving the refinement model through
! diffuse_turbo_mod the way REFINE would (refine/prog/refine_run.f90:491-600,
! the gl_is_der hook :550-552, gl_data(:,:,:,k) of lib_f90/global_data.f90:
! 71-76): the intensity of a model, its derivative columns dI/dtheta_k for
! three REFINE parameters written into gl_data(:,:,:,k) by one JVP call, and
! chi^2 with its gradient by one adjoint call.
! Checks: each column against central differences of dc_model_intensity, and
! d chi^2/d theta_k from the adjoint gradient against 2 sum (I - T) dI/dtheta_k
! from the columns.  Prints PASS/FAIL; exit status 0/1.
!*******************************************************************************
program example_refine
   use iso_c_binding
   use diffuse_turbo_mod
   implicit none
   integer, parameter :: nc = 2, nsite = 4, nat = nc**3 * nsite, nscat = 2, npar = 3
   integer, parameter :: n1 = 11, n2 = 9, n3 = 7, np = n1 * n2 * n3
   real(c_double), parameter :: abc(3) = [4.0d0, 5.0d0, 6.0d0]
   real(c_double), parameter :: site(3, nsite) = reshape([0.11d0, 0.23d0, 0.05d0, 0.52d0, 0.61d0, 0.17d0, &
                                                           0.27d0, 0.74d0, 0.58d0, 0.83d0, 0.36d0, 0.91d0], [3, nsite])
   real(c_double), target :: pos(3, nat), occ(nat), bdw(nat), dparam(5, nat, npar)
   real(c_double), target :: icalc(np), di(np, npar), target_i(np), grad(5, nat), ip(np), im(np), p2(3, nat)
   real(c_double), target :: o2(nat), b2(nat)
   complex(c_double_complex), target :: fcalc(np), cfact(0:9999, nscat)
   integer(c_int), target :: types(nat)
   real(c_double) :: gl_data(n1, n2, n3, npar), chi2, s, h(npar), fd, err_fd, err_dual, dchi_adj, dchi_col
   type(dc_model_desc) :: md
   type(c_ptr) :: ctx
   integer :: i, j, k, l, a, st, q

   ! model: nc^3 orthorhombic cells of 4 sites, 2 types
   a = 0
   do k = 0, nc - 1
      do j = 0, nc - 1
         do i = 0, nc - 1
            do l = 1, nsite
               a = a + 1
               pos(:, a) = site(:, l) + [real(i, c_double), real(j, c_double), real(k, c_double)]
               types(a) = 1 + mod(l, 2)
               occ(a) = 0.85d0 + 0.01d0 * mod(a, 7)
               bdw(a) = 0.4d0 + 0.05d0 * mod(a, 5)
            end do
         end do
      end do
   end do
   ! form factors f_t(s) = c_t + a_t exp(-b_t s^2) on the s = q * 0.001 table
   do q = 0, 9999
      s = (q * 0.001d0)**2
      cfact(q, 1) = cmplx(1.5d0 + 9.0d0 * exp(-12.0d0 * s), 0.0d0, c_double_complex)
      cfact(q, 2) = cmplx(0.8d0 + 5.0d0 * exp(-20.0d0 * s), 0.0d0, c_double_complex)
   end do
   md%pos = c_loc(pos); md%type = c_loc(types); md%occ = c_loc(occ); md%bdw = c_loc(bdw)
   md%natoms = nat; md%nscat = nscat; md%ncfact = 10000; md%cfact = c_loc(cfact)
   md%eck = 0.0d0
   md%eck(1:3) = [-1.5d0, -1.2d0, -0.9d0]
   md%vi = [0.3d0, 0.0d0, 0.0d0, 0.0d0, 0.3d0, 0.0d0, 0.0d0, 0.0d0, 0.3d0]
   md%inc = [n1, n2, n3]
   md%rten = [1.0d0 / abc(1)**2, 0.0d0, 0.0d0, 0.0d0, 1.0d0 / abc(2)**2, 0.0d0, 0.0d0, 0.0d0, 1.0d0 / abc(3)**2]
   ! REFINE parameters: theta_1 shifts the type-1 atoms along a (fractional),
   ! theta_2 scales the type-2 occupancies, theta_3 adds to every B
   dparam = 0.0d0
   do a = 1, nat
      if (types(a) == 1) dparam(1, a, 1) = 1.0d0
      if (types(a) == 2) dparam(4, a, 2) = occ(a)
      dparam(5, a, 3) = 1.0d0
   end do
   h = [1.0d-3, 1.0d-3, 1.0d-2]

   st = dc_create(ctx, -1_c_int)
   if (st /= DC_OK) call fail('dc_create', st)
   st = dc_set_precision(ctx, DC_PREC_FP32)
   st = dc_model_intensity(ctx, md, c_loc(icalc), c_loc(fcalc))
   if (st /= DC_OK) call fail('dc_model_intensity', st)
   ! the derivative columns of all REFINE parameters in one call
   st = dc_model_jvp(ctx, md, c_loc(fcalc), int(npar, c_int), c_loc(dparam), c_loc(di), c_null_ptr)
   if (st /= DC_OK) call fail('dc_model_jvp', st)
   do k = 1, npar
      gl_data(:, :, :, k) = reshape(di(:, k), [n1, n2, n3])
   end do
   ! each column against central differences of the intensity
   err_fd = 0.0d0
   md%pos = c_loc(p2); md%occ = c_loc(o2); md%bdw = c_loc(b2)
   do k = 1, npar
      p2 = pos + h(k) * dparam(1:3, :, k); o2 = occ + h(k) * dparam(4, :, k); b2 = bdw + h(k) * dparam(5, :, k)
      st = dc_model_intensity(ctx, md, c_loc(ip), c_null_ptr)
      p2 = pos - h(k) * dparam(1:3, :, k); o2 = occ - h(k) * dparam(4, :, k); b2 = bdw - h(k) * dparam(5, :, k)
      st = dc_model_intensity(ctx, md, c_loc(im), c_null_ptr)
      fd = 0.0d0
      do i = 1, np
         fd = max(fd, abs((ip(i) - im(i)) / (2.0d0 * h(k))))
      end do
      err_fd = max(err_fd, maxval(abs(reshape(gl_data(:, :, :, k), [np]) - (ip - im) / (2.0d0 * h(k)))) / fd)
   end do
   md%pos = c_loc(pos); md%occ = c_loc(occ); md%bdw = c_loc(bdw)
   ! chi^2 and its gradient by the adjoint; d chi^2 / d theta_k both ways
   target_i = 1.05d0 * icalc
   st = dc_model_chi2(ctx, md, c_loc(target_i), c_null_ptr, chi2, c_null_ptr, c_loc(grad))
   if (st /= DC_OK) call fail('dc_model_chi2', st)
   err_dual = 0.0d0
   do k = 1, npar
      dchi_adj = sum(grad * dparam(:, :, k))
      dchi_col = 2.0d0 * sum((icalc - target_i) * di(:, k))
      err_dual = max(err_dual, abs(dchi_adj - dchi_col) / abs(dchi_col))
   end do
   call dc_destroy(ctx)
   if (err_fd <= 2.0d-3 .and. err_dual <= 1.0d-5) then
      write(*, '(a,es10.3,a,es10.3)') 'PASS refine columns vs central differences ', err_fd, &
         ', adjoint vs columns ', err_dual
   else
      write(*, '(a,es10.3,a,es10.3)') 'FAIL refine columns vs central differences ', err_fd, &
         ', adjoint vs columns ', err_dual
      call exit(1)
   end if

contains

   subroutine fail(what, st)
      character(len=*), intent(in) :: what
      integer, intent(in) :: st
      write(*, '(a,a,a,i6)') 'FAIL ', what, ' status ', st
      call exit(1)
   end subroutine fail

end program example_refine
