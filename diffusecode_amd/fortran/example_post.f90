!*******************************************************************************
! example_post.f90 -- a Fortran host calling the post-processing and NUFFT
! entry points through diffuse_turbo_mod (ISO_C_BINDING descriptors), each
! against a closed form:
!   dc_nufft1      one point: fk(k) = c exp(i k x)
!   dc_four_post   resolution convolution of a delta on an odd line: the
!                  normalised Gaussian profile centred on the delta
!   dc_four_fill   a 2-fold axis on a centred layer: (x + x_rotated) / 2 / 2
!   dc_qval        intensity (value 1) returns dsi
! Prints PASS/FAIL; exit status 0/1.
!*******************************************************************************
program example_post
   use iso_c_binding
   use diffuse_turbo_mod
   implicit none
   integer, parameter :: nk = 9, nl = 21, n1 = 7, n2 = 5
   real(c_double), parameter :: pi = 3.14159265358979323846d0
   type(c_ptr) :: ctx
   type(dc_post_desc) :: pd
   type(dc_symm_desc) :: sd
   type(dc_qval_desc) :: qd
   real(c_double), target :: x(1), y(1), z(1), mats(9)
   complex(c_double_complex), target :: cw(1), fk(nk), csf(nl), f2(n1, n2, 1), csf_q(n1, n2, 1)
   real(c_double), target :: dsi(nl), d2(n1, n2, 1), q(n1, n2, 1)
   complex(c_double_complex), allocatable :: f2ref(:, :, :)
   real(c_double) :: err, s, prof(nl)
   integer :: k, i, j, st, fails

   fails = 0
   st = dc_create(ctx, 0_c_int)
   if (st /= 0) then
      write(*, '(a,i6)') 'FAIL dc_create ', st
      call exit(1)
   end if

   ! --- dc_nufft1: one point
   x = 0.37d0
   y = 0.0d0
   z = 0.0d0
   cw = (2.0d0, -0.5d0)
   st = dc_nufft1(ctx, 1_c_int, 1_c_int64_t, c_loc(x), c_null_ptr, c_null_ptr, c_loc(cw), 1_c_int, 1.0d-9, &
                  int(nk, c_int64_t), 1_c_int64_t, 1_c_int64_t, c_loc(fk))
   err = 0.0d0
   do k = 1, nk
      err = max(err, abs(fk(k) - cw(1) * exp(cmplx(0.0d0, (k - 1 - nk / 2) * x(1), c_double_complex))))
   end do
   call report('dc_nufft1 one point', st, err / abs(cw(1)), 1.0d-8)

   ! --- dc_four_post: Gaussian resolution of a delta on an odd line
   pd%num = [nl, 1, 1]
   pd%eck = 0.0d0
   pd%vi = 0.0d0
   pd%vi(1) = 0.1d0
   pd%rten = [1.0d0, 0.0d0, 0.0d0, 0.0d0, 1.0d0, 0.0d0, 0.0d0, 0.0d0, 1.0d0]
   pd%vr = 1.0d0
   pd%fave = 1.0d0
   pd%nlots = 1
   pd%icc = 1
   pd%ls = 1
   pd%lots_off = 1
   pd%diff_scale = 1.0d0
   pd%res_sigma = [1.0d0, 1.0d0, 1.0d0]
   pd%res_tr = [0.5d0, 0.0d0, 0.0d0, 0.0d0, 0.5d0, 0.0d0, 0.0d0, 0.0d0, 0.5d0]
   pd%filter = 0
   pd%rscale = 1.0d0
   pd%damp = 0.5d0
   pd%width = 4
   dsi = 0.0d0
   dsi(nl / 2 + 1) = 1.0d0
   csf = (0.0d0, 0.0d0)
   st = dc_four_post(ctx, pd, c_loc(csf), c_loc(dsi), c_null_ptr)
   do k = 1, nl
      i = k - 1 - nl / 2
      prof(k) = exp(-0.5d0 * (0.5d0 * i)**2)
   end do
   s = sum(prof)
   err = maxval(abs(dsi - prof / s))
   call report('dc_four_post four_conv delta', st, err, 1.0d-12)

   ! --- dc_four_fill: 2-fold axis on a centred layer
   mats = [-1.0d0, 0.0d0, 0.0d0, 0.0d0, -1.0d0, 0.0d0, 0.0d0, 0.0d0, 1.0d0]
   sd%num = [n1, n2, 1]
   sd%eck = 0.0d0
   sd%eck(1) = -0.1d0 * (n1 / 2)
   sd%eck(2) = -0.1d0 * (n2 / 2)
   sd%vi = 0.0d0
   sd%vi(1) = 0.1d0
   sd%vi(5) = 0.1d0
   sd%fave = 0.0d0
   sd%four_symm = 1
   sd%nops = 2
   sd%mats = c_loc(mats)
   do j = 1, n2
      do i = 1, n1
         f2(i, j, 1) = cmplx(i + 10 * j, i * j, c_double_complex)
         d2(i, j, 1) = i + 0.5d0 * j * j
      end do
   end do
   allocate(f2ref(n1, n2, 1))
   do j = 1, n2
      do i = 1, n1
         f2ref(i, j, 1) = (f2(i, j, 1) + f2(n1 + 1 - i, n2 + 1 - j, 1)) / 2.0d0 / 2.0d0
      end do
   end do
   st = dc_four_fill(ctx, sd, c_loc(f2), c_loc(d2))
   err = maxval(abs(f2 - f2ref))
   call report('dc_four_fill 2-fold', st, err, 1.0d-14)

   ! --- dc_qval: intensity returns dsi
   qd%num = [n1, n2, 1]
   qd%eck = sd%eck
   qd%vi = sd%vi
   qd%rten = pd%rten
   qd%value = 1
   qd%laver = 0
   qd%cfact = c_null_ptr
   qd%cfact_pure = c_null_ptr
   qd%ncfact = 0
   qd%nscat = 0
   qd%amount = c_null_ptr
   qd%n_real_atoms = 1.0d0
   qd%u2aver = 0.0d0
   qd%faver_sign = 1.0d0
   csf_q = (0.0d0, 0.0d0)
   st = dc_qval(ctx, qd, c_loc(csf_q), c_null_ptr, c_loc(d2), c_loc(q))
   err = maxval(abs(q - d2))
   call report('dc_qval intensity', st, err, 0.0d0)

   call dc_destroy(ctx)
   if (fails == 0) then
      write(*, '(a)') 'PASS fortran post/nufft'
   else
      call exit(1)
   end if

contains

   subroutine report(what, st, err, bar)
      character(len=*), intent(in) :: what
      integer, intent(in) :: st
      real(c_double), intent(in) :: err, bar
      if (st /= 0 .or. .not. (err <= bar)) then
         write(*, '(a,a,a,i6,a,es10.3)') 'FAIL ', what, ' status ', st, ' err ', err
         fails = fails + 1
      else
         write(*, '(a,a,a,es10.3)') 'ok   ', what, ' err ', err
      end if
   end subroutine report

end program example_post
