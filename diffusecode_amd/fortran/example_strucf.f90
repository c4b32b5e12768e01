!*******************************************************************************
! example_strucf.f90 -- a Fortran host calling the MI355X path through
! diffuse_turbo_mod (ISO_C_BINDING), the way DISCUS's four_strucf would.
! A perfect N1 x N2 x N3 lattice is transformed on a 2-D layer and checked
! against the closed form F(h) = prod_a exp(i pi h_a (N_a-1)) sin(pi N_a h_a)
! / sin(pi h_a).  Prints PASS/FAIL with max|dI|/max(I); exit status 0/1.
!*******************************************************************************
program example_strucf
   use iso_c_binding
   use diffuse_turbo_mod
   implicit none
   integer, parameter :: n1 = 6, n2 = 5, n3 = 4, na = 31, nb = 21
   integer, parameter :: nat = n1 * n2 * n3
   real(c_double), parameter :: pi = 3.14159265358979323846d0
   real(c_double) :: xat(nat, 3), uin(3), vin(3), win(3), off(3, 3), xm(3), h(3)
   integer :: lmn(6), num(3), fnum(3), istl(na, nb, 1)
   complex(c_double_complex) :: cfact(0:9999, 1), tcsf(na, nb, 1), d
   integer :: i, j, k, a, ier_num, ier_typ, prec
   real(c_double) :: err, imax, bar
   character(len=16) :: arg

   prec = DC_PREC_FP32
   bar = 1.0d-6
   if (command_argument_count() > 0) then
      call get_command_argument(1, arg)
      if (trim(arg) == 'bf16x3') then
         prec = DC_PREC_BF16X3
         bar = 1.0d-5
      else if (trim(arg) == 'fp16x3') then
         prec = DC_PREC_FP16X3
         bar = 1.0d-5
      end if
   end if
   a = 0
   do k = 0, n3 - 1
      do j = 0, n2 - 1
         do i = 0, n1 - 1
            a = a + 1
            xat(a, :) = [real(i, c_double), real(j, c_double), real(k, c_double)]
         end do
      end do
   end do
   xm = [-1.03d0, -0.47d0, 0.2d0]
   uin = [0.037d0, 0.0d0, 0.0d0]
   vin = [0.0d0, 0.051d0, 0.0d0]
   win = 0.0d0
   ! corner as lmn(4:6) times the columns of off_shift (four_strucf_OMP_mod.f90:163-165)
   off = 0.0d0
   off(1, 1) = xm(1); off(2, 2) = xm(2); off(3, 3) = xm(3)
   lmn = [0, 0, 0, 1, 1, 1]
   num = [na, nb, 1]
   fnum = num
   cfact = (1.0d0, 0.0d0)
   istl = 0
   ier_num = 0
   ier_typ = 0
   call dc_four_strucf(1, .false., fnum, xat, nat, uin, vin, win, off, lmn, num, cfact, istl, tcsf, &
                       ier_num, ier_typ, prec)
   if (ier_num /= 0) then
      write(*, '(a,i6,a,i4)') 'FAIL ier_num=', ier_num, ' ier_typ=', ier_typ
      call exit(1)
   end if
   err = 0.0d0
   imax = 0.0d0
   do j = 1, nb
      do i = 1, na
         h = xm + (i - 1) * uin + (j - 1) * vin
         d = exp(cmplx(0.0d0, pi * h(1) * (n1 - 1), c_double_complex)) * sin(pi * n1 * h(1)) / sin(pi * h(1)) &
           * exp(cmplx(0.0d0, pi * h(2) * (n2 - 1), c_double_complex)) * sin(pi * n2 * h(2)) / sin(pi * h(2)) &
           * exp(cmplx(0.0d0, pi * h(3) * (n3 - 1), c_double_complex)) * sin(pi * n3 * h(3)) / sin(pi * h(3))
         err = max(err, abs(abs(tcsf(i, j, 1))**2 - abs(d)**2))
         imax = max(imax, abs(d)**2)
      end do
   end do
   call dc_shutdown()
   if (err / imax <= bar) then
      write(*, '(a,es10.3)') 'PASS fortran drop-in max|dI|/max(I) = ', err / imax
   else
      write(*, '(a,es10.3)') 'FAIL fortran drop-in max|dI|/max(I) = ', err / imax
      call exit(1)
   end if
end program example_strucf
