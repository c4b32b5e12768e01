!*******************************************************************************
! ref_strucf_bind.f90 -- TEST INFRASTRUCTURE ONLY.
!
! A C-callable entry into DISCUS's own four_strucf
! (discus/prog/four_strucf_OMP_mod.f90:19-68, with four_strucf_omp :72-188,
! four_strucf_serial and the 1-D variants, and tcsf_form_generic,
! discus/prog/fourier_form.f90:12-57), compiled from the reference sources by
! oracle/build_ref.py into oracle/_ref/libref_strucf.so.  It sets the
! diffuse_mod globals that the kernel reads (xat, nxat, uin, vin, win,
! off_shift, lmn, num, cfact, istl, four_dbw, cex) from its arguments, calls
! the kernel and copies tcsf out.
!
! The complex exponent table cex is filled as four_cexpt fills it
! (discus/prog/fourier_sup.f90:2602-2632): fourier_sup.f90 itself needs the
! FFTW3 and HDF5 Fortran interfaces, which this image does not have, so that
! one loop is restated here.
!
! variant: 0 = the dispatcher four_strucf (OpenMP or serial, 3-D/2-D or 1-D
! by fnum, as DISCUS chooses), 1 = four_strucf_serial, 2 = four_strucf_omp.
!*******************************************************************************
module ref_strucf_bind
   use iso_c_binding
   implicit none
contains
   subroutine ref_strucf(nxat_c, xat_c, uin_c, vin_c, win_c, off_c, lmn_c, num_c, fnum_c, ncfact, cfact_c, &
                         istl_c, lform_c, nthreads, variant, tcsf_c) bind(C, name="ref_strucf")
      use diffuse_mod
      use four_strucf_mod
      use parallel_mod
      use precision_mod
      use wink_mod, only: zpi
      !$ use omp_lib
      integer(c_int), value, intent(in) :: nxat_c, ncfact, lform_c, nthreads, variant
      real(c_double), intent(in) :: xat_c(nxat_c, 3), uin_c(3), vin_c(3), win_c(3), off_c(3, 3)
      integer(c_int), intent(in) :: lmn_c(6), num_c(3), fnum_c(3)
      complex(c_double_complex), intent(in) :: cfact_c(0:ncfact - 1)
      integer(c_int), intent(in) :: istl_c(num_c(1), num_c(2), num_c(3))
      complex(c_double_complex), intent(out) :: tcsf_c(num_c(1), num_c(2), num_c(3))
      integer :: i
      integer :: fnum(3)
      real(PREC_DP) :: xmult, xarg
      logical :: lform

      do i = 0, int(MASK)
         xmult = (dble(i) * 1.0d0) / dble(I2PI)
         xarg = zpi * xmult
         cex(i) = cmplx(dble(cos(xarg)), dble(sin(xarg)), kind=kind(0.0d0))
      enddo

      nxat = nxat_c
      if (allocated(xat)) deallocate(xat)
      allocate(xat(max(nxat, 1), 3))
      xat = 0.0d0
      if (nxat > 0) xat(1:nxat, :) = xat_c
      uin = uin_c
      vin = vin_c
      win = win_c
      off_shift = off_c
      lmn = lmn_c
      num = num_c
      fnum = fnum_c
      if (allocated(tcsf)) deallocate(tcsf)
      allocate(tcsf(num(1), num(2), num(3)))
      if (allocated(istl)) deallocate(istl)
      allocate(istl(num(1), num(2), num(3)))
      istl = istl_c
      if (allocated(cfact)) deallocate(cfact)
      allocate(cfact(0:ncfact - 1, 1))
      cfact(:, 1) = cfact_c
      if (allocated(four_dbw)) deallocate(four_dbw)
      allocate(four_dbw(1, 1, 1, 1))
      four_dbw = 0.0d0
      lform = lform_c /= 0

      par_omp_use = .true.
      par_omp_maxthreads = max(nthreads, 1)
      par_omp_phys = max(nthreads, 1)
      !$ call omp_set_num_threads(max(nthreads, 1))

      select case (variant)
      case (1)
         call four_strucf_serial(1, lform, .false., .false., 1, 1, fnum)
      case (2)
         call four_strucf_omp(1, lform, .false., .false., 1, 1, fnum)
      case default
         call four_strucf(1, lform, .false., .false., 1, 1, fnum)
      end select
      tcsf_c = tcsf
   end subroutine ref_strucf
end module ref_strucf_bind

!*******************************************************************************
! ref_fourier_lmn -- DISCUS's fourier_lmn (discus/prog/fourier_lmn_mod.f90:4-279):
! the integer corner multiples lmn and the offset vectors off of a grid
! (eck(3,4), vi(3,3), inc(3)).  Returns DISCUS's ier_num (0 = success).
!*******************************************************************************
module ref_lmn_bind
   use iso_c_binding
   implicit none
contains
   integer(c_int) function ref_fourier_lmn(eck_c, vi_c, inc_c, lmn_c, off_c) bind(C, name="ref_fourier_lmn")
      use fourier_lmn_mod
      use errlist_mod
      real(c_double), intent(in) :: eck_c(3, 4), vi_c(3, 3)
      integer(c_int), intent(in) :: inc_c(3)
      integer(c_int), intent(out) :: lmn_c(6)
      real(c_double), intent(out) :: off_c(3, 3)
      integer :: inc(3), lmn(6)
      real(kind=8) :: eck(3, 4), vi(3, 3), off(3, 3)
      eck = eck_c
      vi = vi_c
      inc = inc_c
      ier_num = 0
      lmn = 0
      off = 0.0d0
      call fourier_lmn(eck, vi, inc, lmn, off)
      lmn_c = lmn
      off_c = off
      ref_fourier_lmn = ier_num
   end function ref_fourier_lmn
end module ref_lmn_bind
