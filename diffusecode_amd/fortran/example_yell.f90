!*******************************************************************************
! example_yell.f90 -- a Fortran host writing a Yell 1.0 HDF5 file through
! diffuse_turbo_mod's dc_write_yell (libdiffuse_yell.so), as DISCUS's output
! menu would: q(i,j,k) = i + 10 j + 100 k on a 3 x 2 x 2 grid, a 3D-PDF value
! code (is_direct = 1), eck(:,1) = (0.1, 0.2, 0.3), vi = 2..10 column-major,
! cell 4 5 6 / 90 90 120.  The file name is the first argument.
! Prints PASS/FAIL; exit status 0/1.
!*******************************************************************************
program example_yell
   use iso_c_binding
   use diffuse_turbo_mod, only: dc_write_yell, dc_yell_desc
   implicit none
   type(dc_yell_desc) :: yd
   real(c_double), target :: q(3, 2, 2)
   character(len=512) :: path
   integer :: i, j, k, st

   call get_command_argument(1, path)
   do k = 1, 2
      do j = 1, 2
         do i = 1, 3
            q(i, j, k) = real(i + 10 * j + 100 * k, c_double)
         end do
      end do
   end do
   yd%inc = [3, 2, 2]
   yd%eck = [(0.1d0 * i, i = 1, 12)]
   yd%vi = [(1.0d0 + i, i = 1, 9)]
   yd%extr = [1, 2, 3]
   yd%a0 = [4.0d0, 5.0d0, 6.0d0]
   yd%win = [90.0d0, 90.0d0, 120.0d0]
   yd%value = 15
   yd%valmax = 0.0d0
   st = dc_write_yell(trim(path) // c_null_char, yd, c_loc(q), c_null_ptr)
   if (st /= 0) then
      write(*, '(a,i6)') 'FAIL dc_write_yell status ', st
      call exit(1)
   end if
   write(*, '(a)') 'PASS fortran yell'
end program example_yell
