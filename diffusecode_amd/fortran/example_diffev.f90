!*******************************************************************************
! example_diffev.f90 -- a Fortran host scoring a DIFFEV generation through
! diffuse_turbo_mod in one call (diffev/prog/diffev_kdo.f90:676 runs one
! DISCUS macro per member over MPI instead): three trial structures of
! different sizes, each with its own metric, form factors and target layer,
! concatenated member after member into one dc_batch_desc; chi^2 per member
! and the gradient of every atom from dc_batch_chi2.  Two generations are
! scored (the second replays the captured graph with new positions).
! Check: each member's chi^2 and gradient against dc_model_chi2 of that
! member alone (relative 1e-4 and 1e-3 of max|g|).  Prints PASS/FAIL; exit
! status 0/1.
!*******************************************************************************
program example_diffev
   use iso_c_binding
   use diffuse_turbo_mod
   implicit none
   integer, parameter :: nm = 3, nscat = 2, n1 = 24, n2 = 20, np = n1 * n2
   integer, parameter :: ncell(nm) = [3, 4, 5]
   real(c_double), parameter :: acell(nm) = [4.0d0, 4.5d0, 5.2d0]
   integer(c_int64_t), target :: atom_beg(0:nm)
   real(c_double), allocatable, target :: pos(:, :), occ(:), grad(:, :), gm(:, :)
   integer(c_int), allocatable, target :: types(:)
   real(c_double), target :: rten(3, 3, nm), tgt(np, nm), chi2(nm), icalc(np, nm), im(np), chim
   complex(c_double_complex), target :: cfact(0:9999, nscat, nm)
   type(dc_batch_desc) :: bd
   type(dc_model_desc) :: md
   type(c_ptr) :: ctx
   integer :: m, i, j, k, a, q, st, nat, gen, a0, a1
   integer(c_int64_t) :: seed
   real(c_double) :: s, err_chi, err_grad

   nat = sum(2 * ncell**3)
   allocate(pos(3, nat), occ(nat), types(nat), grad(5, nat), gm(5, nat))
   ! members: cubic cells of 2 sites (bcc-like), sizes 3^3, 4^3, 5^3, small seeded displacements
   seed = 12345
   a = 0
   atom_beg(0) = 0
   do m = 1, nm
      do k = 0, ncell(m) - 1
         do j = 0, ncell(m) - 1
            do i = 0, ncell(m) - 1
               do q = 0, 1
                  a = a + 1
                  pos(:, a) = [i + 0.5d0 * q, j + 0.5d0 * q, k + 0.5d0 * q] + 0.01d0 * [rnd(), rnd(), rnd()]
                  types(a) = 1 + q
                  occ(a) = 0.9d0 + 0.1d0 * rnd()
               end do
            end do
         end do
      end do
      atom_beg(m) = a
      rten(:, :, m) = 0.0d0
      do i = 1, 3
         rten(i, i, m) = 1.0d0 / acell(m)**2
      end do
      do q = 0, 9999
         s = (q * 0.001d0)**2
         cfact(q, 1, m) = cmplx((1.0d0 + 0.2d0 * m) + 8.0d0 * exp(-10.0d0 * s), 0.0d0, c_double_complex)
         cfact(q, 2, m) = cmplx(0.6d0 + 4.0d0 * exp(-15.0d0 * s), 0.0d0, c_double_complex)
      end do
   end do
   bd%nmembers = nm; bd%atom_beg = c_loc(atom_beg); bd%pos = c_loc(pos); bd%type = c_loc(types)
   bd%occ = c_loc(occ); bd%bdw = c_null_ptr; bd%nscat = nscat; bd%ncfact = 10000
   bd%cfact = c_loc(cfact); bd%rten = c_loc(rten)
   bd%eck = 0.0d0
   bd%eck(1:3) = [-2.3d0, -1.9d0, 0.25d0]
   bd%vi = [0.2d0, 0.0d0, 0.0d0, 0.0d0, 0.2d0, 0.0d0, 0.0d0, 0.0d0, 0.0d0]
   bd%inc = [n1, n2, 1]

   st = dc_create(ctx, -1_c_int)
   if (st /= DC_OK) call fail('dc_create', st)
   st = dc_set_precision(ctx, DC_PREC_FP32)
   ! targets: every member's own intensity, scaled per pixel (a measured layer stands in here)
   tgt = 0.0d0
   st = dc_batch_chi2(ctx, bd, c_loc(tgt), c_null_ptr, c_loc(chi2), c_loc(icalc), c_null_ptr)
   if (st /= DC_OK) call fail('dc_batch_chi2 (targets)', st)
   do m = 1, nm
      do i = 1, np
         tgt(i, m) = icalc(i, m) * (1.0d0 + 0.3d0 * sin(0.7d0 * i + m))
      end do
   end do
   err_chi = 0.0d0
   err_grad = 0.0d0
   do gen = 1, 2
      if (gen == 2) pos = pos + 0.002d0
      st = dc_batch_chi2(ctx, bd, c_loc(tgt), c_null_ptr, c_loc(chi2), c_null_ptr, c_loc(grad))
      if (st /= DC_OK) call fail('dc_batch_chi2', st)
      do m = 1, nm
         a0 = int(atom_beg(m - 1)) + 1
         a1 = int(atom_beg(m))
         md%pos = c_loc(pos(1, a0)); md%type = c_loc(types(a0)); md%occ = c_loc(occ(a0)); md%bdw = c_null_ptr
         md%natoms = a1 - a0 + 1; md%nscat = nscat; md%ncfact = 10000; md%cfact = c_loc(cfact(0, 1, m))
         md%eck = bd%eck; md%vi = bd%vi; md%inc = bd%inc; md%rten = reshape(rten(:, :, m), [9])
         im = tgt(:, m)
         st = dc_model_chi2(ctx, md, c_loc(im), c_null_ptr, chim, c_null_ptr, c_loc(gm))
         if (st /= DC_OK) call fail('dc_model_chi2', st)
         err_chi = max(err_chi, abs(chi2(m) - chim) / chim)
         err_grad = max(err_grad, maxval(abs(grad(1:4, a0:a1) - gm(1:4, 1:a1 - a0 + 1))) / &
                                  maxval(abs(gm(1:4, 1:a1 - a0 + 1))))
      end do
   end do
   call dc_destroy(ctx)
   ! batch and single-member runs slice and sum in different orders: both
   ! carry the FP32 forward's rounding (~1e-7 of max I), which the residuals
   ! (30% of I) turn into the relative differences allowed here
   if (err_chi <= 1.0d-4 .and. err_grad <= 1.0d-3) then
      write(*, '(a,es10.3,a,es10.3)') 'PASS diffev batch vs single-member chi2 ', err_chi, ', gradient ', err_grad
   else
      write(*, '(a,es10.3,a,es10.3)') 'FAIL diffev batch vs single-member chi2 ', err_chi, ', gradient ', err_grad
      call exit(1)
   end if

contains

   real(c_double) function rnd()
      seed = mod(seed * 1103515245_c_int64_t + 12345_c_int64_t, 2147483648_c_int64_t)
      rnd = real(seed, c_double) / 2147483648.0d0 - 0.5d0
   end function rnd

   subroutine fail(what, st)
      character(len=*), intent(in) :: what
      integer, intent(in) :: st
      write(*, '(a,a,a,i6)') 'FAIL ', what, ' status ', st
      call exit(1)
   end subroutine fail

end program example_diffev
