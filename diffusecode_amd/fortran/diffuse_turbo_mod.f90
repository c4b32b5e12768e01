!*******************************************************************************
! diffuse_turbo_mod.f90 -- ISO_C_BINDING interface of libdiffuse_turbo.so
! (include/diffuse_turbo.h) for a Fortran host such as DISCUS.
!
! Arrays are passed exactly as DISCUS's diffuse_mod holds them
! (discus/prog/diffuse_mod.f90:52-78): complex(8) tcsf/csf/acsf(num(1),num(2),
! num(3)), real(8) dsi(...), integer istl(...), real(8) xat(NMAX,3),
! complex(8) cfact(0:CFPKT,nscat).
!
! dc_four_strucf is the per-type drop-in: it has the arguments of the
! four_strucf interface (discus/prog/four_strucf_OMP_mod.f90:19) plus the
! module arrays that four_strucf reads implicitly, and maps the library
! status onto DISCUS's ier_num / ier_typ.
!*******************************************************************************
module diffuse_turbo_mod
   use iso_c_binding
   implicit none
   private

   integer(c_int), parameter, public :: DC_OK = 0
   integer(c_int), parameter, public :: DC_PREC_FP32 = 0, DC_PREC_BF16X3 = 1, DC_PREC_FP16X2 = 2, DC_PREC_FP16X3 = 3
   integer(c_int), parameter, public :: DC_LOT_OFF = 1, DC_LOT_BOX = 2, DC_LOT_ELI = 3

   ! dc_run_desc of include/diffuse_turbo.h
   type, bind(C), public :: dc_run_desc
      real(c_double)     :: eck(12)
      real(c_double)     :: vi(9)
      integer(c_int)     :: inc(3)
      real(c_double)     :: rten(9)
      type(c_ptr)        :: cfact
      integer(c_int)     :: ncfact
      integer(c_int)     :: ilots
      integer(c_int)     :: ls(3)
      integer(c_int)     :: nlots
      type(c_ptr)        :: lbeg
      integer(c_int)     :: lperiod
      real(c_double)     :: fave
      type(c_ptr)        :: sel_cell
      integer(c_int)     :: four_friedel
      integer(c_int)     :: ano
      real(c_double)     :: diff_scale
   end type dc_run_desc

   type, bind(C), public :: dc_crystal_desc
      type(c_ptr)        :: cr_pos
      type(c_ptr)        :: cr_iscat
      integer(c_int64_t) :: natoms
      integer(c_int)     :: icc(3)
      integer(c_int)     :: ncatoms
      integer(c_int)     :: nscat
      real(c_double)     :: dim0(3)
      real(c_double)     :: dim_lo(3)
   end type dc_crystal_desc

   ! dc_post_desc / dc_symm_desc / dc_qval_desc of include/diffuse_turbo.h
   type, bind(C), public :: dc_post_desc
      integer(c_int)     :: num(3)
      real(c_double)     :: eck(12)
      real(c_double)     :: vi(9)
      real(c_double)     :: rten(9)
      real(c_double)     :: vr
      real(c_double)     :: fave
      integer(c_int)     :: nlots
      integer(c_int)     :: icc(3)
      integer(c_int)     :: ls(3)
      integer(c_int)     :: lots_off
      real(c_double)     :: diff_scale
      real(c_double)     :: res_sigma(3)
      real(c_double)     :: res_tr(9)
      integer(c_int)     :: filter
      real(c_double)     :: rscale
      real(c_double)     :: damp
      integer(c_int)     :: width
   end type dc_post_desc

   type, bind(C), public :: dc_symm_desc
      integer(c_int)     :: num(3)
      real(c_double)     :: eck(12)
      real(c_double)     :: vi(9)
      real(c_double)     :: fave
      integer(c_int)     :: four_symm
      integer(c_int)     :: nops
      type(c_ptr)        :: mats
   end type dc_symm_desc

   type, bind(C), public :: dc_qval_desc
      integer(c_int)     :: num(3)
      real(c_double)     :: eck(12)
      real(c_double)     :: vi(9)
      real(c_double)     :: rten(9)
      integer(c_int)     :: value
      integer(c_int)     :: laver
      type(c_ptr)        :: cfact
      type(c_ptr)        :: cfact_pure
      integer(c_int)     :: ncfact
      integer(c_int)     :: nscat
      type(c_ptr)        :: amount
      real(c_double)     :: n_real_atoms
      real(c_double)     :: u2aver
      real(c_double)     :: faver_sign
   end type dc_qval_desc

   ! dc_model_desc of include/diffuse_turbo.h: one refinement model (REFINE's
   ! parameters per atom: positions, occupancies, isotropic B)
   type, bind(C), public :: dc_model_desc
      type(c_ptr)        :: pos            ! real(8) pos(3, natoms), fractional
      type(c_ptr)        :: type           ! integer type(natoms), 1-based, 0 = vacancy
      type(c_ptr)        :: occ            ! real(8) occ(natoms)
      type(c_ptr)        :: bdw            ! real(8) bdw(natoms) or c_null_ptr
      integer(c_int64_t) :: natoms
      integer(c_int)     :: nscat
      integer(c_int)     :: ncfact         ! rows of cfact (CFPKT + 1)
      type(c_ptr)        :: cfact          ! complex(8) cfact(0:CFPKT, nscat)
      real(c_double)     :: eck(12)
      real(c_double)     :: vi(9)
      integer(c_int)     :: inc(3)
      real(c_double)     :: rten(9)
   end type dc_model_desc

   ! dc_batch_desc of include/diffuse_turbo.h: a DIFFEV generation, members
   ! concatenated atom after atom, one shared hkl grid
   type, bind(C), public :: dc_batch_desc
      integer(c_int)     :: nmembers
      type(c_ptr)        :: atom_beg       ! integer(8) atom_beg(0:nmembers), atom_beg(0) = 0
      type(c_ptr)        :: pos            ! real(8) pos(3, natoms)
      type(c_ptr)        :: type           ! integer type(natoms), 1-based within nscat
      type(c_ptr)        :: occ            ! real(8) occ(natoms)
      type(c_ptr)        :: bdw            ! real(8) bdw(natoms) or c_null_ptr
      integer(c_int)     :: nscat
      integer(c_int)     :: ncfact
      type(c_ptr)        :: cfact          ! complex(8) cfact(0:CFPKT, nscat, nmembers)
      type(c_ptr)        :: rten           ! real(8) rten(3, 3, nmembers)
      real(c_double)     :: eck(12)
      real(c_double)     :: vi(9)
      integer(c_int)     :: inc(3)
   end type dc_batch_desc

   ! Yell 1.0 HDF5 output (include/diffuse_yell.h, libdiffuse_yell.so):
   ! gen_hdf5_write's arguments (lib_f90/hdf_write.f90:12-16)
   type, bind(C), public :: dc_yell_desc
      integer(c_int)     :: inc(3)
      real(c_double)     :: eck(12)
      real(c_double)     :: vi(9)
      integer(c_int)     :: extr(3)
      real(c_double)     :: a0(3)
      real(c_double)     :: win(3)
      integer(c_int)     :: value
      real(c_double)     :: valmax
   end type dc_yell_desc

   interface
      integer(c_int) function dc_create(ctx, device) bind(C, name='dc_create')
         import :: c_ptr, c_int
         type(c_ptr), intent(out) :: ctx
         integer(c_int), value :: device
      end function dc_create

      subroutine dc_destroy(ctx) bind(C, name='dc_destroy')
         import :: c_ptr
         type(c_ptr), value :: ctx
      end subroutine dc_destroy

      integer(c_int) function dc_set_precision(ctx, prec) bind(C, name='dc_set_precision')
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx
         integer(c_int), value :: prec
      end function dc_set_precision

      integer(c_int) function dc_last_error_type(ctx) bind(C, name='dc_last_error_type')
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx
      end function dc_last_error_type

      integer(c_int) function dc_strucf(ctx, xat, nxat, ld_xat, uin, vin, win, off_shift, lmn, num, fnum, &
                                        cfact_col, ncfact, istl, xm, rten, lform, tcsf) bind(C, name='dc_strucf')
         import :: c_ptr, c_int, c_int64_t, c_double, c_double_complex
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: xat(*)
         integer(c_int64_t), value :: nxat, ld_xat
         real(c_double), intent(in) :: uin(3), vin(3), win(3), off_shift(9)
         integer(c_int), intent(in) :: lmn(6), num(3), fnum(3)
         complex(c_double_complex), intent(in) :: cfact_col(*)
         integer(c_int64_t), value :: ncfact
         integer(c_int), intent(in) :: istl(*)
         real(c_double), intent(in) :: xm(3), rten(9)
         integer(c_int), value :: lform
         complex(c_double_complex), intent(out) :: tcsf(*)
      end function dc_strucf

      integer(c_int) function dc_crystal_set(ctx, cr) bind(C, name='dc_crystal_set')
         import :: c_ptr, c_int, dc_crystal_desc
         type(c_ptr), value :: ctx
         type(dc_crystal_desc), intent(in) :: cr
      end function dc_crystal_set

      integer(c_int) function dc_four_run(ctx, run, csf, acsf, dsi, fmean, dvar) bind(C, name='dc_four_run')
         import :: c_ptr, c_int, dc_run_desc
         type(c_ptr), value :: ctx
         type(dc_run_desc), intent(in) :: run
         type(c_ptr), value :: csf, acsf, dsi, fmean, dvar
      end function dc_four_run

      integer(c_int) function dc_four_post(ctx, post, csf, dsi, dsi3d) bind(C, name='dc_four_post')
         import :: c_ptr, c_int, dc_post_desc
         type(c_ptr), value :: ctx
         type(dc_post_desc), intent(in) :: post
         type(c_ptr), value :: csf, dsi, dsi3d
      end function dc_four_post

      integer(c_int) function dc_four_fill(ctx, symm, csf, dsi) bind(C, name='dc_four_fill')
         import :: c_ptr, c_int, dc_symm_desc
         type(c_ptr), value :: ctx
         type(dc_symm_desc), intent(in) :: symm
         type(c_ptr), value :: csf, dsi
      end function dc_four_fill

      integer(c_int) function dc_qval(ctx, q, csf, acsf, dsi, out) bind(C, name='dc_qval')
         import :: c_ptr, c_int, dc_qval_desc
         type(c_ptr), value :: ctx
         type(dc_qval_desc), intent(in) :: q
         type(c_ptr), value :: csf, acsf, dsi, out
      end function dc_qval

      integer(c_int) function dc_nufft1(ctx, dim, m, x, y, z, c, iflag, tol, n1, n2, n3, fk) bind(C, name='dc_nufft1')
         import :: c_ptr, c_int, c_int64_t, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: dim, iflag
         integer(c_int64_t), value :: m, n1, n2, n3
         type(c_ptr), value :: x, y, z, c, fk
         real(c_double), value :: tol
      end function dc_nufft1

      integer(c_int) function dc_four_strucf_nufft(ctx, dim, nat, xpos, ypos, zpos, occ, inc, lscales, tcsf) &
            bind(C, name='dc_four_strucf_nufft')
         import :: c_ptr, c_int, c_int64_t, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: dim
         integer(c_int64_t), value :: nat
         type(c_ptr), value :: xpos, ypos, zpos, tcsf
         real(c_double), value :: occ
         integer(c_int), intent(in) :: inc(3), lscales(3)
      end function dc_four_strucf_nufft

      integer(c_int) function dc_four_run_nufft(ctx, run, type_occ, csf, dsi) bind(C, name='dc_four_run_nufft')
         import :: c_ptr, c_int, dc_run_desc
         type(c_ptr), value :: ctx
         type(dc_run_desc), intent(in) :: run
         type(c_ptr), value :: type_occ, csf, dsi
      end function dc_four_run_nufft

      ! ---- REFINE: the refinement model (refine/prog/refine_run.f90:491-600) ----
      ! icalc (real) and fcalc (complex) over inc(1)*inc(2)*inc(3), c_null_ptr-able
      integer(c_int) function dc_model_intensity(ctx, model, icalc, fcalc) bind(C, name='dc_model_intensity')
         import :: c_ptr, c_int, dc_model_desc
         type(c_ptr), value :: ctx
         type(dc_model_desc), intent(in) :: model
         type(c_ptr), value :: icalc, fcalc
      end function dc_model_intensity

      ! grad(5, natoms) = dL/d(x, y, z, occ, B) for gi = dL/dI per pixel
      integer(c_int) function dc_model_vjp(ctx, model, fcalc, gi, grad) bind(C, name='dc_model_vjp')
         import :: c_ptr, c_int, dc_model_desc
         type(c_ptr), value :: ctx
         type(dc_model_desc), intent(in) :: model
         type(c_ptr), value :: fcalc, gi, grad
      end function dc_model_vjp

      ! di(P, ndir) [and df(P, ndir) complex] = dI/d(theta_k) for the
      ! directions dparam(5, natoms, ndir): REFINE's derivative columns
      integer(c_int) function dc_model_jvp(ctx, model, fcalc, ndir, dparam, di, df) bind(C, name='dc_model_jvp')
         import :: c_ptr, c_int, dc_model_desc
         type(c_ptr), value :: ctx
         type(dc_model_desc), intent(in) :: model
         type(c_ptr), value :: fcalc
         integer(c_int), value :: ndir
         type(c_ptr), value :: dparam, di, df
      end function dc_model_jvp

      ! chi2 = sum w (I - target)^2 (weight c_null_ptr = 1), icalc and grad(5, natoms) c_null_ptr-able
      integer(c_int) function dc_model_chi2(ctx, model, target, weight, chi2, icalc, grad) bind(C, name='dc_model_chi2')
         import :: c_ptr, c_int, c_double, dc_model_desc
         type(c_ptr), value :: ctx
         type(dc_model_desc), intent(in) :: model
         type(c_ptr), value :: target, weight
         real(c_double), intent(out) :: chi2
         type(c_ptr), value :: icalc, grad
      end function dc_model_chi2

      ! ---- DIFFEV: one generation in one call (diffev/prog/diffev_kdo.f90:676) ----
      ! target, weight, icalc (P, nmembers); chi2(nmembers); grad(5, natoms)
      integer(c_int) function dc_batch_chi2(ctx, batch, target, weight, chi2, icalc, grad) bind(C, name='dc_batch_chi2')
         import :: c_ptr, c_int, dc_batch_desc
         type(c_ptr), value :: ctx
         type(dc_batch_desc), intent(in) :: batch
         type(c_ptr), value :: target, weight, chi2, icalc, grad
      end function dc_batch_chi2

      integer(c_int) function dc_set_autotune(ctx, on) bind(C, name='dc_set_autotune')
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx
         integer(c_int), value :: on
      end function dc_set_autotune

      integer(c_int) function dc_set_graphs(ctx, on) bind(C, name='dc_set_graphs')
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx
         integer(c_int), value :: on
      end function dc_set_graphs

      ! four_stltab on the device (fourier_sup.f90:2637-2687)
      integer(c_int) function dc_four_stltab(ctx, xm, uin, vin, win, num, rten, istl) bind(C, name='dc_four_stltab')
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: xm(3), uin(3), vin(3), win(3), rten(9)
         integer(c_int), intent(in) :: num(3)
         integer(c_int), intent(out) :: istl(*)
      end function dc_four_stltab

      ! in libdiffuse_yell.so (link -ldiffuse_yell where it is called)
      integer(c_int) function dc_write_yell(path, desc, qvalues, sigma) bind(C, name='dc_write_yell')
         import :: c_ptr, c_int, c_char, dc_yell_desc
         character(kind=c_char), dimension(*), intent(in) :: path
         type(dc_yell_desc), intent(in) :: desc
         type(c_ptr), value :: qvalues, sigma
      end function dc_write_yell
   end interface

   public :: dc_create, dc_destroy, dc_set_precision, dc_last_error_type, dc_strucf
   public :: dc_crystal_set, dc_four_run, dc_four_post, dc_four_fill, dc_qval
   public :: dc_nufft1, dc_four_strucf_nufft, dc_four_run_nufft, dc_write_yell
   public :: dc_model_intensity, dc_model_vjp, dc_model_jvp, dc_model_chi2
   public :: dc_batch_chi2, dc_set_autotune, dc_set_graphs, dc_four_stltab
   public :: dc_four_strucf, dc_shutdown

   type(c_ptr), save :: the_ctx = c_null_ptr

contains

   !----------------------------------------------------------------------------
   ! Per-type structure factor, the isotropic branch of four_strucf
   ! (four_strucf_OMP_mod.f90:19-68 with tcsf_form_iso, fourier_form.f90:147-177).
   ! tcsf is overwritten over the whole num grid (zero outside fnum).
   subroutine dc_four_strucf(iscat, lform, fnum, xat, nxat, uin, vin, win, off_shift, lmn, num, &
                             cfact, istl, tcsf, ier_num, ier_typ, precision)
      integer, intent(in) :: iscat
      logical, intent(in) :: lform
      integer, dimension(3), intent(in) :: fnum
      real(c_double), dimension(:, :), intent(in) :: xat          ! xat(NMAX,3)
      integer, intent(in) :: nxat
      real(c_double), dimension(3), intent(in) :: uin, vin, win
      real(c_double), dimension(3, 3), intent(in) :: off_shift
      integer, dimension(6), intent(in) :: lmn
      integer, dimension(3), intent(in) :: num
      complex(c_double_complex), dimension(0:, :), intent(in) :: cfact   ! cfact(0:CFPKT, nscat)
      integer, dimension(:, :, :), intent(in) :: istl
      complex(c_double_complex), dimension(:, :, :), intent(inout) :: tcsf
      integer, intent(inout) :: ier_num, ier_typ
      integer, intent(in), optional :: precision

      integer(c_int) :: st, prec
      integer(c_int) :: c_lmn(6), c_num(3), c_fnum(3)
      real(c_double) :: xm_dummy(3), rten_dummy(9)
      integer(c_int64_t) :: ncf

      if (.not. c_associated(the_ctx)) then
         st = dc_create(the_ctx, -1_c_int)
         if (st /= DC_OK) then
            ier_num = st
            ier_typ = 99
            return
         end if
      end if
      prec = DC_PREC_FP32
      if (present(precision)) prec = int(precision, c_int)
      st = dc_set_precision(the_ctx, prec)
      c_lmn = int(lmn, c_int)
      c_num = int(num, c_int)
      c_fnum = int(fnum, c_int)
      xm_dummy = 0.0_c_double
      rten_dummy = 0.0_c_double
      ncf = int(size(cfact, 1), c_int64_t)
      st = dc_strucf(the_ctx, xat, int(nxat, c_int64_t), int(size(xat, 1), c_int64_t), uin, vin, win, &
                     reshape(off_shift, [9]), c_lmn, c_num, c_fnum, cfact(:, max(iscat, 1)), ncf, istl, &
                     xm_dummy, rten_dummy, merge(1_c_int, 0_c_int, lform), tcsf)
      if (st < 0) then
         ier_num = st
         ier_typ = dc_last_error_type(the_ctx)
      end if
   end subroutine dc_four_strucf

   subroutine dc_shutdown()
      if (c_associated(the_ctx)) call dc_destroy(the_ctx)
      the_ctx = c_null_ptr
   end subroutine dc_shutdown

end module diffuse_turbo_mod
