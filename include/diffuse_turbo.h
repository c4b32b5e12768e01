/*
 * diffuse_turbo.h -- C ABI of the MI355X structure-factor library
 * (libdiffuse_turbo.so).  Plain pointers and sizes only; every array uses
 * the Fortran (column-major) layout of DISCUS's diffuse_mod so that a
 * Fortran host can pass its module arrays directly through ISO_C_BINDING.
 *
 * Layout conventions (identical to discus/prog/diffuse_mod.f90:52-78):
 *   complex(8) arrays  -> interleaved (re, im) doubles
 *   tcsf/csf/acsf(num(1),num(2),num(3)), dsi(num(1),num(2),num(3)),
 *   istl(num(1),num(2),num(3)) int32, abscissa (index 1) fastest
 *   xat(NMAX,3)        -> xat[k + ld_xat*c], c = 0,1,2
 *   cr_pos(3,N)        -> cr_pos[3*i + c]
 *   cfact(0:CFPKT,nscat) complex(8) -> cfact[2*(iq + ncfact*(iscat-1))]
 *   eck(3,4), vi(3,3), off_shift(3,3), rten(3,3) -> a[i + 3*j]
 *
 * Status codes: 0 = success; negative values reuse DISCUS's ier_num
 * numbering where DISCUS has one (the matching ier_typ is reported by
 * dc_last_error_type); -1000 and below are library errors.
 * Threading: one context = one HIP stream; calls on one context are not
 * reentrant, separate contexts are independent.  All host-pointer entry
 * points are synchronous on return.
 */
#ifndef DIFFUSE_TURBO_H
#define DIFFUSE_TURBO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define DC_OK 0
#define DC_WARN_NOCELL 1        /* four_aver: no cell selected, <F> set to 0 (fourier_sup.f90:1319-1327), ER_FOUR */
#define DC_ERR_ISTL (-3)        /* sin(theta)/lambda beyond the table (fourier_sup.f90:2676-2679), ER_FOUR */
#define DC_ERR_CANCEL (-14)     /* cancel flag seen between launches (fourier_sup.f90:237-242), ER_COMM */
#define DC_ERR_ZEROVEC (-132)   /* zero increment vector (fourier_lmn_mod.f90:84-91), ER_APPL */
#define DC_ERR_LOTS (-162)      /* lots on but natoms != cells*sites (fourier.f90:646-651), ER_APPL */
#define DC_ERR_NOATOMS (-198)   /* crystal has no atoms (fourier.f90:640-643), ER_APPL */
#define DC_ERR_ARG (-1000)      /* invalid argument (null pointer, size, range) */
#define DC_ERR_DEVICE (-1001)   /* HIP runtime error */
#define DC_ERR_NOMEM (-1002)    /* device allocation failed */
#define DC_ERR_NODEV (-1003)    /* no usable gfx950 device */
#define DC_ERR_STATE (-1004)    /* call order (e.g. run before dc_crystal_set) */

/* ier_typ values of lib_f90/errlist_mod.f90:28-40 */
#define DC_ER_NONE 0
#define DC_ER_COMM 1
#define DC_ER_APPL 6
#define DC_ER_FOUR 9
#define DC_ER_LIB 99            /* library-side errors (no DISCUS equivalent) */

/* ---- arithmetic of the phase GEMM --------------------------------------- */
#define DC_PREC_FP32 0          /* plain FP32 MFMA (the drop-in)            */
#define DC_PREC_BF16X3 1        /* split BF16 hi/lo, 3 MFMA products; runs with
                                   the <F> subtraction or the cross-lot variance
                                   use the FP16X3 products (dc_four_run*)    */
#define DC_PREC_FP16X2 2        /* split FP16 hi/lo on one side, 2 products */
#define DC_PREC_FP16X3 3        /* split FP16 hi/lo on both sides, 3 MFMA products */

/* ---- lot shapes (discus/prog/diffuse_mod.f90:10-12) --------------------- */
#define DC_LOT_OFF 1
#define DC_LOT_BOX 2
#define DC_LOT_ELI 3

#define DC_CFPKT 9999           /* discus_config_mod.f90:27 */

typedef struct dc_ctx dc_ctx;

/* Context life cycle.  device < 0 picks the current device. */
int dc_create(dc_ctx **out, int device);
void dc_destroy(dc_ctx *ctx);
const char *dc_last_error(const dc_ctx *ctx);
int dc_last_error_type(const dc_ctx *ctx);
int dc_set_precision(dc_ctx *ctx, int precision);
int dc_get_precision(const dc_ctx *ctx);
/* Work-group tile configuration of the phase GEMM: -1 = planned per problem
 * (default), 0..dc_gemm_num_tiles()-1 forces one (used by the autotuner).
 * dc_gemm_tile_shape reports rows x cols x threads of configuration t. */
int dc_set_gemm_tile(dc_ctx *ctx, int tile);
int dc_gemm_num_tiles(void);
int dc_gemm_tile_shape(int tile, int *rows, int *cols, int *threads);
/* Atoms per GEMM slice (one FP32 output plane each, summed in FP64): 0 =
 * planned per problem and precision (default), n > 0 forces n.  No
 * counterpart in DISCUS (whose sum runs atom by atom in FP64); a tuning and
 * test knob, like dc_set_gemm_tile. */
int dc_set_slice_atoms(dc_ctx *ctx, int n);
/* Per-atom phase rotation of the phase GEMM (on by default): row operands
 * carry +psi_n and column operands -psi_n, psi_n a fixed pseudo-random turn
 * per atom, so the products are unchanged and each atom's operand rounding
 * differs.  Off, atoms with equal phases (Bragg peaks of a crystal) add the
 * same rounding error once per atom.  No counterpart in DISCUS; a test knob. */
int dc_set_phase_dither(dc_ctx *ctx, int on);
/* Polled between kernel launches of long runs; a non-zero value aborts the
 * call with DC_ERR_CANCEL, leaving outputs untouched (DISCUS's ier_ctrlc). */
int dc_set_cancel_flag(dc_ctx *ctx, const volatile int *flag);
/* Library version and the device it runs on (gfx name, CU count). */
const char *dc_version(void);
int dc_device_info(const dc_ctx *ctx, char *gfx_name, int len, int *num_cu);

/* Ceilings of the device for the instructions the path relies on, measured
 * with plain streams of them (no counterpart in DISCUS; a measurement aid for
 * roofline figures, which are priced against the published peaks):
 * independent chains of v_mfma_f32_16x16x32_bf16 / _f16 and
 * v_mfma_f32_16x16x4_f32 with operands in registers, and a streaming read
 * and copy of hbm_bytes (0: 1 GiB; a multiple of 16 in [16 MiB, 64 GiB]).
 * Allocates and frees two hbm_bytes device buffers; best of 3 launches each.
 * hbm_copy_gbs counts the bytes read plus the bytes written. */
typedef struct dc_peaks {
    double mfma_bf16_tflops;
    double mfma_f16_tflops;
    double mfma_f32_tflops;
    double hbm_read_gbs;
    double hbm_copy_gbs;
    int64_t hbm_bytes;
    int num_cu;
} dc_peaks;
int dc_measure_peaks(dc_ctx *ctx, int64_t hbm_bytes, dc_peaks *out);

/*
 * dc_strucf -- drop-in for DISCUS's per-type kernel
 *   four_strucf(iscat, lform, ldiscamb, is_anis, isym, ientry, fnum)
 *   (discus/prog/four_strucf_OMP_mod.f90:19-68, isotropic branch), whose
 * inputs are the diffuse_mod globals xat, nxat, uin, vin, win, off_shift,
 * lmn, num, cfact(:,iscat), istl and whose output is tcsf.
 * Computes  tcsf(j,i,h) = cfact(istl(j,i,h)) * sum_k exp(2 pi i phi_k(j,i,h))
 * over j<=fnum(1), i<=fnum(2), h<=fnum(3), with
 *   phi_k = sum_m lmn(m) frac(vec_m . x_k) + (j-1) frac(uin.x_k)
 *           + (i-1) frac(vin.x_k) + (h-1) frac(win.x_k)   (mod 1)
 * vec = (uin, vin, win, off_shift(:,1..3)) -- DISCUS's phase construction
 * (four_strucf_OMP_mod.f90:135-170) carried in 64-bit instead of 22-bit
 * fixed point and evaluated exactly (no 2^16 table).  tcsf is zeroed over
 * the full num grid first, like four_strucf_omp (:129).
 * lform = 0 skips the form factor (cfact_col and istl may be NULL).
 * istl may be NULL when lform != 0: the library then evaluates the bins
 * itself from xm/uin/vin/win and rten with four_stltab's expression
 * (fourier_sup.f90:2663-2674); xm and rten are only read in that case.
 */
int dc_strucf(dc_ctx *ctx, const double *xat, int64_t nxat, int64_t ld_xat,
              const double uin[3], const double vin[3], const double win[3],
              const double off_shift[9], const int lmn[6],
              const int num[3], const int fnum[3],
              const double *cfact_col, int64_t ncfact,
              const int *istl, const double xm[3], const double rten[9],
              int lform, double *tcsf);

/*
 * dc_strucf_dev_noalloc -- dc_strucf on device buffers that allocates
 * nothing.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments and result as dc_strucf,
 * except that xat (ld_xat, 3), cfact_col, istl and tcsf are device pointers
 * (istl full num grid, tcsf complex(8) over num, written over the whole grid).
 * Its scratch -- the atom records, the slice boundaries and the slice planes,
 * processed in batches of slices that fit -- lives in one array of
 * DC_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; a problem
 * whose records plus one slice plane (8 bytes per computed pixel) exceed it
 * returns DC_ERR_ARG.  Every call on a GPU shares that GPU's array (nothing
 * in it needs to start at zero), so calls that overlap on one GPU -- from
 * several contexts, threads or streams -- race.  Synchronous on return.
 */
#define DC_NOALLOC_SCRATCH_BYTES (128 << 20)
int dc_strucf_dev_noalloc(dc_ctx *ctx, const double *d_xat, int64_t nxat, int64_t ld_xat,
                          const double uin[3], const double vin[3], const double win[3],
                          const double off_shift[9], const int lmn[6],
                          const int num[3], const int fnum[3],
                          const double *d_cfact_col, int64_t ncfact,
                          const int *d_istl, const double xm[3], const double rten[9],
                          int lform, double *d_tcsf);

/*
 * dc_four_stltab -- drop-in for four_stltab (discus/prog/fourier_sup.f90:
 * 2637-2687): istl(i,j,l) = NINT(sqrt(skalpro(h,h,cr_rten)/4) / CFINC) over
 * the num grid (abscissa fastest), h(k) = REAL(xm(k) + uin(k)*(i-1) +
 * vin(k)*(j-1) + win(k)*(l-1)) summed in FP64 and rounded to FP32 as
 * DISCUS's default-kind REAL() does.  Computed on the device, bit-identical
 * to the bins every run of this library uses.  Returns DC_ERR_ISTL (-3,
 * ier_typ ER_FOUR) when a bin exceeds CFPKT, as four_stltab does; istl then
 * holds the unclamped bins.
 */
int dc_four_stltab(dc_ctx *ctx, const double xm[3], const double uin[3], const double vin[3],
                   const double win[3], const int num[3], const double rten[9], int *istl);

/* Crystal resident in HBM (crystal_mod.f90:25-138).  cr_iscat is 1-based,
 * 0 = vacancy.  dim0 = cr_dim0(:,1), dim_lo = cr_dim(:,1). */
typedef struct dc_crystal_desc {
    const double *cr_pos;       /* (3, natoms) */
    const int *cr_iscat;        /* (natoms) */
    int64_t natoms;
    int icc[3];
    int ncatoms;
    int nscat;
    double dim0[3];
    double dim_lo[3];
} dc_crystal_desc;

int dc_crystal_set(dc_ctx *ctx, const dc_crystal_desc *crystal);

/* One Fourier run (four_run, fourier_sup.f90:11-323, isotropic IAM). */
typedef struct dc_run_desc {
    double eck[12];             /* corners ll, lr, ul, tl            */
    double vi[9];               /* increment vectors abs, ord, top   */
    int inc[3];                 /* points along abs, ord, top        */
    double rten[9];             /* reciprocal metric tensor cr_rten  */
    const double *cfact;        /* (ncfact, nscat) complex, host     */
    int ncfact;                 /* CFPKT+1                           */
    int ilots;                  /* DC_LOT_*                          */
    int ls[3];                  /* lot size in cells                 */
    int nlots;
    const int *lbeg;            /* (3, nlots) 1-based lot origins    */
    int lperiod;                /* periodic lots                     */
    double fave;                /* 'set aver' fraction, 0 = off      */
    const unsigned char *sel_cell; /* csize(1)*csize(2)*csize(3) mask, abscissa fastest */
    int four_friedel;           /* 'set symm,friedel'                */
    int ano;                    /* anomalous scattering on           */
    double diff_scale;
} dc_run_desc;

/*
 * Outputs, each NULL-able, num(1)*num(2)*num(3) elements:
 *   csf   complex  last lot's F - <F>  (times sqrt(diff_scale))
 *   acsf  complex  DISCUS <F> (four_aver)
 *   dsi   real     dnorm * sum_lots |F_l - acsf|^2 * diff_scale
 *   fmean complex  mean over lots of F_l (FP64)
 *   dvar  real     dnorm * sum_lots |F_l - fmean|^2 * diff_scale
 *                  (cross-lot variance, Welford in FP64)
 * dc_four_run takes host output pointers; dc_four_run_dev takes device
 * pointers (HBM).  Both return when the outputs are written: the launches of
 * a call follow one another on the context's stream with no host wait
 * between them, and the call waits for its stream once, at the end.
 */
int dc_four_run(dc_ctx *ctx, const dc_run_desc *run, double *csf, double *acsf, double *dsi,
                double *fmean, double *dvar);
int dc_four_run_dev(dc_ctx *ctx, const dc_run_desc *run, double *d_csf, double *d_acsf,
                    double *d_dsi, double *d_fmean, double *d_dvar);

/* Statistics of the last call: atom.hkl pairs (full grid), device time of
 * the phase-GEMM kernels and of the whole call (ms, HIP events).  The kernel
 * timings are event pairs read after the call's launches complete:
 * dc_last_stats waits for the context's stream when some are still pending. */
typedef struct dc_stats {
    double pairs;
    double gemm_ms;
    double total_ms;
    double reduce_ms;
    int64_t gemm_launches;
    int64_t slices;
    double forward_ms;          /* refinement model: forward (GEMM + reduction)   */
    double adjoint_ms;          /* refinement model: reverse pass (Q, GEMM, sums) */
    int64_t adjoint_walk;       /* column tiles per work-group of the last reverse
                                   pass (1: one work-group per column tile)       */
    int64_t gemm_scheme;        /* products of the last phase GEMM: DC_PREC_* value,
                                   or 4 = FP32 products with FP64 partial sums    */
} dc_stats;
int dc_last_stats(const dc_ctx *ctx, dc_stats *out);
int dc_synchronize(dc_ctx *ctx);

/*
 * Refinement model with per-atom parameters (REFINE's parameter vector,
 * refine/prog/refine_run.f90:491-600, which today obtains derivatives by
 * re-running the Fourier calculation for every parameter):
 *   F(h) = sum_n f_t(h) occ_n exp(-B_n s^2) exp(2 pi i h.x_n),  I = |F|^2,
 * over the grid h = eck(:,1) + i vi(:,1) + j vi(:,2) + l vi(:,3) (i < inc(1)
 * fastest, like dsi), s = sin(theta)/lambda from rten, f_t = cfact(istl(h), t)
 * with four_stltab's bins (fourier_sup.f90:2663-2674).  cfact carries
 * f + f' + i f'' only: occupancy and B are per atom here.
 * Per-atom B (and every derivative with respect to B) needs the abscissa
 * vi(:,1) orthogonal in the metric to vi(:,2) and vi(:,3), so that
 * s^2 = s^2(i) + s^2(j, l); otherwise DC_ERR_ARG.
 */
typedef struct dc_model_desc {
    const double *pos;          /* (3, natoms) fractional coordinates           */
    const int *type;            /* (natoms) 1-based scattering type, 0 = vacancy */
    const double *occ;          /* (natoms)                                      */
    const double *bdw;          /* (natoms) isotropic B in A^2, NULL = all zero  */
    int64_t natoms;
    int nscat;
    int ncfact;                 /* rows of cfact (CFPKT+1)                       */
    const double *cfact;        /* (ncfact, nscat) complex                       */
    double eck[12];
    double vi[9];
    int inc[3];
    double rten[9];
} dc_model_desc;

/* icalc (real) and fcalc (complex), each inc(1)*inc(2)*inc(3), NULL-able. */
int dc_model_intensity(dc_ctx *ctx, const dc_model_desc *model, double *icalc, double *fcalc);
/*
 * Vector-Jacobian product: given F of the model (fcalc from
 * dc_model_intensity) and gi = dL/dI per pixel, writes grad(5, natoms):
 * dL/dx, dL/dy, dL/dz (fractional), dL/docc, dL/dB for every atom
 * (zero for vacancies).  This is the hand-written adjoint: it regenerates
 * the phase operands instead of storing them.
 */
int dc_model_vjp(dc_ctx *ctx, const dc_model_desc *model, const double *fcalc, const double *gi, double *grad);
/*
 * Jacobian-vector products (forward mode): for ndir directions
 * dparam(5, natoms, ndir) in the layout of grad, writes
 *   di(P, ndir) = dI/dt along each direction, and dF (complex, NULL-able),
 * given F of the model (fcalc from dc_model_intensity).  These are the
 * per-pixel derivative columns REFINE's Levenberg-Marquardt loop builds by
 * finite differences (refine/prog/refine_run.f90:491-600, gl_is_der hook
 * :550-552): one direction per refined parameter.
 */
int dc_model_jvp(dc_ctx *ctx, const dc_model_desc *model, const double *fcalc, int ndir, const double *dparam,
                 double *di, double *df);
/*
 * chi^2 = sum_h w(h) (I(h) - target(h))^2 (weight NULL = 1) and its gradient
 * grad(5, natoms) in one call (forward, residual and adjoint stay on the
 * device).  icalc and grad are NULL-able.
 */
int dc_model_chi2(dc_ctx *ctx, const dc_model_desc *model, const double *target, const double *weight,
                  double *chi2, double *icalc, double *grad);

/*
 * Population scoring (DIFFEV, diffev/prog/diffev_kdo.f90:676 running one
 * macro per member over lib_f90/with_mpi.f90:281-609): nmembers refinement
 * models sharing one grid (eck/vi/inc in each member's reciprocal units),
 * each with its own atoms, form factors and metric, scored by
 *   chi2[m] = sum_h w_m(h) (I_m(h) - target_m(h))^2
 * in one pass: one persistent phase-GEMM launch over all members' atoms, one
 * fused reduction, and (grad != NULL) one adjoint launch.  The first call of
 * a batch shape (member sizes, types, grid, precision, outputs requested)
 * plans, autotunes the GEMM tile and grid size, and captures the step as a
 * hipGraph; later calls of the same shape (new positions, occupancies, B,
 * form factors, metrics, targets) upload and replay it.
 * Arrays: pos (3, total), type/occ/bdw (total) member after member
 * (atom_beg[m] .. atom_beg[m+1]-1); cfact (ncfact, nscat, nmembers);
 * rten (9, nmembers); target/weight/icalc (P, nmembers) with P =
 * inc(1)*inc(2)*inc(3) abscissa fastest; chi2 (nmembers); grad (5, total)
 * as in dc_model_vjp.  Per-atom B and gradients need the abscissa
 * metric-orthogonal to the other axes in every member's metric.
 */
typedef struct dc_batch_desc {
    int nmembers;
    const int64_t *atom_beg;    /* nmembers + 1 offsets, atom_beg[0] = 0 */
    const double *pos;
    const int *type;            /* 1-based within the member's nscat types, 0 = vacancy */
    const double *occ;
    const double *bdw;          /* NULL = all zero */
    int nscat;
    int ncfact;
    const double *cfact;
    const double *rten;
    double eck[12];
    double vi[9];
    int inc[3];
} dc_batch_desc;

int dc_batch_chi2(dc_ctx *ctx, const dc_batch_desc *batch, const double *target, const double *weight,
                  double *chi2, double *icalc, double *grad);
/* The same step with the per-generation arrays in device memory: batch->pos,
 * ->occ, ->bdw (non-NULL = per-atom B), target, weight, icalc and grad are
 * device pointers; atom_beg, type, cfact, rten and chi2 stay host arrays.
 * For a host loop that keeps a generation on the GPU (DIFFEV's trial
 * vectors built on the device); replaces the same per-member macro runs as
 * dc_batch_chi2.  Returns when chi2 is written. */
int dc_batch_chi2_dev(dc_ctx *ctx, const dc_batch_desc *batch, const double *target, const double *weight,
                      double *chi2, double *icalc, double *grad);
/* Autotuner and hipGraph capture of batch calls (both on by default).  The
 * autotuner times the GEMM tiles and persistent grid sizes on the first call
 * of a batch shape, and the planned slice length against half and twice it;
 * a slice length other than the planned one (taken only when it is faster
 * by 2% or more) changes the FP32 summation order, so results can then
 * differ from another context's in the last bits.  dc_set_slice_atoms fixes
 * the length, dc_set_autotune(ctx, 0) the whole plan. */
int dc_set_autotune(dc_ctx *ctx, int on);
int dc_set_graphs(dc_ctx *ctx, int on);
/* Column walk of the reverse pass (every precision) on grids of at most 80
 * abscissa points (dc_model_vjp, dc_model_chi2, dc_batch_chi2): 0 (default)
 * plans it from the grid and the CU count; 1..32 forces that many column
 * tiles of 256 per work-group (1 = one work-group per column tile).  A tuning
 * knob with no counterpart in DISCUS; it changes the FP32 summation order of
 * the gradient partials, not the result beyond rounding.  Reported back in
 * dc_stats.adjoint_walk. */
int dc_set_adjoint_walk(dc_ctx *ctx, int walk);
/* Choice of the last batch shape: tile configuration, persistent work-groups
 * (0 = one per work item; negative = that many work-groups taking the work
 * items, longest first, from a device counter), whether a graph replays it,
 * autotuned GEMM ms. */
int dc_batch_info(const dc_ctx *ctx, int *tile, int *nwg, int *graph, double *tune_ms);

/* ------------------------------------------------------------------------
 * Post-processing of a finished run: the tail of four_run
 * (fourier_sup.f90:279-292) and four_accumulate / qval.
 *
 * Arrays are DISCUS's: num(1)*num(2)*num(3) elements, abscissa fastest,
 * complex as interleaved (re, im) doubles.  The *_dev entry points take
 * device pointers (e.g. the outputs of dc_four_run_dev) and do not
 * synchronise the host; the others take host pointers and return when the
 * outputs are written.  Every step runs on the GPU (hipFFT Z2Z transforms and
 * the library's own kernels).
 * ------------------------------------------------------------------------ */
typedef struct dc_post_desc {
    int num[3];                 /* grid points along abs, ord, top            */
    double eck[12];             /* corners, as dc_run_desc                    */
    double vi[9];               /* increments abs, ord, top = uin, vin, win   */
    double rten[9];             /* reciprocal metric cr_rten                  */
    double vr;                  /* reciprocal cell volume cr_vr               */
    double fave;                /* 'set aver'                                 */
    int nlots;
    int icc[3];                 /* crystal size in unit cells cr_icc          */
    int ls[3];                  /* lot size ls_xyz                            */
    int lots_off;               /* ilots == LOT_OFF: four_conv convolves csf  */
    double diff_scale;          /* already applied to dsi and csf             */
    double res_sigma[3];        /* diff_res(1, 1:3); all < 1e-6: no four_conv */
    double res_tr[9];           /* diff_tr(3,3), column-major                 */
    int filter;                 /* 0 off, 1 Lanczos (four_filter)             */
    double rscale;              /* four_rscale (nint must be 1)               */
    double damp;                /* four_damp                                  */
    int width;                  /* four_width                                 */
} dc_post_desc;

/* four_weight (fourier_sup.f90:1428-1644) into dsi3d, do_four_filter on dsi
 * (:1646-1805, lib_lanczos.f90:258-458), four_conv on csf/dsi/dsi3d
 * (fourier_conv.f90:7-423), do_four_filter on dsi3d when fave == 0: the
 * sequence of fourier_sup.f90:279-290.  dsi3d may be NULL. */
int dc_four_post(dc_ctx *ctx, const dc_post_desc *post, double *csf, double *dsi, double *dsi3d);
int dc_four_post_dev(dc_ctx *ctx, const dc_post_desc *post, double *d_csf, double *d_dsi, double *d_dsi3d);

typedef struct dc_symm_desc {
    int num[3];
    double eck[12];
    double vi[9];
    double fave;                /* < 0: fill Bragg pixels from neighbours      */
    int four_symm;              /* 'set symm' point-group average              */
    int nops;                   /* spc_n / n_center                            */
    const double *mats;         /* (3,3,nops-1): spc_mat(1:3,1:3,2:nops), host */
} dc_symm_desc;

/* four_fill_csf + four_fill_dsi (fourier_sup.f90:3844-4046): point-group
 * average (four_symm_csf/_dsi :4048-4293) when four_symm, then the Bragg
 * fill when fave < 0; in place.  Operations must be integer matrices with
 * determinant +-1 (DC_ERR_ARG otherwise). */
int dc_four_fill(dc_ctx *ctx, const dc_symm_desc *symm, double *csf, double *dsi);
int dc_four_fill_dev(dc_ctx *ctx, const dc_symm_desc *symm, double *d_csf, double *d_dsi);

#define DC_ACCUM_INIT (-1)
#define DC_ACCUM_SINGLE 0
#define DC_ACCUM_ACCUM 1
#define DC_ACCUM_FINISHED 2
#define DC_ERR_NOSUM (-175)     /* four_accumulate finished without sums (fourier_sup.f90:3826-3830) */

/* four_accumulate (fourier_sup.f90:3784-3840) with caller-owned device sums
 * csf_sum/dsi_sum (NULL where DISCUS's are not allocated).  INIT does
 * nothing here (the caller drops its sums); ACCUM adds; FINISHED fills the
 * sums (dc_four_fill) and copies them into csf/dsi; SINGLE applies the fill
 * to csf/dsi when four_symm. */
int dc_four_accumulate_dev(dc_ctx *ctx, int mode, const dc_symm_desc *symm, double *d_csf, double *d_dsi,
                           double *d_csf_sum, double *d_dsi_sum);

/* qval (qval.f90:65-286) over the whole grid.  value: 1 intensity,
 * 2 amplitude, 3 phase (degrees), 4 real, 5 imaginary, 7 S(Q), 9 <f^2>,
 * 10 <f>^2, 11 <f>, 13 I(Q), 17 <f^2> with DBW, 18 <f>^2 with DBW, 19 <f>
 * with DBW.  6 (random phase), 8, 14-16 (PDF) return DC_ERR_ARG. */
typedef struct dc_qval_desc {
    int num[3];
    double eck[12];
    double vi[9];
    double rten[9];             /* for istl (four_stltab)                     */
    int value;
    int laver;                  /* take acsf instead of csf                   */
    const double *cfact;        /* (ncfact, nscat) complex, host              */
    const double *cfact_pure;   /* (ncfact, nscat) complex, host              */
    int ncfact, nscat;
    const double *amount;       /* cr_amount(1:nscat), host                   */
    double n_real_atoms;        /* cr_n_real_atoms                            */
    double u2aver;              /* cr_u2aver                                  */
    double faver_sign;          /* sign <f> takes (DISCUS: left by the last
                                   S(Q) / <f>^2 evaluation; +1 in a fresh run) */
} dc_qval_desc;
int dc_qval(dc_ctx *ctx, const dc_qval_desc *q, const double *csf, const double *acsf, const double *dsi,
            double *out);
int dc_qval_dev(dc_ctx *ctx, const dc_qval_desc *q, const double *d_csf, const double *d_acsf,
                const double *d_dsi, double *d_out);

/* ------------------------------------------------------------------------
 * NUFFT technique (four_run_nufft, fourier_sup.f90:327-690).
 *
 * dc_nufft1 replaces FINUFFT's type-1 calls finufft1d1 / finufft2d1 /
 * finufft3d1 as four_strucf_1d/2d/3d make them (fourier_finufft.f90:62,
 * 124, 210, default options):
 *   fk(k1,k2,k3) = sum_j c_j exp(iflag * i * (k1 x_j + k2 y_j + k3 z_j)),
 *   k_d = -N_d/2 .. (N_d-1)/2 in increasing order, k1 fastest,
 * x_j, y_j, z_j in [-3 pi, 3 pi) (periodic), c and fk complex interleaved,
 * relative accuracy about tol.  dim 1..3; y/z (and n2/n3) unused below dim.
 * ------------------------------------------------------------------------ */
int dc_nufft1(dc_ctx *ctx, int dim, int64_t M, const double *x, const double *y, const double *z,
              const double *c, int iflag, double tol, int64_t n1, int64_t n2, int64_t n3, double *fk);
int dc_nufft1_dev(dc_ctx *ctx, int dim, int64_t M, const double *x, const double *y, const double *z,
                  const double *c, int iflag, double tol, int64_t n1, int64_t n2, int64_t n3, double *fk);

/* four_strucf_1d / _2d / _3d (fourier_finufft.f90:19-245): nat atoms of one
 * type at the scaled positions of four_getatm_anis_iam (2 pi scales pos /
 * ncells), weight occ each, NUFFT on (inc-1)*lscales+1 modes per axis with
 * iflag = +1 and DISCUS's tolerance (1e-9 in 1D/2D, 1e-7 in 3D); every
 * lscales-th mode is kept: tcsf(inc(1), inc(2), inc(3)) complex, abscissa
 * fastest (inc(d) = 1 beyond dim). */
int dc_four_strucf_nufft(dc_ctx *ctx, int dim, int64_t nat, const double *xpos, const double *ypos,
                         const double *zpos, double occ, const int inc[3], const int lscales[3], double *tcsf);
int dc_four_strucf_nufft_dev(dc_ctx *ctx, int dim, int64_t nat, const double *xpos, const double *ypos,
                             const double *zpos, double occ, const int inc[3], const int lscales[3], double *tcsf);

/* four_run_nufft (fourier_sup.f90:327-690) for 3D crystals, isotropic IAM:
 * the crystal of dc_crystal_set, per type the NUFFT of four_strucf_3d with
 * DISCUS's scaling (four_getatm_anis_iam, scales / lscales from
 * fourier_sup.f90:390-483, weight type_occ[t] = cr_occ, NULL = 1), times
 * cfact(istl) (tcsf_form_iso), summed over types; fave > 0 zeros the integer
 * hkl pixels (four_aver_finufft_n :1074-1190); dsi = |F|^2 * diff_scale, csf
 * = F * sqrt(diff_scale).  Increment vectors must run along a*, b*, c*.
 * Like DISCUS, the transform is taken on the modes k * scales / ncells,
 * centred on hkl = 0; eck(:,1) enters only the istl bins and the fave checks
 * (-186, -187, -188, :452-483).  The rest of the tail is dc_four_post. */
int dc_four_run_nufft(dc_ctx *ctx, const dc_run_desc *run, const double *type_occ, double *csf, double *dsi);
int dc_four_run_nufft_dev(dc_ctx *ctx, const dc_run_desc *run, const double *type_occ, double *d_csf, double *d_dsi);

#ifdef __cplusplus
}
#endif
#endif
