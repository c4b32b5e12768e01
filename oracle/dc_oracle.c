/*
 * dc_oracle.c -- CPU oracle for the DISCUS Fourier structure-factor path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product path (diffusecode_amd) never links or calls it.
 *
 * It restates, in plain C, two things:
 *
 *   (1) "turbo" emulation: DISCUS's own integer-phase / sine-table algorithm,
 *       bit-for-bit in its integer arithmetic and in its table semantics:
 *         - per-atom increments NINT(2^22 * (x - INT(x)))
 *           (discus/prog/four_strucf_OMP_mod.f90:135-149)
 *         - corner phase from lmn(1:6) (four_strucf_OMP_mod.f90:163-165)
 *         - table index IAND(ISHFT(iarg,-6), 0xFFFF) into a 2^16-entry
 *           complex table cex (four_strucf_OMP_mod.f90:169,
 *           fourier_sup.f90:2602-2632)
 *         - OpenMP static schedule over abscissa rows with a private row
 *           buffer (four_strucf_OMP_mod.f90:155-182)
 *         - form-factor multiply cfact(istl(h), iscat)
 *           (fourier_form.f90:147-177)
 *       This is also the timed CPU baseline.
 *
 *   (2) "exact": an FP64 direct sum F(h) = sum_j exp(+2 pi i h.r_j) with h.r
 *       formed in long double and reduced mod 1 before sincos, i.e. the
 *       mathematical quantity DISCUS's table approximates.  All tolerance
 *       statements of the GPU path are made against this mode.
 *
 * plus the host-side tables (fourier_lmn_mod.f90:4-280 in __float128,
 * four_stltab fourier_sup.f90:2637-2687, four_formtab/form
 * fourier_sup.f90:2691-2743,2947-2971), lot gathering (four_getatm
 * fourier_sup.f90:1806-1970, four_getav :2440-2510, four_csize :2533-2555),
 * the average structure (four_aver :1194-1357), Friedel halving
 * (four_test_friedel/four_apply_friedel :3645-3761) and the lot loop with
 * Bragg subtraction and |F|^2 accumulation (four_run :11-323).
 *
 * Array conventions follow the reference's Fortran layouts (column-major):
 *   xat(NMAX,3)          -> xat[k + ld*c]
 *   tcsf/csf/acsf(n1,n2,n3) complex(8) -> interleaved re,im, i fastest
 *   dsi(n1,n2,n3) real(8), istl(n1,n2,n3) int32
 *   cfact(0:CFPKT, nscat) complex(8) -> cfact[2*(iq + (CFPKT+1)*(iscat-1))]
 *   cr_pos(3,N)          -> pos[3*i + c]
 *   eck(3,4), vi(3,3), off_shift(3,3), rten(3,3) -> a[i + 3*j]
 *
 * Parity pinning (what checks each function) is listed in the header comment
 * of oracle/oracle.py.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <quadmath.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_I2PI 65536
#define ORC_MASK 65535
#define ORC_CFPKT 9999
#define ORC_CFINC 0.001

#define LOT_OFF 1
#define LOT_BOX 2
#define LOT_ELI 3

#define MODE_EXACT 0
#define MODE_TURBO 1

/* pi as in lib_f90/wink_mod.f90:10-11 */
static const double ORC_ZPI = 2.0 * 3.1415926535897932384626433832795028841971693993751;

static void set_threads(int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

/* Fortran NINT: round half away from zero. */
static long long f_nint(double x) { return llround(x); }

/* ------------------------------------------------------------------------ */
/* four_cexpt (fourier_sup.f90:2602-2632): cex(i) = exp(2 pi i * i/2^16).    */
void orc_cexpt(double *cex)
{
    for (int i = 0; i <= ORC_MASK; ++i) {
        double xmult = (double)i / (double)ORC_I2PI;
        double xarg = ORC_ZPI * xmult;
        cex[2 * i] = cos(xarg);
        cex[2 * i + 1] = sin(xarg);
    }
}

/* ------------------------------------------------------------------------ */
/* fourier_lmn (fourier_lmn_mod.f90:4-280), in __float128 like REAL(16).
 * rnd[0..5] replace the six ran1() draws of the 1-D "line through the
 * origin" branches (lines 185-195); they are an input so that no RNG crosses
 * the boundary.  Returns 0 or -132 (zero increment vector, lines 84-91). */
static void minv3_q(const __float128 a[3][3], __float128 b[3][3])
{
    /* cofactor inverse, a[row][col] */
    __float128 det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1])
                   - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0])
                   + a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) b[i][j] = 0;
    if (det == 0) return;
    __float128 di = 1 / det;
    b[0][0] = +di * (a[1][1] * a[2][2] - a[1][2] * a[2][1]);
    b[1][0] = -di * (a[1][0] * a[2][2] - a[1][2] * a[2][0]);
    b[2][0] = +di * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
    b[0][1] = -di * (a[0][1] * a[2][2] - a[0][2] * a[2][1]);
    b[1][1] = +di * (a[0][0] * a[2][2] - a[0][2] * a[2][0]);
    b[2][1] = -di * (a[0][0] * a[2][1] - a[0][1] * a[2][0]);
    b[0][2] = +di * (a[0][1] * a[1][2] - a[0][2] * a[1][1]);
    b[1][2] = -di * (a[0][0] * a[1][2] - a[0][2] * a[1][0]);
    b[2][2] = +di * (a[0][0] * a[1][1] - a[0][1] * a[1][0]);
}

static long long q_nint(__float128 x)
{
    return (long long)(x >= 0 ? floorq(x + 0.5Q) : -floorq(-x + 0.5Q));
}

int orc_fourier_lmn(const double *eck_in, const double *vi_in, const int *inc,
                    const double *rnd, int *lmn, double *off_in)
{
    const __float128 big = 10000000.0Q;
    __float128 eck[3][4], vi[3][3], off[3][3], mat_a[3][3], mat_i[3][3];
    __float128 vec_r[3], vec_t[3], dummy;
    int direc[3] = {0, 0, 0}, nooff[3] = {0, 0, 0}, dimen = 0, need_work = 1;
    int j, j1, j2;

    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 3; ++r)
            eck[r][c] = (__float128)(int)llround(eck_in[r + 3 * c] * 1.0e7) / big;
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r)
            vi[r][c] = (__float128)(int)llround(vi_in[r + 3 * c] * 1.0e7) / big;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) { mat_a[r][c] = 0; mat_i[r][c] = 0; off[r][c] = 0; }
    for (int i = 0; i < 6; ++i) lmn[i] = 0;

    for (int i = 0; i < 3; ++i) {
        if (inc[i] > 1) {
            dummy = sqrtq(fabsq(vi[0][i] * vi[0][i] + vi[1][i] * vi[1][i] + vi[2][i] * vi[2][i]));
            if (dummy > 1.0e-7Q) {
                direc[dimen++] = i;
                nooff[i] = 1;
                for (int r = 0; r < 3; ++r) mat_a[r][i] = vi[r][i];
            } else {
                return -132;
            }
        }
    }
    if (dimen == 0) {
        off[0][0] = 1; off[1][1] = 1; off[2][2] = 1;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) mat_a[r][c] = off[r][c];
    } else if (dimen == 1) {
        j = direc[0];
        j1 = (j + 1) % 3;           /* MOD(j,3)+1 with 1-based j */
        j2 = (j + 2) % 3;           /* MOD(j+1,3)+1 */
        int axis = -1;
        if (vi[1][j] == 0 && vi[2][j] == 0) axis = 0;
        else if (vi[0][j] == 0 && vi[2][j] == 0) axis = 1;
        else if (vi[0][j] == 0 && vi[1][j] == 0) axis = 2;
        if (axis >= 0) {
            /* lines 103-168: a line parallel to a*, b* or c* */
            __float128 vr = eck[axis][0] / vi[axis][0];
            lmn[0] = (int)q_nint(vr);
            if (fabsq(vr - (__float128)q_nint(vr)) > 1.0e-9Q) {
                lmn[3] = 1;
                off[axis][0] = vr - (__float128)q_nint(vr);
            }
            lmn[4] = 1;
            lmn[5] = 1;
            if (axis == 0) { off[1][1] = eck[1][0]; off[2][2] = eck[2][0]; }
            if (axis == 1) { off[0][1] = eck[0][0]; off[2][2] = eck[2][0]; }
            if (axis == 2) { off[0][1] = eck[0][0]; off[1][2] = eck[1][0]; }
            need_work = 0;
        } else {
            off[0][j2] = eck[1][0] * vi[2][j] - eck[2][0] * vi[1][j];
            off[1][j2] = eck[2][0] * vi[0][j] - eck[0][0] * vi[2][j];
            off[2][j2] = eck[0][0] * vi[1][j] - eck[1][0] * vi[0][j];
            dummy = sqrtq(fabsq(off[0][j2] * off[0][j2] + off[1][j2] * off[1][j2] + off[2][j2] * off[2][j2]));
            if (dummy > 1.0e-7Q) {
                for (int r = 0; r < 3; ++r) off[r][j2] = off[r][j2] * 0.001Q / dummy;
                off[0][j1] = off[1][j2] * vi[2][j] - off[2][j2] * vi[1][j];
                off[1][j1] = off[2][j2] * vi[0][j] - off[0][j2] * vi[2][j];
                off[2][j1] = off[0][j2] * vi[1][j] - off[1][j2] * vi[0][j];
                dummy = sqrtq(fabsq(off[0][j1] * off[0][j1] + off[1][j1] * off[1][j1] + off[2][j1] * off[2][j1]));
                if (dummy > 1.0e-7Q) {
                    for (int r = 0; r < 3; ++r) off[r][j1] = off[r][j1] * 0.001Q / dummy;
                } else {
                    for (int r = 0; r < 3; ++r) off[r][j1] = rnd[r];
                }
            } else {
                for (int r = 0; r < 3; ++r) { off[r][j1] = rnd[r]; off[r][j2] = rnd[3 + r]; }
            }
            for (int r = 0; r < 3; ++r) { mat_a[r][j1] = off[r][j1]; mat_a[r][j2] = off[r][j2]; }
        }
    } else if (dimen == 2) {
        j = direc[0];
        j1 = direc[1];
        j2 = 3 - ((j + 1) + (j1 + 1)) % 3 - 1;  /* 1-based: 3 - MOD(j+j1,3) */
        off[0][j2] = vi[1][j] * vi[2][j1] - vi[2][j] * vi[1][j1];
        off[1][j2] = vi[2][j] * vi[0][j1] - vi[0][j] * vi[2][j1];
        off[2][j2] = vi[0][j] * vi[1][j1] - vi[1][j] * vi[0][j1];
        dummy = sqrtq(off[0][j2] * off[0][j2] + off[1][j2] * off[1][j2] + off[2][j2] * off[2][j2]);
        for (int r = 0; r < 3; ++r) { off[r][j2] = off[r][j2] * 0.001Q / dummy; mat_a[r][j2] = off[r][j2]; }
    }

    if (need_work) {
        minv3_q(mat_a, mat_i);
        for (int r = 0; r < 3; ++r)
            vec_r[r] = mat_i[r][0] * eck[0][0] + mat_i[r][1] * eck[1][0] + mat_i[r][2] * eck[2][0];
        for (int r = 0; r < 3; ++r)
            if (fabsq(vec_r[r] - (__float128)q_nint(vec_r[r])) < 1.0e-9Q) vec_r[r] = (__float128)q_nint(vec_r[r]);
        __float128 fr[3];
        for (int r = 0; r < 3; ++r) {
            lmn[r] = (int)q_nint(vec_r[r]);
            fr[r] = vec_r[r] - (__float128)q_nint(vec_r[r]);
        }
        for (int r = 0; r < 3; ++r)
            vec_t[r] = mat_a[r][0] * fr[0] + mat_a[r][1] * fr[1] + mat_a[r][2] * fr[2];
        for (int c = 0; c < 3; ++c) {
            if (!nooff[c]) {
                lmn[c + 3] = lmn[c];
                lmn[c] = 0;
            } else {
                off[c][c] = off[c][c] + vec_t[c];
                lmn[3 + c] = 1;
                if (fabsq(off[c][c]) < 1.0e-9Q) {
                    off[c][c] = 0;
                    lmn[3 + c] = 0;
                }
            }
        }
    }
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) off_in[r + 3 * c] = (double)off[r][c];
    return 0;
}

/* ------------------------------------------------------------------------ */
/* sin(theta)/lambda bin of one reciprocal point, the expression of
 * four_stltab (fourier_sup.f90:2663-2674): h(k) = REAL(xm(k) + uin(k)*(i-1)
 * + vin(k)*(j-1) + win(k)*(l-1)) -- the sum in FP64, then REAL() of default
 * kind, i.e. rounded to FP32 (the reference builds without
 * -fdefault-real-8, CMakeLists.txt:129), stored back into the FP64 h --
 * and skalpro = dot(h, rten*h) (discus/prog/metric.f90:90-105).
 * Evaluated left to right without contraction so that the GPU reproduces
 * it bit for bit. */
#pragma GCC push_options
#pragma GCC optimize("fp-contract=off")
int orc_istl_point(const double xm[3], const double uin[3], const double vin[3],
                   const double win[3], int i, int j, int l, const double rten[9])
{
    double h[3], t[3];
    for (int k = 0; k < 3; ++k)
        h[k] = (double)(float)(xm[k] + uin[k] * (double)i + vin[k] * (double)j + win[k] * (double)l);
    for (int r = 0; r < 3; ++r)
        t[r] = rten[r + 0] * h[0] + rten[r + 3] * h[1] + rten[r + 6] * h[2];
    double q2 = (h[0] * t[0] + h[1] * t[1] + h[2] * t[2]) / 4.0;
    return (int)f_nint(sqrt(q2) * (1.0 / ORC_CFINC));
}
#pragma GCC pop_options

/* four_stltab (fourier_sup.f90:2637-2687).  Returns 0 or -3 (istl>CFPKT). */
int orc_stltab(const double xm[3], const double uin[3], const double vin[3], const double win[3],
               const int num[3], const double rten[9], int *istl)
{
    for (int l = 0; l < num[2]; ++l)
        for (int j = 0; j < num[1]; ++j)
            for (int i = 0; i < num[0]; ++i) {
                int v = orc_istl_point(xm, uin, vin, win, i, j, l, rten);
                if (v > ORC_CFPKT) return -3;
                istl[i + (long)num[0] * (j + (long)num[1] * l)] = v;
            }
    return 0;
}

/* form (fourier_sup.f90:2947-2971): f = c + sum_i a_i exp(-b_i h2).
 * scat holds 11 coefficients per type: scat[0]=c, scat[2i-1]=a_i, scat[2i]=b_i. */
static double form(const double *scat, int lxray, double h2, int power)
{
    double f = scat[0];
    if (lxray)
        for (int i = 1; i <= power; ++i) f += scat[2 * i - 1] * exp(-scat[2 * i] * h2);
    return f;
}

/* four_formtab (fourier_sup.f90:2691-2743): cfact = sb*(f+f') + i sb*f'',
 * sb = occ*exp(-B s^2) when the Debye-Waller term is on. */
void orc_formtab(int nscat, const double *scat, int lxray, int power, const double *occ,
                 const double *dw, int ldbw, int ano, const double *delfr, const double *delfi,
                 double *cfact)
{
    for (int t = 0; t < nscat; ++t) {
        for (int iq = 0; iq <= ORC_CFPKT; ++iq) {
            double q2 = ((double)iq * ORC_CFINC) * ((double)iq * ORC_CFINC);
            double sf = form(scat + 11 * t, lxray, q2, power);
            double sfp = ano ? delfr[t] : 0.0, sfpp = ano ? delfi[t] : 0.0;
            double sb = ldbw ? exp(-(dw[t] * q2)) * occ[t] : occ[t];
            long o = 2L * (iq + (long)(ORC_CFPKT + 1) * t);
            cfact[o] = sb * (sf + sfp);
            cfact[o + 1] = sb * sfpp;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Form-factor multiply, tcsf_form_iso (fourier_form.f90:147-177). */
static void form_multiply(const int num[3], const int fnum[3], const double *cfact_col,
                          const int *istl, double *tcsf)
{
    for (int h = 0; h < fnum[2]; ++h)
        for (int j = 0; j < fnum[1]; ++j)
            for (int i = 0; i < fnum[0]; ++i) {
                long p = i + (long)num[0] * (j + (long)num[1] * h);
                const double *c = cfact_col + 2L * istl[p];
                double re = tcsf[2 * p], im = tcsf[2 * p + 1];
                tcsf[2 * p] = re * c[0] - im * c[1];
                tcsf[2 * p + 1] = re * c[1] + im * c[0];
            }
}

/* four_strucf_omp / four_strucf_omp_1D (four_strucf_OMP_mod.f90:72-188,
 * 290-396): the turbo algorithm.  tcsf is zeroed over the full grid.
 * Compiled with fast-math like DISCUS (CMakeLists.txt:129); its integer
 * phase arithmetic is unaffected, only the FP64 summation order may vary. */
__attribute__((optimize("fast-math")))
void orc_strucf_turbo(const double *xat, long nxat, long ld, const double uin[3],
                      const double vin[3], const double win[3], const double off[9],
                      const int lmn[6], const int num[3], const int fnum[3], const double *cex,
                      const double *cfact_col, const int *istl, int lform, double *tcsf,
                      int nthreads)
{
    long ntot = (long)num[0] * num[1] * num[2];
    memset(tcsf, 0, sizeof(double) * 2 * ntot);
    if (nxat <= 0) {
        if (lform) form_multiply(num, fnum, cfact_col, istl, tcsf);
        return;
    }
    long long *inc = (long long *)malloc(sizeof(long long) * 6 * nxat);
    const double *vec[6] = {uin, vin, win, off, off + 3, off + 6};
    set_threads(nthreads);
#pragma omp parallel for schedule(static)
    for (long k = 0; k < nxat; ++k) {
        double x0 = xat[k], x1 = xat[k + ld], x2 = xat[k + 2 * ld];
        for (int m = 0; m < 6; ++m) {
            const double *v = vec[m];
            double x = v[0] * x0 + v[1] * x1 + v[2] * x2;
            inc[6 * k + m] = f_nint(64.0 * ORC_I2PI * (x - (double)(long long)x + 0.0));
        }
    }
    const long n2 = fnum[1], n3 = fnum[2];
    /* lmn(1..3) pair with iinc{u,v,w}; lmn(4..6) with jinc{u,v,w} */
#pragma omp parallel
    {
        double *row = (double *)malloc(sizeof(double) * 2 * n2 * n3);
#pragma omp for schedule(static)
        for (long j = 0; j < fnum[0]; ++j) {
            memset(row, 0, sizeof(double) * 2 * n2 * n3);
            for (long k = 0; k < nxat; ++k) {
                const long long *c = inc + 6 * k;
                long long iarg0 = (long long)lmn[0] * c[0] + (long long)lmn[1] * c[1] +
                                  (long long)lmn[2] * c[2] + (long long)lmn[3] * c[3] +
                                  (long long)lmn[4] * c[4] + (long long)lmn[5] * c[5] + j * c[0];
                for (long i = 0; i < n2; ++i) {
                    long long iarg = iarg0 + c[1] * i;
                    for (long h = 0; h < n3; ++h) {
                        unsigned idx = (unsigned)(((unsigned long long)iarg >> 6) & ORC_MASK);
                        row[2 * (i + n2 * h)] += cex[2 * idx];
                        row[2 * (i + n2 * h) + 1] += cex[2 * idx + 1];
                        iarg += c[2];
                    }
                }
            }
            for (long h = 0; h < n3; ++h)
                for (long i = 0; i < n2; ++i) {
                    long p = j + (long)num[0] * (i + (long)num[1] * h);
                    tcsf[2 * p] += row[2 * (i + n2 * h)];
                    tcsf[2 * p + 1] += row[2 * (i + n2 * h) + 1];
                }
        }
        free(row);
    }
    free(inc);
    if (lform) form_multiply(num, fnum, cfact_col, istl, tcsf);
}

/* Exact FP64 direct sum over the same (fnum-limited) grid:
 *   h(j,i,k) = xm + j*uin + i*vin + k*win, F = sum_n exp(+2 pi i h.r_n),
 * h.r formed in long double and reduced mod 1 before sincos. */
void orc_strucf_exact(const double *xat, long nxat, long ld, const double xm[3],
                      const double uin[3], const double vin[3], const double win[3],
                      const int num[3], const int fnum[3], const double *cfact_col,
                      const int *istl, int lform, double *tcsf, int nthreads)
{
    long ntot = (long)num[0] * num[1] * num[2];
    memset(tcsf, 0, sizeof(double) * 2 * ntot);
    set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) collapse(2)
    for (long h = 0; h < fnum[2]; ++h)
        for (long i = 0; i < fnum[1]; ++i) {
            for (long j = 0; j < fnum[0]; ++j) {
                long double hv[3];
                for (int c = 0; c < 3; ++c)
                    hv[c] = (long double)xm[c] + (long double)j * uin[c] + (long double)i * vin[c] +
                            (long double)h * win[c];
                double sr = 0.0, si = 0.0;
                for (long k = 0; k < nxat; ++k) {
                    long double ph = hv[0] * xat[k] + hv[1] * xat[k + ld] + hv[2] * xat[k + 2 * ld];
                    ph -= floorl(ph);
                    double s, c;
                    sincos(ORC_ZPI * (double)ph, &s, &c);
                    sr += c;
                    si += s;
                }
                long p = j + (long)num[0] * (i + (long)num[1] * h);
                tcsf[2 * p] = sr;
                tcsf[2 * p + 1] = si;
            }
        }
    if (lform) form_multiply(num, fnum, cfact_col, istl, tcsf);
}

/* ------------------------------------------------------------------------ */
/* four_test_friedel (fourier_sup.f90:3645-3728).  Returns the halved axis
 * (1..3) or 0; fnum receives the reduced sizes. */
int orc_test_friedel(const double eck[12], const double vi[9], const int inc[3], int ano,
                     int four_friedel, int fnum[3])
{
    fnum[0] = inc[0]; fnum[1] = inc[1]; fnum[2] = inc[2];
    if (!four_friedel || ano) return 0;
    double point[3];
    for (int r = 0; r < 3; ++r) {
        double rut = eck[r] + vi[r] * (fnum[0] - 1) + vi[r + 3] * (fnum[1] - 1) + vi[r + 6] * (fnum[2] - 1);
        point[r] = eck[r] + rut;
    }
    const double eps = (double)1.E-6f;  /* REAL(PREC_DP) :: EPS = 1.E-6 (single literal) */
    if ((fnum[0] % 2) && (fnum[1] % 2) && (fnum[2] % 2)) {
        if (fabs(point[0]) < eps && fabs(point[2]) < eps && fabs(point[1]) < eps) {
            if (fnum[0] > 1) { fnum[0] = (fnum[0] + 1) / 2; return 1; }
            if (fnum[1] > 1) { fnum[1] = (fnum[1] + 1) / 2; return 2; }
            if (fnum[2] > 1) { fnum[2] = (fnum[2] + 1) / 2; return 3; }
        }
    }
    return 0;
}

/* four_apply_friedel (fourier_sup.f90:3732-3761), the sequential in-place
 * loop exactly as written (i outer, k inner). */
void orc_apply_friedel(const int num[3], const int fnum[3], double *acsf, double *csf, double *dsi)
{
    for (long i = 0; i < fnum[0]; ++i)
        for (long j = 0; j < fnum[1]; ++j)
            for (long k = 0; k < fnum[2]; ++k) {
                long s = i + (long)num[0] * (j + (long)num[1] * k);
                long t = (num[0] - 1 - i) + (long)num[0] * ((num[1] - 1 - j) + (long)num[1] * (num[2] - 1 - k));
                if (acsf) { acsf[2 * t] = acsf[2 * s]; acsf[2 * t + 1] = -acsf[2 * s + 1]; }
                if (csf) { csf[2 * t] = csf[2 * s]; csf[2 * t + 1] = -csf[2 * s + 1]; }
                if (dsi) dsi[t] = dsi[s];
            }
}

/* ------------------------------------------------------------------------ */
/* Crystal + run description for the run-level restatement. */
typedef struct {
    const double *cr_pos;      /* (3, natoms) */
    const int *cr_iscat;       /* (natoms), 1-based types, 0 = void */
    long natoms;
    int icc[3];
    int ncatoms;
    double dim0[3];            /* cr_dim0(:,1) */
    double dim_lo[3];          /* cr_dim(:,1)  */
    int nscat;
    double eck[12];
    double vi[9];
    int inc[3];
    double rten[9];
    const double *cfact;       /* (CFPKT+1, nscat) complex */
    int ilots;
    int ls[3];
    int nlots;
    const int *lbeg;           /* (3, nlots), 1-based cells */
    int lperiod;
    double fave;
    const unsigned char *sel_cell; /* csize(1)*csize(2)*csize(3), i fastest */
    int four_friedel;
    int ano;
    int mode;                  /* MODE_EXACT / MODE_TURBO */
    double rnd[6];
    double diff_scale;
} orc_run_t;

/* four_csize (fourier_sup.f90:2533-2555). */
void orc_csize(const int icc[3], int lperiod, const int ls[3], int csize[3])
{
    for (int i = 0; i < 3; ++i) {
        if (lperiod) csize[i] = icc[i];
        else { csize[i] = icc[i] - ls[i] + 1; if (csize[i] < 1) csize[i] = 1; }
    }
}

/* four_getatm (fourier_sup.f90:1806-1970) for isotropic crystals.
 * Returns nxat; writes xat(ld,3); *ncell receives the lot cell count. */
long orc_getatm(const orc_run_t *p, int iscat, const int lbeg[3], double *xat, long ld, long *ncell)
{
    long nxat = 0;
    long cr_end = (long)p->ncatoms * p->icc[0] * p->icc[1] * p->icc[2]; /* 0-based first extra */
    *ncell = 0;
    if (p->ilots == LOT_OFF) {
        *ncell = (long)p->icc[0] * p->icc[1] * p->icc[2];
        for (long i = 0; i < p->natoms; ++i)
            if (p->cr_iscat[i] == iscat) {
                for (int c = 0; c < 3; ++c) xat[nxat + ld * c] = p->cr_pos[3 * i + c];
                ++nxat;
            }
        return nxat;
    }
    double x0[3];
    for (int c = 0; c < 3; ++c) x0[c] = (double)p->ls[c] / 2.0;
    for (int kk = 0; kk < p->ls[2]; ++kk) {
        int icell[3];
        double offset[3], xtest[3];
        icell[2] = kk + lbeg[2];
        offset[2] = p->dim0[2] + lbeg[2] - 1;
        xtest[2] = ((double)kk - x0[2] + 0.5) * ((double)kk - x0[2] + 0.5) / (x0[2] * x0[2]);
        if (icell[2] > p->icc[2]) { icell[2] -= p->icc[2]; offset[2] -= (double)p->icc[2]; }
        for (int jj = 0; jj < p->ls[1]; ++jj) {
            icell[1] = jj + lbeg[1];
            offset[1] = p->dim0[1] + lbeg[1] - 1;
            xtest[1] = ((double)jj - x0[1] + 0.5) * ((double)jj - x0[1] + 0.5) / (x0[1] * x0[1]);
            if (icell[1] > p->icc[1]) { icell[1] -= p->icc[1]; offset[1] -= (double)p->icc[1]; }
            for (int ii = 0; ii < p->ls[0]; ++ii) {
                icell[0] = ii + lbeg[0];
                offset[0] = p->dim0[0] + lbeg[0] - 1;
                xtest[0] = ((double)ii - x0[0] + 0.5) * ((double)ii - x0[0] + 0.5) / (x0[0] * x0[0]);
                if (icell[0] > p->icc[0]) { icell[0] -= p->icc[0]; offset[0] -= (double)p->icc[0]; }
                if (p->ilots == LOT_ELI && !(xtest[0] + xtest[1] + xtest[2] <= 1.0)) continue;
                ++*ncell;
                for (int is = 1; is <= p->ncatoms; ++is) {
                    /* celltoindex (celltoindex.f90:21), 0-based */
                    long ia = (((long)(icell[2] - 1) * p->icc[0] * p->icc[1] +
                                (long)(icell[1] - 1) * p->icc[0] + (icell[0] - 1)) * p->ncatoms + is) - 1;
                    if (p->cr_iscat[ia] == iscat) {
                        for (int c = 0; c < 3; ++c) xat[nxat + ld * c] = p->cr_pos[3 * ia + c] - offset[c];
                        ++nxat;
                    }
                }
                for (long ir = cr_end; ir < p->natoms; ++ir) {
                    int ok = p->cr_iscat[ir] == iscat;
                    for (int c = 0; c < 3 && ok; ++c) {
                        int jcell = icell[c] + (int)llround(p->dim_lo[c]) - 1;
                        ok = (int)p->cr_pos[3 * ir + c] == jcell;
                    }
                    if (ok) {
                        for (int c = 0; c < 3; ++c) xat[nxat + ld * c] = p->cr_pos[3 * ir + c] - offset[c];
                        ++nxat;
                    }
                }
            }
        }
    }
    return nxat;
}

/* four_getav (fourier_sup.f90:2440-2510): lattice points of the lot shape. */
long orc_getav(const orc_run_t *p, double *xat, long ld)
{
    long n = 0;
    if (p->ilots == LOT_OFF) {
        for (int ii = 0; ii < p->icc[0]; ++ii)
            for (int jj = 0; jj < p->icc[1]; ++jj)
                for (int kk = 0; kk < p->icc[2]; ++kk) {
                    xat[n] = (double)ii + p->dim0[0];
                    xat[n + ld] = (double)jj + p->dim0[1];
                    xat[n + 2 * ld] = (double)kk + p->dim0[2];
                    ++n;
                }
        return n;
    }
    double x0[3];
    for (int c = 0; c < 3; ++c) x0[c] = (double)p->ls[c] / 2.0;
    for (int kk = 0; kk < p->ls[2]; ++kk) {
        double xt2 = ((double)kk - x0[2] + 0.5) * ((double)kk - x0[2] + 0.5) / (x0[2] * x0[2]);
        for (int jj = 0; jj < p->ls[1]; ++jj) {
            double xt1 = ((double)jj - x0[1] + 0.5) * ((double)jj - x0[1] + 0.5) / (x0[1] * x0[1]);
            for (int ii = 0; ii < p->ls[0]; ++ii) {
                double xt0 = ((double)ii - x0[0] + 0.5) * ((double)ii - x0[0] + 0.5) / (x0[0] * x0[0]);
                if (p->ilots == LOT_ELI && !(xt0 + xt1 + xt2 <= 1.0)) continue;
                xat[n] = ii; xat[n + ld] = jj; xat[n + 2 * ld] = kk;
                ++n;
            }
        }
    }
    return n;
}

typedef struct {
    const orc_run_t *p;
    double xm[3], uin[3], vin[3], win[3], off[9];
    int lmn[6], num[3];
    double *cex;
    int *istl;
    int nthreads;
} orc_ctx_t;

static void strucf(const orc_ctx_t *c, const double *xat, long nxat, long ld, int iscat, int lform,
                   const int fnum[3], double *tcsf)
{
    const double *col = lform ? c->p->cfact + 2L * (ORC_CFPKT + 1) * (iscat - 1) : NULL;
    if (c->p->mode == MODE_TURBO)
        orc_strucf_turbo(xat, nxat, ld, c->uin, c->vin, c->win, c->off, c->lmn, c->num, fnum,
                         c->cex, col, c->istl, lform, tcsf, c->nthreads);
    else
        orc_strucf_exact(xat, nxat, ld, c->xm, c->uin, c->vin, c->win, c->num, fnum, col, c->istl,
                         lform, tcsf, c->nthreads);
}

/* four_aver (fourier_sup.f90:1194-1357).  acsf must be zeroed by the caller.
 * Returns 0, or +1 (no cell selected: norm 0, lines 1319-1327). */
static int aver(const orc_ctx_t *c, const int fnum[3], double *acsf, double *xat, long ld, double *tcsf)
{
    const orc_run_t *p = c->p;
    if (p->fave == 0.0) return 0;
    int csize[3];
    orc_csize(p->icc, p->lperiod, p->ls, csize);
    long ntot = (long)c->num[0] * c->num[1] * c->num[2];
    long ncell = 0;
    for (int iscat = 1; iscat <= p->nscat; ++iscat) {
        long nxat = 0;
        for (int ii = 1; ii <= csize[0]; ++ii)
            for (int jj = 1; jj <= csize[1]; ++jj)
                for (int kk = 1; kk <= csize[2]; ++kk) {
                    if (!p->sel_cell[(ii - 1) + (long)csize[0] * ((jj - 1) + (long)csize[1] * (kk - 1))]) continue;
                    ++ncell;
                    int icell[3] = {ii, jj, kk};
                    for (int is = 1; is <= p->ncatoms; ++is) {
                        long ia = (((long)(kk - 1) * p->icc[0] * p->icc[1] + (long)(jj - 1) * p->icc[0] +
                                    (ii - 1)) * p->ncatoms + is) - 1;
                        if (p->cr_iscat[ia] == iscat) {
                            for (int d = 0; d < 3; ++d)
                                xat[nxat + ld * d] = p->cr_pos[3 * ia + d] - (double)(icell[d] - 1) - p->dim0[d];
                            ++nxat;
                        }
                    }
                }
        strucf(c, xat, nxat, ld, iscat, 1, fnum, tcsf);
        for (long q = 0; q < 2 * ntot; ++q) acsf[q] += tcsf[q];
    }
    ncell = ncell / p->nscat;
    long nl = orc_getav(p, xat, ld);
    strucf(c, xat, nl, ld, 0, 0, c->num, tcsf);
    double norm = ncell > 0 ? 1.0 / (double)ncell : 0.0;
    for (long q = 0; q < ntot; ++q) {
        double ar = acsf[2 * q], ai = acsf[2 * q + 1], tr = tcsf[2 * q], ti = tcsf[2 * q + 1];
        acsf[2 * q] = (ar * tr - ai * ti) * norm;
        acsf[2 * q + 1] = (ar * ti + ai * tr) * norm;
    }
    return ncell > 0 ? 0 : 1;
}

/* four_run (fourier_sup.f90:11-323) for isotropic crystals, up to and
 * including the lot normalisation and diff_scale (lines 273-282).  The
 * post-processing that follows (four_weight, filters, convolution,
 * accumulate, qinfo) is not part of this path.
 * Optional outputs: fmean (complex, mean of F over lots) and dvar (real,
 * dnorm * sum_l |F_l - <F>|^2, Welford in FP64), both NULL-able.
 * Returns 0 or a DISCUS error code (-3, -132). */
int orc_four_run(const orc_run_t *p, double *csf, double *acsf, double *dsi, double *fmean,
                 double *dvar, int nthreads)
{
    orc_ctx_t c;
    memset(&c, 0, sizeof(c));
    c.p = p;
    c.nthreads = nthreads;
    for (int r = 0; r < 3; ++r) {
        c.xm[r] = p->eck[r];
        c.uin[r] = p->vi[r];
        c.vin[r] = p->vi[r + 3];
        c.win[r] = p->vi[r + 6];
        c.num[r] = p->inc[r];
    }
    long ntot = (long)c.num[0] * c.num[1] * c.num[2];
    memset(csf, 0, sizeof(double) * 2 * ntot);
    memset(acsf, 0, sizeof(double) * 2 * ntot);
    memset(dsi, 0, sizeof(double) * ntot);
    int st = orc_fourier_lmn(p->eck, p->vi, p->inc, p->rnd, c.lmn, c.off);
    if (st) return st;
    c.cex = (double *)malloc(sizeof(double) * 2 * ORC_I2PI);
    orc_cexpt(c.cex);
    c.istl = (int *)malloc(sizeof(int) * ntot);
    st = orc_stltab(c.xm, c.uin, c.vin, c.win, c.num, p->rten, c.istl);
    if (st) { free(c.cex); free(c.istl); return st; }
    int fnum[3];
    orc_test_friedel(p->eck, p->vi, p->inc, p->ano, p->four_friedel, fnum);
    int friedel = fnum[0] != c.num[0] || fnum[1] != c.num[1] || fnum[2] != c.num[2];

    long ld = p->natoms > 0 ? p->natoms : 1;
    long nlat = p->ilots == LOT_OFF ? (long)p->icc[0] * p->icc[1] * p->icc[2] : (long)p->ls[0] * p->ls[1] * p->ls[2];
    if (nlat > ld) ld = nlat;
    double *xat = (double *)malloc(sizeof(double) * 3 * ld);
    double *tcsf = (double *)malloc(sizeof(double) * 2 * ntot);
    aver(&c, fnum, acsf, xat, ld, tcsf);
    double *wm = NULL, *wv = NULL;
    if (fmean || dvar) {
        wm = (double *)calloc(2 * ntot, sizeof(double));
        wv = (double *)calloc(ntot, sizeof(double));
    }
    long ncell = 0;
    for (int nlot = 0; nlot < p->nlots; ++nlot) {
        memset(csf, 0, sizeof(double) * 2 * ntot);
        const int *lb = p->lbeg + 3 * nlot;
        for (int iscat = 1; iscat <= p->nscat; ++iscat) {
            long nxat = orc_getatm(p, iscat, lb, xat, ld, &ncell);
            if (nxat > 0) {
                strucf(&c, xat, nxat, ld, iscat, 1, fnum, tcsf);
                for (long q = 0; q < 2 * ntot; ++q) csf[q] += tcsf[q];
            }
        }
        if (wm) {
            /* Welford update over the computed (fnum) region, FP64 */
            double inv = 1.0 / (double)(nlot + 1);
            for (long q = 0; q < ntot; ++q) {
                double fr = csf[2 * q], fi = csf[2 * q + 1];
                double dr = fr - wm[2 * q], di = fi - wm[2 * q + 1];
                wm[2 * q] += dr * inv;
                wm[2 * q + 1] += di * inv;
                wv[q] += dr * (fr - wm[2 * q]) + di * (fi - wm[2 * q + 1]);
            }
        }
        for (long q = 0; q < ntot; ++q) {
            csf[2 * q] -= acsf[2 * q];
            csf[2 * q + 1] -= acsf[2 * q + 1];
            dsi[q] += csf[2 * q] * csf[2 * q] + csf[2 * q + 1] * csf[2 * q + 1];
        }
        if (friedel) orc_apply_friedel(c.num, fnum, acsf, csf, dsi);
        if (friedel && wm) orc_apply_friedel(c.num, fnum, NULL, wm, wv);
    }
    double dnorm = 1.0;
    if (p->nlots != 1) {
        dnorm = (double)((long)p->icc[0] * p->icc[1] * p->icc[2]) / (double)(ncell * p->nlots);
        for (long q = 0; q < ntot; ++q) dsi[q] *= dnorm;
    }
    double sq = sqrt(p->diff_scale);
    for (long q = 0; q < 2 * ntot; ++q) csf[q] *= sq;
    for (long q = 0; q < ntot; ++q) dsi[q] *= p->diff_scale;
    if (fmean) memcpy(fmean, wm, sizeof(double) * 2 * ntot);
    if (dvar) {
        double s = (p->nlots != 1 ? dnorm : 1.0) * p->diff_scale;
        for (long q = 0; q < ntot; ++q) dvar[q] = wv[q] * s;
    }
    free(wm); free(wv);
    free(xat); free(tcsf); free(c.cex); free(c.istl);
    return 0;
}

/* Size helper for the Python wrapper. */
long orc_sizeof_run(void) { return (long)sizeof(orc_run_t); }

/* ------------------------------------------------------------------------ */
/* Refinement model with per-atom occupancy and isotropic B (the product's
 * dc_model_* entry points; REFINE varies these per atom,
 * refine/prog/refine_run.f90:491-600):
 *   F(h) = sum_n f_t(bin h) occ_n exp(-B_n s^2) exp(2 pi i h.x_n),  I = |F|^2
 * over h = eck(:,1) + i vi(:,1) + j vi(:,2) + l vi(:,3), s^2 = h.rten.h / 4,
 * bin = four_stltab's nearest bin (orc_istl_point).  pos is (3,N), cfact
 * (ncfact, nscat) complex, type 1-based with 0 = vacancy. */
typedef struct {
    long double h[3];
    double s2;
    int bin;
} orc_pix_t;

static orc_pix_t *model_pixels(const double eck[12], const double vi[9], const int inc[3], const double rten[9],
                               long *np)
{
    long P = (long)inc[0] * inc[1] * inc[2];
    orc_pix_t *px = (orc_pix_t *)malloc(sizeof(orc_pix_t) * (P > 0 ? P : 1));
    for (long p = 0; p < P; ++p) {
        int i = (int)(p % inc[0]), j = (int)((p / inc[0]) % inc[1]), l = (int)(p / ((long)inc[0] * inc[1]));
        double hd[3];
        for (int c = 0; c < 3; ++c) {
            px[p].h[c] = (long double)eck[c] + (long double)i * vi[c] + (long double)j * vi[3 + c] +
                         (long double)l * vi[6 + c];
            hd[c] = (double)px[p].h[c];
        }
        double q = 0.0;
        for (int r = 0; r < 3; ++r)
            for (int k = 0; k < 3; ++k) q += hd[r] * rten[r + 3 * k] * hd[k];
        px[p].s2 = q / 4.0;
        px[p].bin = orc_istl_point(eck, vi, vi + 3, vi + 6, i, j, l, rten);
    }
    *np = P;
    return px;
}

/* one atom's E = exp(-B s^2) exp(2 pi i h.x) at one pixel */
static void atom_e(const orc_pix_t *p, const double *x, double b, double *er, double *ei)
{
    long double ph = p->h[0] * x[0] + p->h[1] * x[1] + p->h[2] * x[2];
    ph -= floorl(ph);
    double s, c;
    sincos(ORC_ZPI * (double)ph, &s, &c);
    double d = exp(-b * p->s2);
    *er = c * d;
    *ei = s * d;
}

void orc_model_f(const double *pos, const int *type, const double *occ, const double *bdw, long n,
                 const double *cfact, int ncfact, const double eck[12], const double vi[9], const int inc[3],
                 const double rten[9], double *F, int nthreads)
{
    long P;
    orc_pix_t *px = model_pixels(eck, vi, inc, rten, &P);
    set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
    for (long p = 0; p < P; ++p) {
        double fr = 0.0, fi = 0.0;
        for (long k = 0; k < n; ++k) {
            if (type[k] <= 0) continue;
            double er, ei;
            atom_e(&px[p], pos + 3 * k, bdw ? bdw[k] : 0.0, &er, &ei);
            const double *u = cfact + 2 * ((long)px[p].bin + (long)ncfact * (type[k] - 1));
            double ar = occ[k] * (u[0] * er - u[1] * ei), ai = occ[k] * (u[0] * ei + u[1] * er);
            fr += ar;
            fi += ai;
        }
        F[2 * p] = fr;
        F[2 * p + 1] = fi;
    }
    free(px);
}

/* Direct FP64 gradient of L with dL/dI = G (per pixel), given F of the model:
 * grad(5,N) = (dL/dx, dL/dy, dL/dz, dL/docc, dL/dB); vacancies get zeros. */
void orc_model_grad(const double *pos, const int *type, const double *occ, const double *bdw, long n,
                    const double *cfact, int ncfact, const double eck[12], const double vi[9], const int inc[3],
                    const double rten[9], const double *F, const double *G, double *grad, int nthreads)
{
    long P;
    orc_pix_t *px = model_pixels(eck, vi, inc, rten, &P);
    set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4)
    for (long k = 0; k < n; ++k) {
        double *g = grad + 5 * k;
        for (int c = 0; c < 5; ++c) g[c] = 0.0;
        if (type[k] <= 0) continue;
        double s0 = 0.0, sh[3] = {0.0, 0.0, 0.0}, ss = 0.0;
        for (long p = 0; p < P; ++p) {
            double er, ei;
            atom_e(&px[p], pos + 3 * k, bdw ? bdw[k] : 0.0, &er, &ei);
            const double *u = cfact + 2 * ((long)px[p].bin + (long)ncfact * (type[k] - 1));
            /* Q = 2 G conj(F) u */
            double cr = 2.0 * G[p] * F[2 * p], ci = -2.0 * G[p] * F[2 * p + 1];
            double qr = cr * u[0] - ci * u[1], qi = cr * u[1] + ci * u[0];
            double tr = qr * er - qi * ei, ti = qr * ei + qi * er;
            s0 += tr;
            for (int c = 0; c < 3; ++c) sh[c] += (double)px[p].h[c] * ti;
            ss += px[p].s2 * tr;
        }
        for (int c = 0; c < 3; ++c) g[c] = -ORC_ZPI * occ[k] * sh[c];
        g[3] = s0;
        g[4] = -occ[k] * ss;
    }
    free(px);
}

/* Central differences of chi^2 = sum w (|F|^2 - T)^2 for nparam single
 * parameters (atom[q], kind[q] in 0..4 = x, y, z, occ, B), step h, from the
 * model's exact F0.  Only the perturbed atom's term changes, so each
 * difference is formed per pixel as w (I+ - I-)(I+ + I- - 2T) with
 * I+ - I- = 2 Re(conj F0 (d+ - d-)) + |d+|^2 - |d-|^2, d = change of the
 * atom's term: no cancellation between two large chi^2 sums. */
void orc_model_fd(const double *pos, const int *type, const double *occ, const double *bdw, long n,
                  const double *cfact, int ncfact, const double eck[12], const double vi[9], const int inc[3],
                  const double rten[9], const double *F0, const double *target, const double *weight,
                  int nparam, const int *atom, const int *kind, double h, double *out, int nthreads)
{
    long P;
    orc_pix_t *px = model_pixels(eck, vi, inc, rten, &P);
    (void)n;
    set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
    for (int q = 0; q < nparam; ++q) {
        const long k = atom[q];
        const int kd = kind[q];
        if (type[k] <= 0) { out[q] = 0.0; continue; }
        double x0[3] = {pos[3 * k], pos[3 * k + 1], pos[3 * k + 2]};
        double o0 = occ[k], b0 = bdw ? bdw[k] : 0.0;
        double acc = 0.0;
        for (long p = 0; p < P; ++p) {
            const double *u = cfact + 2 * ((long)px[p].bin + (long)ncfact * (type[k] - 1));
            double t[3][2];   /* atom term at theta, theta+h, theta-h */
            for (int v = 0; v < 3; ++v) {
                double x[3] = {x0[0], x0[1], x0[2]}, o = o0, b = b0;
                double dlt = v == 0 ? 0.0 : (v == 1 ? h : -h);
                if (kd < 3) x[kd] += dlt; else if (kd == 3) o += dlt; else b += dlt;
                double er, ei;
                atom_e(&px[p], x, b, &er, &ei);
                t[v][0] = o * (u[0] * er - u[1] * ei);
                t[v][1] = o * (u[0] * ei + u[1] * er);
            }
            double dpr = t[1][0] - t[0][0], dpi = t[1][1] - t[0][1];
            double dmr = t[2][0] - t[0][0], dmi = t[2][1] - t[0][1];
            double fr = F0[2 * p], fi = F0[2 * p + 1];
            double di = 2.0 * (fr * (dpr - dmr) + fi * (dpi - dmi)) + (dpr * dpr + dpi * dpi) - (dmr * dmr + dmi * dmi);
            double i0 = fr * fr + fi * fi;
            double ip = (fr + dpr) * (fr + dpr) + (fi + dpi) * (fi + dpi);
            double im = (fr + dmr) * (fr + dmr) + (fi + dmi) * (fi + dmi);
            (void)i0;
            double w = weight ? weight[p] : 1.0;
            acc += w * di * (ip + im - 2.0 * target[p]);
        }
        out[q] = acc / (2.0 * h);
    }
    free(px);
}
