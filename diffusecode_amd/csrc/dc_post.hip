// dc_post.hip -- post-processing of a finished Fourier run on the GPU:
// four_weight, the Lanczos filter, the resolution convolution four_conv, the
// point-group average and Bragg fill of four_accumulate, and qval.
//
// Convolutions follow DISCUS's FFT recipe (fourier_conv.f90:7-423,
// lib_conv.f90:10-269): the field is shifted into FFT order by
// maptofftfd (lib_f90/map_1dtofield.f90), transformed forward, multiplied by
// the kernel's spectrum, transformed FORWARD again and read back by
// fftfdtoline.  The second forward transform reverses the index, which the
// reverse map undoes.  The maps differ per dimension and, in four_conv_2D,
// per axis order; every variant reduces to two per-axis parity terms:
//   forward   FFT index = (i + 1 + m/2 - lf) mod m          (natural i)
//   reverse   natural i <- FFT index (m - 1 - i - m/2 + lr) mod m
// with lf = (m-1) mod 2 everywhere, and lr = lf (3D, 2D Lanczos), 0 (1D), or
// the other axis's parity (four_conv_2D when num(2) > num(1), where
// fftfdtoline_22D overwrites l in natural order and reads it in FFT order,
// map_1dtofield.f90:575-582).  The transforms run on natural axes: a DFT
// over all axes does not depend on their order.
//
// Layout: a field is `batch` sub-fields of dims m[0..2] (m[0] fastest) at
// natural strides s[0..2], batch stride bs.  The work buffer holds each
// sub-field contiguously in FFT order, abscissa fastest (hipFFT dims
// reversed).  Sums are FP64 in a fixed order (per-block trees, then one
// thread per sub-field), so results repeat bit for bit.
#include <hipfft/hipfft.h>

#include <cmath>
#include <map>
#include <tuple>

#include "dc_host.h"

using namespace dc;

namespace {

constexpr int SUM_BLOCKS = 512;
constexpr int SUM_THREADS = 256;

struct Field {
    int m[3];                  // sub-field dims (1 for unused axes)
    long long s[3];            // natural strides of the sub-field axes (elements)
    long long bs;              // natural stride between sub-fields
    int batch;
    __host__ __device__ long long sub() const { return (long long)m[0] * m[1] * m[2]; }
    __host__ __device__ long long total() const { return sub() * batch; }
};

struct MapAxes {
    int lf[3], lr[3];
};

__device__ inline int pmod(int a, int m) {
    int r = a % m;
    return r < 0 ? r + m : r;
}

// natural offset of element (b, i, j, k)
__device__ inline long long nat_off(const Field &f, int b, int i, int j, int k) {
    return (long long)b * f.bs + (long long)i * f.s[0] + (long long)j * f.s[1] + (long long)k * f.s[2];
}

__device__ inline void unflat(const Field &f, long long e, int &b, int &i, int &j, int &k) {
    const long long sub = f.sub();
    b = (int)(e / sub);
    long long r = e - (long long)b * sub;
    i = (int)(r % f.m[0]);
    r /= f.m[0];
    j = (int)(r % f.m[1]);
    k = (int)(r / f.m[1]);
}

// Gather into FFT order: FFT element (b, ii, jj, kk) takes natural
// i = (ii - 1 - m/2 + lf) mod m along each axis.
template <bool CPLX>
__global__ void map_fwd_kernel(Field f, MapAxes a, const double *__restrict__ src, double2 *__restrict__ work) {
    const long long n = f.total();
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        int b, ii, jj, kk;
        unflat(f, e, b, ii, jj, kk);
        const int i = pmod(ii - 1 - f.m[0] / 2 + a.lf[0], f.m[0]);
        const int j = pmod(jj - 1 - f.m[1] / 2 + a.lf[1], f.m[1]);
        const int k = pmod(kk - 1 - f.m[2] / 2 + a.lf[2], f.m[2]);
        const long long o = nat_off(f, b, i, j, k);
        double2 v;
        if (CPLX) {
            v.x = src[2 * o];
            v.y = src[2 * o + 1];
        } else {
            v.x = src[o];
            v.y = 0.0;
        }
        work[e] = v;
    }
}

// Gather back: natural (b, i, j, k) <- FFT element (m - 1 - i - m/2 + lr) mod m,
// divided by `div` (DISCUS's in_pattern / N, or lib_conv's weight).
template <bool CPLX>
__global__ void map_rev_kernel(Field f, MapAxes a, const double2 *__restrict__ work, double div, double *__restrict__ dst) {
    const long long n = f.total();
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        int b, i, j, k;
        unflat(f, e, b, i, j, k);
        const int ii = pmod(f.m[0] - 1 - i - f.m[0] / 2 + a.lr[0], f.m[0]);
        const int jj = pmod(f.m[1] - 1 - j - f.m[1] / 2 + a.lr[1], f.m[1]);
        const int kk = pmod(f.m[2] - 1 - k - f.m[2] / 2 + a.lr[2], f.m[2]);
        const double2 v = work[(long long)b * f.sub() + ii + (long long)f.m[0] * (jj + (long long)f.m[1] * kk)];
        const long long o = nat_off(f, b, i, j, k);
        if (CPLX) {
            dst[2 * o] = v.x / div;
            dst[2 * o + 1] = v.y / div;
        } else {
            dst[o] = v.x / div;
        }
    }
}

// work[b][x] = (spec[x] * sscale) * work[b][x]
__global__ void spec_mul_kernel(long long sub, long long n, const double2 *__restrict__ spec, double sscale,
                                double2 *__restrict__ work) {
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        const double2 t0 = spec[e % sub];
        const double tr = t0.x * sscale, ti = t0.y * sscale;
        const double2 w = work[e];
        work[e] = make_double2(tr * w.x - ti * w.y, tr * w.y + ti * w.x);
    }
}

// four_conv's profile at FFT index = pixel offset (fourier_conv.f90:96-104,
// 197-212, 336-355): offsets above m/2 wrap to negative, posit = diff_tr x
// offset, product of 1/sqrt(2 pi)/sigma exp(-posit^2/2) over the dimensions.
struct ProfArgs {
    int m[3];
    int nd;
    double tr[9];
    double sig[3];
};

__global__ void conv_profile_kernel(ProfArgs p, double2 *__restrict__ out) {
    const long long n = (long long)p.m[0] * p.m[1] * p.m[2];
    const double zpi = 6.283185307179586476925286766559;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        int idx[3];
        long long r = e;
        idx[0] = (int)(r % p.m[0]);
        r /= p.m[0];
        idx[1] = (int)(r % p.m[1]);
        idx[2] = (int)(r / p.m[1]);
        double v[3];
        for (int a = 0; a < 3; ++a) v[a] = (double)(idx[a] > p.m[a] / 2 ? idx[a] - p.m[a] : idx[a]);
        double val = 1.0;
        for (int rr = 0; rr < p.nd; ++rr) {
            const double pos = p.tr[rr] * v[0] + p.tr[rr + 3] * v[1] + p.tr[rr + 6] * v[2];
            val = val * (1.0 / sqrt(zpi) / p.sig[rr] * exp(-0.5 * pos * pos));
        }
        out[e] = make_double2(val, 0.0);
    }
}

// Lanczos sinc table of do_lanczos_*_fft (lib_lanczos.f90:258-458), written
// directly in FFT order: FFT element -> natural index by the forward map,
// table(i) = wsinc(q, g, rcut) * lqgm(q, g, m) for i (1-based) inside
// [hl, hh], q = centre / nscale, g = i / nscale, centre = (N + 1) / 2.
struct TabArgs {
    int m[3];
    int nd;
    int lf[3];
    int lo[3], hi[3], cen[3];
    double nscale, rcut;
    int width;
};

__device__ inline double sinc_d(double x) { return fabs(x) < 1.0e-7 ? 1.0 : sin(x) / x; }

__global__ void lanczos_table_kernel(TabArgs t, double2 *__restrict__ out) {
    const long long n = (long long)t.m[0] * t.m[1] * t.m[2];
    const double pi = 3.14159265358979323846264338328;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        int ff[3];
        long long r = e;
        ff[0] = (int)(r % t.m[0]);
        r /= t.m[0];
        ff[1] = (int)(r % t.m[1]);
        ff[2] = (int)(r / t.m[1]);
        double val = 0.0;
        bool in = true;
        double qg[3] = {0.0, 0.0, 0.0};
        for (int a = 0; a < t.nd; ++a) {
            const int i1 = pmod(ff[a] - 1 - t.m[a] / 2 + t.lf[a], t.m[a]) + 1;
            if (i1 < t.lo[a] || i1 > t.hi[a]) in = false;
            qg[a] = (double)t.cen[a] / t.nscale - (double)i1 / t.nscale;
        }
        if (in) {
            double w = 1.0, lq = 1.0;
            bool inside = true;
            for (int a = 0; a < t.nd; ++a) {
                w = w * sinc_d(pi * qg[a] * t.rcut);
                inside = inside && fabs(qg[a]) < (double)t.width;
                lq = lq * sinc_d(pi * qg[a] / (double)t.width);
            }
            if (t.nd == 1) w = w * w;   // wsinc_1d is sinc^2 (lib_lanczos.f90:523)
            val = inside ? w * lq : 0.0;
        }
        out[e] = make_double2(val, 0.0);
    }
}

// Separable circular stencil along one axis of a (batched) field:
// dst(p) = sum_u src(p + shift - r_u) v_u, indices mod m along the axis.
constexpr int MAXTAP = 64;
struct TapArgs {
    Field f;
    int axis, ntap, shift;
    int r[MAXTAP];
    double v[MAXTAP];
};

__global__ void axis_conv_kernel(TapArgs t, const double *__restrict__ src, double *__restrict__ dst) {
    const long long n = t.f.total();
    const int m = t.f.m[t.axis];
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        int b, c3[3];
        unflat(t.f, e, b, c3[0], c3[1], c3[2]);
        const int q = c3[t.axis];
        double acc = 0.0;
        for (int u = 0; u < t.ntap; ++u) {
            int cc[3] = {c3[0], c3[1], c3[2]};
            cc[t.axis] = pmod(q + t.shift - t.r[u], m);
            acc += src[nat_off(t.f, b, cc[0], cc[1], cc[2])] * t.v[u];
        }
        dst[nat_off(t.f, b, c3[0], c3[1], c3[2])] = acc;
    }
}

// FP64 sums of real sub-fields, fixed order: block g of sub-field b takes
// the contiguous chunk [g*chunk, (g+1)*chunk) of its elements (threads stride
// inside it); tree in LDS; part[b][g].  Contiguous sub-fields skip the
// index arithmetic.
__global__ __launch_bounds__(SUM_THREADS) void sum_part_kernel(Field f, int contiguous, const double *__restrict__ x,
                                                               double *__restrict__ part) {
    __shared__ double s[SUM_THREADS];
    const int b = blockIdx.y;
    const long long sub = f.sub();
    const long long chunk = (sub + SUM_BLOCKS - 1) / SUM_BLOCKS;
    const long long e0 = (long long)blockIdx.x * chunk;
    const long long e1 = e0 + chunk < sub ? e0 + chunk : sub;
    double acc = 0.0;
    if (contiguous) {
        const double *xb = x + (long long)b * f.bs;
        for (long long e = e0 + threadIdx.x; e < e1; e += SUM_THREADS) acc += xb[e];
    } else {
        for (long long e = e0 + threadIdx.x; e < e1; e += SUM_THREADS) {
            long long r = e;
            const int i = (int)(r % f.m[0]);
            r /= f.m[0];
            const int j = (int)(r % f.m[1]);
            const int k = (int)(r / f.m[1]);
            acc += x[nat_off(f, b, i, j, k)];
        }
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = SUM_THREADS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(long long)b * SUM_BLOCKS + blockIdx.x] = s[0];
}

__global__ void sum_finish_kernel(int batch, const double *__restrict__ part, double *__restrict__ out) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch) return;
    double acc = 0.0;
    for (int g = 0; g < SUM_BLOCKS; ++g) acc += part[(long long)b * SUM_BLOCKS + g];
    out[b] = acc;
}

// x[b] = (x[b] * num[b]) / den[b]          (four_conv: csf*weight/weight2)
// x[b] = x[b] * (num[b] / den[b] * mul)    (do_lanczos: outfield*factor)
template <bool CPLX, bool LANCZOS>
__global__ void rescale_kernel(Field f, const double *__restrict__ num, const double *__restrict__ den, double mul,
                               double *__restrict__ x) {
    const long long n = f.total();
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        int b, i, j, k;
        unflat(f, e, b, i, j, k);
        const long long o = nat_off(f, b, i, j, k);
        if (LANCZOS) {
            const double fac = num[b] / den[b] * mul;
            x[o] = x[o] * fac;
        } else if (CPLX) {
            x[2 * o] = x[2 * o] * num[b] / den[b];
            x[2 * o + 1] = x[2 * o + 1] * num[b] / den[b];
        } else {
            x[o] = x[o] * num[b] / den[b];
        }
    }
}

// ---- four_weight: Bragg pixels of dsi3d divided by scalef -----------------
struct WeightArgs {
    int n[3];
    double e[3];
    double v[9];
    double inv_ds;      // 1 / diff_scale applied first (dsi3d is taken before it)
    double scalef;
    int apply;          // fave == 0 and |scalef - 1| > 1e-5
};

__device__ inline double nint_d(double x) { return x >= 0.0 ? floor(x + 0.5) : -floor(-x + 0.5); }

__global__ void weight_kernel(WeightArgs a, const double *__restrict__ dsi, double *__restrict__ d3) {
#pragma clang fp contract(off)
    const long long n = (long long)a.n[0] * a.n[1] * a.n[2];
    const double tol = (double)1.0e-4f;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(e % a.n[0]);
        const int j = (int)((e / a.n[0]) % a.n[1]);
        const int k = (int)(e / ((long long)a.n[0] * a.n[1]));
        double v = dsi[e] * a.inv_ds;
        if (a.apply) {
            bool bragg = true;
            for (int r = 0; r < 3; ++r) {
                const double h = a.e[r] + a.v[r] * (double)i + a.v[r + 3] * (double)j + a.v[r + 6] * (double)k;
                bragg = bragg && fabs(h - nint_d(h)) < tol;
            }
            if (bragg) v = v / a.scalef;
        }
        d3[e] = v;
    }
}

// ---- point-group average (gather form of four_symm_csf/_dsi) -------------
constexpr int MAXOPS = 192;
struct SymmArgs {
    int n[3];
    int iz[3];
    int nop;                 // operations 2..nops
    double nops;             // spc_n / n_center
    int m[MAXOPS][9];        // exact integer inverses of the operations
};

template <bool CPLX>
__global__ void symm_kernel(const SymmArgs *__restrict__ a, const double *__restrict__ x, double *__restrict__ out) {
    const long long n = (long long)a->n[0] * a->n[1] * a->n[2];
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        const int t0 = (int)(e % a->n[0]) - a->iz[0];
        const int t1 = (int)((e / a->n[0]) % a->n[1]) - a->iz[1];
        const int t2 = (int)(e / ((long long)a->n[0] * a->n[1])) - a->iz[2];
        double sr = CPLX ? x[2 * e] : x[e];
        double si = CPLX ? x[2 * e + 1] : 0.0;
        int w = 1;
        for (int o = 0; o < a->nop; ++o) {
            const int *M = a->m[o];
            // source offset s = t . S^-1 (row vector times matrix), column-major M
            const int s0 = t0 * M[0] + t1 * M[1] + t2 * M[2] + a->iz[0];
            const int s1 = t0 * M[3] + t1 * M[4] + t2 * M[5] + a->iz[1];
            const int s2 = t0 * M[6] + t1 * M[7] + t2 * M[8] + a->iz[2];
            if (s0 < 0 || s0 >= a->n[0] || s1 < 0 || s1 >= a->n[1] || s2 < 0 || s2 >= a->n[2]) continue;
            const long long so = s0 + (long long)a->n[0] * (s1 + (long long)a->n[1] * s2);
            if (CPLX) {
                sr += x[2 * so];
                si += x[2 * so + 1];
            } else {
                sr += x[so];
            }
            ++w;
        }
        const double wd = (double)w;
        if (CPLX) {
            out[2 * e] = sr / wd / a->nops;
            out[2 * e + 1] = si / wd / a->nops;
        } else {
            out[e] = sr / wd / a->nops;
        }
    }
}

// ---- Bragg fill (fave < 0) ---------------------------------------------------
struct FillArgs {
    int n[3];
    double e[3];
    double v[9];
    double step[3];
};

__device__ inline bool is_bragg(const FillArgs &a, int h, int k, int l) {
#pragma clang fp contract(off)
    bool ok = true;
    for (int r = 0; r < 3; ++r) {
        const double x = a.e[r] + (double)h * a.v[r] + (double)k * a.v[r + 3] + (double)l * a.v[r + 6];
        ok = ok && fabs(x - nint_d(x)) < a.step[r];
    }
    return ok;
}

template <bool CPLX>
__device__ inline void fill_one(const FillArgs &a, int h, int k, int l, const double *src, double *dst) {
    double sr = 0.0, si = 0.0;
    int np = 0;
    const int nb[6][3] = {{h - 1, k, l}, {h + 1, k, l}, {h, k - 1, l}, {h, k + 1, l}, {h, k, l - 1}, {h, k, l + 1}};
    for (int q = 0; q < 6; ++q) {
        const int x = nb[q][0], y = nb[q][1], z = nb[q][2];
        if (x < 0 || x >= a.n[0] || y < 0 || y >= a.n[1] || z < 0 || z >= a.n[2]) continue;
        const long long o = x + (long long)a.n[0] * (y + (long long)a.n[1] * z);
        if (CPLX) {
            sr += src[2 * o];
            si += src[2 * o + 1];
        } else {
            sr += src[o];
        }
        ++np;
    }
    if (np > 0) {
        const long long o = h + (long long)a.n[0] * (k + (long long)a.n[1] * l);
        if (CPLX) {
            dst[2 * o] = sr / (double)np;
            dst[2 * o + 1] = si / (double)np;
        } else {
            dst[o] = sr / (double)np;
        }
    }
}

// No Bragg pixel has a Bragg neighbour: every fill reads original values, in
// any order.  src is a copy of the field, dst the field.
template <bool CPLX>
__global__ void fill_parallel_kernel(FillArgs a, const double *__restrict__ src, double *__restrict__ dst) {
    const long long n = (long long)a.n[0] * a.n[1] * a.n[2];
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        const int h = (int)(e % a.n[0]);
        const int k = (int)((e / a.n[0]) % a.n[1]);
        const int l = (int)(e / ((long long)a.n[0] * a.n[1]));
        if (is_bragg(a, h, k, l)) fill_one<CPLX>(a, h, k, l, src, dst);
    }
}

// Bragg pixels may be neighbours: DISCUS's sequential order (h outer, l
// inner) with updated values.  One wave: the 64 lanes test 64 consecutive
// pixels of the loop order, lane 0 fills the Bragg ones in order.
template <bool CPLX>
__global__ __launch_bounds__(64) void fill_serial_kernel(FillArgs a, double *__restrict__ x) {
    const long long n = (long long)a.n[0] * a.n[1] * a.n[2];
    __shared__ int flag[64];
    for (long long base = 0; base < n; base += 64) {
        const long long e = base + threadIdx.x;    // loop order index: l fastest, then k, then h
        int h = 0, k = 0, l = 0;
        bool b = false;
        if (e < n) {
            l = (int)(e % a.n[2]);
            k = (int)((e / a.n[2]) % a.n[1]);
            h = (int)(e / ((long long)a.n[2] * a.n[1]));
            b = is_bragg(a, h, k, l);
        }
        flag[threadIdx.x] = b ? 1 : 0;
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int q = 0; q < 64 && base + q < n; ++q) {
                if (!flag[q]) continue;
                const long long eq = base + q;
                const int lq = (int)(eq % a.n[2]);
                const int kq = (int)((eq / a.n[2]) % a.n[1]);
                const int hq = (int)(eq / ((long long)a.n[2] * a.n[1]));
                fill_one<CPLX>(a, hq, kq, lq, x, x);
            }
        }
        __syncthreads();
    }
}

__global__ void axpy_kernel(long long n, const double *__restrict__ x, double *__restrict__ y) {
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x)
        y[e] = y[e] + x[e];
}

// ---- qval ---------------------------------------------------------------------
struct QvalArgs {
    Grid g;                  // corner eck(:,1), increments, metric (istl)
    int value, laver, nscat, ncfact;
    double n_real, u2aver, faver_sign;
};

__global__ void qval_kernel(QvalArgs a, const double2 *__restrict__ csf, const double2 *__restrict__ acsf,
                            const double *__restrict__ dsi, const double2 *__restrict__ cf,
                            const double2 *__restrict__ cp, const double *__restrict__ amount, double *__restrict__ out) {
    const long long n = (long long)a.g.n[0] * a.g.n[1] * a.g.n[2];
    const double zpi = 6.283185307179586476925286766559;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(e % a.g.n[0]);
        const int j = (int)((e / a.g.n[0]) % a.g.n[1]);
        const double2 f = a.laver ? acsf[e] : csf[e];
        const double ff = f.x * f.x + f.y * f.y;
        double q = 0.0;
        switch (a.value) {
        case 1: q = a.laver ? ff : dsi[e]; break;
        case 2: q = sqrt(ff); break;
        case 3: q = (f.x == 0.0 && f.y == 0.0) ? 0.0 : atan2(f.y, f.x) * (180.0 / 3.14159265358979323846264338328); break;
        case 4: q = f.x; break;
        case 5: q = f.y; break;
        case 13: q = (a.laver ? ff : dsi[e]) / a.n_real; break;
        default: {
            // form-factor averages at istl(i, j, 1) (qval.f90:187-285)
            const int bin = istl_point(a.g, i, j, 0);
            const bool pure = a.value == 7 || a.value == 9 || a.value == 10 || a.value == 11;
            const double2 *tab = pure ? cp : cf;
            double s = 0.0;
            for (int t = 0; t < a.nscat; ++t) {
                const double2 c = tab[(long long)t * a.ncfact + bin];
                const double a2 = c.x * c.x + c.y * c.y;
                if (a.value == 9 || a.value == 17) {
                    s = s + a2 * amount[t];
                } else {
                    double sg = a.faver_sign;
                    if (a.value == 7 || a.value == 10 || a.value == 18) sg = cp[(long long)t * a.ncfact + 1].x < 0.0 ? -1.0 : 1.0;
                    s = s + sqrt(a2) * amount[t] * sg;
                }
            }
            if (a.value == 7) {
                const double fa2 = s * s;
                const int b1 = istl_point(a.g, i, 0, 0);   // istl(i, 1, 1) (qval.f90:204)
                const double arg = 2.0 * (double)b1 * CFINC;
                const float q2 = (float)(zpi * zpi * (arg * arg));
                q = (a.laver ? ff : dsi[e]) / fa2 / a.n_real + 1.0 - exp(-(double)q2 * a.u2aver);
            } else if (a.value == 10 || a.value == 18) {
                q = s / a.n_real;
                q = q * q;
            } else {
                q = s / a.n_real;
            }
        }
        }
        out[e] = q;
    }
}

inline int grid_for(long long n) {
    long long g = (n + 255) / 256;
    if (g > 8192) g = 8192;
    return (int)(g < 1 ? 1 : g);
}

}  // namespace

// ---- host state --------------------------------------------------------------
struct PostState {
    std::map<std::tuple<int, int, int, int>, hipfftHandle> plans;   // (m0, m1, m2, batch)
    DevBuf work, spec, part, sums, copy, sym, stage[3], q_cf, q_cp, q_amt;
    SymmArgs h_sym;            // host copy of the operation table (source of its async upload)
};

namespace dc {
void post_release(dc_ctx *c) {
    if (!c->post) return;
    for (auto &kv : c->post->plans) (void)hipfftDestroy(kv.second);
    for (DevBuf *b : {&c->post->work, &c->post->spec, &c->post->part, &c->post->sums, &c->post->copy, &c->post->sym,
                      &c->post->stage[0], &c->post->stage[1], &c->post->stage[2], &c->post->q_cf, &c->post->q_cp,
                      &c->post->q_amt})
        if (b->p) (void)hipFree(b->p);
    delete c->post;
    c->post = nullptr;
}
}  // namespace dc

namespace {

PostState &post_state(dc_ctx *c) {
    if (!c->post) c->post = new PostState();
    return *c->post;
}

int fft_plan(dc_ctx *c, const Field &f, hipfftHandle &out) {
    PostState &ps = post_state(c);
    auto key = std::make_tuple(f.m[0], f.m[1], f.m[2], f.batch);
    auto it = ps.plans.find(key);
    if (it != ps.plans.end()) {
        out = it->second;
    } else {
        int nd = f.m[2] > 1 ? 3 : (f.m[1] > 1 ? 2 : 1);
        int dims[3];
        for (int a = 0; a < nd; ++a) dims[a] = f.m[nd - 1 - a];   // row-major: slowest first
        hipfftHandle h;
        hipfftResult r = hipfftPlanMany(&h, nd, dims, nullptr, 1, (int)f.sub(), nullptr, 1, (int)f.sub(), HIPFFT_Z2Z,
                                        f.batch);
        if (r != HIPFFT_SUCCESS) return set_err(c, DC_ERR_DEVICE, DC_ER_LIB, "hipfftPlanMany failed (" + std::to_string((int)r) + ")");
        ps.plans[key] = h;
        out = h;
    }
    if (hipfftSetStream(out, c->st) != HIPFFT_SUCCESS) return set_err(c, DC_ERR_DEVICE, DC_ER_LIB, "hipfftSetStream failed");
    return 0;
}

int fft_fwd(dc_ctx *c, const Field &f, double2 *buf) {
    hipfftHandle h;
    int st = fft_plan(c, f, h);
    if (st) return st;
    hipfftResult r = hipfftExecZ2Z(h, reinterpret_cast<hipfftDoubleComplex *>(buf), reinterpret_cast<hipfftDoubleComplex *>(buf),
                                   HIPFFT_FORWARD);
    if (r != HIPFFT_SUCCESS) return set_err(c, DC_ERR_DEVICE, DC_ER_LIB, "hipfftExecZ2Z failed (" + std::to_string((int)r) + ")");
    return 0;
}

// sums[b] of a real field into `out` (device, batch doubles)
int field_sums(dc_ctx *c, const Field &f, const double *x, double *out) {
    PostState &ps = post_state(c);
    int st = ensure(c, ps.part, sizeof(double) * SUM_BLOCKS * (size_t)f.batch);
    if (st) return st;
    const int contiguous = f.s[0] == 1 && f.s[1] == f.m[0] && (f.m[2] == 1 || f.s[2] == (long long)f.m[0] * f.m[1]);
    sum_part_kernel<<<dim3(SUM_BLOCKS, f.batch), SUM_THREADS, 0, c->st>>>(f, contiguous, x, ptr<double>(ps.part));
    sum_finish_kernel<<<(f.batch + 63) / 64, 64, 0, c->st>>>(f.batch, ptr<double>(ps.part), out);
    HIPCHK(hipGetLastError());
    return 0;
}

// One field through map -> FFT -> * spec -> FFT -> reverse map (in place on x).
int convolve(dc_ctx *c, const Field &f, const MapAxes &ax, bool cplx, double *x, const double2 *spec, double sscale,
             double div) {
    PostState &ps = post_state(c);
    int st = ensure(c, ps.work, sizeof(double2) * (size_t)f.total());
    if (st) return st;
    double2 *w = ptr<double2>(ps.work);
    const int gr = grid_for(f.total());
    if (cplx)
        map_fwd_kernel<true><<<gr, 256, 0, c->st>>>(f, ax, x, w);
    else
        map_fwd_kernel<false><<<gr, 256, 0, c->st>>>(f, ax, x, w);
    HIPCHK(hipGetLastError());
    if ((st = fft_fwd(c, f, w))) return st;
    spec_mul_kernel<<<gr, 256, 0, c->st>>>(f.sub(), f.total(), spec, sscale, w);
    HIPCHK(hipGetLastError());
    if ((st = fft_fwd(c, f, w))) return st;
    if (cplx)
        map_rev_kernel<true><<<gr, 256, 0, c->st>>>(f, ax, w, div, x);
    else
        map_rev_kernel<false><<<gr, 256, 0, c->st>>>(f, ax, w, div, x);
    HIPCHK(hipGetLastError());
    return 0;
}

int dsort_first(const int num[3]) {   // MAXLOC: first axis of the largest size
    int a = 0;
    for (int k = 1; k < 3; ++k)
        if (num[k] > num[a]) a = k;
    return a;
}

Field whole_field(const int num[3], int nd) {
    Field f;
    f.m[0] = num[0];
    f.m[1] = nd >= 2 ? num[1] : 1;
    f.m[2] = nd >= 3 ? num[2] : 1;
    f.s[0] = 1;
    f.s[1] = num[0];
    f.s[2] = (long long)num[0] * num[1];
    f.bs = (long long)num[0] * num[1] * num[2];
    f.batch = 1;
    return f;
}

// four_conv (fourier_conv.f90:7-423)
int run_conv(dc_ctx *c, const dc_post_desc *d, double *csf, double *dsi, double *d3) {
    const float EPS = 1.0e-6f;
    if (d->res_sigma[0] < (double)EPS && d->res_sigma[1] < (double)EPS && d->res_sigma[2] < (double)EPS) return 0;
    const int *num = d->num;
    int isdim = (num[0] != 1) + (num[1] != 1) + (num[2] != 1);
    if (isdim == 0) return 0;
    if (isdim == 1 && num[0] == 1)
        return set_err(c, DC_ERR_ARG, DC_ER_LIB, "four_conv: a 1D pattern must run along the abscissa (four_conv_1D reads dsi(:,1,1))");
    if (isdim == 2 && num[2] != 1)
        return set_err(c, DC_ERR_ARG, DC_ER_LIB, "four_conv: a 2D pattern needs num(3) == 1 (four_conv_2D reads dsi(:,:,1))");
    for (int a = 0; a < isdim; ++a)
        if (!(d->res_sigma[a] > 0.0)) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "four_conv: zero sigma on a used axis");
    Field f = whole_field(num, isdim);
    MapAxes ax{};
    for (int a = 0; a < 3; ++a) {
        ax.lf[a] = (f.m[a] - 1) % 2;
        ax.lr[a] = isdim == 1 ? 0 : ax.lf[a];
    }
    if (isdim == 2 && dsort_first(num) == 1) {   // transposed FFT field (num(2) > num(1))
        ax.lr[0] = (num[1] - 1) % 2;
        ax.lr[1] = (num[0] - 1) % 2;
    }
    PostState &ps = post_state(c);
    int st = ensure(c, ps.spec, sizeof(double2) * (size_t)f.sub());
    if (st) return st;
    if ((st = ensure(c, ps.sums, sizeof(double) * 4))) return st;
    ProfArgs pa{};
    for (int a = 0; a < 3; ++a) {
        pa.m[a] = f.m[a];
        pa.sig[a] = d->res_sigma[a];
    }
    pa.nd = isdim;
    for (int q = 0; q < 9; ++q) pa.tr[q] = d->res_tr[q];
    double2 *spec = ptr<double2>(ps.spec);
    conv_profile_kernel<<<grid_for(f.sub()), 256, 0, c->st>>>(pa, spec);
    HIPCHK(hipGetLastError());
    if ((st = fft_fwd(c, f, spec))) return st;
    const double ntot = (double)f.sub();
    const double sscale = isdim == 1 ? 1.0 / ntot : 1.0 / std::sqrt(ntot);
    // in_pattern / N: 1D in double, 2D/3D through REAL() (single precision)
    const double div = isdim == 1 ? ntot : (double)(float)((long long)num[0] * num[1] * num[2]);
    double *sums = ptr<double>(ps.sums);
    if (d->lots_off && csf && (st = convolve(c, f, ax, true, csf, spec, sscale, div))) return st;
    if ((st = field_sums(c, f, dsi, sums))) return st;                     // weight
    if ((st = convolve(c, f, ax, false, dsi, spec, sscale, div))) return st;
    if ((st = field_sums(c, f, dsi, sums + 1))) return st;                 // weight2
    const Field whole = whole_field(num, 3);   // csf and dsi are rescaled over the whole array
    const int gr = grid_for(whole.total());
    if (csf) rescale_kernel<true, false><<<gr, 256, 0, c->st>>>(whole, sums, sums + 1, 1.0, csf);
    rescale_kernel<false, false><<<gr, 256, 0, c->st>>>(whole, sums, sums + 1, 1.0, dsi);
    HIPCHK(hipGetLastError());
    if (d->fave == 0.0 && d3) {
        if ((st = field_sums(c, f, d3, sums + 2))) return st;
        if ((st = convolve(c, f, ax, false, d3, spec, sscale, div))) return st;
        if ((st = field_sums(c, f, d3, sums + 3))) return st;
        rescale_kernel<false, false><<<gr, 256, 0, c->st>>>(whole, sums + 2, sums + 3, 1.0, d3);
        HIPCHK(hipGetLastError());
    }
    return 0;
}

// do_lanczos_*_fft over a batch of equal sub-fields (lib_lanczos.f90:258-458)
// Per-axis factor of the Lanczos table of do_lanczos_*_fft as circular taps
// at FFT-index offsets.  The 2D/3D table wsinc * lqgm is a product over the
// axes of sinc(pi qg rcut) sinc(pi qg / m) [|qg| < m] (wsinc_1d squares its
// sinc), each clipped to the natural range 1..N like the table.  A tap at
// FFT index r of the table convolves as the natural offset r (the maps are
// shifts); the 1D reverse map adds shift = lf - lr.  false: more taps than
// MAXTAP (the transform path is used).
bool lanczos_taps(const Field &f, int nd, double nscale, double rcut, int m, int axis, TapArgs &t) {
    const int N = f.m[axis];
    const int lf = (N - 1) % 2, lr = nd == 1 ? 0 : lf;
    const int cen = (N + 1) / 2;
    const int lo = (int)std::lround(-nscale * m - 1.0) + cen, hi = (int)std::lround(nscale * m + 1.0) + cen;
    const double pi = 3.14159265358979323846264338328;
    auto sinc = [](double v) { return std::fabs(v) < 1.0e-7 ? 1.0 : std::sin(v) / v; };
    t.f = f;
    t.axis = axis;
    t.shift = lf - lr;
    t.ntap = 0;
    for (int i = std::max(1, lo); i <= std::min(N, hi); ++i) {
        const double qg = (double)cen / nscale - (double)i / nscale;
        if (!(std::fabs(qg) < (double)m)) continue;
        double w = sinc(pi * qg * rcut);
        if (nd == 1) w = w * w;
        const double v = w * sinc(pi * qg / (double)m);
        if (v == 0.0) continue;
        if (t.ntap == MAXTAP) return false;
        t.r[t.ntap] = ((i + N / 2 - lf) % N + N) % N;
        t.v[t.ntap] = v;
        ++t.ntap;
    }
    return true;
}

int lanczos_batch(dc_ctx *c, const Field &f, int nd, double nscale, double rcut, int m, double *x) {
    PostState &ps = post_state(c);
    {
        // separable direct form: one short circular stencil per axis
        TapArgs taps[3];
        bool ok = true;
        for (int a = 0; a < nd; ++a) ok = ok && lanczos_taps(f, nd, nscale, rcut, m, a, taps[a]);
        if (ok) {
            int st = ensure(c, ps.sums, sizeof(double) * 2 * (size_t)f.batch);
            if (st) return st;
            const long long span = (long long)(f.batch - 1) * f.bs + (long long)(f.m[0] - 1) * f.s[0] +
                                   (long long)(f.m[1] - 1) * f.s[1] + (long long)(f.m[2] - 1) * f.s[2] + 1;
            if ((st = ensure(c, ps.copy, sizeof(double) * (size_t)span))) return st;
            double *sums = ptr<double>(ps.sums);
            if ((st = field_sums(c, f, x, sums))) return st;                    // sum(infield)
            double *bufs[2] = {x, ptr<double>(ps.copy)};
            int cur = 0;
            const int gr = grid_for(f.total());
            for (int a = 0; a < nd; ++a) {
                axis_conv_kernel<<<gr, 256, 0, c->st>>>(taps[a], bufs[cur], bufs[1 - cur]);
                cur = 1 - cur;
            }
            HIPCHK(hipGetLastError());
            if (cur == 1) HIPCHK(hipMemcpyAsync(x, ps.copy.p, sizeof(double) * (size_t)span, hipMemcpyDeviceToDevice, c->st));
            if ((st = field_sums(c, f, x, sums + f.batch))) return st;         // sum(outfield)
            rescale_kernel<false, true><<<gr, 256, 0, c->st>>>(f, sums, sums + f.batch, nscale, x);
            HIPCHK(hipGetLastError());
            return 0;
        }
    }
    int st = ensure(c, ps.spec, sizeof(double2) * (size_t)f.sub());
    if (st) return st;
    if ((st = ensure(c, ps.sums, sizeof(double) * 2 * (size_t)f.batch))) return st;
    MapAxes ax{};
    TabArgs t{};
    t.nd = nd;
    t.nscale = nscale;
    t.rcut = rcut;
    t.width = m;
    for (int a = 0; a < 3; ++a) {
        t.m[a] = f.m[a];
        ax.lf[a] = (f.m[a] - 1) % 2;
        ax.lr[a] = nd == 1 ? 0 : ax.lf[a];
        t.lf[a] = ax.lf[a];
        t.cen[a] = (f.m[a] + 1) / 2;
        t.lo[a] = (int)std::lround(-nscale * m - 1.0) + t.cen[a];
        t.hi[a] = (int)std::lround(nscale * m + 1.0) + t.cen[a];
    }
    double2 *spec = ptr<double2>(ps.spec);
    Field one = f;
    one.batch = 1;
    lanczos_table_kernel<<<grid_for(f.sub()), 256, 0, c->st>>>(t, spec);
    HIPCHK(hipGetLastError());
    if ((st = fft_fwd(c, one, spec))) return st;
    double *sums = ptr<double>(ps.sums);
    if ((st = field_sums(c, f, x, sums))) return st;                        // sum(infield)
    double div = 1.0;                                                       // lib_conv's weight
    for (int a = 0; a < nd; ++a) div *= std::sqrt(std::pow((double)f.m[a], 3.0));
    if ((st = convolve(c, f, ax, false, x, spec, 1.0, div))) return st;
    if ((st = field_sums(c, f, x, sums + f.batch))) return st;             // sum(outfield)
    rescale_kernel<false, true><<<grid_for(f.total()), 256, 0, c->st>>>(f, sums, sums + f.batch, nscale, x);
    HIPCHK(hipGetLastError());
    return 0;
}

// do_four_filter (fourier_sup.f90:1646-1805)
int run_filter(dc_ctx *c, const dc_post_desc *d, double *x) {
    if (d->filter != 1) return 0;
    const int *num = d->num;
    if (std::lround(d->rscale) != 1)
        return set_err(c, DC_ERR_ARG, DC_ER_LIB, "Lanczos filter: nint(four_rscale) must be 1 (the output grid is the input grid)");
    if (d->width < 1) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "Lanczos filter: four_width must be >= 1");
    const int isd = (num[0] > 1) + (num[1] > 1) + (num[2] > 1);
    const int lim = 2 * d->width + 1;
    const long long s1 = num[0], s2 = (long long)num[0] * num[1];
    Field f{};
    int nd = 0;
    if (isd == 3) {
        if (num[0] > lim && num[1] > lim && num[2] > lim) {
            f = whole_field(num, 3);
            nd = 3;
        } else if (num[2] <= lim) {             // layers k
            f = Field{{num[0], num[1], 1}, {1, s1, 0}, s2, num[2]};
            nd = 2;
        } else if (num[1] <= lim) {             // layers j: (i, k)
            f = Field{{num[0], num[2], 1}, {1, s2, 0}, s1, num[1]};
            nd = 2;
        } else {                                 // layers i: (j, k)
            f = Field{{num[1], num[2], 1}, {s1, s2, 0}, 1, num[0]};
            nd = 2;
        }
    } else if (isd == 2) {
        if (num[2] != 1)
            return set_err(c, DC_ERR_ARG, DC_ER_LIB, "Lanczos filter: a 2D pattern needs num(3) == 1 (do_four_filter reads dsi(:,:,1))");
        f = whole_field(num, 2);
        nd = 2;
    } else if (isd == 1) {
        if (num[0] == 1)
            return set_err(c, DC_ERR_ARG, DC_ER_LIB, "Lanczos filter: a 1D pattern must run along the abscissa (do_four_filter reads dsi(:,1,1))");
        f = whole_field(num, 1);
        nd = 1;
    } else {
        return 0;
    }
    return lanczos_batch(c, f, nd, d->rscale, d->damp, d->width, x);
}

// four_weight's scale factor (fourier_sup.f90:1464-1609) on the host.
int weight_scalef(dc_ctx *c, const dc_post_desc *d, double &scalef) {
    scalef = 1.0;
    if (d->fave != 0.0) return 0;
    const int *num = d->num;
    const double *vi = d->vi;
    const int *cells = d->nlots == 1 ? d->icc : d->ls;
    const int isdim = (num[0] != 1) + (num[1] != 1) + (num[2] != 1);
    double b[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};   // b[col][row]
    long long ncells = 1;
    auto unit_pair = [&](int col) {                          // fourier_sup.f90:1492-1540
        const int o0 = col == 0 ? 1 : 0, o1 = col == 2 ? 1 : 2;
        if (b[col][0] != 0.0) {
            b[o0][1] = 1.0;
            b[o1][2] = 1.0;
        } else if (b[col][1] != 0.0) {
            b[o0][2] = 1.0;
            b[o1][0] = 1.0;
        } else if (b[col][2] != 0.0) {
            b[o0][0] = 1.0;
            b[o1][1] = 1.0;
        }
    };
    if (isdim == 1) {
        int col, ax;
        if (num[0] != 1) {
            col = 0;
            ax = 0;
        } else if (num[1] == 1) {                            // the line runs along top, DISCUS takes vin
            col = 1;
            ax = 1;
        } else {
            col = 2;
            ax = 2;
        }
        for (int r = 0; r < 3; ++r) b[col][r] = vi[3 * col + r];
        unit_pair(col);
        ncells = cells[ax];
    } else if (isdim == 2) {
        if (num[2] == 1) {
            for (int r = 0; r < 3; ++r) {
                b[0][r] = vi[r];
                b[1][r] = vi[3 + r];
                b[2][r] = (b[0][r] != 0.0 || b[1][r] != 0.0) ? 0.0 : 1.0;
            }
            ncells = (long long)cells[0] * cells[1];
        } else if (num[1] == 1) {
            for (int r = 0; r < 3; ++r) {
                b[0][r] = vi[r];
                b[2][r] = vi[6 + r];
                b[1][r] = (b[0][r] != 0.0 || b[2][r] != 0.0) ? 0.0 : 1.0;
            }
            ncells = (long long)cells[0] * cells[2];
        } else {
            return set_err(c, DC_ERR_ARG, DC_ER_LIB, "four_weight: DISCUS leaves the bases undefined for a 2D grid with num(1) == 1");
        }
    } else if (isdim == 3) {
        for (int q = 0; q < 3; ++q)
            for (int r = 0; r < 3; ++r) b[q][r] = vi[3 * q + r];
        ncells = (long long)cells[0] * cells[1] * cells[2];
    } else {
        return set_err(c, DC_ERR_ARG, DC_ER_LIB, "four_weight: DISCUS leaves the bases undefined for a single point");
    }
    // do_volume(.false., u, v, w) (metric.f90:576-605, do_bang :35-82)
    const double *g = d->rten;
    auto skal = [&](const double *x, const double *y) {
        double s = 0.0;
        for (int r = 0; r < 3; ++r)
            for (int q = 0; q < 3; ++q) s += x[r] * g[r + 3 * q] * y[q];
        return s;
    };
    const float rad_sp = (float)(3.14159265358979323846 / 180.0);
    auto cosd_bang = [&](const double *x, const double *y) {
        const double xx = std::sqrt(skal(x, x)), yy = std::sqrt(skal(y, y));
        double deg = 0.0;
        if (xx > 0.0 && yy > 0.0) {
            const double arg = skal(x, y) / (xx * yy);
            if (std::fabs(arg) < 1.0)
                deg = std::acos(arg) / (double)rad_sp;
            else
                deg = arg >= 1.0 ? 0.0 : 180.0;
        }
        return std::cos(deg * (3.14159265358979323846 / 180.0));
    };
    const double L0 = std::sqrt(skal(b[0], b[0])), L1 = std::sqrt(skal(b[1], b[1])), L2 = std::sqrt(skal(b[2], b[2]));
    const double c1 = cosd_bang(b[1], b[2]), c2 = cosd_bang(b[0], b[2]), c3 = cosd_bang(b[0], b[1]);
    const double vox = L0 * L1 * L2 * std::sqrt(1.0 - c1 * c1 - c2 * c2 - c3 * c3 + 2.0 * c1 * c2 * c3);
    if (vox > 0.0) scalef = (double)ncells / d->vr * vox;
    return 0;
}

int check_grid(dc_ctx *c, const int num[3]) {
    for (int a = 0; a < 3; ++a)
        if (num[a] < 1) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "grid size must be >= 1 on every axis");
    if ((long long)num[0] * num[1] * num[2] > (1LL << 31))
        return set_err(c, DC_ERR_ARG, DC_ER_LIB, "grid larger than 2^31 points");
    return 0;
}

// integer inverse of an operation with determinant +-1, column-major
bool int_inverse(const double *m, int out[9]) {
    int a[9];
    for (int q = 0; q < 9; ++q) {
        if (m[q] != std::nearbyint(m[q]) || std::fabs(m[q]) > 8.0) return false;
        a[q] = (int)m[q];
    }
    auto A = [&](int r, int col) { return a[r + 3 * col]; };
    const int det = A(0, 0) * (A(1, 1) * A(2, 2) - A(1, 2) * A(2, 1)) - A(0, 1) * (A(1, 0) * A(2, 2) - A(1, 2) * A(2, 0)) +
                    A(0, 2) * (A(1, 0) * A(2, 1) - A(1, 1) * A(2, 0));
    if (det != 1 && det != -1) return false;
    int inv[3][3];
    inv[0][0] = A(1, 1) * A(2, 2) - A(1, 2) * A(2, 1);
    inv[0][1] = A(0, 2) * A(2, 1) - A(0, 1) * A(2, 2);
    inv[0][2] = A(0, 1) * A(1, 2) - A(0, 2) * A(1, 1);
    inv[1][0] = A(1, 2) * A(2, 0) - A(1, 0) * A(2, 2);
    inv[1][1] = A(0, 0) * A(2, 2) - A(0, 2) * A(2, 0);
    inv[1][2] = A(0, 2) * A(1, 0) - A(0, 0) * A(1, 2);
    inv[2][0] = A(1, 0) * A(2, 1) - A(1, 1) * A(2, 0);
    inv[2][1] = A(0, 1) * A(2, 0) - A(0, 0) * A(2, 1);
    inv[2][2] = A(0, 0) * A(1, 1) - A(0, 1) * A(1, 0);
    for (int r = 0; r < 3; ++r)
        for (int col = 0; col < 3; ++col) out[r + 3 * col] = inv[r][col] * det;
    return true;
}

// pixel of hkl = 0 (four_symm_csf, fourier_sup.f90:4088-4109)
void symm_izero(const dc_symm_desc *d, int iz[3]) {
    double m[3][3];   // m[row][col]
    for (int a = 0; a < 3; ++a)
        for (int r = 0; r < 3; ++r) m[r][a] = d->num[a] > 1 ? d->vi[3 * a + r] : (r == a ? 1.0 : 0.0);
    const double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                       m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    double inv[3][3];
    inv[0][0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) / det;
    inv[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) / det;
    inv[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) / det;
    inv[1][0] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) / det;
    inv[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) / det;
    inv[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) / det;
    inv[2][0] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) / det;
    inv[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) / det;
    inv[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) / det;
    for (int a = 0; a < 3; ++a) {
        const double u = -(inv[a][0] * d->eck[0] + inv[a][1] * d->eck[1] + inv[a][2] * d->eck[2]);
        const double v = 1.0 + u;
        iz[a] = d->num[a] > 1 ? (int)(v >= 0 ? std::floor(v + 0.5) : -std::floor(-v + 0.5)) - 1 : 0;
    }
}

template <bool CPLX>
int fill_field(dc_ctx *c, const dc_symm_desc *d, const SymmArgs *d_sa, const FillArgs &fa, bool adjacent, double *x) {
    if (!x) return 0;
    PostState &ps = post_state(c);
    const long long n = (long long)d->num[0] * d->num[1] * d->num[2];
    const size_t bytes = sizeof(double) * (CPLX ? 2 : 1) * (size_t)n;
    const int gr = grid_for(n);
    if (d->four_symm) {
        int st = ensure(c, ps.copy, bytes);
        if (st) return st;
        symm_kernel<CPLX><<<gr, 256, 0, c->st>>>(d_sa, x, ptr<double>(ps.copy));
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(x, ps.copy.p, bytes, hipMemcpyDeviceToDevice, c->st));
    }
    if (d->fave < 0.0) {
        if (adjacent) {
            fill_serial_kernel<CPLX><<<1, 64, 0, c->st>>>(fa, x);
        } else {
            int st = ensure(c, ps.copy, bytes);
            if (st) return st;
            HIPCHK(hipMemcpyAsync(ps.copy.p, x, bytes, hipMemcpyDeviceToDevice, c->st));
            fill_parallel_kernel<CPLX><<<gr, 256, 0, c->st>>>(fa, ptr<double>(ps.copy), x);
        }
        HIPCHK(hipGetLastError());
    }
    return 0;
}

int check_post(dc_ctx *c, const dc_post_desc *d) {
    if (!c) return DC_ERR_ARG;
    if (!d) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "null post descriptor");
    return check_grid(c, d->num);
}

}  // namespace

extern "C" {

int dc_four_post_dev(dc_ctx *c, const dc_post_desc *d, double *csf, double *dsi, double *d3) {
    int st = check_post(c, d);
    if (st) return st;
    if (!dsi) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "dsi is required (csf and dsi3d may be NULL)");
    HIPCHK(hipSetDevice(c->device));
    const int *num = d->num;
    const long long n = (long long)num[0] * num[1] * num[2];
    // four_weight (dsi3d is taken before diff_scale)
    if (d3) {
        double scalef = 1.0;
        if ((st = weight_scalef(c, d, scalef))) return st;
        WeightArgs wa{};
        for (int a = 0; a < 3; ++a) {
            wa.n[a] = num[a];
            wa.e[a] = d->eck[a];
        }
        for (int q = 0; q < 9; ++q) wa.v[q] = d->vi[q];
        wa.inv_ds = d->diff_scale != 0.0 && d->diff_scale != 1.0 ? 1.0 / d->diff_scale : 1.0;
        wa.scalef = scalef;
        wa.apply = d->fave == 0.0 && std::fabs(scalef - 1.0) > (double)1.0e-5f;
        weight_kernel<<<grid_for(n), 256, 0, c->st>>>(wa, dsi, d3);
        HIPCHK(hipGetLastError());
    }
    if ((st = run_filter(c, d, dsi))) return st;
    if ((st = run_conv(c, d, csf, dsi, d3))) return st;
    if (d->fave == 0.0 && d3 && (st = run_filter(c, d, d3))) return st;
    return DC_OK;
}

int dc_four_fill_dev(dc_ctx *c, const dc_symm_desc *d, double *csf, double *dsi) {
    if (!c) return DC_ERR_ARG;
    if (!d) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "null symmetry descriptor");
    int st = check_grid(c, d->num);
    if (st) return st;
    if (d->fave == 0.0 && !d->four_symm) return DC_OK;   // fourier_sup.f90:3862, 3964
    HIPCHK(hipSetDevice(c->device));
    PostState &ps = post_state(c);
    SymmArgs sa{};
    if (d->four_symm) {
        if (d->nops < 1 || d->nops - 1 > MAXOPS) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "nops out of range");
        if (d->nops > 1 && !d->mats) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "null symmetry matrices");
        for (int a = 0; a < 3; ++a) sa.n[a] = d->num[a];
        symm_izero(d, sa.iz);
        sa.nop = d->nops - 1;
        sa.nops = (double)d->nops;
        for (int o = 0; o < sa.nop; ++o)
            if (!int_inverse(d->mats + 9 * o, sa.m[o]))
                return set_err(c, DC_ERR_ARG, DC_ER_LIB, "symmetry operation " + std::to_string(o + 2) +
                                                             " is not an integer matrix with determinant +-1");
        HIPCHK(hipStreamSynchronize(c->st));   // the previous upload of h_sym has been consumed
        ps.h_sym = sa;
        if ((st = upload(c, ps.sym, reinterpret_cast<const unsigned char *>(&ps.h_sym), sizeof(SymmArgs)))) return st;
    }
    FillArgs fa{};
    bool adjacent = false;
    if (d->fave < 0.0) {
        const float f005 = 0.005f;
        for (int a = 0; a < 3; ++a) {
            fa.n[a] = d->num[a];
            fa.e[a] = d->eck[a];
        }
        for (int q = 0; q < 9; ++q) fa.v[q] = d->vi[q];
        for (int r = 0; r < 3; ++r) {
            fa.step[r] = std::max(std::max(d->vi[r], d->vi[r + 3]), d->vi[r + 6]) * (double)f005;
            if (d->num[r] == 1) fa.step[r] = (double)0.0001f;
        }
        // Two Bragg pixels can be neighbours along axis a only if vi(:,a) is
        // within 2*step of an integer vector in every component.
        for (int a = 0; a < 3; ++a) {
            if (d->num[a] < 2) continue;
            bool near = true;
            for (int r = 0; r < 3; ++r) {
                const double x = d->vi[3 * a + r];
                near = near && std::fabs(x - std::nearbyint(x)) < 2.0 * std::fabs(fa.step[r]) + 1e-12;
            }
            adjacent = adjacent || near;
        }
    }
    const SymmArgs *dsa = ptr<SymmArgs>(ps.sym);
    if ((st = fill_field<true>(c, d, dsa, fa, adjacent, csf))) return st;
    if ((st = fill_field<false>(c, d, dsa, fa, adjacent, dsi))) return st;
    return DC_OK;
}

int dc_four_accumulate_dev(dc_ctx *c, int mode, const dc_symm_desc *d, double *csf, double *dsi, double *csf_sum,
                           double *dsi_sum) {
    if (!c) return DC_ERR_ARG;
    if (!d) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "null symmetry descriptor");
    int st = check_grid(c, d->num);
    if (st) return st;
    HIPCHK(hipSetDevice(c->device));
    const long long n = (long long)d->num[0] * d->num[1] * d->num[2];
    switch (mode) {
    case DC_ACCUM_SINGLE:
        if (!d->four_symm) return DC_OK;
        return dc_four_fill_dev(c, d, csf, dsi);
    case DC_ACCUM_INIT:
        return DC_OK;
    case DC_ACCUM_ACCUM:
        if (!csf_sum || !dsi_sum || !csf || !dsi) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "accumulate needs csf, dsi and both sums");
        axpy_kernel<<<grid_for(2 * n), 256, 0, c->st>>>(2 * n, csf, csf_sum);
        axpy_kernel<<<grid_for(n), 256, 0, c->st>>>(n, dsi, dsi_sum);
        HIPCHK(hipGetLastError());
        return DC_OK;
    case DC_ACCUM_FINISHED:
        if (!csf_sum || !dsi_sum) return set_err(c, DC_ERR_NOSUM, DC_ER_APPL, "four_accumulate finished without sums");
        if ((st = dc_four_fill_dev(c, d, csf_sum, dsi_sum))) return st;
        if (csf) HIPCHK(hipMemcpyAsync(csf, csf_sum, sizeof(double) * 2 * n, hipMemcpyDeviceToDevice, c->st));
        if (dsi) HIPCHK(hipMemcpyAsync(dsi, dsi_sum, sizeof(double) * n, hipMemcpyDeviceToDevice, c->st));
        return DC_OK;
    default:
        return set_err(c, DC_ERR_ARG, DC_ER_LIB, "unknown accumulate mode");
    }
}

int dc_qval_dev(dc_ctx *c, const dc_qval_desc *q, const double *csf, const double *acsf, const double *dsi, double *out) {
    if (!c) return DC_ERR_ARG;
    if (!q || !out) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "null qval descriptor or output");
    int st = check_grid(c, q->num);
    if (st) return st;
    const int v = q->value;
    const bool known = (v >= 1 && v <= 5) || v == 7 || v == 9 || v == 10 || v == 11 || v == 13 || (v >= 17 && v <= 19);
    if (!known) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "qval value " + std::to_string(v) + " is not supported (random phase and PDF values are outside this library)");
    const bool needs_f = v == 2 || v == 3 || v == 4 || v == 5 || (q->laver && (v == 1 || v == 7 || v == 13));
    const bool needs_d = !q->laver && (v == 1 || v == 7 || v == 13);
    const bool needs_tab = v == 7 || v == 9 || v == 10 || v == 11 || v >= 17;
    if (needs_f && !(q->laver ? acsf : csf)) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "qval needs csf (acsf with laver)");
    if (needs_d && !dsi) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "qval needs dsi");
    if (needs_tab && (!q->cfact || !q->cfact_pure || !q->amount || q->nscat < 1 || q->ncfact < 2))
        return set_err(c, DC_ERR_ARG, DC_ER_LIB, "qval needs cfact, cfact_pure and amount");
    HIPCHK(hipSetDevice(c->device));
    PostState &ps = post_state(c);
    QvalArgs a{};
    for (int k = 0; k < 3; ++k) {
        a.g.n[k] = q->num[k];
        a.g.xm[k] = q->eck[k];
        a.g.uin[k] = q->vi[k];
        a.g.vin[k] = q->vi[3 + k];
        a.g.win[k] = q->vi[6 + k];
    }
    for (int k = 0; k < 9; ++k) a.g.rten[k] = q->rten[k];
    a.value = v;
    a.laver = q->laver;
    a.nscat = needs_tab ? q->nscat : 0;
    a.ncfact = q->ncfact;
    a.n_real = q->n_real_atoms;
    a.u2aver = q->u2aver;
    a.faver_sign = q->faver_sign;
    if (needs_tab) {
        const size_t nt = (size_t)q->ncfact * q->nscat * 2;
        if ((st = upload(c, ps.q_cf, q->cfact, nt))) return st;
        if ((st = upload(c, ps.q_cp, q->cfact_pure, nt))) return st;
        if ((st = upload(c, ps.q_amt, q->amount, (size_t)q->nscat))) return st;
        // istl of the grid corners must stay inside the table
        Grid g = a.g;
        if (max_corner_bin(g) > q->ncfact - 1)
            return set_err(c, DC_ERR_ISTL, DC_ER_FOUR, "sin(theta)/lambda beyond the form-factor table");
    }
    const long long n = (long long)q->num[0] * q->num[1] * q->num[2];
    qval_kernel<<<grid_for(n), 256, 0, c->st>>>(a, reinterpret_cast<const double2 *>(csf), reinterpret_cast<const double2 *>(acsf),
                                               dsi, ptr<double2>(ps.q_cf), ptr<double2>(ps.q_cp), ptr<double>(ps.q_amt), out);
    HIPCHK(hipGetLastError());
    return DC_OK;
}

// ---- host-pointer variants: stage through the context's buffers -------------
int dc_four_post(dc_ctx *c, const dc_post_desc *d, double *csf, double *dsi, double *d3) {
    int st = check_post(c, d);
    if (st) return st;
    if (!dsi) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "dsi is required");
    HIPCHK(hipSetDevice(c->device));
    PostState &ps = post_state(c);
    const size_t n = (size_t)d->num[0] * d->num[1] * d->num[2];
    double *dc = nullptr, *dd = nullptr, *d3d = nullptr;
    if (csf) {
        if ((st = upload(c, ps.stage[0], csf, 2 * n))) return st;
        dc = ptr<double>(ps.stage[0]);
    }
    if ((st = upload(c, ps.stage[1], dsi, n))) return st;
    dd = ptr<double>(ps.stage[1]);
    if (d3) {
        if ((st = ensure(c, ps.stage[2], sizeof(double) * n))) return st;
        d3d = ptr<double>(ps.stage[2]);
    }
    if ((st = dc_four_post_dev(c, d, dc, dd, d3d))) return st;
    if (csf) HIPCHK(hipMemcpyAsync(csf, dc, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemcpyAsync(dsi, dd, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
    if (d3) HIPCHK(hipMemcpyAsync(d3, d3d, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return DC_OK;
}

int dc_four_fill(dc_ctx *c, const dc_symm_desc *d, double *csf, double *dsi) {
    if (!c) return DC_ERR_ARG;
    if (!d) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "null symmetry descriptor");
    int st = check_grid(c, d->num);
    if (st) return st;
    HIPCHK(hipSetDevice(c->device));
    PostState &ps = post_state(c);
    const size_t n = (size_t)d->num[0] * d->num[1] * d->num[2];
    double *dc = nullptr, *dd = nullptr;
    if (csf) {
        if ((st = upload(c, ps.stage[0], csf, 2 * n))) return st;
        dc = ptr<double>(ps.stage[0]);
    }
    if (dsi) {
        if ((st = upload(c, ps.stage[1], dsi, n))) return st;
        dd = ptr<double>(ps.stage[1]);
    }
    if ((st = dc_four_fill_dev(c, d, dc, dd))) return st;
    if (csf) HIPCHK(hipMemcpyAsync(csf, dc, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, c->st));
    if (dsi) HIPCHK(hipMemcpyAsync(dsi, dd, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return DC_OK;
}

int dc_qval(dc_ctx *c, const dc_qval_desc *q, const double *csf, const double *acsf, const double *dsi, double *out) {
    if (!c) return DC_ERR_ARG;
    if (!q || !out) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "null qval descriptor or output");
    int st = check_grid(c, q->num);
    if (st) return st;
    HIPCHK(hipSetDevice(c->device));
    PostState &ps = post_state(c);
    const size_t n = (size_t)q->num[0] * q->num[1] * q->num[2];
    const double *dc = nullptr, *da = nullptr, *dd = nullptr;
    if (csf) {
        if ((st = upload(c, ps.stage[0], csf, 2 * n))) return st;
        dc = ptr<double>(ps.stage[0]);
    }
    if (acsf) {
        if ((st = upload(c, ps.copy, acsf, 2 * n))) return st;
        da = ptr<double>(ps.copy);
    }
    if (dsi) {
        if ((st = upload(c, ps.stage[1], dsi, n))) return st;
        dd = ptr<double>(ps.stage[1]);
    }
    if ((st = ensure(c, ps.stage[2], sizeof(double) * n))) return st;
    if ((st = dc_qval_dev(c, q, dc, da, dd, ptr<double>(ps.stage[2])))) return st;
    HIPCHK(hipMemcpyAsync(out, ps.stage[2].p, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return DC_OK;
}

}  // extern "C"
