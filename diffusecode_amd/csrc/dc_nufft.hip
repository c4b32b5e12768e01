// dc_nufft.hip -- type-1 non-uniform FFT on the GPU: DISCUS's second Fourier
// technique (four_run_nufft, fourier_sup.f90:327-690), which calls FINUFFT's
// finufft1d1/2d1/3d1 through four_strucf_1d/2d/3d (fourier_finufft.f90:19-245):
//
//   fk(k) = sum_j c_j exp(iflag * i * k . x_j),   k_d = -N_d/2 .. (N_d-1)/2
//
// Method: the exponential-of-semicircle kernel of Barnett, Magland and
// af Klinteberg (SIAM J. Sci. Comput. 41 (2019) C479-C504), written here:
//   phi(z) = exp(beta (sqrt(1 - z^2) - 1)), |z| <= 1, width w grid points,
//   w = ceil(log10(1/tol)) + 1 rounded up to even (2..16), beta = 2.30 w,
//   oversampling 2;
// 1. spread: every point adds c_j phi(2 (l - u_j)/w) to the w^d fine-grid
//    points l around u_j = x_j n / (2 pi) (periodic), FP64 vector atomics;
// 2. transform the fine grid (hipFFT Z2Z, backward for iflag > 0);
// 3. deconvolve: fk(k) = B(k mod n) / prod_d phihat_d(k_d), phihat from a
//    Gauss-Legendre rule (Newton on the three-term recurrence) on the host.
// Everything is FP64; the atomics make the last bits order-dependent.
#include <hipfft/hipfft.h>

#include <algorithm>
#include <array>
#include <cstring>
#include <cmath>
#include <map>
#include <tuple>
#include <vector>

#include "dc_host.h"

using namespace dc;

namespace {

constexpr int MAXW = 16;

struct NuParams {
    int dim;
    int w;
    double beta;
    long long n[3];      // fine grid
    long long N[3];      // modes
};

struct SpreadArgs {
    int dim, w;
    double beta, half_w, two_over_w;
    long long n[3];
    double scale[3];     // n / (2 pi)
};

inline int grid_for(long long n) {
    long long g = (n + 255) / 256;
    if (g > 16384) g = 16384;
    return (int)(g < 1 ? 1 : g);
}

__device__ inline long long wrapi(long long l, long long n) {
    long long r = l % n;
    return r < 0 ? r + n : r;
}

// One thread per point: kernel values per axis (W compile-time, in
// registers), then the W^DIM products added with FP64 vector atomics.
template <int DIM, int W>
__global__ __launch_bounds__(256) void spread_kernel(SpreadArgs a, long long M, const double *__restrict__ x,
                                                     const double *__restrict__ y, const double *__restrict__ z,
                                                     const double2 *__restrict__ c, double *__restrict__ grid) {
    for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < M; j += (long long)gridDim.x * blockDim.x) {
        // axes 1 and 2 keep their W kernel values in registers; the third
        // axis evaluates its value inside its (not unrolled) loop
        constexpr int KD = DIM < 3 ? DIM : 2;
        double ker[KD][W];
        long long l0[DIM];
        double uu[DIM], lov[DIM];
        const double *xs[3] = {x, y, z};
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            double u = xs[d][j] * a.scale[d];
            u = u - floor(u / (double)a.n[d]) * (double)a.n[d];     // into [0, n)
            const double lo = ceil(u - a.half_w);
            l0[d] = (long long)lo;
            uu[d] = u;
            lov[d] = lo;
            if (d < KD) {
#pragma unroll
                for (int i = 0; i < W; ++i) {
                    const double zz = (lo + (double)i - u) * a.two_over_w;
                    const double s = 1.0 - zz * zz;
                    ker[d < KD ? d : 0][i] = s > 0.0 ? exp(a.beta * (sqrt(s) - 1.0)) : 0.0;
                }
            }
        }
        const double2 cj = c[j];
        if constexpr (DIM == 1) {
#pragma unroll
            for (int i = 0; i < W; ++i) {
                const long long g = wrapi(l0[0] + i, a.n[0]);
                unsafeAtomicAdd(grid + 2 * g, cj.x * ker[0][i]);
                unsafeAtomicAdd(grid + 2 * g + 1, cj.y * ker[0][i]);
            }
        } else if constexpr (DIM == 2) {
            long long g1[W];
#pragma unroll
            for (int i = 0; i < W; ++i) g1[i] = wrapi(l0[0] + i, a.n[0]);
            for (int i2 = 0; i2 < W; ++i2) {
                const long long g2 = wrapi(l0[1] + i2, a.n[1]) * a.n[0];
                const double k2r = cj.x * ker[1][i2], k2i = cj.y * ker[1][i2];
#pragma unroll
                for (int i1 = 0; i1 < W; ++i1) {
                    const long long g = g2 + g1[i1];
                    unsafeAtomicAdd(grid + 2 * g, k2r * ker[0][i1]);
                    unsafeAtomicAdd(grid + 2 * g + 1, k2i * ker[0][i1]);
                }
            }
        } else {
            long long g1[W];
#pragma unroll
            for (int i = 0; i < W; ++i) g1[i] = wrapi(l0[0] + i, a.n[0]);
            for (int i3 = 0; i3 < W; ++i3) {
                const long long g3 = wrapi(l0[DIM - 1] + i3, a.n[2]) * a.n[1];
                const double z3 = (lov[DIM - 1] + (double)i3 - uu[DIM - 1]) * a.two_over_w;
                const double s3 = 1.0 - z3 * z3;
                const double k3 = s3 > 0.0 ? exp(a.beta * (sqrt(s3) - 1.0)) : 0.0;
#pragma unroll
                for (int i2 = 0; i2 < W; ++i2) {
                    const long long g2 = (g3 + wrapi(l0[1] + i2, a.n[1])) * a.n[0];
                    const double k23 = k3 * ker[1][i2];
                    const double vr = cj.x * k23, vi = cj.y * k23;
#pragma unroll
                    for (int i1 = 0; i1 < W; ++i1) {
                        const long long g = g2 + g1[i1];
                        unsafeAtomicAdd(grid + 2 * g, vr * ker[0][i1]);
                        unsafeAtomicAdd(grid + 2 * g + 1, vi * ker[0][i1]);
                    }
                }
            }
        }
    }
}

// 3D with coalesced atomics: a team of 2W lanes spreads one point; lane t
// takes abscissa offset t/2 and component t%2, so each atomic instruction of
// the team covers 2W consecutive doubles (one or two 128-byte lines) instead
// of one double per lane.  64/(2W) points per wave.
template <int W>
__global__ __launch_bounds__(256) void spread3_team_kernel(SpreadArgs a, long long M, const double *__restrict__ x,
                                                           const double *__restrict__ y, const double *__restrict__ z,
                                                           const double2 *__restrict__ c, double *__restrict__ grid) {
    constexpr int T = 2 * W;
    constexpr int PER = 64 / T;
    const int lane = threadIdx.x & 63;
    const int team = lane / T;
    const int t = lane - team * T;
    const int i1 = t >> 1, comp = t & 1;
    const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long long nwaves = ((long long)gridDim.x * blockDim.x) >> 6;
    for (long long base = wave * PER; base < M; base += nwaves * PER) {
        const long long j = base + team;
        if (team >= PER || j >= M) continue;
        double u[3], lo[3];
        const double *xs[3] = {x, y, z};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double v = xs[d][j] * a.scale[d];
            v = v - floor(v / (double)a.n[d]) * (double)a.n[d];
            u[d] = v;
            lo[d] = ceil(v - a.half_w);
        }
        double k2[W], k3[W];
#pragma unroll
        for (int i = 0; i < W; ++i) {
            const double z2 = (lo[1] + (double)i - u[1]) * a.two_over_w;
            const double z3 = (lo[2] + (double)i - u[2]) * a.two_over_w;
            const double s2 = 1.0 - z2 * z2, s3 = 1.0 - z3 * z3;
            k2[i] = s2 > 0.0 ? exp(a.beta * (sqrt(s2) - 1.0)) : 0.0;
            k3[i] = s3 > 0.0 ? exp(a.beta * (sqrt(s3) - 1.0)) : 0.0;
        }
        const double z1 = (lo[0] + (double)i1 - u[0]) * a.two_over_w;
        const double s1 = 1.0 - z1 * z1;
        const double k1 = s1 > 0.0 ? exp(a.beta * (sqrt(s1) - 1.0)) : 0.0;
        const double cv = (comp ? c[j].y : c[j].x) * k1;
        const long long g1 = wrapi((long long)lo[0] + i1, a.n[0]);
        const long long l1 = (long long)lo[1], l2 = (long long)lo[2];
        for (int i3 = 0; i3 < W; ++i3) {
            const long long g3 = wrapi(l2 + i3, a.n[2]) * a.n[1];
            const double v3 = cv * k3[i3];
#pragma unroll
            for (int i2 = 0; i2 < W; ++i2) {
                const long long g = (g3 + wrapi(l1 + i2, a.n[1])) * a.n[0] + g1;
                unsafeAtomicAdd(grid + 2 * g + comp, v3 * k2[i2]);
            }
        }
    }
}

// 3D, one real strength sv for all points (the occupancy of one atom type):
// a team of W lanes spreads one point into component `comp` of the complex
// grid (0 real, 1 imaginary), lane t taking abscissa offset t.  Two types
// share one grid, one in each component, so one transform serves both and
// each point costs W^3 atomics instead of the 2 W^3 of a complex strength.
template <int W>
__global__ __launch_bounds__(256) void spread3_real_kernel(SpreadArgs a, long long M, const double *__restrict__ x,
                                                           const double *__restrict__ y, const double *__restrict__ z,
                                                           double sv, int comp, double *__restrict__ grid) {
    constexpr int PER = 64 / W;
    const int lane = threadIdx.x & 63;
    const int team = lane / W;
    const int i1 = lane - team * W;
    const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long long nwaves = ((long long)gridDim.x * blockDim.x) >> 6;
    for (long long base = wave * PER; base < M; base += nwaves * PER) {
        const long long j = base + team;
        if (team >= PER || j >= M) continue;
        double u[3], lo[3];
        const double *xs[3] = {x, y, z};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double v = xs[d][j] * a.scale[d];
            v = v - floor(v / (double)a.n[d]) * (double)a.n[d];
            u[d] = v;
            lo[d] = ceil(v - a.half_w);
        }
        double k2[W], k3[W];
#pragma unroll
        for (int i = 0; i < W; ++i) {
            const double z2 = (lo[1] + (double)i - u[1]) * a.two_over_w;
            const double z3 = (lo[2] + (double)i - u[2]) * a.two_over_w;
            const double s2 = 1.0 - z2 * z2, s3 = 1.0 - z3 * z3;
            k2[i] = s2 > 0.0 ? exp(a.beta * (sqrt(s2) - 1.0)) : 0.0;
            k3[i] = s3 > 0.0 ? exp(a.beta * (sqrt(s3) - 1.0)) : 0.0;
        }
        const double z1 = (lo[0] + (double)i1 - u[0]) * a.two_over_w;
        const double s1 = 1.0 - z1 * z1;
        const double cv = sv * (s1 > 0.0 ? exp(a.beta * (sqrt(s1) - 1.0)) : 0.0);
        const long long g1 = wrapi((long long)lo[0] + i1, a.n[0]);
        const long long l1 = (long long)lo[1], l2 = (long long)lo[2];
        for (int i3 = 0; i3 < W; ++i3) {
            const long long g3 = wrapi(l2 + i3, a.n[2]) * a.n[1];
            const double v3 = cv * k3[i3];
#pragma unroll
            for (int i2 = 0; i2 < W; ++i2) {
                const long long g = (g3 + wrapi(l1 + i2, a.n[1])) * a.n[0] + g1;
                unsafeAtomicAdd(grid + 2 * g + comp, v3 * k2[i2]);
            }
        }
    }
}

void launch_spread_real(int w, hipStream_t st, const SpreadArgs &a, long long M, const double *x, const double *y,
                        const double *z, double sv, int comp, double *g) {
    const int tg = grid_for(M * w);   // w threads per point
    switch (w) {
    case 2: spread3_real_kernel<2><<<tg, 256, 0, st>>>(a, M, x, y, z, sv, comp, g); return;
    case 4: spread3_real_kernel<4><<<tg, 256, 0, st>>>(a, M, x, y, z, sv, comp, g); return;
    case 6: spread3_real_kernel<6><<<tg, 256, 0, st>>>(a, M, x, y, z, sv, comp, g); return;
    case 8: spread3_real_kernel<8><<<tg, 256, 0, st>>>(a, M, x, y, z, sv, comp, g); return;
    case 10: spread3_real_kernel<10><<<tg, 256, 0, st>>>(a, M, x, y, z, sv, comp, g); return;
    case 12: spread3_real_kernel<12><<<tg, 256, 0, st>>>(a, M, x, y, z, sv, comp, g); return;
    case 14: spread3_real_kernel<14><<<tg, 256, 0, st>>>(a, M, x, y, z, sv, comp, g); return;
    default: spread3_real_kernel<16><<<tg, 256, 0, st>>>(a, M, x, y, z, sv, comp, g); return;
    }
}

template <int DIM>
void launch_spread(int w, int grid, hipStream_t st, const SpreadArgs &a, long long M, const double *x, const double *y,
                   const double *z, const double2 *c, double *g) {
    if constexpr (DIM == 3) {
        const int tg = grid_for(M * 2 * w);   // ~2w threads per point
        switch (w) {
        case 2: spread3_team_kernel<2><<<tg, 256, 0, st>>>(a, M, x, y, z, c, g); return;
        case 4: spread3_team_kernel<4><<<tg, 256, 0, st>>>(a, M, x, y, z, c, g); return;
        case 6: spread3_team_kernel<6><<<tg, 256, 0, st>>>(a, M, x, y, z, c, g); return;
        case 8: spread3_team_kernel<8><<<tg, 256, 0, st>>>(a, M, x, y, z, c, g); return;
        case 10: spread3_team_kernel<10><<<tg, 256, 0, st>>>(a, M, x, y, z, c, g); return;
        case 12: spread3_team_kernel<12><<<tg, 256, 0, st>>>(a, M, x, y, z, c, g); return;
        case 14: spread3_team_kernel<14><<<tg, 256, 0, st>>>(a, M, x, y, z, c, g); return;
        default: spread3_team_kernel<16><<<tg, 256, 0, st>>>(a, M, x, y, z, c, g); return;
        }
    }
    switch (w) {
    case 2: spread_kernel<DIM, 2><<<grid, 256, 0, st>>>(a, M, x, y, z, c, g); break;
    case 4: spread_kernel<DIM, 4><<<grid, 256, 0, st>>>(a, M, x, y, z, c, g); break;
    case 6: spread_kernel<DIM, 6><<<grid, 256, 0, st>>>(a, M, x, y, z, c, g); break;
    case 8: spread_kernel<DIM, 8><<<grid, 256, 0, st>>>(a, M, x, y, z, c, g); break;
    case 10: spread_kernel<DIM, 10><<<grid, 256, 0, st>>>(a, M, x, y, z, c, g); break;
    case 12: spread_kernel<DIM, 12><<<grid, 256, 0, st>>>(a, M, x, y, z, c, g); break;
    case 14: spread_kernel<DIM, 14><<<grid, 256, 0, st>>>(a, M, x, y, z, c, g); break;
    default: spread_kernel<DIM, 16><<<grid, 256, 0, st>>>(a, M, x, y, z, c, g); break;
    }
}

// fk(m) = B(k mod n) * corr1(m1) corr2(m2) corr3(m3), CMCL order (k1 fastest).
struct DeconvArgs {
    long long N[3], n[3];
    long long sub[3];        // output stride per axis (every sub-th mode kept)
    long long out[3];        // output counts
};

__global__ void deconv_kernel(DeconvArgs a, const double2 *__restrict__ B, const double *__restrict__ c1,
                              const double *__restrict__ c2, const double *__restrict__ c3, double2 *__restrict__ fk) {
    const long long total = a.out[0] * a.out[1] * a.out[2];
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long long)gridDim.x * blockDim.x) {
        const long long o1 = e % a.out[0];
        const long long o2 = (e / a.out[0]) % a.out[1];
        const long long o3 = e / (a.out[0] * a.out[1]);
        const long long m1 = o1 * a.sub[0], m2 = o2 * a.sub[1], m3 = o3 * a.sub[2];
        const long long k1 = m1 - a.N[0] / 2, k2 = m2 - a.N[1] / 2, k3 = m3 - a.N[2] / 2;
        const long long g = wrapi(k1, a.n[0]) + a.n[0] * (wrapi(k2, a.n[1]) + a.n[1] * wrapi(k3, a.n[2]));
        const double f = c1[m1] * c2[m2] * c3[m3];
        const double2 b = B[g];
        fk[e] = make_double2(b.x * f, b.y * f);
    }
}

// Pruned 3D transform.  After the full FFT along x, only the x columns the
// output needs (list L0, c0 of them) go on: tr_xy gathers them and
// transposes so that y is contiguous, (y, i, z); after the y pass tr_yz
// gathers the needed y rows (L1, c1) with z contiguous, (z, j, i), for the z
// pass.  Tiles of 32 x 32 through LDS keep both sides coalesced.
constexpr int TR = 32, TRY = 8;

// out[y + n1 (i + c0 z)] = g[L0[i] + n0 (y + n1 z)]
__global__ __launch_bounds__(TR *TRY) void tr_xy_kernel(const double2 *__restrict__ g, const int *__restrict__ L0,
                                                         long long n0, long long n1, long long n2, int c0,
                                                         double2 *__restrict__ out) {
    __shared__ double2 tile[TR][TR + 1];
    const int i0 = blockIdx.x * TR, y0 = blockIdx.y * TR;
    const long long z = blockIdx.z;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int i = i0 + tx;
    const int xi = i < c0 ? L0[i] : 0;
#pragma unroll
    for (int r = 0; r < TR; r += TRY) {
        const long long y = y0 + ty + r;
        if (i < c0 && y < n1) tile[ty + r][tx] = g[xi + n0 * (y + n1 * z)];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < TR; r += TRY) {
        const int ii = i0 + ty + r;
        const long long y = y0 + tx;
        if (ii < c0 && y < n1) out[y + n1 * (ii + (long long)c0 * z)] = tile[tx][ty + r];
    }
}

// out[z + n2 (j + c1 i)] = in[L1[j] + n1 (i + c0 z)]
__global__ __launch_bounds__(TR *TRY) void tr_yz_kernel(const double2 *__restrict__ in, const int *__restrict__ L1,
                                                         long long n1, long long n2, int c0, int c1,
                                                         double2 *__restrict__ out) {
    __shared__ double2 tile[TR][TR + 1];
    const int j0 = blockIdx.x * TR;
    const long long z0 = (long long)blockIdx.y * TR;
    const long long i = blockIdx.z;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int j = j0 + tx;
    const int yj = j < c1 ? L1[j] : 0;
#pragma unroll
    for (int r = 0; r < TR; r += TRY) {
        const long long z = z0 + ty + r;
        if (j < c1 && z < n2) tile[ty + r][tx] = in[yj + n1 * (i + (long long)c0 * z)];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < TR; r += TRY) {
        const int jj = j0 + ty + r;
        const long long z = z0 + tx;
        if (jj < c1 && z < n2) out[z + n2 * (jj + (long long)c1 * i)] = tile[tx][ty + r];
    }
}

// Output modes of a 3D transform: the value of mode (o1, o2, o3) sits at
// Z[P0[o1] s0 + P1[o2] s1 + P2[o3] s2] and its mirror -k at the M tables
// (pruned layout (z, j, i): strides n2 c1, n2, 1 on compact positions; full
// grid: 1, n0, n0 n1 on fine indices).
struct PruneArgs {
    long long out[3];
    long long s[3];
    const int *P[3], *M[3];
    const double *corr[3];
    long long sub[3];
};

__global__ void deconv3_kernel(PruneArgs a, const double2 *__restrict__ Z, double2 *__restrict__ fa,
                               double2 *__restrict__ fb) {
    const long long total = a.out[0] * a.out[1] * a.out[2];
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long long)gridDim.x * blockDim.x) {
        const long long o1 = e % a.out[0];
        const long long o2 = (e / a.out[0]) % a.out[1];
        const long long o3 = e / (a.out[0] * a.out[1]);
        const double f = a.corr[0][o1 * a.sub[0]] * a.corr[1][o2 * a.sub[1]] * a.corr[2][o3 * a.sub[2]];
        const double2 zp = Z[a.P[0][o1] * a.s[0] + a.P[1][o2] * a.s[1] + a.P[2][o3] * a.s[2]];
        if (!fb) {
            fa[e] = make_double2(zp.x * f, zp.y * f);
        } else {
            // two real grids in one complex transform: Hermitian separation
            const double2 zm = Z[a.M[0][o1] * a.s[0] + a.M[1][o2] * a.s[1] + a.M[2][o3] * a.s[2]];
            const double h = 0.5 * f;
            fa[e] = make_double2((zp.x + zm.x) * h, (zp.y - zm.y) * h);
            fb[e] = make_double2((zp.y + zm.y) * h, (zm.x - zp.x) * h);
        }
    }
}

__global__ void zero_kernel(long long n, double *__restrict__ p) {
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) p[e] = 0.0;
}

__global__ void occ_weights_kernel(long long M, double occ, double2 *__restrict__ c) {
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < M; e += (long long)gridDim.x * blockDim.x)
        c[e] = make_double2(occ, 0.0);
}


// smallest even 2^a 3^b 5^c >= n
long long next_smooth(long long n) {
    if (n < 2) n = 2;
    for (long long m = n + (n & 1);; m += 2) {
        long long r = m;
        for (long long p : {2LL, 3LL, 5LL})
            while (r % p == 0) r /= p;
        if (r == 1) return m;
    }
}

// Gauss-Legendre nodes/weights on [-1, 1]: Newton on P_q from the recurrence.
void gauss_legendre(int q, std::vector<double> &xs, std::vector<double> &ws) {
    xs.assign(q, 0.0);
    ws.assign(q, 0.0);
    const double pi = 3.14159265358979323846;
    for (int i = 0; i < (q + 1) / 2; ++i) {
        double x = std::cos(pi * (i + 0.75) / (q + 0.5));
        double dp = 0.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = x;
            for (int k = 2; k <= q; ++k) {
                const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = q * (x * p1 - p0) / (x * x - 1.0);
            const double dx = p1 / dp;
            x -= dx;
            if (std::fabs(dx) < 1e-16) break;
        }
        double p0 = 1.0, p1 = x;
        for (int k = 2; k <= q; ++k) {
            const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        dp = q * (x * p1 - p0) / (x * x - 1.0);
        const double wgt = 2.0 / ((1.0 - x * x) * dp * dp);
        xs[i] = -x;
        xs[q - 1 - i] = x;
        ws[i] = wgt;
        ws[q - 1 - i] = wgt;
    }
}

// 1 / phihat(k) for k = -N/2 .. (N-1)/2, phihat(k) = int_{-w/2}^{w/2} phi(2t/w) cos(2 pi k t / n) dt
std::vector<double> correction(long long N, long long n, int w, double beta) {
    const int q = 4 * w + 24;
    std::vector<double> gx, gw;
    gauss_legendre(q, gx, gw);
    const double h = 0.5 * w;
    const double pi = 3.14159265358979323846;
    std::vector<double> out(N);
    for (long long m = 0; m < N; ++m) {
        const double k = (double)(m - N / 2);
        double s = 0.0;
        for (int i = 0; i < q; ++i) {
            const double t = h * gx[i];                       // t in [-w/2, w/2]
            const double zz = gx[i];
            const double ph = std::exp(beta * (std::sqrt(std::max(0.0, 1.0 - zz * zz)) - 1.0));
            s += gw[i] * ph * std::cos(2.0 * pi * k * t / (double)n);
        }
        out[m] = 1.0 / (s * h);
    }
    return out;
}

}  // namespace

// Index tables of the pruned 3D transform for one (grid, modes, subsampling)
// shape: kept fine indices along x and y (the output modes k and their
// mirrors -k) and, per output mode, the positions of k and -k.
struct PruneTables {
    int c0 = 0, c1 = 0;
    DevBuf L0, L1, P[3], M[3];
    std::vector<std::vector<int>> host;   // sources of the uploads, kept alive
};

struct NufftState {
    std::map<std::tuple<long long, long long, long long>, hipfftHandle> plans;
    std::map<std::pair<long long, long long>, hipfftHandle> plans1d;   // (length, batch): contiguous lines
    std::map<std::array<long long, 13>, PruneTables> prune;
    DevBuf work;                                                     // the y pass of the pruned transform
    std::map<std::tuple<long long, long long, int>, DevBuf> corr;   // (N, n, w) -> 1 / phihat on the device
    std::vector<std::vector<double>> corr_host;                      // sources of the uploads, kept alive
    DevBuf grid, pts[3], c, fk, one;
    DevBuf pts2[3], fk2;                                             // the second type of a pair (run level)
};

namespace dc {
void nufft_release(dc_ctx *c) {
    if (!c->nufft) return;
    for (auto &kv : c->nufft->plans) (void)hipfftDestroy(kv.second);
    for (auto &kv : c->nufft->plans1d) (void)hipfftDestroy(kv.second);
    NufftState *s = c->nufft;
    for (auto &kv : s->prune) {
        PruneTables &t = kv.second;
        for (DevBuf *b : {&t.L0, &t.L1, &t.P[0], &t.P[1], &t.P[2], &t.M[0], &t.M[1], &t.M[2]})
            if (b->p) (void)hipFree(b->p);
    }
    if (s->work.p) (void)hipFree(s->work.p);
    for (DevBuf *b : {&s->grid, &s->pts[0], &s->pts[1], &s->pts[2], &s->c, &s->fk, &s->one, &s->pts2[0], &s->pts2[1],
                      &s->pts2[2], &s->fk2})
        if (b->p) (void)hipFree(b->p);
    for (auto &kv : s->corr)
        if (kv.second.p) (void)hipFree(kv.second.p);
    delete s;
    c->nufft = nullptr;
}
}  // namespace dc

namespace {

NufftState &nu_state(dc_ctx *c) {
    if (!c->nufft) c->nufft = new NufftState();
    return *c->nufft;
}

int nu_params(dc_ctx *c, int dim, double tol, const long long N[3], NuParams &p) {
    if (dim < 1 || dim > 3) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "NUFFT dimension must be 1, 2 or 3");
    if (!(tol > 0.0)) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "NUFFT tolerance must be > 0");
    p.dim = dim;
    p.w = (int)std::ceil(std::log10(1.0 / tol)) + 1;
    p.w = std::min(MAXW, std::max(2, p.w));
    p.w += p.w & 1;                       // even widths only (compiled kernels); accuracy only improves
    p.beta = 2.30 * p.w;
    for (int d = 0; d < 3; ++d) {
        p.N[d] = d < dim ? N[d] : 1;
        if (p.N[d] < 1) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "NUFFT mode counts must be >= 1");
        p.n[d] = d < dim ? next_smooth(std::max(2 * p.N[d], (long long)(2 * p.w))) : 1;
    }
    if ((double)p.n[0] * p.n[1] * p.n[2] > 2.0e9) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "NUFFT fine grid too large");
    return 0;
}

// Correction tables 1/phihat per axis (cached per (N, n, w)) and the FFT plan
// of the fine grid (cached per grid), on the context's stream.
int nu_setup(dc_ctx *c, const NuParams &p, const double *corr[3], hipfftHandle &h, bool need_plan) {
    NufftState &s = nu_state(c);
    int st;
    for (int d = 0; d < 3; ++d) {
        auto key = std::make_tuple(d < p.dim ? p.N[d] : 1LL, d < p.dim ? p.n[d] : 1LL, d < p.dim ? p.w : 0);
        auto it = s.corr.find(key);
        if (it == s.corr.end()) {
            s.corr_host.push_back(d < p.dim ? correction(p.N[d], p.n[d], p.w, p.beta) : std::vector<double>(1, 1.0));
            DevBuf b;
            if ((st = upload(c, b, s.corr_host.back().data(), s.corr_host.back().size()))) return st;
            it = s.corr.emplace(key, b).first;
        }
        corr[d] = reinterpret_cast<const double *>(it->second.p);
    }
    if (!need_plan) return 0;
    // fine-grid transform (row-major dims: slowest first)
    auto key = std::make_tuple(p.n[0], p.n[1], p.n[2]);
    auto it = s.plans.find(key);
    if (it == s.plans.end()) {
        int dims[3];
        for (int d = 0; d < p.dim; ++d) dims[d] = (int)p.n[p.dim - 1 - d];
        hipfftResult r = hipfftPlanMany(&h, p.dim, dims, nullptr, 1, 0, nullptr, 1, 0, HIPFFT_Z2Z, 1);
        if (r != HIPFFT_SUCCESS) return set_err(c, DC_ERR_DEVICE, DC_ER_LIB, "hipfftPlanMany failed (" + std::to_string((int)r) + ")");
        s.plans[key] = h;
    } else {
        h = it->second;
    }
    if (hipfftSetStream(h, c->st) != HIPFFT_SUCCESS) return set_err(c, DC_ERR_DEVICE, DC_ER_LIB, "hipfftSetStream failed");
    return 0;
}

SpreadArgs spread_args(const NuParams &p) {
    SpreadArgs a{};
    a.dim = p.dim;
    a.w = p.w;
    a.beta = p.beta;
    a.half_w = 0.5 * p.w;
    a.two_over_w = 2.0 / p.w;
    for (int d = 0; d < 3; ++d) {
        a.n[d] = p.n[d];
        a.scale[d] = (double)p.n[d] / (2.0 * 3.14159265358979323846);
    }
    return a;
}

DeconvArgs deconv_args(const NuParams &p, const long long sub[3], const long long out[3]) {
    DeconvArgs da{};
    for (int d = 0; d < 3; ++d) {
        da.N[d] = p.N[d];
        da.n[d] = p.n[d];
        da.sub[d] = sub[d];
        da.out[d] = out[d];
    }
    return da;
}

int nu_exec(dc_ctx *c, hipfftHandle h, double *g, int iflag) {
    hipfftResult r = hipfftExecZ2Z(h, reinterpret_cast<hipfftDoubleComplex *>(g), reinterpret_cast<hipfftDoubleComplex *>(g),
                                   iflag > 0 ? HIPFFT_BACKWARD : HIPFFT_FORWARD);
    if (r != HIPFFT_SUCCESS) return set_err(c, DC_ERR_DEVICE, DC_ER_LIB, "hipfftExecZ2Z failed (" + std::to_string((int)r) + ")");
    return 0;
}

int plan1d(dc_ctx *c, long long n, long long batch, hipfftHandle &h) {
    NufftState &s = nu_state(c);
    auto key = std::make_pair(n, batch);
    auto it = s.plans1d.find(key);
    if (it == s.plans1d.end()) {
        int nn = (int)n;
        hipfftResult r = hipfftPlanMany(&h, 1, &nn, nullptr, 1, nn, nullptr, 1, nn, HIPFFT_Z2Z, (int)batch);
        if (r != HIPFFT_SUCCESS) return set_err(c, DC_ERR_DEVICE, DC_ER_LIB, "hipfftPlanMany failed (" + std::to_string((int)r) + ")");
        s.plans1d[key] = h;
    } else {
        h = it->second;
    }
    if (hipfftSetStream(h, c->st) != HIPFFT_SUCCESS) return set_err(c, DC_ERR_DEVICE, DC_ER_LIB, "hipfftSetStream failed");
    return 0;
}

int prune_tables(dc_ctx *c, const NuParams &p, const long long sub[3], const long long out[3], bool compact,
                 PruneTables *&tp) {
    NufftState &s = nu_state(c);
    const std::array<long long, 13> key = {p.n[0], p.n[1], p.n[2], p.N[0], p.N[1], p.N[2],
                                           sub[0], sub[1], sub[2], out[0], out[1], out[2], compact ? 1 : 0};
    auto it = s.prune.find(key);
    if (it != s.prune.end()) {
        tp = &it->second;
        return 0;
    }
    PruneTables t;
    std::vector<int> L[2];
    int st;
    for (int d = 0; d < 3; ++d) {
        std::vector<int> kp(out[d]), km(out[d]);
        for (long long o = 0; o < out[d]; ++o) {
            const long long k = o * sub[d] - p.N[d] / 2;
            kp[o] = (int)(((k % p.n[d]) + p.n[d]) % p.n[d]);
            km[o] = (int)(((-k % p.n[d]) + p.n[d]) % p.n[d]);
        }
        if (compact && d < 2) {
            std::vector<int> all(kp);
            all.insert(all.end(), km.begin(), km.end());
            std::sort(all.begin(), all.end());
            all.erase(std::unique(all.begin(), all.end()), all.end());
            for (long long o = 0; o < out[d]; ++o) {
                kp[o] = (int)(std::lower_bound(all.begin(), all.end(), kp[o]) - all.begin());
                km[o] = (int)(std::lower_bound(all.begin(), all.end(), km[o]) - all.begin());
            }
            L[d] = all;
        }
        t.host.push_back(kp);
        t.host.push_back(km);
        if ((st = upload(c, t.P[d], t.host[t.host.size() - 2].data(), out[d]))) return st;
        if ((st = upload(c, t.M[d], t.host.back().data(), out[d]))) return st;
    }
    t.c0 = (int)L[0].size();
    t.c1 = (int)L[1].size();
    t.host.push_back(L[0]);
    if ((st = upload(c, t.L0, t.host.back().data(), L[0].size()))) return st;
    t.host.push_back(L[1]);
    if ((st = upload(c, t.L1, t.host.back().data(), L[1].size()))) return st;
    tp = &s.prune.emplace(key, std::move(t)).first->second;
    return 0;
}

// 3D fine grid g -> output modes, pruned: full FFT along x, the kept x
// columns along y, the kept (x, y) lines along z, then the correction (and,
// with fb, the separation of two real grids).  g is overwritten.
int nu_transform3(dc_ctx *c, const NuParams &p, const double *const corr[3], double *g, int iflag, const long long sub[3],
                  const long long out[3], double2 *fa, double2 *fb) {
    NufftState &s = nu_state(c);
    const long long n0 = p.n[0], n1 = p.n[1], n2 = p.n[2];
    // the transposes put n2 and the kept x count (<= n0) on grid z
    const bool prunable = n2 <= 65535 && n0 <= 65535;
    PruneTables *t = nullptr;
    int st = prune_tables(c, p, sub, out, prunable, t);
    if (st) return st;
    PruneArgs a{};
    for (int d = 0; d < 3; ++d) {
        a.out[d] = out[d];
        a.sub[d] = sub[d];
        a.P[d] = ptr<int>(t->P[d]);
        a.M[d] = ptr<int>(t->M[d]);
        a.corr[d] = corr[d];
    }
    if (!prunable) {
        const double *cr[3];
        hipfftHandle h;
        if ((st = nu_setup(c, p, cr, h, true)) || (st = nu_exec(c, h, g, iflag))) return st;
        a.s[0] = 1;
        a.s[1] = n0;
        a.s[2] = n0 * n1;
        deconv3_kernel<<<grid_for(out[0] * out[1] * out[2]), 256, 0, c->st>>>(a, reinterpret_cast<const double2 *>(g), fa, fb);
        HIPCHK(hipGetLastError());
        return 0;
    }
    const int c0 = t->c0, c1 = t->c1;
    if ((st = ensure(c, s.work, sizeof(double2) * (size_t)(n1 * c0 * n2)))) return st;
    hipfftHandle hx, hy, hz;
    if ((st = plan1d(c, n0, n1 * n2, hx)) || (st = plan1d(c, n1, (long long)c0 * n2, hy)) ||
        (st = plan1d(c, n2, (long long)c1 * c0, hz)))
        return st;
    double2 *w = ptr<double2>(s.work);
    if ((st = nu_exec(c, hx, g, iflag))) return st;
    tr_xy_kernel<<<dim3((c0 + TR - 1) / TR, (unsigned)((n1 + TR - 1) / TR), (unsigned)n2), dim3(TR, TRY), 0, c->st>>>(
        reinterpret_cast<const double2 *>(g), ptr<int>(t->L0), n0, n1, n2, c0, w);
    HIPCHK(hipGetLastError());
    if ((st = nu_exec(c, hy, reinterpret_cast<double *>(w), iflag))) return st;
    tr_yz_kernel<<<dim3((c1 + TR - 1) / TR, (unsigned)((n2 + TR - 1) / TR), (unsigned)c0), dim3(TR, TRY), 0, c->st>>>(
        w, ptr<int>(t->L1), n1, n2, c0, c1, reinterpret_cast<double2 *>(g));
    HIPCHK(hipGetLastError());
    if ((st = nu_exec(c, hz, g, iflag))) return st;
    a.s[0] = n2 * c1;
    a.s[1] = n2;
    a.s[2] = 1;
    deconv3_kernel<<<grid_for(out[0] * out[1] * out[2]), 256, 0, c->st>>>(a, reinterpret_cast<const double2 *>(g), fa, fb);
    HIPCHK(hipGetLastError());
    return 0;
}

// Spread, transform, deconvolve; points/weights already on the device.
int nufft_run(dc_ctx *c, const NuParams &p, long long M, const double *dx, const double *dy, const double *dz,
              const double2 *dc, int iflag, const long long sub[3], const long long out[3], double2 *dfk) {
    NufftState &s = nu_state(c);
    const long long ng = p.n[0] * p.n[1] * p.n[2];
    int st = ensure(c, s.grid, sizeof(double2) * (size_t)ng);
    if (st) return st;
    const double *corr[3];
    hipfftHandle h;
    if ((st = nu_setup(c, p, corr, h, p.dim < 3))) return st;
    double *g = ptr<double>(s.grid);
    zero_kernel<<<grid_for(2 * ng), 256, 0, c->st>>>(2 * ng, g);
    const SpreadArgs a = spread_args(p);
    if (M > 0) {
        if (p.dim == 1)
            launch_spread<1>(p.w, grid_for(M), c->st, a, M, dx, dy, dz, dc, g);
        else if (p.dim == 2)
            launch_spread<2>(p.w, grid_for(M), c->st, a, M, dx, dy, dz, dc, g);
        else
            launch_spread<3>(p.w, grid_for(M), c->st, a, M, dx, dy, dz, dc, g);
    }
    HIPCHK(hipGetLastError());
    if (p.dim == 3) return nu_transform3(c, p, corr, g, iflag, sub, out, dfk, nullptr);
    if ((st = nu_exec(c, h, g, iflag))) return st;
    const DeconvArgs da = deconv_args(p, sub, out);
    deconv_kernel<<<grid_for(out[0] * out[1] * out[2]), 256, 0, c->st>>>(da, reinterpret_cast<const double2 *>(g), corr[0], corr[1],
                                                                          corr[2], dfk);
    HIPCHK(hipGetLastError());
    return 0;
}

// Two 3D transforms of real constant strengths (sa on the points xa, sb on
// xb) as one complex transform: A spread into the real part, B into the
// imaginary part, separated in the deconvolution.  mb = 0: A alone (plain
// deconvolution).
int nufft_run_real_pair(dc_ctx *c, const NuParams &p, long long ma, const double *const xa[3], double sa, long long mb,
                        const double *const xb[3], double sb, int iflag, const long long sub[3], const long long out[3],
                        double2 *fa, double2 *fb) {
    NufftState &s = nu_state(c);
    const long long ng = p.n[0] * p.n[1] * p.n[2];
    int st = ensure(c, s.grid, sizeof(double2) * (size_t)ng);
    if (st) return st;
    const double *corr[3];
    hipfftHandle h;
    if ((st = nu_setup(c, p, corr, h, false))) return st;
    double *g = ptr<double>(s.grid);
    zero_kernel<<<grid_for(2 * ng), 256, 0, c->st>>>(2 * ng, g);
    const SpreadArgs a = spread_args(p);
    if (ma > 0) launch_spread_real(p.w, c->st, a, ma, xa[0], xa[1], xa[2], sa, 0, g);
    if (mb > 0) launch_spread_real(p.w, c->st, a, mb, xb[0], xb[1], xb[2], sb, 1, g);
    HIPCHK(hipGetLastError());
    return nu_transform3(c, p, corr, g, iflag, sub, out, fa, mb > 0 ? fb : nullptr);
}

int check_points(dc_ctx *c, int dim, long long M, const double *x, const double *y, const double *z) {
    if (M < 0) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "negative point count");
    if (M > 0 && (!x || (dim > 1 && !y) || (dim > 2 && !z))) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "null coordinate array");
    if (M > (1LL << 31)) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "more than 2^31 points");
    return 0;
}

}  // namespace

extern "C" {

int dc_nufft1_dev(dc_ctx *c, int dim, int64_t M, const double *x, const double *y, const double *z, const double *cw,
                  int iflag, double tol, int64_t n1, int64_t n2, int64_t n3, double *fk) {
    if (!c) return DC_ERR_ARG;
    int st = check_points(c, dim, M, x, y, z);
    if (st) return st;
    if (!fk || (M > 0 && !cw)) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "null weights or output");
    const long long N[3] = {n1, n2, n3};
    NuParams p;
    if ((st = nu_params(c, dim, tol, N, p))) return st;
    HIPCHK(hipSetDevice(c->device));
    const long long sub[3] = {1, 1, 1};
    const long long out[3] = {p.N[0], p.N[1], p.N[2]};
    return nufft_run(c, p, M, x, y, z, reinterpret_cast<const double2 *>(cw), iflag, sub, out, reinterpret_cast<double2 *>(fk));
}

int dc_nufft1(dc_ctx *c, int dim, int64_t M, const double *x, const double *y, const double *z, const double *cw,
              int iflag, double tol, int64_t n1, int64_t n2, int64_t n3, double *fk) {
    if (!c) return DC_ERR_ARG;
    int st = check_points(c, dim, M, x, y, z);
    if (st) return st;
    if (!fk || (M > 0 && !cw)) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "null weights or output");
    const long long N[3] = {n1, n2, n3};
    NuParams p;
    if ((st = nu_params(c, dim, tol, N, p))) return st;
    HIPCHK(hipSetDevice(c->device));
    NufftState &s = nu_state(c);
    const double *xs[3] = {x, y, z};
    for (int d = 0; d < dim; ++d)
        if ((st = upload(c, s.pts[d], xs[d], (size_t)M))) return st;
    if ((st = upload(c, s.c, cw, 2 * (size_t)M))) return st;
    const long long nout = p.N[0] * p.N[1] * p.N[2];
    if ((st = ensure(c, s.fk, sizeof(double2) * (size_t)nout))) return st;
    const long long sub[3] = {1, 1, 1};
    const long long out[3] = {p.N[0], p.N[1], p.N[2]};
    if ((st = nufft_run(c, p, M, ptr<double>(s.pts[0]), dim > 1 ? ptr<double>(s.pts[1]) : nullptr,
                        dim > 2 ? ptr<double>(s.pts[2]) : nullptr, ptr<double2>(s.c), iflag, sub, out, ptr<double2>(s.fk))))
        return st;
    HIPCHK(hipMemcpyAsync(fk, s.fk.p, sizeof(double2) * (size_t)nout, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return DC_OK;
}

int dc_four_strucf_nufft_dev(dc_ctx *c, int dim, int64_t nat, const double *xpos, const double *ypos, const double *zpos,
                             double occ, const int inc[3], const int lscales[3], double *tcsf) {
    if (!c) return DC_ERR_ARG;
    int st = check_points(c, dim, nat, xpos, ypos, zpos);
    if (st) return st;
    if (!inc || !lscales || !tcsf) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "null inc, lscales or tcsf");
    long long N[3], sub[3], out[3];
    for (int d = 0; d < 3; ++d) {
        if (d < dim && (inc[d] < 1 || lscales[d] < 1)) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "inc and lscales must be >= 1");
        sub[d] = d < dim ? lscales[d] : 1;
        out[d] = d < dim ? inc[d] : 1;
        N[d] = d < dim ? (long long)(inc[d] - 1) * lscales[d] + 1 : 1;   // fourier_finufft.f90:56, 113-114, 199-201
    }
    NuParams p;
    const double tol = dim == 3 ? 1.0e-7 : 1.0e-9;                      // fourier_finufft.f90:59, 121, 207
    if ((st = nu_params(c, dim, tol, N, p))) return st;
    HIPCHK(hipSetDevice(c->device));
    NufftState &s = nu_state(c);
    if ((st = ensure(c, s.c, sizeof(double2) * (size_t)std::max<int64_t>(nat, 1)))) return st;
    occ_weights_kernel<<<grid_for(nat), 256, 0, c->st>>>(nat, occ, ptr<double2>(s.c));
    HIPCHK(hipGetLastError());
    return nufft_run(c, p, nat, xpos, ypos, zpos, ptr<double2>(s.c), 1, sub, out, reinterpret_cast<double2 *>(tcsf));
}

int dc_four_strucf_nufft(dc_ctx *c, int dim, int64_t nat, const double *xpos, const double *ypos, const double *zpos,
                         double occ, const int inc[3], const int lscales[3], double *tcsf) {
    if (!c) return DC_ERR_ARG;
    int st = check_points(c, dim, nat, xpos, ypos, zpos);
    if (st) return st;
    if (!inc || !lscales || !tcsf) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "null inc, lscales or tcsf");
    HIPCHK(hipSetDevice(c->device));
    NufftState &s = nu_state(c);
    const double *xs[3] = {xpos, ypos, zpos};
    for (int d = 0; d < dim; ++d)
        if ((st = upload(c, s.pts[d], xs[d], (size_t)nat))) return st;
    long long nout = 1;
    for (int d = 0; d < dim; ++d) nout *= inc[d];
    if ((st = ensure(c, s.fk, sizeof(double2) * (size_t)nout))) return st;
    if ((st = dc_four_strucf_nufft_dev(c, dim, nat, ptr<double>(s.pts[0]), dim > 1 ? ptr<double>(s.pts[1]) : nullptr,
                                       dim > 2 ? ptr<double>(s.pts[2]) : nullptr, occ, inc, lscales, ptr<double>(s.fk))))
        return st;
    HIPCHK(hipMemcpyAsync(tcsf, s.fk.p, sizeof(double2) * (size_t)nout, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return DC_OK;
}

}  // extern "C"

// ---- run level: four_run_nufft (fourier_sup.f90:327-690), isotropic IAM ------
namespace {

__global__ void gather_scaled_kernel(long long n, const int *__restrict__ idx, const double *__restrict__ pos, double fx,
                                     double fy, double fz, double *__restrict__ x, double *__restrict__ y,
                                     double *__restrict__ z) {
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        const long long a = idx[e];
        x[e] = pos[3 * a] * fx;
        y[e] = pos[3 * a + 1] * fy;
        z[e] = pos[3 * a + 2] * fz;
    }
}

// csf (+)= tcsf * cfact(istl(i,j,h), t)  (tcsf_form_iso, fourier_form.f90)
__global__ void form_acc_kernel(Grid g, int t, int ncfact, int first, const double2 *__restrict__ cfact,
                                const double2 *__restrict__ fk, double2 *__restrict__ csf) {
    const long long n = (long long)g.n[0] * g.n[1] * g.n[2];
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(e % g.n[0]);
        const int j = (int)((e / g.n[0]) % g.n[1]);
        const int l = (int)(e / ((long long)g.n[0] * g.n[1]));
        const double2 cf = cfact[(long long)t * ncfact + istl_point(g, i, j, l)];
        const double2 f = fk[e];
        const double2 v = make_double2(f.x * cf.x - f.y * cf.y, f.x * cf.y + f.y * cf.x);
        csf[e] = first ? v : make_double2(csf[e].x + v.x, csf[e].y + v.y);
    }
}

// four_aver_finufft_n (fourier_sup.f90:1074-1190): F = 0 at the pixels of
// integer hkl reached through vi^-1 (skipping pixels outside the grid).
struct AverArgs {
    int n[3];
    int h0[3], nh[3];
    double e[3];
    double inv[9];      // row-major
};

__global__ void aver_zero_kernel(AverArgs a, double2 *__restrict__ csf) {
    const long long tot = (long long)a.nh[0] * a.nh[1] * a.nh[2];
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long long)gridDim.x * blockDim.x) {
        const int h = a.h0[0] + (int)(e % a.nh[0]);
        const int k = a.h0[1] + (int)((e / a.nh[0]) % a.nh[1]);
        const int l = a.h0[2] + (int)(e / ((long long)a.nh[0] * a.nh[1]));
        const double d[3] = {h - a.e[0], k - a.e[1], l - a.e[2]};
        long long ii[3];
        bool ok = true;
        for (int r = 0; r < 3; ++r) {
            const double v = a.inv[3 * r] * d[0] + a.inv[3 * r + 1] * d[1] + a.inv[3 * r + 2] * d[2] + 1.0;
            const double nv = v >= 0.0 ? floor(v + 0.5) : -floor(-v + 0.5);
            ok = ok && fabs(v - nv) < 1.0e-9;
            ii[r] = (long long)nv - 1;
            ok = ok && ii[r] >= 0 && ii[r] < a.n[r];
        }
        if (ok) csf[ii[0] + a.n[0] * (ii[1] + (long long)a.n[1] * ii[2])] = make_double2(0.0, 0.0);
    }
}

__global__ void nufft_finish_kernel(long long n, double sq, double ds, const double2 *__restrict__ f, double *__restrict__ csf,
                                    double *__restrict__ dsi) {
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        const double2 v = f[e];
        if (dsi) dsi[e] = fmax(0.0, v.x * v.x + v.y * v.y) * ds;
        if (csf) {
            csf[2 * e] = v.x * sq;
            csf[2 * e + 1] = v.y * sq;
        }
    }
}

double fint(double x) { return x < 0 ? std::ceil(x) : std::floor(x); }                 // Fortran INT
long long fnint(double x) { return (long long)(x >= 0 ? std::floor(x + 0.5) : -std::floor(-x + 0.5)); }

// Per-axis transform parameters of four_run_nufft (fourier_sup.f90:390-483).
int nufft_run_params(dc_ctx *c, const dc_run_desc *r, double scales[3], int lscales[3], int ncells[3]) {
    const double EPS = 1.0e-8, EPS2 = 1.0e-4, MAXSCALE = 3.0;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int64_t i = 0; i < c->natoms; ++i)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], c->h_pos[3 * i + d]);
            hi[d] = std::max(hi[d], c->h_pos[3 * i + d]);
        }
    int is_dim = 0;
    for (int i = 0; i < 3; ++i) {
        const double vii = r->vi[4 * i];
        for (int j = 0; j < 3; ++j)
            if (j != i && r->vi[i + 3 * j] != 0.0)
                return set_err(c, DC_ERR_ARG, DC_ER_LIB, "NUFFT technique: increment vectors must run along a*, b*, c*");
        if ((hi[i] - lo[i] > 1.0e-8 && c->icc[i] > 1) || std::fabs(vii) > 1.0e-8) ++is_dim;
        if (c->icc[i] == 1 && std::fabs(vii) > 1.0e-8)
            return set_err(c, DC_ERR_ARG, DC_ER_LIB, "NUFFT technique: crystals one cell thick along an axis are not supported");
    }
    if (is_dim != 3) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "NUFFT technique: only 3D crystals are supported");
    for (int j = 0; j < 3; ++j) {
        scales[j] = fint(std::fabs(r->vi[4 * j] * (hi[j] - lo[j])) - 0.01) + 1.0;
        lscales[j] = 1;
        if (scales[j] > MAXSCALE) {
            const long long ii = fnint(scales[j]);
            for (long long l = 2; l <= ii; ++l) {
                const double q = scales[j] / (double)l;
                if (q <= MAXSCALE && std::fabs(q - fint(q)) < 0.01) {
                    lscales[j] = (int)l;
                    scales[j] = q;
                    break;
                }
            }
        }
        ncells[j] = c->icc[j];
    }
    if (r->fave != 0.0) {
        static const char *axes[3] = {"abscissa", "ordinate", "top axis"};
        for (int i = 0; i < 3; ++i) {
            const double vii = r->vi[4 * i];
            const double t = std::fabs(vii * (double)c->icc[i]);
            if (std::fabs(t - (double)fnint(t)) > 0.0) {
                if (std::fabs(scales[i] - (double)fnint(scales[i])) > EPS)
                    return set_err(c, -186, DC_ER_APPL, std::string("Increment vector along ") + axes[i]);
                if (std::fabs(1.0 / vii - (double)fnint(1.0 / vii)) > EPS2)
                    return set_err(c, -187, DC_ER_APPL, std::string("Increment vector along ") + axes[i]);
                const int jj = r->inc[i] % 2 == 1 ? 1 : 0;
                if (std::fabs(r->eck[i] + (double)((r->inc[i] - jj) / 2) * vii) > EPS)
                    return set_err(c, -188, DC_ER_APPL, std::string("Increment vector along ") + axes[i]);
            }
        }
    }
    return 0;
}

}  // namespace

extern "C" {

int dc_four_run_nufft_dev(dc_ctx *c, const dc_run_desc *r, const double *type_occ, double *d_csf, double *d_dsi) {
    if (!c) return DC_ERR_ARG;
    if (!r || !r->cfact || r->ncfact < 1) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "dc_four_run_nufft: bad descriptor");
    if (!c->have_crystal) return set_err(c, DC_ERR_STATE, DC_ER_LIB, "dc_four_run_nufft: no crystal (dc_crystal_set)");
    if (c->natoms < 1) return set_err(c, DC_ERR_NOATOMS, DC_ER_APPL, "No fourier will be calculated");
    for (int k = 0; k < 3; ++k)
        if (r->inc[k] < 1) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "dc_four_run_nufft: inc < 1");
    double scales[3];
    int lscales[3], ncells[3];
    int st = nufft_run_params(c, r, scales, lscales, ncells);
    if (st) return st;
    Grid g{};
    for (int k = 0; k < 3; ++k) {
        g.n[k] = r->inc[k];
        g.f[k] = r->inc[k];
        g.xm[k] = r->eck[k];
        g.uin[k] = r->vi[k];
        g.vin[k] = r->vi[3 + k];
        g.win[k] = r->vi[6 + k];
    }
    for (int k = 0; k < 9; ++k) g.rten[k] = r->rten[k];
    if (max_corner_bin(g) > std::min(CFPKT, r->ncfact - 1))
        return set_err(c, DC_ERR_ISTL, DC_ER_FOUR, "sin(theta)/lambda beyond the form factor table");
    HIPCHK(hipSetDevice(c->device));
    stats_begin(c);
    HIPCHK(hipEventRecord(c->ev[0], c->st));
    NufftState &s = nu_state(c);
    const long long ntot = (long long)g.n[0] * g.n[1] * g.n[2];
    if ((st = upload(c, c->cfact, r->cfact, 2 * (size_t)r->ncfact * c->nscat))) return st;
    if ((st = ensure(c, s.fk, sizeof(double2) * (size_t)ntot))) return st;
    if ((st = ensure(c, c->csf, sizeof(double2) * (size_t)ntot))) return st;   // the sum over types
    double2 *acc = ptr<double2>(c->csf);
    const double pi = 3.14159265358979323846;
    double fac[3];
    for (int d = 0; d < 3; ++d) fac[d] = 2.0 * pi * scales[d] / (double)ncells[d];   // four_getatm_anis_iam
    long long N[3];
    for (int d = 0; d < 3; ++d) N[d] = (long long)(r->inc[d] - 1) * lscales[d] + 1;
    NuParams p;
    if ((st = nu_params(c, 3, 1.0e-7, N, p))) return st;                 // fourier_finufft.f90:207
    const long long sub[3] = {lscales[0], lscales[1], lscales[2]};
    const long long out[3] = {r->inc[0], r->inc[1], r->inc[2]};
    bool first = true;
    double pairs = 0.0;
    // the types with atoms, two per transform (real strengths cr_occ: one in
    // each component of the complex grid)
    std::vector<int> types;
    for (int t = 0; t < c->nscat; ++t)
        if (c->type_beg[t + 1] > c->type_beg[t]) types.push_back(t);
    if ((st = ensure(c, s.fk2, sizeof(double2) * (size_t)ntot))) return st;
    for (size_t q = 0; q < types.size(); q += 2) {
        if (cancelled(c)) return set_err(c, DC_ERR_CANCEL, DC_ER_COMM, "cancelled");
        const int ta = types[q], tb = q + 1 < types.size() ? types[q + 1] : -1;
        const long long na = c->type_beg[ta + 1] - c->type_beg[ta];
        const long long nb = tb >= 0 ? c->type_beg[tb + 1] - c->type_beg[tb] : 0;
        const double *xa[3], *xb[3] = {nullptr, nullptr, nullptr};
        for (int d = 0; d < 3; ++d) {
            if ((st = ensure(c, s.pts[d], sizeof(double) * (size_t)na))) return st;
            xa[d] = ptr<double>(s.pts[d]);
        }
        gather_scaled_kernel<<<grid_for(na), 256, 0, c->st>>>(na, ptr<int>(c->type_idx) + c->type_beg[ta], ptr<double>(c->pos),
                                                            fac[0], fac[1], fac[2], ptr<double>(s.pts[0]), ptr<double>(s.pts[1]),
                                                            ptr<double>(s.pts[2]));
        if (nb > 0) {
            for (int d = 0; d < 3; ++d) {
                if ((st = ensure(c, s.pts2[d], sizeof(double) * (size_t)nb))) return st;
                xb[d] = ptr<double>(s.pts2[d]);
            }
            gather_scaled_kernel<<<grid_for(nb), 256, 0, c->st>>>(nb, ptr<int>(c->type_idx) + c->type_beg[tb], ptr<double>(c->pos),
                                                                fac[0], fac[1], fac[2], ptr<double>(s.pts2[0]),
                                                                ptr<double>(s.pts2[1]), ptr<double>(s.pts2[2]));
        }
        HIPCHK(hipGetLastError());
        const double sa = type_occ ? type_occ[ta] : 1.0, sb = (tb >= 0 && type_occ) ? type_occ[tb] : 1.0;
        if ((st = nufft_run_real_pair(c, p, na, xa, sa, nb, xb, sb, 1, sub, out, ptr<double2>(s.fk), ptr<double2>(s.fk2))))
            return st;
        form_acc_kernel<<<grid_for(ntot), 256, 0, c->st>>>(g, ta, r->ncfact, first ? 1 : 0, ptr<double2>(c->cfact),
                                                          ptr<double2>(s.fk), acc);
        first = false;
        pairs += (double)na * (double)ntot;
        if (nb > 0) {
            form_acc_kernel<<<grid_for(ntot), 256, 0, c->st>>>(g, tb, r->ncfact, 0, ptr<double2>(c->cfact), ptr<double2>(s.fk2),
                                                              acc);
            pairs += (double)nb * (double)ntot;
        }
        HIPCHK(hipGetLastError());
    }
    if (first) zero_kernel<<<grid_for(2 * ntot), 256, 0, c->st>>>(2 * ntot, reinterpret_cast<double *>(acc));
    if (r->fave > 0.0) {
        // four_aver_finufft_n: vi with unit diagonal on flat axes, integer hkl box from INT(eck)
        double m[3][3];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) m[a][b] = r->vi[a + 3 * b];
        for (int a = 0; a < 3; ++a)
            if (r->inc[a] == 1) m[a][a] = 1.0;
        const double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                           m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
        if (det == 0.0) return set_err(c, DC_ERR_ARG, DC_ER_APPL, "Increment vectors cannot be inverted");
        AverArgs aa{};
        aa.inv[0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) / det;
        aa.inv[1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) / det;
        aa.inv[2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) / det;
        aa.inv[3] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) / det;
        aa.inv[4] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) / det;
        aa.inv[5] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) / det;
        aa.inv[6] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) / det;
        aa.inv[7] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) / det;
        aa.inv[8] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) / det;
        const int corner[3] = {1, 2, 3};   // eck(:,2) lr, eck(:,3) ul, eck(:,4) tl
        for (int a = 0; a < 3; ++a) {
            aa.n[a] = r->inc[a];
            aa.e[a] = r->eck[a];
            const long long h0 = fnint(fint(r->eck[a])), h1 = fnint(fint(r->eck[a + 3 * corner[a]]));
            aa.h0[a] = (int)h0;
            aa.nh[a] = (int)std::max(0LL, h1 - h0 + 1);
        }
        if ((long long)aa.nh[0] * aa.nh[1] * aa.nh[2] > 0)
            aver_zero_kernel<<<grid_for((long long)aa.nh[0] * aa.nh[1] * aa.nh[2]), 256, 0, c->st>>>(aa, acc);
        HIPCHK(hipGetLastError());
    }
    nufft_finish_kernel<<<grid_for(ntot), 256, 0, c->st>>>(ntot, std::sqrt(r->diff_scale), r->diff_scale, acc, d_csf, d_dsi);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(c->ev[1], c->st));
    c->stats.pairs = pairs;
    return DC_OK;
}

int dc_four_run_nufft(dc_ctx *c, const dc_run_desc *r, const double *type_occ, double *csf, double *dsi) {
    if (!c) return DC_ERR_ARG;
    if (!r) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "dc_four_run_nufft: bad descriptor");
    HIPCHK(hipSetDevice(c->device));
    const size_t n = (size_t)std::max(1, r->inc[0]) * std::max(1, r->inc[1]) * std::max(1, r->inc[2]);
    int st = 0;
    if ((st = ensure(c, c->out[0], sizeof(double) * 2 * n))) return st;
    if ((st = ensure(c, c->out[1], sizeof(double) * n))) return st;
    if ((st = dc_four_run_nufft_dev(c, r, type_occ, csf ? ptr<double>(c->out[0]) : nullptr, dsi ? ptr<double>(c->out[1]) : nullptr)))
        return st;
    if (csf) HIPCHK(hipMemcpyAsync(csf, c->out[0].p, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, c->st));
    if (dsi) HIPCHK(hipMemcpyAsync(dsi, c->out[1].p, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->ev[0], c->ev[1]) == hipSuccess) c->stats.total_ms = ms;
    return DC_OK;
}

}  // extern "C"
