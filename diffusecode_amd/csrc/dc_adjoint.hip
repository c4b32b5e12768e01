// dc_adjoint.hip -- reverse-mode derivative of the structure-factor path.
//
// Model (per atom n of scattering type t, fractional position x_n,
// occupancy o_n, isotropic displacement B_n):
//     F(p) = sum_n f_t(p) o_n exp(-B_n s_p^2) exp(2 pi i h_p . x_n),   I = |F|^2.
// For a loss L(I) with G_p = dL/dI_p put Q_t(p) = 2 G_p conj(F_p) f_t(p) and
// E_n(p) = exp(-B_n s_p^2) exp(2 pi i h_p . x_n).  Then
//     dL/do_n     = Re sum_p Q E
//     dL/dx_n,a   = -2 pi o_n Im sum_p h_p,a Q E
//     dL/dB_n     = -o_n Re sum_p s_p^2 Q E.
// REFINE obtains these by re-running the Fourier calculation for every
// parameter (finite differences, refine/prog/refine_run.f90:491-600); here
// they come from one pass that reuses the forward's phase generation.
//
// GEMM form (grid rows r = abscissa, columns c = (ordinate, top) flattened
// with period C0; the grid must be metric-orthogonal between rows and
// columns so that s^2 = s_r^2 + s_c^2 and E splits as A(r,n) B(c,n)):
//     Z_w(n, c) = sum_r w(r) A(r, n) Q(r, c),   w in {1, r - rc, s_r^2}
// on MFMA (FP32, K = grid rows, A generated per stage, Q streamed from HBM),
// then an epilogue contracts each Z with the generated column factor B(c, n)
// and the column weights (1, j - jc, m - mc, s_c^2).  h_p,a is affine in
// (r, j, m), so the three row weights and four column weights give every
// sum above; the weights are centred to keep the FP32 terms small.
#include <algorithm>

#include "dc_internal.h"

namespace dc {

typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int WA = 2, WC = 2, NWA = 2, NWC = 4;   // 16x16 tiles per wave, waves per work-group
constexpr int NA = 16 * WA * NWA;                  // atoms per work-group
constexpr int NC = 16 * WC * NWC;                  // grid columns per work-group
constexpr int NT = 64 * NWA * NWC;
constexpr int KR = 8;                              // grid rows per LDS stage
constexpr int KP = 2 * KR;                         // real K per stage (re, im interleaved)
constexpr float LOG2E = 1.4426950408889634f;

struct AdSmem {
    float a[2][KP][NA];       // A(k', atom), k' = 2 rr + (0 re, 1 im)
    float2 q[2][KR][NC];      // Q(rr, column)
    float sr[2][KR];          // s^2 row parts of the stage rows
    AtomRec at[NA];
    float red[NWC][NA][6];    // per column-wave partial sums (fixed-order reduction)
};

__device__ __forceinline__ float2 cis(uint64_t ph) {
    float c, s;
    sincos_turn((uint32_t)((ph + 0x80000000ull) >> 32), c, s);
    return make_float2(c, s);
}

// A of stage rows r0..r0+KR-1 (one element per thread) and the Q rows.
__device__ __forceinline__ void load_stage(AdSmem &sm, int buf, const AdjArgs &a, const float *__restrict__ sq,
                                           int r0, int col0, int na, long long qbase) {
    const int t = threadIdx.x;
    {
        const int at = t & (NA - 1), rr = t / NA;      // NT == NA * KR
        const int r = r0 + rr;
        float2 v = make_float2(0.f, 0.f);
        if (r < a.R && at < na) {
            const AtomRec &ar = sm.at[at];
            v = cis((uint64_t)r * ar.ra);
            const float d = __builtin_amdgcn_exp2f(-ar.pad * LOG2E * sq[r]);
            v = make_float2(v.x * d, v.y * d);
        }
        sm.a[buf][2 * rr][at] = v.x;
        sm.a[buf][2 * rr + 1][at] = v.y;
    }
#pragma unroll
    for (int u = 0; u < (KR * NC) / NT; ++u) {
        const int e = t + u * NT;
        const int rr = e / NC, cc = e % NC;
        const int r = r0 + rr, col = col0 + cc;
        sm.q[buf][rr][cc] = (r < a.R && col < a.C) ? a.q[qbase + col + (long long)a.C * r] : make_float2(0.f, 0.f);
    }
    if (t < KR) sm.sr[buf][t] = r0 + t < a.R ? sq[r0 + t] : 0.0f;
}

__global__ void __launch_bounds__(NT) adjoint_kernel(AdjArgs a) {
    __shared__ AdSmem sm;
    const int4 tile = a.tiles[blockIdx.y];
    const int atom0 = tile.x, na = tile.y - tile.x;
    const long long qbase = (long long)tile.z * a.plane;
    const float *sq = a.sq + (long long)tile.w * a.sq_stride;
    const int col0 = blockIdx.x * NC;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lg = lane >> 4;
    const int wa0 = (wave % NWA) * 16 * WA, wc0 = (wave / NWA) * 16 * WC;

    for (int x = threadIdx.x; x < NA; x += NT) {
        AtomRec r{};
        if (x < na) r = a.rec[atom0 + x];
        sm.at[x] = r;
    }
    __syncthreads();

    f32x4 accR[3][WA][WC], accI[3][WA][WC];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int i = 0; i < WA; ++i)
#pragma unroll
            for (int j = 0; j < WC; ++j) {
                accR[k][i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
                accI[k][i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
            }

    const int nst = (a.R + KR - 1) / KR;
    load_stage(sm, 0, a, sq, 0, col0, na, qbase);
    __syncthreads();
    for (int s = 0; s < nst; ++s) {
        const int buf = s & 1;
        if (s + 1 < nst) load_stage(sm, buf ^ 1, a, sq, (s + 1) * KR, col0, na, qbase);
        const bool im = (lg & 1) != 0;
#pragma unroll
        for (int ks = 0; ks < KR / 2; ++ks) {
            const int rr = 2 * ks + (lg >> 1);
            const float wr = (float)(s * KR + rr) - a.rc;
            const float ws = sm.sr[buf][rr];
            float av[WA][3];
#pragma unroll
            for (int i = 0; i < WA; ++i) {
                const float x = sm.a[buf][4 * ks + lg][wa0 + 16 * i + li];
                av[i][0] = x;
                av[i][1] = x * wr;
                av[i][2] = x * ws;
            }
#pragma unroll
            for (int j = 0; j < WC; ++j) {
                const float2 b = sm.q[buf][rr][wc0 + 16 * j + li];
                const float bR = im ? -b.y : b.x;
                const float bI = im ? b.x : b.y;
#pragma unroll
                for (int i = 0; i < WA; ++i)
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        accR[k][i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[i][k], bR, accR[k][i][j], 0, 0, 0);
                        accI[k][i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[i][k], bI, accI[k][i][j], 0, 0, 0);
                    }
            }
        }
        __syncthreads();
    }

    // epilogue: lane holds atoms 4 lg + v of column li of every 16x16 tile
    float acc[WA][4][6];
#pragma unroll
    for (int i = 0; i < WA; ++i)
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int k = 0; k < 6; ++k) acc[i][v][k] = 0.0f;
#pragma unroll
    for (int j = 0; j < WC; ++j) {
        const int col = col0 + wc0 + 16 * j + li;
        const int jj = col % a.C0, mm = col / a.C0;
        const bool valid = col < a.C && jj < a.f1;
        const float sc = valid ? sq[a.R + col] : 0.0f;
        const float wj = (float)jj - a.jc, wm = (float)mm - a.mc;
#pragma unroll
        for (int i = 0; i < WA; ++i)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const AtomRec &ar = sm.at[wa0 + 16 * i + 4 * lg + v];
                float2 b = cis(ar.p0 + (uint64_t)jj * ar.ca + (uint64_t)mm * ar.ca1);
                const float d = valid ? __builtin_amdgcn_exp2f(-ar.pad * LOG2E * sc) : 0.0f;
                b = make_float2(b.x * d, b.y * d);
                const float z0r = accR[0][i][j][v], z0i = accI[0][i][j][v];
                const float tr = z0r * b.x - z0i * b.y;
                const float ti = z0r * b.y + z0i * b.x;
                const float u1 = accR[1][i][j][v] * b.y + accI[1][i][j][v] * b.x;
                const float u2 = accR[2][i][j][v] * b.x - accI[2][i][j][v] * b.y;
                acc[i][v][0] += tr;
                acc[i][v][1] += ti;
                acc[i][v][2] += u1;
                acc[i][v][3] += wj * ti;
                acc[i][v][4] += wm * ti;
                acc[i][v][5] += u2 + sc * tr;
            }
    }
    // sum over the 16 columns of a lane group, then over the waves of a row of tiles
#pragma unroll
    for (int i = 0; i < WA; ++i)
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                float x = acc[i][v][k];
                x += __shfl_xor(x, 1);
                x += __shfl_xor(x, 2);
                x += __shfl_xor(x, 4);
                x += __shfl_xor(x, 8);
                if (li == 0) sm.red[wave / NWA][wa0 + 16 * i + 4 * lg + v][k] = x;
            }
    __syncthreads();
    float *out = a.part + ((long long)blockIdx.x * a.natoms + atom0) * 6;
    for (int x = threadIdx.x; x < na * 6; x += NT) {
        float v = 0.0f;
#pragma unroll
        for (int w = 0; w < NWC; ++w) v += (&sm.red[w][0][0])[x];
        out[x] = v;
    }
}

// Column-tile partials -> gradients in the caller's atom order (FP64).
__global__ void adjoint_finish_kernel(AdjFinishArgs a) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= a.natoms) return;
    double s[6] = {0, 0, 0, 0, 0, 0};
    for (int t = 0; t < a.ntile_c; ++t) {
        const float *p = a.part + ((long long)t * a.natoms + n) * 6;
        for (int k = 0; k < 6; ++k) s[k] += (double)p[k];
    }
    const long long ci = a.idx[n];
    const double o = a.occ[ci];
    const double twopi = 6.283185307179586476925;
    double *g = a.grad + 5 * ci;
    for (int c = 0; c < 3; ++c)
        g[c] = -twopi * o * (a.hc[c] * s[1] + a.v[c] * s[2] + a.v[3 + c] * s[3] + a.v[6 + c] * s[4]);
    g[3] = s[0];
    g[4] = -o * s[5];
}

// Q planes: Q_t(p) = 2 G_p conj(F_p) f_t(bin p), column-fastest layout.
__global__ void model_q_kernel(QArgs a) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.g.P) return;
    const int i = (int)(p % a.g.f[0]);
    const long long q = p / a.g.f[0];
    const int j = (int)(q % a.g.f[1]), m = (int)(q / a.g.f[1]);
    int bin = istl_point(a.g, i, j, m);
    bin = bin >= a.ncfact ? a.ncfact - 1 : bin;
    const double fr = a.f[2 * p], fi = a.f[2 * p + 1], g2 = 2.0 * a.gi[p];
    const double rr = g2 * fr, ri = -g2 * fi;                 // 2 G conj(F)
    const long long o = (long long)j + (long long)a.C0 * m + (long long)a.C * i;
    for (int t = 0; t < a.nq; ++t) {
        const double2 u = a.cfact[(long long)a.qtype[t] * a.ncfact + bin];
        a.q[t * a.plane + o] = make_float2((float)(rr * u.x - ri * u.y), (float)(rr * u.y + ri * u.x));
    }
}

// G = 2 w (I - T) and per-block sums of w (I - T)^2 (FP64).
__global__ void chi2_resid_kernel(const double *icalc, const double *tgt, const double *wt, long long n, double *gi,
                                  double *blocksum) {
    __shared__ double red[256];
    double acc = 0.0;
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long long)gridDim.x * blockDim.x) {
        const double w = wt ? wt[p] : 1.0;
        const double d = icalc[p] - tgt[p];
        gi[p] = 2.0 * w * d;
        acc += w * d * d;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) blocksum[blockIdx.x] = red[0];
}

// One member per blockIdx.y; grid-stride over the member's pixels.
__global__ void batch_reduce_kernel(BatchReduceArgs a) {
    __shared__ double red[256];
    const int m = blockIdx.y;
    Grid g = a.g;
    for (int k = 0; k < 9; ++k) g.rten[k] = a.rten[9 * m + k];
    const long long P = g.P;
    const int s0 = a.mslice[m], s1 = a.mslice[m + 1];
    const double2 *cf = a.cfact + (long long)m * a.nscat * a.ncfact;
    double acc = 0.0;
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < P; p += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(p % g.f[0]);
        const long long q = p / g.f[0];
        const int j = (int)(q % g.f[1]), l = (int)(q / g.f[1]);
        int bin = istl_point(g, i, j, l);
        bin = bin >= a.ncfact ? a.ncfact - 1 : bin;
        const long long po = i * g.ps[0] + j * g.ps[1] + l * g.ps[2];
        double fr = 0.0, fi = 0.0;
        for (int s = s0; s < s1; ++s) {
            const float2 v = a.partial[(long long)s * g.plane + po];
            const double2 u = cf[(long long)a.slice_type[s] * a.ncfact + bin];
            fr += u.x * (double)v.x - u.y * (double)v.y;
            fi += u.x * (double)v.y + u.y * (double)v.x;
        }
        const long long mp = (long long)m * P + p;
        const double I = fr * fr + fi * fi;
        if (a.icalc) a.icalc[mp] = I;
        const double w = a.weight ? a.weight[mp] : 1.0;
        const double d = I - a.target[mp];
        acc += w * d * d;
        if (a.q) {
            const double g2 = 4.0 * w * d;                       // 2 G, G = 2 w (I - T)
            const double rr = g2 * fr, ri = -g2 * fi;
            const long long o = (long long)j + (long long)a.C0 * l + (long long)a.C * i;
            for (int t = 0; t < a.nscat; ++t) {
                const double2 u = cf[(long long)t * a.ncfact + bin];
                a.q[((long long)m * a.nscat + t) * a.plane + o] =
                    make_float2((float)(rr * u.x - ri * u.y), (float)(rr * u.y + ri * u.x));
            }
        }
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) a.chi_part[(long long)m * gridDim.x + blockIdx.x] = red[0];
}

__global__ void fill_zero_kernel(double *p, long long n) {
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x)
        p[k] = 0.0;
}

}  // namespace

hipError_t launch_batch_reduce(const BatchReduceArgs &a, int nblocks, hipStream_t st) {
    if (a.nmembers <= 0) return hipSuccess;
    batch_reduce_kernel<<<dim3(nblocks, a.nmembers), 256, 0, st>>>(a);
    return hipGetLastError();
}

hipError_t launch_fill_zero(double *p, long long n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const long long nb = std::min<long long>((n + 255) / 256, 4096);
    fill_zero_kernel<<<(unsigned)nb, 256, 0, st>>>(p, n);
    return hipGetLastError();
}

int adjoint_tile_atoms() { return NA; }
int adjoint_tile_cols() { return NC; }

hipError_t launch_adjoint(const AdjArgs &a, int ntiles, hipStream_t st) {
    if (ntiles <= 0 || a.R <= 0 || a.C <= 0) return hipSuccess;
    dim3 grid((a.C + NC - 1) / NC, ntiles);
    adjoint_kernel<<<grid, dim3(NT), 0, st>>>(a);
    return hipGetLastError();
}

hipError_t launch_adjoint_finish(const AdjFinishArgs &a, hipStream_t st) {
    if (a.natoms <= 0) return hipSuccess;
    adjoint_finish_kernel<<<(a.natoms + 255) / 256, 256, 0, st>>>(a);
    return hipGetLastError();
}

hipError_t launch_model_q(const QArgs &a, hipStream_t st) {
    if (a.g.P <= 0) return hipSuccess;
    model_q_kernel<<<(unsigned)((a.g.P + 255) / 256), 256, 0, st>>>(a);
    return hipGetLastError();
}

hipError_t launch_chi2_resid(const double *icalc, const double *tgt, const double *wt, long long n, double *gi,
                             double *blocksum, int nblocks, hipStream_t st) {
    chi2_resid_kernel<<<nblocks, 256, 0, st>>>(icalc, tgt, wt, n, gi, blocksum);
    return hipGetLastError();
}

}  // namespace dc
