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
//
// Split variant (BF16x3, v_mfma_f32_16x16x32_bf16): the weighted A elements
// are split into BF16 hi + lo in FP32 and stored in LDS in the two forms of
// the real embedding; Q comes pre-split (q_split_kernel) in a layout whose
// 16-byte words are exactly one lane's MFMA operand, loaded from L2 into
// registers one stage ahead.  Products hi.hi + hi.lo + lo.hi, FP32
// accumulation, the same epilogue.
#include <algorithm>
#include <atomic>

#include "dc_internal.h"

namespace dc {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

namespace {

constexpr int WA = 2, WC = 2, NWA = 1, NWC = 8;   // 16x16 tiles per wave, waves per work-group
constexpr int NA = 16 * WA * NWA;                  // atoms per work-group
constexpr int NC = 16 * WC * NWC;                  // grid columns per work-group
constexpr int NT = 64 * NWA * NWC;
constexpr int KR = 16;                             // grid rows per LDS stage
static_assert(NT == NA * KR, "one A element per thread and stage");
static_assert(KR == 16, "row tables hold 16 rows per stage");
constexpr int MAXST = 64;                          // stages with tabulated row factors (R <= 1024)
constexpr int KP = 2 * KR;                         // real K per stage (re, im interleaved)
constexpr float LOG2E = 1.4426950408889634f;

// Column factors of the epilogue, exact two-level: column col0 + 16h + l of
// atom n has phase TH[n][h] + l ca + delta w (delta = 1 past the wrap of the
// flattened columns, at l >= wpos[h]; w = ca1 - C0 ca).
struct ColTables {
    float2 th[NA][NC / 16];
    float2 tl[NA][2][16];
    int wpos[NC / 16];
};

// Row factors, exact two-level: row 16 s + rr of atom n has phase
// 16 s ra + rr ra, so A = RH[s][n] * RL[n][rr] (grids of up to S stages).
template <int S>
struct RowTablesT {
    float2 rh[S][NA];
    // row-major with a padded row (8 float2): the stage builders read 4 rows
    // x 8 atoms per half-wave, which then fall on distinct banks
    float2 rl[16][NA + 8];
};
using RowTables = RowTablesT<MAXST>;

struct AdSmem {
    float a[2][KP][NA];       // A(k', atom), k' = 2 rr + (0 re, 1 im)
    float2 q[2][KR][NC];      // Q(rr, column)
    float sr[2][KR];          // s^2 row parts of the stage rows
    AtomRec at[NA];
    float red[NWC][NA][6];    // per column-wave partial sums (fixed-order reduction)
    ColTables ct;
    RowTables rt;
};

__device__ __forceinline__ float2 cis(uint64_t ph) {
    float c, s;
    sincos_turn((uint32_t)((ph + 0x80000000ull) >> 32), c, s);
    return make_float2(c, s);
}

template <class RT>
__device__ __forceinline__ void build_row_tables(RT &rt, const AtomRec *at, int nst) {
    for (int x = threadIdx.x; x < NA * (16 + nst); x += NT) {
        const int n = x % NA, e = x / NA;
        if (e < 16)
            rt.rl[e][n] = cis((uint64_t)e * at[n].ra);
        else
            rt.rh[e - 16][n] = cis((uint64_t)(16 * (e - 16)) * at[n].ra);
    }
}

// A element of atom n at row r = 16 s + rr (r < R): row phase and exp(-B s_r^2)
template <class RT>
__device__ __forceinline__ float2 row_element(const RT &rt, const AtomRec &ar, int n, int s, int rr, int r, bool tab,
                                              float s2) {
    float2 v;
    if (tab) {
        const float2 h = rt.rh[s][n], l = rt.rl[rr][n];
        v = make_float2(h.x * l.x - h.y * l.y, h.x * l.y + h.y * l.x);
    } else {
        v = cis((uint64_t)r * ar.ra);
    }
    const float d = __builtin_amdgcn_exp2f(-ar.pad * LOG2E * s2);
    return make_float2(v.x * d, v.y * d);
}

// A of stage rows r0..r0+KR-1 (one element per thread) and the Q rows.
__device__ __forceinline__ void load_stage(AdSmem &sm, int buf, const AdjArgs &a, const float *__restrict__ sq,
                                           int st, int col0, int na, long long qbase, bool tab) {
    const int t = threadIdx.x;
    const int r0 = st * KR;
    {
        const int at = t & (NA - 1), rr = t / NA;      // NT == NA * KR
        const int r = r0 + rr;
        float2 v = make_float2(0.f, 0.f);
        if (r < a.R && at < na) v = row_element(sm.rt, sm.at[at], at, st, rr, r, tab, sq[r]);
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

// Epilogue shared by the FP32 and split kernels: lane holds atoms 4 lg + v of
// column li of every 16x16 tile; contract with the generated column factor
// and weights, reduce over columns in a fixed order, write the partials.
// Three steps: column tables of the tile at col0 (epi_tables), contraction
// into per-lane column sums (epi_contract, adds to acc), reduction over lanes
// and waves into partial set `set` (epi_reduce).
__device__ __forceinline__ void epi_tables(const AdjArgs &a, const AtomRec *at, ColTables &ct, int col0) {
    // two-level column tables of the work-group's atoms (caller synchronised the LDS)
    constexpr int NH = NC / 16, NE = NH + 32;
    for (int x = threadIdx.x; x < NA * NE; x += NT) {
        const int n = x / NE, e = x % NE;
        const AtomRec &ar = at[n];
        if (e < NH) {
            const int c0 = col0 + 16 * e;
            ct.th[n][e] = cis(ar.p0 + (uint64_t)(c0 % a.C0) * ar.ca + (uint64_t)(c0 / a.C0) * ar.ca1);
            if (n == 0) ct.wpos[e] = a.C0 - c0 % a.C0;
        } else {
            const int d = (e - NH) >> 4, l = (e - NH) & 15;
            const uint64_t w = ar.ca1 - (uint64_t)a.C0 * ar.ca;
            ct.tl[n][d][l] = cis((uint64_t)l * ar.ca + (d ? w : 0ull));
        }
    }
}

__device__ __forceinline__ void epi_contract(const AdjArgs &a, const float *__restrict__ sq, const AtomRec *at,
                                             const ColTables &ct, const f32x4 (&accR)[3][WA][WC],
                                             const f32x4 (&accI)[3][WA][WC], int col0, float (&acc)[WA][4][6]) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lg = lane >> 4;
    const int wa0 = (wave % NWA) * 16 * WA, wc0 = (wave / NWA) * 16 * WC;
#pragma unroll
    for (int j = 0; j < WC; ++j) {
        const int col = col0 + wc0 + 16 * j + li;
        const int jj = col % a.C0, mm = col / a.C0;
        const bool valid = col < a.C && jj < a.f1;
        const float sc = valid ? sq[a.R + col] : 0.0f;
        const float wj = (float)jj - a.jc, wm = (float)mm - a.mc;
        const int h = wc0 / 16 + j;
        const int dl = li >= ct.wpos[h] ? 1 : 0;
#pragma unroll
        for (int i = 0; i < WA; ++i)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int n = wa0 + 16 * i + 4 * lg + v;
                const float2 th = ct.th[n][h], tl = ct.tl[n][dl][li];
                float2 b = make_float2(th.x * tl.x - th.y * tl.y, th.x * tl.y + th.y * tl.x);
                const float d = valid ? __builtin_amdgcn_exp2f(-at[n].pad * LOG2E * sc) : 0.0f;
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
}

// Sum of the per-wave partials (red, synchronised by the caller) in wave
// order, written as partial set `set`.
__device__ __forceinline__ void epi_write(const AdjArgs &a, const float (&red)[NWC][NA][6], int set, int atom0,
                                          int na) {
    float *out = a.part + ((long long)set * a.natoms + atom0) * 6;
    for (int x = threadIdx.x; x < na * 6; x += NT) {
        float v = 0.0f;
#pragma unroll
        for (int w = 0; w < NWC; ++w) v += (&red[w][0][0])[x];
        out[x] = v;
    }
}

__device__ __forceinline__ void epi_reduce(const AdjArgs &a, float (&red)[NWC][NA][6], float (&acc)[WA][4][6],
                                           int set, int atom0, int na) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lg = lane >> 4;
    const int wa0 = (wave % NWA) * 16 * WA;
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
                if (li == 0) red[wave / NWA][wa0 + 16 * i + 4 * lg + v][k] = x;
            }
    __syncthreads();
    epi_write(a, red, set, atom0, na);
}

// Row-local lane exchange (data-parallel primitives): lane li of each 16-lane
// row reads lane src(li) of the same row.
template <int CTRL>
__device__ __forceinline__ float dpp(float x) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), CTRL, 0xF, 0xF, false));
}
constexpr int DPP_ROW_MIRROR = 0x140;        // src = 15 - li
constexpr int DPP_ROW_HALF_MIRROR = 0x141;   // src = (li & 8) | (7 - (li & 7))
constexpr int DPP_QUAD_XOR2 = 0x4E;          // quad_perm [2,3,0,1]: src = li ^ 2
constexpr int DPP_QUAD_XOR1 = 0xB1;          // quad_perm [1,0,3,2]: src = li ^ 1

// Sums the 48 per-lane column sums of a 16-lane row over its lanes and
// scatters the totals: four halving exchanges (partners 15 - li, the mirror
// inside the half-row, li ^ 2, li ^ 1; bit 8, 4, 2, 1 of li picks the half a
// lane keeps) leave lane li with the row totals of flat entries 3 li, 3 li + 1,
// 3 li + 2 of acc (flat = (4 i + v) 6 + k).  Fixed order, so deterministic.
__device__ __forceinline__ void row_reduce_scatter(const float (&acc)[WA][4][6], float (&out)[3], int li) {
    static_assert(WA * 4 * 6 == 48, "48 column sums per lane");
    const float *x = &acc[0][0][0];
    float y[24], z[12], w[6];
#pragma unroll
    for (int j = 0; j < 24; ++j) {
        const bool h = (li & 8) != 0;
        y[j] = (h ? x[24 + j] : x[j]) + dpp<DPP_ROW_MIRROR>(h ? x[j] : x[24 + j]);
    }
#pragma unroll
    for (int j = 0; j < 12; ++j) {
        const bool h = (li & 4) != 0;
        z[j] = (h ? y[12 + j] : y[j]) + dpp<DPP_ROW_HALF_MIRROR>(h ? y[j] : y[12 + j]);
    }
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        const bool h = (li & 2) != 0;
        w[j] = (h ? z[6 + j] : z[j]) + dpp<DPP_QUAD_XOR2>(h ? z[j] : z[6 + j]);
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const bool h = (li & 1) != 0;
        out[j] = (h ? w[3 + j] : w[j]) + dpp<DPP_QUAD_XOR1>(h ? w[j] : w[3 + j]);
    }
}

__device__ __forceinline__ void adjoint_epilogue(const AdjArgs &a, const float *__restrict__ sq,
                                                 const AtomRec *at, float (&red)[NWC][NA][6], ColTables &ct,
                                                 f32x4 (&accR)[3][WA][WC], f32x4 (&accI)[3][WA][WC], int col0,
                                                 int atom0, int na) {
    epi_tables(a, at, ct, col0);
    __syncthreads();
    float acc[WA][4][6];
#pragma unroll
    for (int i = 0; i < WA; ++i)
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int k = 0; k < 6; ++k) acc[i][v][k] = 0.0f;
    epi_contract(a, sq, at, ct, accR, accI, col0, acc);
    epi_reduce(a, red, acc, blockIdx.x, atom0, na);
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
    const bool tab = nst <= MAXST;
    if (tab) build_row_tables(sm.rt, sm.at, nst);
    __syncthreads();
    load_stage(sm, 0, a, sq, 0, col0, na, qbase, tab);
    __syncthreads();
    for (int s = 0; s < nst; ++s) {
        const int buf = s & 1;
        if (s + 1 < nst) load_stage(sm, buf ^ 1, a, sq, s + 1, col0, na, qbase, tab);
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

    adjoint_epilogue(a, sq, sm.at, sm.red, sm.ct, accR, accI, col0, atom0, na);
}

// ---------------------------------------------------------------------------
// Split BF16x3 adjoint.
constexpr int SKR = 16;   // grid rows per stage: one 16x16x32 k-step (K' = 32)

__device__ __forceinline__ uint32_t bpack(float x, float y) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){x, y}, bf16x2));
}
// (x, y) = hi + lo, both packed BF16 pairs (x in the low half)
__device__ __forceinline__ void bsplit(float x, float y, uint32_t &hi, uint32_t &lo) {
    hi = bpack(x, y);
    lo = bpack(x - __builtin_bit_cast(float, hi << 16), y - __builtin_bit_cast(float, hi & 0xFFFF0000u));
}
__device__ __forceinline__ uint32_t rot16(uint32_t w) { return (w >> 16) | (w << 16); }

// One stage of split A in LDS: [form (re,-im) | (im,re)][weight][hi | lo][row group][atom]
typedef u32x4 StageA[2][3][2][4][NA];

struct AdSmemS {
    StageA a[2];                  // double-buffered stages
    AtomRec at[NA];
    float red[NWC][NA][6];
    ColTables ct;
    RowTables rt;
};

// A of stage st (rows 16 st .. 16 st + 15): thread (atom, row group g, e)
// makes row 4g + e, all three weights, both forms, hi and lo.
template <class RT>
__device__ __forceinline__ void gen_stage_split(StageA &dst, const RT &rt, const AtomRec *ats, const AdjArgs &a,
                                                const float *__restrict__ sq, int st, int na, bool tab) {
    const int t = threadIdx.x;
    // e fastest: a wave's 4-byte stores of one plane cover 64 consecutive words
    // (with the atom fastest they met on the same banks four at a time;
    // measured with the padded row tables: population forward + adjoint
    // 11.73 -> 11.43 ms, identical results)
    const int e = t & 3, at = (t >> 2) & (NA - 1), g = t / (4 * NA);
    const int rr = 4 * g + e, r = st * SKR + rr;
    float2 v = make_float2(0.f, 0.f);
    float s2 = 0.0f;
    if (r < a.R && at < na) {
        s2 = sq[r];
        v = row_element(rt, ats[at], at, st, rr, r, tab, s2);
    }
    const float wt[3] = {1.0f, (float)r - a.rc, s2};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        uint32_t hi, lo;
        bsplit(v.x * wt[k], v.y * wt[k], hi, lo);
        reinterpret_cast<uint32_t *>(&dst[0][k][0][g][at])[e] = hi ^ 0x80000000u;
        reinterpret_cast<uint32_t *>(&dst[1][k][0][g][at])[e] = rot16(hi);
        reinterpret_cast<uint32_t *>(&dst[0][k][1][g][at])[e] = lo ^ 0x80000000u;
        reinterpret_cast<uint32_t *>(&dst[1][k][1][g][at])[e] = rot16(lo);
    }
}

// B operand words of stage st: lane (column li, row group lg) of each of the
// wave's 16-column tiles starting at column cw (zero past the padded columns).
__device__ __forceinline__ void load_bwords(const uint32_t *__restrict__ qh, const uint32_t *__restrict__ ql,
                                            const AdjArgs &a, int cw, int st, u32x4 (&bh)[WC], u32x4 (&bl)[WC]) {
    const int lane = threadIdx.x & 63, li = lane & 15, lg = lane >> 4;
#pragma unroll
    for (int j = 0; j < WC; ++j) {
        const int col = cw + 16 * j + li;
        const long long o = ((long long)(st * 4 + lg) * a.C + col) * 4;
        if (col < a.C) {
            bh[j] = *reinterpret_cast<const u32x4 *>(qh + o);
            bl[j] = *reinterpret_cast<const u32x4 *>(ql + o);
        } else {
            bh[j] = u32x4{0u, 0u, 0u, 0u};
            bl[j] = u32x4{0u, 0u, 0u, 0u};
        }
    }
}

// The 72 MFMAs of one stage of one wave: hi.hi + hi.lo + lo.hi for the three
// row weights, both forms of the real embedding, WA x WC tiles.
__device__ __forceinline__ void split_mfma_stage(const StageA &sa, const u32x4 (&bh)[WC], const u32x4 (&bl)[WC],
                                                 f32x4 (&accR)[3][WA][WC], f32x4 (&accI)[3][WA][WC], int wa0) {
    const int lane = threadIdx.x & 63, li = lane & 15, lg = lane >> 4;
#pragma unroll
    for (int i = 0; i < WA; ++i) {
        u32x4 A[2][3][2];
#pragma unroll
        for (int f = 0; f < 2; ++f)
#pragma unroll
            for (int k = 0; k < 3; ++k)
#pragma unroll
                for (int p = 0; p < 2; ++p) A[f][k][p] = sa[f][k][p][lg][wa0 + 16 * i + li];
#pragma unroll
        for (int j = 0; j < WC; ++j)
#pragma unroll
            for (int t = 0; t < 3; ++t)          // hi.hi, hi.lo, lo.hi
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const int p = t == 2 ? 1 : 0;
                    const u32x4 y = t == 1 ? bl[j] : bh[j];
                    accR[k][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                        __builtin_bit_cast(bf16x8, A[0][k][p]), __builtin_bit_cast(bf16x8, y), accR[k][i][j], 0, 0, 0);
                    accI[k][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                        __builtin_bit_cast(bf16x8, A[1][k][p]), __builtin_bit_cast(bf16x8, y), accI[k][i][j], 0, 0, 0);
                }
    }
}

__global__ void __launch_bounds__(NT) adjoint_split_kernel(AdjArgs a) {
    __shared__ AdSmemS sm;
    const int4 tile = a.tiles[blockIdx.y];
    const int atom0 = tile.x, na = tile.y - tile.x;
    const float *sq = a.sq + (long long)tile.w * a.sq_stride;
    const uint32_t *qh = a.qs + (long long)tile.z * 2 * a.Rg * a.C * 4;
    const uint32_t *ql = qh + (long long)a.Rg * a.C * 4;
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

    const int nst = (a.R + SKR - 1) / SKR;
    const bool tab = nst <= MAXST;
    if (tab) build_row_tables(sm.rt, sm.at, nst);
    __syncthreads();
    u32x4 bh[WC], bl[WC], nbh[WC], nbl[WC];
    load_bwords(qh, ql, a, col0 + wc0, 0, bh, bl);
    gen_stage_split(sm.a[0], sm.rt, sm.at, a, sq, 0, na, tab);
    __syncthreads();
    for (int s = 0; s < nst; ++s) {
        const int buf = s & 1;
        if (s + 1 < nst) {
            load_bwords(qh, ql, a, col0 + wc0, s + 1, nbh, nbl);
            gen_stage_split(sm.a[buf ^ 1], sm.rt, sm.at, a, sq, s + 1, na, tab);
        }
        split_mfma_stage(sm.a[buf], bh, bl, accR, accI, wa0);
        if (s + 1 < nst) {
#pragma unroll
            for (int j = 0; j < WC; ++j) {
                bh[j] = nbh[j];
                bl[j] = nbl[j];
            }
        }
        __syncthreads();
    }
    adjoint_epilogue(a, sq, sm.at, sm.red, sm.ct, accR, accI, col0, atom0, na);
}

// Split adjoint for short grids (R <= 16 RES_ST rows): the work-group makes
// A for every stage once, keeps it in LDS, and walks a range of a.ctpg column
// tiles.  The stages of one column tile run without barriers (A is read-only,
// B streams into registers one stage ahead, also across column tiles); the
// epilogue's column sums accumulate per lane over the whole range, so the
// partial set is the column range (blockIdx.x), not the column tile.
constexpr int RES_ST = 5;
constexpr int RES_MAXCT = 32;   // column tiles per work-group at most (bounds the FP32 column sums)

// Epilogue of one column tile of a walk: column tables, contraction, and the
// row totals added into the lane's running sums tot (entries 3 li .. 3 li + 2).
__device__ __forceinline__ void walk_tile_epilogue(const AdjArgs &a, const float *__restrict__ sq, const AtomRec *at,
                                                   ColTables &ct, const f32x4 (&accR)[3][WA][WC],
                                                   const f32x4 (&accI)[3][WA][WC], int col0, float (&tot)[3]) {
    const int li = threadIdx.x & 15;
    __syncthreads();   // the previous tile's contraction has read the column tables
    epi_tables(a, at, ct, col0);
    __syncthreads();
    float acc[WA][4][6];
#pragma unroll
    for (int i = 0; i < WA; ++i)
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int k = 0; k < 6; ++k) acc[i][v][k] = 0.0f;
    epi_contract(a, sq, at, ct, accR, accI, col0, acc);
    float part[3];
    row_reduce_scatter(acc, part, li);
#pragma unroll
    for (int q = 0; q < 3; ++q) tot[q] += part[q];
}

// End of a walk: the lanes' totals into the per-wave slots, summed over the
// waves in order and written as partial set blockIdx.x.
__device__ __forceinline__ void walk_finish(const AdjArgs &a, float (&red)[NWC][NA][6], const float (&tot)[3], int atom0,
                                            int na) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lg = lane >> 4;
    const int wa0 = (wave % NWA) * 16 * WA;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const int f = 3 * li + q, i = f / 24, v = (f / 6) % 4, k = f % 6;
        red[wave / NWA][wa0 + 16 * i + 4 * lg + v][k] = tot[q];
    }
    __syncthreads();
    epi_write(a, red, blockIdx.x, atom0, na);
}

struct AdSmemR {
    StageA a[RES_ST];
    AtomRec at[NA];
    float red[NWC][NA][6];
    ColTables ct;
    RowTablesT<RES_ST> rt;
};

__global__ void __launch_bounds__(NT) adjoint_split_res_kernel(AdjArgs a) {
    __shared__ AdSmemR sm;
    const int4 tile = a.tiles[blockIdx.y];
    const int atom0 = tile.x, na = tile.y - tile.x;
    const float *sq = a.sq + (long long)tile.w * a.sq_stride;
    const uint32_t *qh = a.qs + (long long)tile.z * 2 * a.Rg * a.C * 4;
    const uint32_t *ql = qh + (long long)a.Rg * a.C * 4;
    const int ntc = (a.C + NC - 1) / NC;
    const int ct0 = blockIdx.x * a.ctpg, ct1 = min(ntc, ct0 + a.ctpg);
    const int wave = threadIdx.x >> 6;
    const int wa0 = (wave % NWA) * 16 * WA, wc0 = (wave / NWA) * 16 * WC;

    for (int x = threadIdx.x; x < NA; x += NT) {
        if (x < na)
            sm.at[x] = a.rec[atom0 + x];
        else
            sm.at[x] = AtomRec{};
    }
    __syncthreads();
    const int nst = (a.R + SKR - 1) / SKR;   // <= RES_ST, checked by the launcher
    build_row_tables(sm.rt, sm.at, nst);
    __syncthreads();
    for (int s = 0; s < nst; ++s) gen_stage_split(sm.a[s], sm.rt, sm.at, a, sq, s, na, true);
    u32x4 bh[WC], bl[WC], nbh[WC], nbl[WC];
    if (ct0 < ct1) load_bwords(qh, ql, a, ct0 * NC + wc0, 0, bh, bl);
    __syncthreads();

    float tot[3] = {0.0f, 0.0f, 0.0f};   // row totals of flat entries 3 li .. 3 li + 2, over the column range
    for (int ct = ct0; ct < ct1; ++ct) {
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
        for (int s = 0; s < nst; ++s) {
            // next (column tile, stage) of the walk
            const int nct = s + 1 < nst ? ct : ct + 1, ns = s + 1 < nst ? s + 1 : 0;
            const bool more = nct < ct1;
            if (more) load_bwords(qh, ql, a, nct * NC + wc0, ns, nbh, nbl);
            split_mfma_stage(sm.a[s], bh, bl, accR, accI, wa0);
            if (more) {
#pragma unroll
                for (int j = 0; j < WC; ++j) {
                    bh[j] = nbh[j];
                    bl[j] = nbl[j];
                }
            }
        }
        walk_tile_epilogue(a, sq, sm.at, sm.ct, accR, accI, ct * NC, tot);
    }
    walk_finish(a, sm.red, tot, atom0, na);
}

// FP32 adjoint for short grids, the same walk: the FP32 A elements of every
// stage (20 KB for 5 stages) stay in LDS, and each lane loads its Q operands
// (row 2 ks + lg / 2 of the stage, column li of each 16-column tile) from L2
// into registers: each k-step's slot is refilled with the next (column tile,
// stage) as soon as its MFMAs have read it, a stage ahead of use.
struct AdSmemRF {
    float a[RES_ST][KP][NA];   // A(k', atom), k' = 2 rr + (0 re, 1 im)
    float sr[RES_ST][KR];
    AtomRec at[NA];
    float red[NWC][NA][6];
    ColTables ct;
    RowTablesT<RES_ST> rt;
};

__device__ __forceinline__ void load_qslot(const AdjArgs &a, long long qbase, int cw, int st, int ks, float2 (&q)[WC]) {
    const int lane = threadIdx.x & 63, li = lane & 15, lg = lane >> 4;
#pragma unroll
    for (int j = 0; j < WC; ++j) {
        const int r = st * KR + 2 * ks + (lg >> 1), col = cw + 16 * j + li;
        q[j] = (r < a.R && col < a.C) ? a.q[qbase + col + (long long)a.C * r] : make_float2(0.f, 0.f);
    }
}

__global__ void __launch_bounds__(NT) adjoint_res_kernel(AdjArgs a) {
    __shared__ AdSmemRF sm;
    const int4 tile = a.tiles[blockIdx.y];
    const int atom0 = tile.x, na = tile.y - tile.x;
    const long long qbase = (long long)tile.z * a.plane;
    const float *sq = a.sq + (long long)tile.w * a.sq_stride;
    const int ntc = (a.C + NC - 1) / NC;
    const int ct0 = blockIdx.x * a.ctpg, ct1 = min(ntc, ct0 + a.ctpg);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lg = lane >> 4;
    const int wa0 = (wave % NWA) * 16 * WA, wc0 = (wave / NWA) * 16 * WC;
    const bool im = (lg & 1) != 0;

    for (int x = threadIdx.x; x < NA; x += NT) {
        if (x < na)
            sm.at[x] = a.rec[atom0 + x];
        else
            sm.at[x] = AtomRec{};
    }
    __syncthreads();
    const int nst = (a.R + KR - 1) / KR;   // <= RES_ST, checked by the launcher
    build_row_tables(sm.rt, sm.at, nst);
    __syncthreads();
    for (int s = 0; s < nst; ++s) {
        const int t = threadIdx.x, at = t & (NA - 1), rr = t / NA;   // NT == NA * KR
        const int r = s * KR + rr;
        float2 v = make_float2(0.f, 0.f);
        if (r < a.R && at < na) v = row_element(sm.rt, sm.at[at], at, s, rr, r, true, sq[r]);
        sm.a[s][2 * rr][at] = v.x;
        sm.a[s][2 * rr + 1][at] = v.y;
        if (t < KR) sm.sr[s][t] = s * KR + t < a.R ? sq[s * KR + t] : 0.0f;
    }
    float2 qr[KR / 2][WC];
    if (ct0 < ct1)
#pragma unroll
        for (int ks = 0; ks < KR / 2; ++ks) load_qslot(a, qbase, ct0 * NC + wc0, 0, ks, qr[ks]);
    __syncthreads();

    float tot[3] = {0.0f, 0.0f, 0.0f};
    for (int ct = ct0; ct < ct1; ++ct) {
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
        for (int s = 0; s < nst; ++s) {
            const int nct = s + 1 < nst ? ct : ct + 1, ns = s + 1 < nst ? s + 1 : 0;
            const bool more = nct < ct1;
#pragma unroll
            for (int ks = 0; ks < KR / 2; ++ks) {
                const int rr = 2 * ks + (lg >> 1);
                const float wr = (float)(s * KR + rr) - a.rc;
                const float ws = sm.sr[s][rr];
                float av[WA][3];
#pragma unroll
                for (int i = 0; i < WA; ++i) {
                    const float x = sm.a[s][4 * ks + lg][wa0 + 16 * i + li];
                    av[i][0] = x;
                    av[i][1] = x * wr;
                    av[i][2] = x * ws;
                }
#pragma unroll
                for (int j = 0; j < WC; ++j) {
                    const float2 b = qr[ks][j];
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
                if (more) load_qslot(a, qbase, nct * NC + wc0, ns, ks, qr[ks]);
            }
        }
        walk_tile_epilogue(a, sq, sm.at, sm.ct, accR, accI, ct * NC, tot);
    }
    walk_finish(a, sm.red, tot, atom0, na);
}

// FP32 complex Q planes -> split words: for plane t, part (hi, lo), row
// group g (rows 4g..4g+3) and column c one 16-byte word of 4 packed pairs.
__global__ void q_split_kernel(const float2 *__restrict__ q, int nq, int R, int C, long long plane, int Rg,
                               uint32_t *__restrict__ qs) {
    const long long n = (long long)nq * Rg * C;
    for (long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (long long)gridDim.x * blockDim.x) {
        const int c = (int)(x % C);
        const long long y = x / C;
        const int g = (int)(y % Rg);
        const long long t = y / Rg;
        u32x4 hi, lo;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int r = 4 * g + e;
            const float2 v = r < R ? q[t * plane + c + (long long)C * r] : make_float2(0.f, 0.f);
            uint32_t h, l;
            bsplit(v.x, v.y, h, l);
            hi[e] = h;
            lo[e] = l;
        }
        uint32_t *base = qs + t * 2 * Rg * (long long)C * 4;
        *reinterpret_cast<u32x4 *>(base + ((long long)g * C + c) * 4) = hi;
        *reinterpret_cast<u32x4 *>(base + ((long long)(Rg + g) * C + c) * 4) = lo;
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

// One member per blockIdx.y; grid-stride over the member's pixels.  A
// pixel's planes are loaded BR_BATCH at a time, all issued before they are
// summed (planes past the member's last slice repeat it, so every load is
// in bounds); CF_LDS: the pixel's form factor of each of the a.nscat
// (<= BR_LDS_TYPES) types is read once into LDS, so that the sums and the Q
// planes wait on no further global load.  Same arithmetic, same order, with
// or without.
constexpr int BR_BATCH = 8;
constexpr int BR_BLOCK = 256;
constexpr int BR_LDS_TYPES = 8;
template <bool CF_LDS>
__global__ __launch_bounds__(BR_BLOCK) void batch_reduce_kernel(BatchReduceArgs a) {
    __shared__ double red[BR_BLOCK];
    extern __shared__ double2 cf_lds[];   // CF_LDS: [type][thread]
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
        if constexpr (CF_LDS)
            for (int t = 0; t < a.nscat; ++t) cf_lds[t * BR_BLOCK + threadIdx.x] = cf[(long long)t * a.ncfact + bin];
        auto cfv = [&](int t) {
            if constexpr (CF_LDS) return cf_lds[t * BR_BLOCK + threadIdx.x];
            else return cf[(long long)t * a.ncfact + bin];
        };
        // the member's slices come in runs of one type: sum a run's planes,
        // then apply that type's form factor once
        double fr = 0.0, fi = 0.0, sr = 0.0, si = 0.0;
        for (int sb = s0; sb < s1; sb += BR_BATCH) {
            // the batch's slice types first, then its planes: the sums wait
            // for the types without waiting for the planes behind them
            int tt[BR_BATCH + 1];
#pragma unroll
            for (int u = 0; u <= BR_BATCH; ++u) tt[u] = a.slice_type[min(sb + u, s1 - 1)];
            float2 v[BR_BATCH];
#pragma unroll
            for (int u = 0; u < BR_BATCH; ++u) v[u] = a.partial[(long long)min(sb + u, s1 - 1) * g.plane + po];
#pragma unroll
            for (int u = 0; u < BR_BATCH; ++u) {
                const int s = sb + u;
                if (s >= s1) break;
                sr += (double)v[u].x;
                si += (double)v[u].y;
                const int t = tt[u];
                if (s + 1 == s1 || tt[u + 1] != t) {
                    const double2 w = cfv(t);
                    fr += w.x * sr - w.y * si;
                    fi += w.x * si + w.y * sr;
                    sr = si = 0.0;
                }
            }
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
                const double2 u = cfv(t);
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

__global__ void jvp_reduce_kernel(JvpArgs a) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.g.P) return;
    const Grid &g = a.g;
    const int i = (int)(p % g.f[0]);
    const long long q = p / g.f[0];
    const int j = (int)(q % g.f[1]), l = (int)(q / g.f[1]);
    int bin = istl_point(g, i, j, l);
    bin = bin >= a.ncfact ? a.ncfact - 1 : bin;
    double h[3];
    for (int k = 0; k < 3; ++k) h[k] = g.xm[k] + g.uin[k] * i + g.vin[k] * j + g.win[k] * l;
    double s2 = 0.0;
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) s2 += h[r] * g.rten[r + 3 * k] * h[k];
    s2 *= 0.25;
    const long long po = i * g.ps[0] + j * g.ps[1] + l * g.ps[2];
    const double fr = a.f[2 * p], fi = a.f[2 * p + 1];
    const double twopi = 6.283185307179586476925;
    for (int d = 0; d < a.ndir; ++d) {
        double gr[5], gi[5];
        for (int c = 0; c < 5; ++c) {
            gr[c] = 0.0;
            gi[c] = 0.0;
            for (int sl = a.dslice[5 * d + c]; sl < a.dslice[5 * d + c + 1]; ++sl) {
                const float2 v = a.partial[(long long)sl * g.plane + po];
                const double2 u = a.cfact[(long long)a.slice_type[sl] * a.ncfact + bin];
                gr[c] += u.x * (double)v.x - u.y * (double)v.y;
                gi[c] += u.x * (double)v.y + u.y * (double)v.x;
            }
        }
        const double hr = h[0] * gr[1] + h[1] * gr[2] + h[2] * gr[3];
        const double hi = h[0] * gi[1] + h[1] * gi[2] + h[2] * gi[3];
        const double dr = gr[0] - twopi * hi - s2 * gr[4];
        const double di = gi[0] + twopi * hr - s2 * gi[4];
        a.di[(long long)d * g.P + p] = 2.0 * (fr * dr + fi * di);
        if (a.df) {
            a.df[2 * ((long long)d * g.P + p)] = dr;
            a.df[2 * ((long long)d * g.P + p) + 1] = di;
        }
    }
}

__global__ void fill_zero_kernel(double *p, long long n) {
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x)
        p[k] = 0.0;
}

}  // namespace

hipError_t launch_batch_reduce(const BatchReduceArgs &a, int nblocks, hipStream_t st) {
    if (a.nmembers <= 0) return hipSuccess;
    if (a.nscat >= 1 && a.nscat <= BR_LDS_TYPES)
        batch_reduce_kernel<true>
            <<<dim3(nblocks, a.nmembers), BR_BLOCK, sizeof(double2) * BR_BLOCK * a.nscat, st>>>(a);
    else
        batch_reduce_kernel<false><<<dim3(nblocks, a.nmembers), BR_BLOCK, 0, st>>>(a);
    return hipGetLastError();
}

hipError_t launch_jvp_reduce(const JvpArgs &a, hipStream_t st) {
    if (a.g.P <= 0 || a.ndir <= 0) return hipSuccess;
    jvp_reduce_kernel<<<(unsigned)((a.g.P + 255) / 256), 256, 0, st>>>(a);
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

// Short grids (R <= 16 RES_ST, every scheme) walk column ranges with A
// resident in LDS; one such work-group fills a CU, so the walk length is
// chosen to balance rounds of work-groups over the CUs against the per-group
// prologue (about half a column tile of work).
int adjoint_max_walk() { return RES_MAXCT; }

int device_cu_count() {
    static std::atomic<int> cus[64];
    int dev = 0, ncu = 256;
    if (hipGetDevice(&dev) == hipSuccess && dev >= 0 && dev < 64) {
        ncu = cus[dev].load(std::memory_order_relaxed);
        if (ncu <= 0) {
            if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0)
                ncu = 256;
            cus[dev].store(ncu, std::memory_order_relaxed);
        }
    }
    return ncu;
}

int adjoint_plan(AdjArgs &a, int ntiles, bool split, int force) {
    const int ntc = std::max(1, (a.C + NC - 1) / NC);
    a.ctpg = 1;
    (void)split;
    if (ntiles <= 0 || a.R <= 0 || (a.R + SKR - 1) / SKR > RES_ST) return ntc;
    if (force > 0) {
        a.ctpg = std::min({force, ntc, RES_MAXCT});
        return (ntc + a.ctpg - 1) / a.ctpg;
    }
    const int ncu = device_cu_count();
    double best = 0.0;
    int bestc = 0;
    for (int ctpg = std::min(ntc, RES_MAXCT); ctpg >= 1; --ctpg) {
        const long long groups = (ntc + ctpg - 1) / ctpg;
        const double rounds = (double)((groups * ntiles + ncu - 1) / ncu);
        const double cost = rounds * (ctpg + 0.5);
        if (bestc == 0 || cost < best) {
            best = cost;
            bestc = ctpg;
        }
    }
    a.ctpg = bestc;
    return (ntc + bestc - 1) / bestc;
}

hipError_t launch_adjoint(const AdjArgs &a, int ntiles, hipStream_t st, bool split) {
    if (ntiles <= 0 || a.R <= 0 || a.C <= 0) return hipSuccess;
    const int ntc = (a.C + NC - 1) / NC;
    if (a.ctpg < 1 || a.ctpg > RES_MAXCT || (a.ctpg > 1 && (a.R + SKR - 1) / SKR > RES_ST)) return hipErrorInvalidValue;
    dim3 grid(ntc, ntiles);
    if (split && (!a.qs || a.Rg * 4 < a.R || a.Rg % 4)) return hipErrorInvalidValue;
    if (a.ctpg > 1) {
        grid.x = (ntc + a.ctpg - 1) / a.ctpg;
        if (split)
            adjoint_split_res_kernel<<<grid, dim3(NT), 0, st>>>(a);
        else
            adjoint_res_kernel<<<grid, dim3(NT), 0, st>>>(a);
    } else if (split) {
        adjoint_split_kernel<<<grid, dim3(NT), 0, st>>>(a);
    } else {
        adjoint_kernel<<<grid, dim3(NT), 0, st>>>(a);
    }
    return hipGetLastError();
}

hipError_t launch_q_split(const float2 *q, int nq, int R, int C, long long plane, int Rg, uint32_t *qs,
                          hipStream_t st) {
    const long long n = (long long)nq * Rg * C;
    if (n <= 0) return hipSuccess;
    const long long nb = std::min<long long>((n + 255) / 256, 8192);
    q_split_kernel<<<(unsigned)nb, 256, 0, st>>>(q, nq, R, C, plane, Rg, qs);
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
