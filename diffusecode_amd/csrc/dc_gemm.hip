// dc_gemm.hip -- the phase GEMM of the structure-factor path on gfx950.
//
// For one reciprocal grid the phase of atom n at output element (r, c)
// factorises, exp(2 pi i (p0_n + phi_row_n(r) + phi_col_n(c))), so a block of
// the grid is the complex GEMM
//     F(r, c) = sum_n A(r, n) * B(c, n),
//     A(r, n) = exp(2 pi i phi_row_n(r)),   B(c, n) = w_n exp(2 pi i (p0_n + phi_col_n(c)))
// (the GEMM form of the reference's triple loop four_strucf_OMP_mod.f90:160-179,
// which walks the same phases one table lookup at a time).  A and B are never
// stored in HBM: every work-group builds its KS-atom slices of A and B in LDS.
//
// Operand generation (exact, two-level).  Within a tile, row rho = 16*rh + rl
// of side A has the phase
//     P(r0) + rho*s + k(rho)*w      (64-bit fixed point, modulo one turn)
// with s the per-atom step along the fast axis, w the correction applied when
// the fast axis wraps (period R0) and k(rho) the number of wraps since the
// tile start.  It splits exactly into
//     [P(r0) + 16 rh s + k(16 rh) w]  +  [rl s + delta w],  delta in {0, 1},
// so each element is the product of two table entries, TH[rh] * TL_delta[rl],
// whose phases are exact integers turned into (cos, sin) by sincos_turn.  Per
// 16-atom stage a work-group evaluates 16 x (64 + RT/16 + CT/16) sincos and
// (RT + CT) x 16 complex products instead of (RT + CT) x 16 sincos.
//
// The complex product is embedded as a real GEMM with the atom dimension
// interleaved (re, im):  Cr = [Ar Ai].[Br -Bi],  Ci = [Ar Ai].[Bi Br].
//
// Schemes (template SCHEME):
//   SCH_FP32   v_mfma_f32_16x16x4_f32, FP32 operands: the plain drop-in.
//   SCH_BF16X3 v_mfma_f32_16x16x32_bf16, x = xh + xl split on both sides,
//              products xh*yh + xh*yl + xl*yh, FP32 accumulation.
//   SCH_FP16X2 v_mfma_f32_16x16x32_f16, A split (ah + al), B single half,
//              products ah*b + al*b.
//
// Pipeline: LDS holds two element buffers and two table buffers; iteration s
// builds the elements of stage s+1 and the tables of stage s+2 while the
// MFMAs consume stage s, with one barrier per stage.
//
// Output: partial[slice * P + r + R * c] (float2), one plane per slice
// (a slice = a contiguous atom range of one type of one lot).
#include "dc_internal.h"

namespace dc {

constexpr int SCH_FP32 = 0;
constexpr int SCH_BF16X3 = 1;
constexpr int SCH_FP16X2 = 2;

constexpr int KS = 16;            // atoms per LDS stage
constexpr int LSTR = KS + 4;      // u32 words per row of a packed plane (80 B, conflict-free b128 reads)

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t pack2(float a, float b, std::integral_constant<int, SCH_BF16X3>) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){a, b}, bf16x2));
}
__device__ __forceinline__ uint32_t pack2(float a, float b, std::integral_constant<int, SCH_FP16X2>) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){a, b}, f16x2));
}
__device__ __forceinline__ f32x2 unpack2(uint32_t w, std::integral_constant<int, SCH_BF16X3>) {
    return (f32x2){__builtin_bit_cast(float, w << 16), __builtin_bit_cast(float, w & 0xFFFF0000u)};
}
__device__ __forceinline__ f32x2 unpack2(uint32_t w, std::integral_constant<int, SCH_FP16X2>) {
    return __builtin_convertvector(__builtin_bit_cast(f16x2, w), f32x2);
}

// Round the 64-bit phase to the 32-bit turn fraction used by sincos_turn.
__device__ __forceinline__ uint32_t phase_top(uint64_t p) { return (uint32_t)((p + 0x80000000ull) >> 32); }

template <int WR_, int WC_, int NWR_, int NWC_>
struct TileCfg {
    static constexpr int WR = WR_, WC = WC_, NWR = NWR_, NWC = NWC_;
    static constexpr int RT = 16 * WR * NWR, CT = 16 * WC * NWC, NT = 64 * NWR * NWC;
    static constexpr int RH = RT / 16, CH = CT / 16;
    static constexpr int FSA = RT + 16, FSB = CT + 16;   // FP32 plane strides
};

template <int SCHEME, class T>
struct Elems {
    uint32_t a_hi[T::RT * LSTR], a_lo[T::RT * LSTR];
    uint32_t b_hi[T::CT * LSTR], b_lo[(SCHEME == SCH_BF16X3 ? T::CT : 1) * LSTR];
};
template <class T>
struct Elems<SCH_FP32, T> {
    float a[2 * KS * T::FSA];
    float b[2 * KS * T::FSB];
};

// Per-stage tables, [index][atom] layout.
template <class T>
struct Tables {
    float2 al[2][16 * KS], ah[T::RH * KS];
    float2 bl[2][16 * KS], bh[T::CH * KS];
};

template <int SCHEME, class T>
struct Smem {
    Elems<SCHEME, T> e[2];
    Tables<T> t[2];
    int a_k16[T::RH], a_wpos[T::RH];   // wraps before group rh, position of the wrap inside it
    int b_k16[T::CH], b_wpos[T::CH];
};

// Side geometry of the tile: start index, offset along the fast axis, period.
struct Side {
    int start, a0, hi0, period;
};

// ---------------------------------------------------------------------------
// Tables of stage [k0, k0+KS): thread t owns atom t % KS; table rows are
// enumerated as [A lo0 (16), A lo1 (16), A hi (RH), B lo0 (16), B lo1 (16), B hi (CH)].
template <class T>
__device__ __forceinline__ void build_tables(Tables<T> &tb, const int *__restrict__ a_k16,
                                             const int *__restrict__ b_k16, const AtomRec *__restrict__ rec,
                                             int k0, int kend, const Side &sa, const Side &sb) {
    constexpr int NROW = 64 + T::RH + T::CH;
    const int a = threadIdx.x & (KS - 1);
    const int k = k0 + a;
    const bool live = k < kend;
    const AtomRec ar = rec[live ? k : kend - 1];
    const float w = live ? ar.w : 0.0f;
    const uint64_t wa = ar.ra1 - (uint64_t)sa.period * ar.ra;
    const uint64_t wb = ar.ca1 - (uint64_t)sb.period * ar.ca;
    const uint64_t pa = (uint64_t)sa.a0 * ar.ra + (uint64_t)sa.hi0 * ar.ra1;
    const uint64_t pb = ar.p0 + (uint64_t)sb.a0 * ar.ca + (uint64_t)sb.hi0 * ar.ca1;
#pragma unroll
    for (int q0 = 0; q0 < NROW; q0 += T::NT / KS) {
        const int q = q0 + (int)(threadIdx.x >> 4);
        if (q >= NROW) break;
        uint64_t ph;
        float amp = 1.0f;
        float2 *dst;
        if (q < 32) {
            const int d = q >> 4, l = q & 15;
            ph = (uint64_t)l * ar.ra + (d ? wa : 0ull);
            dst = &tb.al[d][l * KS + a];
        } else if (q < 32 + T::RH) {
            const int h = q - 32;
            ph = pa + (uint64_t)(16 * h) * ar.ra + (uint64_t)a_k16[h] * wa;
            dst = &tb.ah[h * KS + a];
        } else if (q < 64 + T::RH) {
            const int qq = q - 32 - T::RH;
            const int d = qq >> 4, l = qq & 15;
            ph = (uint64_t)l * ar.ca + (d ? wb : 0ull);
            dst = &tb.bl[d][l * KS + a];
        } else {
            const int h = q - 64 - T::RH;
            ph = pb + (uint64_t)(16 * h) * ar.ca + (uint64_t)b_k16[h] * wb;
            amp = w;
            dst = &tb.bh[h * KS + a];
        }
        float c, s;
        sincos_turn(phase_top(ph), c, s);
        *dst = make_float2(c * amp, s * amp);
    }
}

__device__ __forceinline__ float2 cmul(float2 x, float2 y) {
    return make_float2(__builtin_fmaf(x.x, y.x, -x.y * y.y), __builtin_fmaf(x.x, y.y, x.y * y.x));
}

// Elements of one stage from its tables.  Thread t owns atom t % KS.
template <int SCHEME, class T>
__device__ __forceinline__ void build_elems(Elems<SCHEME, T> &el, const Tables<T> &tb, const int *__restrict__ a_wpos,
                                            const int *__restrict__ b_wpos) {
    const int a = threadIdx.x & (KS - 1);
    const int g = threadIdx.x >> 4;
    constexpr int NG = T::NT / KS;
#pragma unroll
    for (int rho = g; rho < T::RT; rho += NG) {
        const int h = rho >> 4, l = rho & 15;
        const float2 v = cmul(tb.ah[h * KS + a], tb.al[l >= a_wpos[h]][l * KS + a]);
        if constexpr (SCHEME == SCH_FP32) {
            el.a[(2 * a) * T::FSA + rho] = v.x;
            el.a[(2 * a + 1) * T::FSA + rho] = v.y;
        } else {
            constexpr auto S = std::integral_constant<int, SCHEME>{};
            const uint32_t hi = pack2(v.x, v.y, S);
            const f32x2 hv = unpack2(hi, S);
            el.a_hi[rho * LSTR + a] = hi;
            el.a_lo[rho * LSTR + a] = pack2(v.x - hv.x, v.y - hv.y, S);
        }
    }
#pragma unroll
    for (int gam = g; gam < T::CT; gam += NG) {
        const int h = gam >> 4, l = gam & 15;
        const float2 v = cmul(tb.bh[h * KS + a], tb.bl[l >= b_wpos[h]][l * KS + a]);
        if constexpr (SCHEME == SCH_FP32) {
            el.b[(2 * a) * T::FSB + gam] = v.x;
            el.b[(2 * a + 1) * T::FSB + gam] = v.y;
        } else if constexpr (SCHEME == SCH_BF16X3) {
            constexpr auto S = std::integral_constant<int, SCHEME>{};
            const uint32_t hi = pack2(v.x, v.y, S);
            const f32x2 hv = unpack2(hi, S);
            el.b_hi[gam * LSTR + a] = hi;
            el.b_lo[gam * LSTR + a] = pack2(v.x - hv.x, v.y - hv.y, S);
        } else {
            el.b_hi[gam * LSTR + a] = pack2(v.x, v.y, std::integral_constant<int, SCH_FP16X2>{});
        }
    }
}

// MFMAs of one stage for this wave's (WR x WC) 16x16 tiles; dead tiles skipped.
template <int SCHEME, class T>
__device__ __forceinline__ void mfma_stage(const Elems<SCHEME, T> &el, f32x4 (&accR)[T::WR][T::WC],
                                           f32x4 (&accI)[T::WR][T::WC], int wr0, int wc0, int lrows, int lcols) {
    const int lane = threadIdx.x & 63, li = lane & 15, lg = lane >> 4;
    if constexpr (SCHEME == SCH_FP32) {
#pragma unroll
        for (int ks = 0; ks < KS / 2; ++ks) {
            float av[T::WR], bR[T::WC], bI[T::WC];
#pragma unroll
            for (int i = 0; i < T::WR; ++i) av[i] = el.a[(4 * ks + lg) * T::FSA + wr0 + 16 * i + li];
            const int atom = 2 * ks + (lane >> 5);
            const bool im = (lg & 1) != 0;
#pragma unroll
            for (int j = 0; j < T::WC; ++j) {
                const float br = el.b[(2 * atom) * T::FSB + wc0 + 16 * j + li];
                const float bi = el.b[(2 * atom + 1) * T::FSB + wc0 + 16 * j + li];
                bR[j] = im ? -bi : br;
                bI[j] = im ? br : bi;
            }
#pragma unroll
            for (int j = 0; j < T::WC; ++j) {
                if (16 * j >= lcols) continue;
#pragma unroll
                for (int i = 0; i < T::WR; ++i) {
                    if (16 * i >= lrows) continue;
                    accR[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[i], bR[j], accR[i][j], 0, 0, 0);
                    accI[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[i], bI[j], accI[i][j], 0, 0, 0);
                }
            }
        }
    } else {
        u32x4 ah[T::WR], al[T::WR];
#pragma unroll
        for (int i = 0; i < T::WR; ++i) {
            const int o = (wr0 + 16 * i + li) * LSTR + 4 * lg;
            ah[i] = *reinterpret_cast<const u32x4 *>(&el.a_hi[o]);
            al[i] = *reinterpret_cast<const u32x4 *>(&el.a_lo[o]);
        }
#pragma unroll
        for (int j = 0; j < T::WC; ++j) {
            if (16 * j >= lcols) continue;
            const int o = (wc0 + 16 * j + li) * LSTR + 4 * lg;
            const u32x4 bh = *reinterpret_cast<const u32x4 *>(&el.b_hi[o]);
            // B'_r = (re, -im) ; B'_i = (im, re)
            const u32x4 brh = bh ^ 0x80000000u;
            const u32x4 bih = (bh >> 16) | (bh << 16);
            if constexpr (SCHEME == SCH_BF16X3) {
                const u32x4 bl = *reinterpret_cast<const u32x4 *>(&el.b_lo[o]);
                const u32x4 brl = bl ^ 0x80000000u;
                const u32x4 bil = (bl >> 16) | (bl << 16);
#pragma unroll
                for (int i = 0; i < T::WR; ++i) {
                    if (16 * i >= lrows) continue;
                    const bf16x8 a_h = __builtin_bit_cast(bf16x8, ah[i]);
                    const bf16x8 a_l = __builtin_bit_cast(bf16x8, al[i]);
                    accR[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_h, __builtin_bit_cast(bf16x8, brh), accR[i][j], 0, 0, 0);
                    accI[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_h, __builtin_bit_cast(bf16x8, bih), accI[i][j], 0, 0, 0);
                    accR[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_h, __builtin_bit_cast(bf16x8, brl), accR[i][j], 0, 0, 0);
                    accI[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_h, __builtin_bit_cast(bf16x8, bil), accI[i][j], 0, 0, 0);
                    accR[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_l, __builtin_bit_cast(bf16x8, brh), accR[i][j], 0, 0, 0);
                    accI[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_l, __builtin_bit_cast(bf16x8, bih), accI[i][j], 0, 0, 0);
                }
            } else {
#pragma unroll
                for (int i = 0; i < T::WR; ++i) {
                    if (16 * i >= lrows) continue;
                    const f16x8 a_h = __builtin_bit_cast(f16x8, ah[i]);
                    const f16x8 a_l = __builtin_bit_cast(f16x8, al[i]);
                    accR[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a_h, __builtin_bit_cast(f16x8, brh), accR[i][j], 0, 0, 0);
                    accI[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a_h, __builtin_bit_cast(f16x8, bih), accI[i][j], 0, 0, 0);
                    accR[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a_l, __builtin_bit_cast(f16x8, brh), accR[i][j], 0, 0, 0);
                    accI[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a_l, __builtin_bit_cast(f16x8, bih), accI[i][j], 0, 0, 0);
                }
            }
        }
    }
}

template <int SCHEME, class T>
__global__ void __launch_bounds__(T::NT)
phase_gemm_kernel(const AtomRec *__restrict__ rec, const int *__restrict__ slice_beg, float2 *__restrict__ partial,
                  GemmGeom g, int slice0) {
    __shared__ Smem<SCHEME, T> sm;
    const int slice = slice0 + blockIdx.z;
    const int kbeg = slice_beg[slice], kend = slice_beg[slice + 1];
    const int row0 = blockIdx.y * T::RT, col0 = blockIdx.x * T::CT;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wr0 = (wave % T::NWR) * 16 * T::WR, wc0 = (wave / T::NWR) * 16 * T::WC;
    const int lrows = g.R - row0 - wr0, lcols = g.C - col0 - wc0;   // live extent of this wave's tiles

    Side sa{row0, row0 % g.R0, row0 / g.R0, g.R0};
    Side sb{col0, col0 % g.C0, col0 / g.C0, g.C0};
    // wrap bookkeeping of the 16-groups of the tile (periods >= 16 checked by the host)
    for (int h = threadIdx.x; h < T::RH + T::CH; h += T::NT) {
        const bool isa = h < T::RH;
        const Side &s = isa ? sa : sb;
        const int hh = isa ? h : h - T::RH;
        const int x = s.a0 + 16 * hh;
        const int k16 = x / s.period;
        const int wpos = s.period - (x % s.period);   // first wrap inside the group at offset wpos
        (isa ? sm.a_k16 : sm.b_k16)[hh] = k16;
        (isa ? sm.a_wpos : sm.b_wpos)[hh] = wpos;
    }

    f32x4 accR[T::WR][T::WC], accI[T::WR][T::WC];
#pragma unroll
    for (int i = 0; i < T::WR; ++i)
#pragma unroll
        for (int j = 0; j < T::WC; ++j) {
            accR[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
            accI[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
        }

    const int nst = (kend - kbeg + KS - 1) / KS;
    __syncthreads();
    build_tables<T>(sm.t[0], sm.a_k16, sm.b_k16, rec, kbeg, kend, sa, sb);
    __syncthreads();
    build_elems<SCHEME, T>(sm.e[0], sm.t[0], sm.a_wpos, sm.b_wpos);
    if (nst > 1) build_tables<T>(sm.t[1], sm.a_k16, sm.b_k16, rec, kbeg + KS, kend, sa, sb);
    __syncthreads();
    for (int s = 0; s < nst; ++s) {
        const int cur = s & 1, nxt = cur ^ 1;
        if (s + 1 < nst) build_elems<SCHEME, T>(sm.e[nxt], sm.t[nxt], sm.a_wpos, sm.b_wpos);
        if (s + 2 < nst) build_tables<T>(sm.t[cur], sm.a_k16, sm.b_k16, rec, kbeg + (s + 2) * KS, kend, sa, sb);
        if (lrows > 0 && lcols > 0) mfma_stage<SCHEME, T>(sm.e[cur], accR, accI, wr0, wc0, lrows, lcols);
        __syncthreads();
    }

    // epilogue: lane holds rows 4*lg + v (v = 0..3) of column li of each 16x16 tile
    const int lane = threadIdx.x & 63, li = lane & 15, lg = lane >> 4;
    float2 *out = partial + (long long)(slice - slice0) * g.P;
#pragma unroll
    for (int i = 0; i < T::WR; ++i)
#pragma unroll
        for (int j = 0; j < T::WC; ++j) {
            const int c = col0 + wc0 + 16 * j + li;
            if (c >= g.C) continue;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = row0 + wr0 + 16 * i + 4 * lg + v;
                if (r < g.R) out[r + (long long)g.R * c] = make_float2(accR[i][j][v], accI[i][j][v]);
            }
        }
}

// ---------------------------------------------------------------------------
// Tile configurations (WR x WC MFMA tiles per wave, NWR x NWC waves).
using CfgA = TileCfg<4, 4, 2, 4>;   // 128 x 256, 8 waves
using CfgB = TileCfg<4, 4, 4, 2>;   // 256 x 128, 8 waves
using CfgC = TileCfg<3, 4, 2, 4>;   //  96 x 256, 8 waves
using CfgD = TileCfg<4, 4, 2, 2>;   // 128 x 128, 4 waves
using CfgE = TileCfg<2, 4, 2, 4>;   //  64 x 256, 8 waves
using CfgF = TileCfg<1, 4, 1, 4>;   //  16 x 256, 4 waves (lines / few rows)

static const TileInfo kTiles[] = {
    {CfgA::RT, CfgA::CT, CfgA::NT}, {CfgB::RT, CfgB::CT, CfgB::NT}, {CfgC::RT, CfgC::CT, CfgC::NT},
    {CfgD::RT, CfgD::CT, CfgD::NT}, {CfgE::RT, CfgE::CT, CfgE::NT}, {CfgF::RT, CfgF::CT, CfgF::NT},
};

int gemm_num_tiles() { return (int)(sizeof(kTiles) / sizeof(kTiles[0])); }
TileInfo gemm_tile_info(int t) { return kTiles[t]; }

template <int SCHEME, class T>
static void launch_one(const AtomRec *rec, const int *slice_beg, float2 *partial, const GemmGeom &g, int slice0,
                       int nslice, hipStream_t st) {
    dim3 grid((g.C + T::CT - 1) / T::CT, (g.R + T::RT - 1) / T::RT, nslice);
    phase_gemm_kernel<SCHEME, T><<<grid, dim3(T::NT), 0, st>>>(rec, slice_beg, partial, g, slice0);
}

template <int SCHEME>
static hipError_t launch_scheme(int tile, const AtomRec *rec, const int *slice_beg, float2 *partial,
                                const GemmGeom &g, int slice0, int nslice, hipStream_t st) {
    switch (tile) {
    case 0: launch_one<SCHEME, CfgA>(rec, slice_beg, partial, g, slice0, nslice, st); break;
    case 1: launch_one<SCHEME, CfgB>(rec, slice_beg, partial, g, slice0, nslice, st); break;
    case 2: launch_one<SCHEME, CfgC>(rec, slice_beg, partial, g, slice0, nslice, st); break;
    case 3: launch_one<SCHEME, CfgD>(rec, slice_beg, partial, g, slice0, nslice, st); break;
    case 4: launch_one<SCHEME, CfgE>(rec, slice_beg, partial, g, slice0, nslice, st); break;
    case 5: launch_one<SCHEME, CfgF>(rec, slice_beg, partial, g, slice0, nslice, st); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_phase_gemm(int scheme, int tile, const AtomRec *rec, const int *slice_beg, float2 *partial,
                             const GemmGeom &g, int slice0, int nslice, hipStream_t st) {
    // the two-level operand tables need wrap periods of at least 16
    if (g.R0 < 16 && g.R0 != g.R) return hipErrorInvalidValue;
    if (g.C0 < 16 && g.C0 != g.C) return hipErrorInvalidValue;
    switch (scheme) {
    case SCH_FP32: return launch_scheme<SCH_FP32>(tile, rec, slice_beg, partial, g, slice0, nslice, st);
    case SCH_BF16X3: return launch_scheme<SCH_BF16X3>(tile, rec, slice_beg, partial, g, slice0, nslice, st);
    case SCH_FP16X2: return launch_scheme<SCH_FP16X2>(tile, rec, slice_beg, partial, g, slice0, nslice, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace dc
