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
// about (RT + CT) x 16 complex products instead of (RT + CT) x 16 sincos.
//
// Per-atom phase rotation (dither_turn).  Every A element of atom n carries
// an extra pseudo-random phase +psi_n and every B element -psi_n (the low
// tables a second one, psi2_n, which the high tables take back), so the
// products are unchanged in exact arithmetic.  Without it, atoms of equal
// phase -- a crystal at its Bragg peaks -- round identically, and those
// errors add coherently over the whole slice (shifted fcc 40^3 crystal,
// 256,000 atoms in phase: BF16X3 2.3e-5 -> 5.0e-6, FP16X3 over 4,096-atom
// slices 4.9e-5 -> 4.0e-7, FP16X2 2.7e-4 -> 7.5e-7, FP32 3.2e-7 -> 3.9e-8
// in max|dI|/max(I); measured cost 0.3-3% of the launch).
//
// The complex product is embedded as a real GEMM with the atom dimension
// interleaved (re, im):  Cr = [Ar -Ai].[Br Bi],  Ci = [Ai Ar].[Br Bi]; each
// wave forms both A variants from one (re, im) load by sign/half swaps.
//
// Schemes (template SCHEME):
//   SCH_FP32   v_mfma_f32_16x16x4_f32, FP32 operands: the plain drop-in.
//   SCH_FP32D  the same products with FP64 partial sums every 4 atoms, for
//              runs whose result is a catastrophic difference (the <F>
//              subtraction, the cross-lot variance); each slice is written
//              as two FP32 planes (hi, lo).
//   SCH_BF16X3 v_mfma_f32_16x16x32_bf16, x = xh + xl split on both sides,
//              products xh*yh + xh*yl + xl*yh, FP32 accumulation.
//   SCH_FP16X2 v_mfma_f32_16x16x32_f16, A split (ah + al), B single half,
//              products ah*b + al*b.
//   SCH_FP16X3 v_mfma_f32_16x16x32_f16, both sides split into FP16 hi + lo,
//              products xh*yh + xh*yl + xl*yh: the same MFMA count as
//              BF16X3, but 11-bit halves leave 2^-22 instead of 2^-16 per
//              product, which coherent sums (Bragg peaks of crystals whose
//              sites are not at special positions) add up over every atom.
//              The FP16 schemes scale the B amplitudes of each slice by a
//              power of two (max |w| into [0.5, 1)) and undo it exactly in
//              the epilogue, so any weight range keeps FP16's precision.
//
// Pipeline: LDS holds two element buffers and two table buffers.  In
// iteration s every wave builds its share of the elements of stage s+1 and
// of the tables of stage s+2, then runs its MFMAs of stage s; one barrier per
// stage.  The atom records of stage s+3 are loaded during stage s.
//
// Output: partial[slice * P + r + R * c] (float2), one plane per slice
// (a slice = a contiguous atom range of one type of one lot).
#include <type_traits>

#include "dc_internal.h"

namespace dc {

constexpr int SCH_FP32 = 0;
constexpr int SCH_BF16X3 = 1;
constexpr int SCH_FP16X2 = 2;
constexpr int SCH_FP16X3 = 3;
constexpr int SCH_FP32D = SCHEME_FP32_ACC64;
// FP32 operands on v_mfma_f32_16x16x4_f32 (FP32D: partial sums kept in FP64)
__host__ __device__ constexpr bool fp32ops(int s) { return s == SCH_FP32 || s == SCH_FP32D; }
// three products (hi.hi, hi.lo, lo.hi) with both sides split
__host__ __device__ constexpr bool split3(int s) { return s == SCH_BF16X3 || s == SCH_FP16X3; }
// FP16 operands (v_mfma_f32_16x16x32_f16)
__host__ __device__ constexpr bool f16ops(int s) { return s == SCH_FP16X2 || s == SCH_FP16X3; }

constexpr int KS = 16;            // atoms per LDS stage

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template <int SCHEME>
__device__ __forceinline__ uint32_t pack2(float a, float b) {
    if constexpr (f16ops(SCHEME))
        return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){a, b}, f16x2));
    else
        return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){a, b}, bf16x2));
}
template <int SCHEME>
__device__ __forceinline__ f32x2 unpack2(uint32_t w) {
    if constexpr (f16ops(SCHEME))
        return __builtin_convertvector(__builtin_bit_cast(f16x2, w), f32x2);
    else
        return (f32x2){__builtin_bit_cast(float, w << 16), __builtin_bit_cast(float, w & 0xFFFF0000u)};
}

// Low FP16 halves of a split: given h = (f16(x), f16(y)) packed, returns
// (f16(x - h.x), f16(y - h.y)) packed.  x - h.x is exact in FP32, so each
// half is one rounding, as with a conversion of the FP32 difference; the
// mixed-precision FMAs take 2 instructions where unpacking h, subtracting
// and packing again takes 5.
__device__ __forceinline__ uint32_t lo_halves_f16(uint32_t h, float x, float y) {
    uint32_t r;
    asm("v_fma_mixlo_f16 %0, -1.0, %1, %2 op_sel_hi:[0,1,0]\n\t"
        "v_fma_mixhi_f16 %0, -1.0, %1, %3 op_sel:[0,1,0] op_sel_hi:[0,1,0]"
        : "=&v"(r)
        : "v"(h), "v"(x), "v"(y));
    return r;
}

// Round the 64-bit phase to the 32-bit turn fraction used by sincos_turn.
__device__ __forceinline__ uint32_t phase_top(uint64_t p) { return (uint32_t)((p + 0x80000000ull) >> 32); }

// Work-group layout: NWR x NWC waves, each owning a (16 WR) x (16 WC) block of
// the (16 WR NWR) x (16 WC NWC) tile.  Every wave both builds a share of the
// operand tables/elements of the coming stages and runs its MFMAs: on gfx950
// the VALU work of one wave only partly overlaps another wave's matrix-core
// work, so splitting waves into producers and consumers does not pay.
template <int WR_, int WC_, int NWR_, int NWC_>
struct TileCfg {
    static constexpr int WR = WR_, WC = WC_, NWR = NWR_, NWC = NWC_;
    static constexpr int NW = NWR * NWC, NT = 64 * NW, NPT = NT;
    static constexpr int RT = 16 * WR * NWR, CT = 16 * WC * NWC;
    static constexpr int RH = RT / 16, CH = CT / 16;
};

// Packed planes hold one u32 (re | im << 16) per (row, atom): 16 words per
// row, the four 16-byte atom quads of a row XOR-swizzled by bits 2..3 of the
// row so that 16-byte accesses of the element builders (4 atom quads x 4 rows) and
// of the MFMA operand reads (16 rows, one quad) fall on distinct banks.
__device__ __forceinline__ int pk_off(int row, int quad) { return row * 16 + 4 * (quad ^ ((row >> 2) & 3)); }
// FP32 planes: (row, k') with k' = 2*atom + (0 re, 1 im), 36 floats per row.
constexpr int FST = 36;

// The planes and tables are 16-byte aligned so that the 16-byte element
// stores and MFMA operand loads compile to ds_write_b128 / ds_read_b128: with
// only the 8-byte alignment of float2 the compiler splits them into pairs of
// 8-byte accesses, whose lanes then meet on the same banks two at a time
// (SQ_LDS_BANK_CONFLICT; measured A/B: headline 256 x 176 FP16X3 11.13 ->
// 10.94 ms, population 256 x 128 2.82 -> 2.68 ms, identical results).
template <int SCHEME, class T>
struct alignas(16) Elems {
    alignas(16) uint32_t a_hi[T::RT * 16];
    alignas(16) uint32_t a_lo[T::RT * 16];
    alignas(16) uint32_t b_hi[T::CT * 16];
    alignas(16) uint32_t b_lo[(split3(SCHEME) ? T::CT : 1) * 16];
};
template <class T>
struct alignas(16) ElemsF32 {
    alignas(16) float a[T::RT * FST];
    alignas(16) float b[T::CT * FST];
};
template <class T>
struct Elems<SCH_FP32, T> : ElemsF32<T> {};
template <class T>
struct Elems<SCH_FP32D, T> : ElemsF32<T> {};

// Per-stage tables, rows of KS atoms: TL_d[l] (l < 16, wrap state d) and
// TH[h] for both sides, laid out al[2][16] | ah[RH] | bl[2][16] | bh[CH].
template <class T>
struct alignas(16) Tables {
    alignas(16) float2 al[2][16 * KS];
    alignas(16) float2 ah[T::RH * KS];
    alignas(16) float2 bl[2][16 * KS];
    alignas(16) float2 bh[T::CH * KS];
    float dmp[KS];                      // B_n log2(e) of the stage's atoms (damped variant)
};

template <int SCHEME, class T>
struct alignas(16) Smem {
    Elems<SCHEME, T> e[2];
    Tables<T> t[2];
    int a_k16[T::RH], a_wpos[T::RH];   // wraps before group h, position of the wrap inside it
    int b_k16[T::CH], b_wpos[T::CH];
    float s_a[T::RT], s_b[T::CT];      // s^2 row / column parts of the tile (damped variant)
    float wmax[T::NW];                 // per-wave max |w| of the slice (FP16 schemes)
};

// Side geometry of the tile: start index, offset along the fast axis, period.
struct Side {
    int start, a0, hi0, period;
};

__device__ __forceinline__ float2 cmul(float2 x, float2 y) {
    return make_float2(__builtin_fmaf(x.x, y.x, -x.y * y.y), __builtin_fmaf(x.x, y.y, x.y * y.x));
}
__device__ __forceinline__ float2 cis(uint64_t ph) {
    float c, s;
    sincos_turn(phase_top(ph), c, s);
    return make_float2(c, s);
}

// ---------------------------------------------------------------------------
// Tables of one stage, built by all threads (pt = thread index).  Each atom
// a = pt % 16 (index katom of the record array) has NPT/16 helper threads
// r = pt / 16 (waves 0 and 1 build the low tables, 3 sincos per thread, the
// other waves the high rows; spreading the same work evenly over all waves
// measured 3-5% slower, 11.2 vs 10.7 ms on the headline):
//   r < 8      low tables of side r/4, entries l = m, m+4, m+8, m+12
//              (m = r % 4): TL0[l] = cis(l s + psi2), TL1[l] = cis(l s + w +
//              psi2).  Split schemes (SPLIT): helper m evaluates
//              cis(m s + psi2), cis(4m s) and cis(4m s + w) from their exact
//              phases; the four helpers of an atom and side (lanes a + 16 m
//              of one wave) exchange the last two by lane shuffles, and
//              TL_d[m + 4j] is the product cis(m s + psi2) * cis(4j s + d w):
//              one FP32 rounding more, far below the split's 2^-22, in place
//              of 8 sincos per thread (headline FP16X3 GEMM 10.70 -> 10.62
//              ms, 4 interleaved A/B pairs; coherent sums of the shifted
//              crystal 8.7e-7 -> 8.4e-7, BF16X3 3.3e-6 both).  The FP32
//              schemes evaluate all 8 entries from their exact phases: their
//              Bragg-subtracted runs hold 1e-6, which the extra rounding
//              exceeds (1.1e-6 on the lots + <F> test).
//              Each store of a wave covers 4 consecutive rows (512
//              contiguous bytes); with rows 4m .. 4m+3 per thread its 4 row
//              groups sat 512 bytes apart on the same banks (measured A/B:
//              headline 10.75 -> 10.52 ms, population GEMM 2.78 -> 2.63 ms,
//              identical results)
//   r >= 8     high-table rows h = r-8, r-8+NPT/16-8, ... of [A (RH) | B (CH)],
//              A rows rotated by psi - psi2 and B rows by -psi - psi2
// so every A element carries +psi and every B element -psi (dither_turn):
// the products are unchanged, their rounding errors differ per atom.
template <class T, bool SPLIT>
__device__ __forceinline__ void build_tables(Tables<T> &tb, const int *__restrict__ a_k16,
                                             const int *__restrict__ b_k16, const AtomRec &ar, bool live,
                                             const Side &sa, const Side &sb, int pt, float wscale, uint32_t katom,
                                             uint32_t salt, bool padh) {
    const int a = pt & (KS - 1), r = pt >> 4;
    constexpr int NR = T::NPT / KS;           // helper threads per atom
    static_assert(NR > 8, "the low tables take waves 0 and 1, the high rows need more threads");
    // padh: the record carries the hash (dither_hash32 of katom, written by
    // the record prep for undamped launches), so the 512 threads of the
    // work-group do not each evaluate it every stage; same bits either way
    uint64_t psi;
    if constexpr (SPLIT) {
        if (padh)
            psi = (uint64_t)__builtin_bit_cast(uint32_t, ar.pad) << 32;
        else
            psi = dither_turn32(katom, salt);
    } else {
        psi = dither_turn(katom, salt);
    }
    const uint64_t psi2 = psi * 0xD6E8FEB86659FD93ull;
    if (r == 0) tb.dmp[a] = live ? ar.pad * 1.4426950408889634f : 0.0f;
    if (r < 8) {
        const bool sideb = r >= 4;
        const int m = r & 3;
        const uint64_t step = sideb ? ar.ca : ar.ra;
        const uint64_t wrp = sideb ? ar.ca1 - (uint64_t)sb.period * ar.ca : ar.ra1 - (uint64_t)sa.period * ar.ra;
        float2 *t0 = sideb ? tb.bl[0] : tb.al[0];
        float2 *t1 = sideb ? tb.bl[1] : tb.al[1];
        if constexpr (SPLIT) {
            const float2 base = cis((uint64_t)m * step + psi2);
            const uint64_t hs = (uint64_t)(4 * m) * step;
            const float2 g0 = cis(hs), g1 = cis(hs + wrp);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                // helper j of this atom and side: lane a + 16 j of this wave
                // (waves 0 and 1 are exactly the threads r < 8)
                const int src = a + 16 * j;
                const float2 h0 = make_float2(__shfl(g0.x, src), __shfl(g0.y, src));
                const float2 h1 = make_float2(__shfl(g1.x, src), __shfl(g1.y, src));
                t0[(m + 4 * j) * KS + a] = cmul(base, h0);
                t1[(m + 4 * j) * KS + a] = cmul(base, h1);
            }
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint64_t ph = (uint64_t)(m + 4 * j) * step + psi2;
                t0[(m + 4 * j) * KS + a] = cis(ph);
                t1[(m + 4 * j) * KS + a] = cis(ph + wrp);
            }
        }
    } else {
        const float w = live ? ar.w * wscale : 0.0f;   // wscale: power of two, exact
        const uint64_t wa = ar.ra1 - (uint64_t)sa.period * ar.ra;
        const uint64_t wb = ar.ca1 - (uint64_t)sb.period * ar.ca;
        const uint64_t pa = (uint64_t)sa.a0 * ar.ra + (uint64_t)sa.hi0 * ar.ra1 + psi - psi2;
        const uint64_t pb = ar.p0 + (uint64_t)sb.a0 * ar.ca + (uint64_t)sb.hi0 * ar.ca1 - psi - psi2;
        for (int q = r - 8; q < T::RH + T::CH; q += NR - 8) {
            const bool sideb = q >= T::RH;
            const int h = sideb ? q - T::RH : q;
            const uint64_t step = sideb ? ar.ca : ar.ra;
            const uint64_t wrp = sideb ? wb : wa;
            const int k16 = sideb ? b_k16[h] : a_k16[h];
            const uint64_t ph = (sideb ? pb : pa) + (uint64_t)(16 * h) * step + (uint64_t)(uint32_t)k16 * wrp;
            const float amp = sideb ? w : 1.0f;
            const float2 v = cis(ph);
            (sideb ? tb.bh : tb.ah)[h * KS + a] = make_float2(v.x * amp, v.y * amp);
        }
    }
}

// Load the 4 table entries of atoms 4g..4g+3 (32 bytes).
__device__ __forceinline__ void ld4(const float2 *p, float2 (&v)[4]) {
    const f32x4 x = *reinterpret_cast<const f32x4 *>(p);
    const f32x4 y = *reinterpret_cast<const f32x4 *>(p + 2);
    v[0] = make_float2(x[0], x[1]);
    v[1] = make_float2(x[2], x[3]);
    v[2] = make_float2(y[0], y[1]);
    v[3] = make_float2(y[2], y[3]);
}

// One 16-row group h of side A (isA) or B: lane (l = row low bits, g = atom
// quad) forms the 4 elements TH[h] * TL_d[l] and stores them.
template <int SCHEME, class T, bool isA, bool DAMP, bool SEL>
__device__ __forceinline__ void elems_group(Elems<SCHEME, T> &el, const float2 *__restrict__ th, const float2 (&tl0)[4],
                                            const float2 (&tl1)[4], const f32x4 &dm, const float *__restrict__ s_tab,
                                            int h, int wpos, int l, int g) {
    float2 hv[4];
    ld4(th + h * KS + 4 * g, hv);
    const bool d = l >= wpos;
    float2 v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = cmul(hv[q], (SEL && d) ? tl1[q] : tl0[q]);
    const int row = 16 * h + l;
    if constexpr (DAMP) {
        // per-atom isotropic displacement factor exp(-B_n s^2) of this row / column
        const float sv = s_tab[row];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float f = __builtin_amdgcn_exp2f(-dm[q] * sv);
            v[q] = make_float2(v[q].x * f, v[q].y * f);
        }
    }
    if constexpr (fp32ops(SCHEME)) {
        float *dst = (isA ? el.a : el.b) + row * FST + 8 * g;
        *reinterpret_cast<f32x4 *>(dst) = f32x4{v[0].x, v[0].y, v[1].x, v[1].y};
        *reinterpret_cast<f32x4 *>(dst + 4) = f32x4{v[2].x, v[2].y, v[3].x, v[3].y};
    } else {
        u32x4 hi, lo;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            hi[q] = pack2<SCHEME>(v[q].x, v[q].y);
            if constexpr (f16ops(SCHEME)) {
                lo[q] = lo_halves_f16(hi[q], v[q].x, v[q].y);
            } else {
                const f32x2 hvv = unpack2<SCHEME>(hi[q]);
                lo[q] = pack2<SCHEME>(v[q].x - hvv.x, v[q].y - hvv.y);
            }
        }
        const int o = pk_off(row, g);
        if constexpr (isA) {
            *reinterpret_cast<u32x4 *>(&el.a_hi[o]) = hi;
            *reinterpret_cast<u32x4 *>(&el.a_lo[o]) = lo;
        } else {
            *reinterpret_cast<u32x4 *>(&el.b_hi[o]) = hi;
            if constexpr (split3(SCHEME)) *reinterpret_cast<u32x4 *>(&el.b_lo[o]) = lo;
        }
    }
}

// Elements of one stage, built by wave pw: its contiguous run of the RH + CH
// 16-row groups.  The low tables of a side are loaded once.  A side whose
// fast axis has a single period over the grid (awrap / bwrap false: R0 >= R,
// C0 >= C) has no wrap inside any 16-group that holds grid rows, so it skips
// the per-lane choice between TL_0 and TL_1 (groups past the grid edge hold
// padding, which is never stored).  Population GEMM (256 x 128, both sides
// single-period) 2.605 -> 2.543 ms, headline (rows wrap at 81) unchanged,
// 4 interleaved A/B pairs, identical results.
template <int SCHEME, class T, bool DAMP>
__device__ __forceinline__ void build_elems(Elems<SCHEME, T> &el, const Tables<T> &tb, const int *__restrict__ a_wpos,
                                            const int *__restrict__ b_wpos, const float *__restrict__ s_a,
                                            const float *__restrict__ s_b, int pw, bool awrap, bool bwrap) {
    constexpr int NJOB = T::RH + T::CH;
    const int lane = threadIdx.x & 63, g = lane & 3, l = lane >> 2;
    const int j0 = (pw * NJOB) / T::NW, j1 = ((pw + 1) * NJOB) / T::NW;
    float2 tl0[4], tl1[4];
    f32x4 dm = f32x4{0.f, 0.f, 0.f, 0.f};
    if constexpr (DAMP) dm = *reinterpret_cast<const f32x4 *>(&tb.dmp[4 * g]);
    if (j0 < T::RH) {
        ld4(&tb.al[0][l * KS + 4 * g], tl0);
        const int je = j1 < T::RH ? j1 : T::RH;
        if (awrap) {
            ld4(&tb.al[1][l * KS + 4 * g], tl1);
            for (int j = j0; j < je; ++j) elems_group<SCHEME, T, true, DAMP, true>(el, tb.ah, tl0, tl1, dm, s_a, j, a_wpos[j], l, g);
        } else {
            for (int j = j0; j < je; ++j) elems_group<SCHEME, T, true, DAMP, false>(el, tb.ah, tl0, tl0, dm, s_a, j, 16, l, g);
        }
    }
    if (j1 > T::RH) {
        ld4(&tb.bl[0][l * KS + 4 * g], tl0);
        if (bwrap) {
            ld4(&tb.bl[1][l * KS + 4 * g], tl1);
            for (int j = (j0 > T::RH ? j0 : T::RH); j < j1; ++j)
                elems_group<SCHEME, T, false, DAMP, true>(el, tb.bh, tl0, tl1, dm, s_b, j - T::RH, b_wpos[j - T::RH], l, g);
        } else {
            for (int j = (j0 > T::RH ? j0 : T::RH); j < j1; ++j)
                elems_group<SCHEME, T, false, DAMP, false>(el, tb.bh, tl0, tl0, dm, s_b, j - T::RH, 16, l, g);
        }
    }
}

// MFMAs of one stage for this wave's (WR x WC) 16x16 tiles (waves whose
// whole block lies outside the grid skip the stage).  A is read once as
// (re, im) and permuted in registers into the two forms of the real
// embedding: (re, -im) for Cr and (im, re) for Ci.
template <int SCHEME, class T>
__device__ __forceinline__ void mfma_stage(const Elems<SCHEME, T> &el, f32x4 (&accR)[T::WR][T::WC],
                                           f32x4 (&accI)[T::WR][T::WC], int wr0, int wc0) {
    const int lane = threadIdx.x & 63, li = lane & 15, lg = lane >> 4;
    if constexpr (fp32ops(SCHEME)) {
#pragma unroll
        for (int ks = 0; ks < KS / 2; ++ks) {
            float av[T::WR], bR[T::WC], bI[T::WC];
#pragma unroll
            for (int i = 0; i < T::WR; ++i) av[i] = el.a[(wr0 + 16 * i + li) * FST + 4 * ks + lg];
            const int atom = 2 * ks + (lane >> 5);
            const bool im = (lg & 1) != 0;
#pragma unroll
            for (int j = 0; j < T::WC; ++j) {
                const float2 b = *reinterpret_cast<const float2 *>(&el.b[(wc0 + 16 * j + li) * FST + 2 * atom]);
                bR[j] = im ? -b.y : b.x;
                bI[j] = im ? b.x : b.y;
            }
#pragma unroll
            for (int j = 0; j < T::WC; ++j)
#pragma unroll
                for (int i = 0; i < T::WR; ++i) {
                    accR[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[i], bR[j], accR[i][j], 0, 0, 0);
                    accI[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[i], bI[j], accI[i][j], 0, 0, 0);
                }
        }
    } else {
        // operands of the whole stage in registers, then each split term sweeps
        // all WR x WC tiles so consecutive MFMAs never share an accumulator;
        // wide column blocks (WC > 4) load B per 16-column tile instead
        constexpr bool JOUTER = T::WC > 4;
        u32x4 arh[T::WR], arl[T::WR], aih[T::WR], ail[T::WR], bh[JOUTER ? 1 : T::WC], bl[JOUTER ? 1 : T::WC];
#pragma unroll
        for (int i = 0; i < T::WR; ++i) {
            const int o = pk_off(wr0 + 16 * i + li, lg);
            const u32x4 h = *reinterpret_cast<const u32x4 *>(&el.a_hi[o]);
            const u32x4 l = *reinterpret_cast<const u32x4 *>(&el.a_lo[o]);
            arh[i] = h ^ 0x80000000u;
            aih[i] = (h >> 16) | (h << 16);
            arl[i] = l ^ 0x80000000u;
            ail[i] = (l >> 16) | (l << 16);
        }
        constexpr int NTERM = split3(SCHEME) ? 3 : 2;
        if constexpr (JOUTER) {
            // B of column tile j+1 is read from LDS while the MFMAs of tile j
            // run: the hi word at the top of tile j, the lo word once the
            // (a_hi, b_lo) term of tile j has been issued.  The scheduling
            // barriers keep the compiler from sinking the reads back next to
            // their first use, which left one LDS latency exposed per column
            // tile (headline 256 x 176 BF16X3: 13.35 -> 12.71 ms, FP16X3
            // 13.79 -> 13.36 ms, measured A/B, identical results).
            u32x4 yh = *reinterpret_cast<const u32x4 *>(&el.b_hi[pk_off(wc0 + li, lg)]);
            u32x4 yl = yh;
            if constexpr (split3(SCHEME)) yl = *reinterpret_cast<const u32x4 *>(&el.b_lo[pk_off(wc0 + li, lg)]);
#pragma unroll
            for (int j = 0; j < T::WC; ++j) {
                const int on = pk_off(wc0 + 16 * (j + 1 < T::WC ? j + 1 : j) + li, lg);
                u32x4 yh_n = yh, yl_n = yl;
                if (j + 1 < T::WC) yh_n = *reinterpret_cast<const u32x4 *>(&el.b_hi[on]);
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int t = 0; t < NTERM; ++t) {
                    if (split3(SCHEME) && t == 2 && j + 1 < T::WC) {
                        yl_n = *reinterpret_cast<const u32x4 *>(&el.b_lo[on]);
                        __builtin_amdgcn_sched_barrier(0);
                    }
#pragma unroll
                    for (int i = 0; i < T::WR; ++i) {
                        const bool alo = (t == NTERM - 1);
                        const u32x4 xr = alo ? arl[i] : arh[i];
                        const u32x4 xi = alo ? ail[i] : aih[i];
                        const u32x4 y = (split3(SCHEME) && t == 1) ? yl : yh;
                        if constexpr (!f16ops(SCHEME)) {
                            accR[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, xr), __builtin_bit_cast(bf16x8, y), accR[i][j], 0, 0, 0);
                            accI[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, xi), __builtin_bit_cast(bf16x8, y), accI[i][j], 0, 0, 0);
                        } else {
                            accR[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, xr), __builtin_bit_cast(f16x8, y), accR[i][j], 0, 0, 0);
                            accI[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, xi), __builtin_bit_cast(f16x8, y), accI[i][j], 0, 0, 0);
                        }
                    }
                }
                yh = yh_n;
                yl = yl_n;
            }
            return;
        }
#pragma unroll
        for (int j = 0; j < (JOUTER ? 0 : T::WC); ++j) {
            const int o = pk_off(wc0 + 16 * j + li, lg);
            bh[j] = *reinterpret_cast<const u32x4 *>(&el.b_hi[o]);
            if constexpr (split3(SCHEME)) bl[j] = *reinterpret_cast<const u32x4 *>(&el.b_lo[o]);
        }
#pragma unroll
        for (int t = 0; t < NTERM; ++t) {
#pragma unroll
            for (int j = 0; j < (JOUTER ? 0 : T::WC); ++j)
#pragma unroll
                for (int i = 0; i < T::WR; ++i) {
                    // terms: (a_hi, b_hi), (a_hi, b_lo) [three-product schemes], (a_lo, b_hi)
                    const bool alo = (t == NTERM - 1);
                    const u32x4 xr = alo ? arl[i] : arh[i];
                    const u32x4 xi = alo ? ail[i] : aih[i];
                    const u32x4 y = (split3(SCHEME) && t == 1) ? bl[j] : bh[j];
                    if constexpr (!f16ops(SCHEME)) {
                        accR[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, xr), __builtin_bit_cast(bf16x8, y), accR[i][j], 0, 0, 0);
                        accI[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, xi), __builtin_bit_cast(bf16x8, y), accI[i][j], 0, 0, 0);
                    } else {
                        accR[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, xr), __builtin_bit_cast(f16x8, y), accR[i][j], 0, 0, 0);
                        accI[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, xi), __builtin_bit_cast(f16x8, y), accI[i][j], 0, 0, 0);
                    }
                }
        }
    }
}

// FP32D: the MFMAs of one stage with FP64 partial sums.  v_mfma_f32_16x16x4_f32
// adds each of its 4 products to the accumulator with its own rounding, so
// a chain's error grows with the size of its running sum; here a chain
// covers two k-steps (4 atoms, 8 products per output) and is then added
// into FP64 sums.  On coherent sums (the Bragg pixels that the <F>
// subtraction and the cross-lot variance cancel) the error of a slice
// stays near one rounding of a 4-atom sum, not of the whole slice.
template <class T>
__device__ __forceinline__ void mfma_stage_d(const Elems<SCH_FP32D, T> &el, f32x4 (&accR)[T::WR][T::WC],
                                             f32x4 (&accI)[T::WR][T::WC], double (&sR)[T::WR][T::WC][4],
                                             double (&sI)[T::WR][T::WC][4], int wr0, int wc0) {
    const int lane = threadIdx.x & 63, li = lane & 15, lg = lane >> 4;
#pragma unroll
    for (int ks = 0; ks < KS / 2; ++ks) {
        float av[T::WR], bR[T::WC], bI[T::WC];
#pragma unroll
        for (int i = 0; i < T::WR; ++i) av[i] = el.a[(wr0 + 16 * i + li) * FST + 4 * ks + lg];
        const int atom = 2 * ks + (lane >> 5);
        const bool im = (lg & 1) != 0;
#pragma unroll
        for (int j = 0; j < T::WC; ++j) {
            const float2 b = *reinterpret_cast<const float2 *>(&el.b[(wc0 + 16 * j + li) * FST + 2 * atom]);
            bR[j] = im ? -b.y : b.x;
            bI[j] = im ? b.x : b.y;
        }
        const bool first = (ks & 1) == 0;
#pragma unroll
        for (int j = 0; j < T::WC; ++j)
#pragma unroll
            for (int i = 0; i < T::WR; ++i) {
                const f32x4 zr = first ? f32x4{0.f, 0.f, 0.f, 0.f} : accR[i][j];
                const f32x4 zi = first ? f32x4{0.f, 0.f, 0.f, 0.f} : accI[i][j];
                accR[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[i], bR[j], zr, 0, 0, 0);
                accI[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[i], bI[j], zi, 0, 0, 0);
            }
        if (!first) {
#pragma unroll
            for (int j = 0; j < T::WC; ++j)
#pragma unroll
                for (int i = 0; i < T::WR; ++i)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        sR[i][j][v] += (double)accR[i][j][v];
                        sI[i][j][v] += (double)accI[i][j][v];
                    }
        }
    }
}

template <int SCHEME, class T, bool DAMP>
__global__ void __launch_bounds__(T::NT)
phase_gemm_kernel(const AtomRec *__restrict__ rec, const int *__restrict__ slice_beg, float2 *__restrict__ partial,
                  const float *__restrict__ sq, const int *__restrict__ slice_sq, GemmGeom g, int slice0,
                  int nitems) {
    __shared__ Smem<SCHEME, T> sm;
    const int nct = (g.C + T::CT - 1) / T::CT, nrt = (g.R + T::RT - 1) / T::RT;
    const bool padh = !DAMP && g.pad_hash != 0;   // rotation hash read from the records
    // work items (column tile fastest, then row tile, then slice); a launch
    // with fewer work-groups than items walks them with a grid stride.  The
    // slices are taken in g.order when given (longest first, so that uneven
    // slices leave no long item for the last round); planes stay in slice order
    // One work-group per item and no launch order: the blocks are dispatched
    // to the 8 XCDs round robin (block b to XCD b % 8), so block b takes an
    // item of a contiguous chunk per XCD and the row tiles of a slice share one
    // XCD's L2 for the slice's atom records (headline FP16X3: HBM reads 225 ->
    // 72 MB per launch, GEMM -0.3%, measured A/B, identical results).
    int item0 = blockIdx.x;
    if (!g.queue && !g.order && (int)gridDim.x == nitems) {
        constexpr int NX = 8;
        const int q = nitems / NX, r = nitems % NX, x = blockIdx.x % NX, k = blockIdx.x / NX;
        item0 = (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + k;
    }
    // with a work-item counter every work-group takes the next item when it
    // is done with one (uniform across the work-group through LDS)
    __shared__ int s_item;
    auto next_item = [&](int cur) -> int {
        if (!g.queue) return cur + (int)gridDim.x;
        __syncthreads();
        if (threadIdx.x == 0) s_item = atomicAdd(g.queue, 1);
        __syncthreads();
        return s_item;
    };
    for (int item = g.queue ? next_item(0) : item0; item < nitems; item = next_item(item)) {
        const int si = item / (nct * nrt);
        const int slice = slice0 + (g.order ? g.order[si] : si);
        const int kbeg = slice_beg[slice], kend = slice_beg[slice + 1];
        const int row0 = ((item / nct) % nrt) * T::RT, col0 = (item % nct) * T::CT;
        const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        const int wr0 = (wave % T::NWR) * 16 * T::WR, wc0 = (wave / T::NWR) * 16 * T::WC;
        const bool live = g.R - row0 - wr0 > 0 && g.C - col0 - wc0 > 0;   // this wave's block touches the grid
        const int pt = threadIdx.x;

        Side sa{row0, row0 % g.R0, row0 / g.R0, g.R0};
        Side sb{col0, col0 % g.C0, col0 / g.C0, g.C0};
        // wrap bookkeeping of the 16-groups of the tile (periods >= 16 when a side wraps)
        for (int h = threadIdx.x; h < T::RH + T::CH; h += T::NT) {
            const bool isa = h < T::RH;
            const Side &s = isa ? sa : sb;
            const int hh = isa ? h : h - T::RH;
            const int x = s.a0 + 16 * hh;
            (isa ? sm.a_k16 : sm.b_k16)[hh] = x / s.period;
            (isa ? sm.a_wpos : sm.b_wpos)[hh] = s.period - (x % s.period);   // first wrap inside the group
        }
        if constexpr (DAMP) {
            // sq = [s^2 of the R rows | s^2 of the C columns] of the slice's
            // metric (slice_sq: offset per slice); padding reads 0
            const float *sqs = sq + (slice_sq ? slice_sq[slice] : 0);
            for (int x = threadIdx.x; x < T::RT + T::CT; x += T::NT) {
                const bool isa = x < T::RT;
                const int o = isa ? row0 + x : col0 + x - T::RT;
                const int lim = isa ? g.R : g.C;
                (isa ? sm.s_a[x] : sm.s_b[x - T::RT]) = o < lim ? sqs[(isa ? 0 : g.R) + o] : 0.0f;
            }
        }

        f32x4 accR[T::WR][T::WC], accI[T::WR][T::WC];
        // FP32 scheme: two-level accumulation.  The 16x16x4 MFMA rounds its
        // FP32 accumulator after every product; on coherent sums (Bragg peaks
        // of displaced crystals) that error grows with the size of the
        // partial sums, so the MFMAs accumulate 64 atoms (4 stages) at a time
        // and the groups are added into a second set of FP32 sums.  (A flush
        // after every stage measured slower, 60.0 vs 54.6 ms on the 161^3
        // headline, and less accurate on its coherent sums, 3.5e-7 vs 3.9e-8:
        // the second level then adds 64 nearly equal stage sums per slice,
        // whose rounding no longer averages out.)
        constexpr bool TWO = SCHEME == SCH_FP32;
        constexpr int SR = TWO ? T::WR : 1, SC = TWO ? T::WC : 1;
        f32x4 sumR[SR][SC], sumI[SR][SC];
        // FP32D: FP64 partial sums (see mfma_stage_d)
        constexpr bool D64 = SCHEME == SCH_FP32D;
        constexpr int DR = D64 ? T::WR : 1, DC = D64 ? T::WC : 1;
        double dR[DR][DC][4], dI[DR][DC][4];
#pragma unroll
        for (int i = 0; i < DR; ++i)
#pragma unroll
            for (int j = 0; j < DC; ++j)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    dR[i][j][v] = 0.0;
                    dI[i][j][v] = 0.0;
                }
#pragma unroll
        for (int i = 0; i < T::WR; ++i)
#pragma unroll
            for (int j = 0; j < T::WC; ++j) {
                accR[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
                accI[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
            }
#pragma unroll
        for (int i = 0; i < SR; ++i)
#pragma unroll
            for (int j = 0; j < SC; ++j) {
                sumR[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
                sumI[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
            }

        const int nst = (kend - kbeg + KS - 1) / KS;
        // records of this thread's atom for the stages whose tables come next,
        // loaded ahead of use so the global-memory latency is hidden
        const int at = pt & (KS - 1);
        auto fetch = [&](int st) {
            const int k = kbeg + st * KS + at;
            return rec[k < kend ? k : kend - 1];
        };
        // wide column blocks are register-bound: they load the records of a
        // stage's tables at the end of the stage before (after its last
        // spill reload, ahead of the barrier), the others three stages
        // ahead.  Reloads of spilled registers wait for every outstanding
        // vector-memory load (vmcnt(0)); a record load issued at the top of
        // the stage made the first reload of the element builder wait for it
        // (headline 256 x 176 FP16X3 GEMM 10.63 -> 10.46 ms, population
        // 2.61 -> 2.59 ms, 4 interleaved A/B pairs, identical results)
        constexpr bool DEEP = T::WC <= 4 && !D64;
        const AtomRec r0 = fetch(0);
        const AtomRec r1 = fetch(nst > 1 ? 1 : 0);
        AtomRec rn = fetch(nst > 2 ? 2 : 0);
        // FP16 schemes: B amplitudes of the slice scaled by 2^-e, max |w| = f 2^e
        // with f in [0.5, 1); the epilogue multiplies by 2^e (both exact)
        float wscale = 1.0f, wunscale = 1.0f;
        if constexpr (f16ops(SCHEME)) {
            float m = 0.0f;
            for (int k = kbeg + pt; k < kend; k += T::NT) m = fmaxf(m, fabsf(rec[k].w));
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
            if ((pt & 63) == 0) sm.wmax[wave] = m;
        }
        __syncthreads();
        if constexpr (f16ops(SCHEME)) {
            float m = 0.0f;
#pragma unroll
            for (int w = 0; w < T::NW; ++w) m = fmaxf(m, sm.wmax[w]);
            if (m > 0.0f && m <= 3.0e38f) {
                int e = 0;
                (void)frexpf(m, &e);
                wscale = ldexpf(1.0f, -e);
                wunscale = ldexpf(1.0f, e);
            }
        }
        build_tables<T, !fp32ops(SCHEME)>(sm.t[0], sm.a_k16, sm.b_k16, r0, kbeg + at < kend, sa, sb, pt, wscale, kbeg + at, g.dither, padh);
        __syncthreads();
        const bool awrap = g.R0 < g.R, bwrap = g.C0 < g.C;
        build_elems<SCHEME, T, DAMP>(sm.e[0], sm.t[0], sm.a_wpos, sm.b_wpos, sm.s_a, sm.s_b, wave, awrap, bwrap);
        if (nst > 1) build_tables<T, !fp32ops(SCHEME)>(sm.t[1], sm.a_k16, sm.b_k16, r1, kbeg + KS + at < kend, sa, sb, pt, wscale, kbeg + KS + at, g.dither, padh);
        __syncthreads();
        for (int s = 0; s < nst; ++s) {
            const int cur = s & 1, nxt = cur ^ 1;
            AtomRec rt;
            if constexpr (DEEP) {
                rt = rn;
                if (s + 3 < nst) rn = fetch(s + 3);
            } else {
                rt = rn;
            }
            // The three steps of a stage touch disjoint buffers (elements of
            // s+1 from tables of s+1, tables of s+2, MFMAs on elements of s),
            // so their order is free; the tables come before the elements
            // (headline FP16X3 GEMM -0.6%, population -1.3%, measured A/B,
            // identical results).  Three-product schemes: waves 4..7 issue
            // the MFMAs first, so the two waves sharing a SIMD (w and w+4)
            // overlap one's operand generation with the other's matrix work
            // instead of both generating, then both multiplying.  Measured on
            // the headline's 256 x 176 tile: BF16X3 2% faster, FP16X3 1-1.6%
            // faster (with B read one column tile ahead), FP16X2 unchanged,
            // FP32 2% slower; on the narrower tiles the second copy of the
            // MFMA block costs registers.  So BF16X3 and FP16X3 on the
            // wide-column tiles.
            const bool mfma_first = split3(SCHEME) && T::WC > 4 && ((wave >> 2) & 1) != 0;
            // Split schemes raise the wave's issue priority while it issues
            // its MFMAs, so that the SIMD's other wave, generating operands,
            // does not delay them (population GEMM, 256 x 128 damped FP16X3:
            // 2.665 -> 2.617 ms, 6 interleaved A/B pairs; headline unchanged)
            constexpr bool PRIO = !fp32ops(SCHEME);
            if (mfma_first && live) {
                if constexpr (PRIO) __builtin_amdgcn_s_setprio(1);
                mfma_stage<SCHEME, T>(sm.e[cur], accR, accI, wr0, wc0);
                if constexpr (PRIO) __builtin_amdgcn_s_setprio(0);
            }
            if (s + 2 < nst) build_tables<T, !fp32ops(SCHEME)>(sm.t[cur], sm.a_k16, sm.b_k16, rt, kbeg + (s + 2) * KS + at < kend, sa, sb, pt, wscale,
                                              kbeg + (s + 2) * KS + at, g.dither, padh);
            if (s + 1 < nst) build_elems<SCHEME, T, DAMP>(sm.e[nxt], sm.t[nxt], sm.a_wpos, sm.b_wpos, sm.s_a, sm.s_b, wave, awrap, bwrap);
            if (!mfma_first && live) {
                if constexpr (D64) {
                    mfma_stage_d<T>(sm.e[cur], accR, accI, dR, dI, wr0, wc0);
                } else {
                    if constexpr (PRIO) __builtin_amdgcn_s_setprio(1);
                    mfma_stage<SCHEME, T>(sm.e[cur], accR, accI, wr0, wc0);
                    if constexpr (PRIO) __builtin_amdgcn_s_setprio(0);
                }
            }
            if constexpr (TWO) {
                if ((s & 3) == 3 || s + 1 == nst) {
#pragma unroll
                    for (int i = 0; i < SR; ++i)
#pragma unroll
                        for (int j = 0; j < SC; ++j) {
                            sumR[i][j] += accR[i][j];
                            sumI[i][j] += accI[i][j];
                            accR[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
                            accI[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
                        }
                }
            }
            if constexpr (!DEEP) rn = fetch(s + 3 < nst ? s + 3 : 0);
            __syncthreads();
        }
        if constexpr (TWO) {
#pragma unroll
            for (int i = 0; i < SR; ++i)
#pragma unroll
                for (int j = 0; j < SC; ++j) {
                    accR[i][j] = sumR[i][j];
                    accI[i][j] = sumI[i][j];
                }
        }

        // epilogue: lane holds rows 4*lg + v (v = 0..3) of column li of each 16x16 tile.
        // With an even plane stride a lane's four rows are 32 contiguous,
        // 16-byte aligned bytes, stored as two 16-byte writes (population
        // GEMM -1.2%, step -3.5%, headline -0.7%, measured A/B, identical
        // results); the last, partial row group and odd strides store by row.
        const int lane = threadIdx.x & 63, li = lane & 15, lg = lane >> 4;
        // FP32D writes each slice as two planes, hi = fl(s) and lo = fl(s - hi),
        // which the FP64 reductions add like two slices of the same type
        float2 *out = partial + (long long)(slice - slice0) * g.P * (D64 ? 2 : 1);
        const bool wide = !D64 && (g.Rs & 1) == 0;
#pragma unroll
        for (int i = 0; i < T::WR; ++i)
#pragma unroll
            for (int j = 0; j < T::WC; ++j) {
                const int c = col0 + wc0 + 16 * j + li;
                if (c >= g.C) continue;
                if constexpr (!D64) {
                    const int r4 = row0 + wr0 + 16 * i + 4 * lg;
                    if (wide && r4 + 3 < g.R) {
                        const float s = f16ops(SCHEME) ? wunscale : 1.0f;
                        float4 *p = reinterpret_cast<float4 *>(out + r4 + (long long)g.Rs * c);
                        p[0] = make_float4(accR[i][j][0] * s, accI[i][j][0] * s, accR[i][j][1] * s, accI[i][j][1] * s);
                        p[1] = make_float4(accR[i][j][2] * s, accI[i][j][2] * s, accR[i][j][3] * s, accI[i][j][3] * s);
                        continue;
                    }
                }
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = row0 + wr0 + 16 * i + 4 * lg + v;
                    if (r < g.R) {
                        const long long o = r + (long long)g.Rs * c;
                        if constexpr (D64) {
                            const double xr = dR[i][j][v], xi = dI[i][j][v];
                            const float hr = (float)xr, hi = (float)xi;
                            out[o] = make_float2(hr, hi);
                            out[o + g.P] = make_float2((float)(xr - (double)hr), (float)(xi - (double)hi));
                        } else if constexpr (f16ops(SCHEME)) {
                            out[o] = make_float2(accR[i][j][v] * wunscale, accI[i][j][v] * wunscale);
                        } else {
                            out[o] = make_float2(accR[i][j][v], accI[i][j][v]);
                        }
                    }
                }
            }
    }
}

// ---------------------------------------------------------------------------
// Tile configurations (WR x WC MFMA tiles per wave, NWR x NWC waves).
using CfgA = TileCfg<4, 4, 2, 4>;   // 128 x 256, 8 waves
using CfgB = TileCfg<3, 4, 2, 4>;   //  96 x 256
using CfgC = TileCfg<2, 4, 2, 4>;   //  64 x 256
using CfgD = TileCfg<4, 4, 2, 2>;   // 128 x 128, 4 waves
using CfgE = TileCfg<4, 2, 2, 4>;   // 128 x 128, 8 waves
using CfgF = TileCfg<1, 4, 1, 4>;   //  16 x 256 (lines / few rows)
using CfgG = TileCfg<2, 11, 8, 1>;  // 256 x 176 (flattened rows x one short axis, e.g. 161)
using CfgH = TileCfg<1, 11, 8, 1>;  // 128 x 176
using CfgI = TileCfg<2, 8, 8, 1>;   // 256 x 128 (column counts that are multiples of 128, e.g. 256^2 layers)

// split_only: tiles whose FP32 variant (two accumulator sets) spills so much
// that it is slower than the 96x256 tile (which spills a little and is still
// the fastest FP32 tile, measured); the planner does not offer them to FP32.
// eff: split-scheme time per padded (output x atom) relative to 256 x 176,
// measured with every tile forced (BF16X3 and FP16X3, MI355X) on 3D blocks of
// 101^3 and 121^3 points (256k atoms), 8 layers of 401^2 with 64 lots and the
// 161^3 block: 128x256 1.08-1.27, 96x256 1.05-1.32, 64x256 1.57-1.70,
// 128x128/4w 2.14-2.43, 128x128/8w 1.54-1.76, 16x256 6.2-6.6, 128x176
// 1.24-1.39.  The smaller tiles generate more operand elements per product
// and have more work items per launch, each with its own set-up.
static const TileInfo kTiles[] = {
    {CfgA::RT, CfgA::CT, CfgA::NT, 1, 1.17f, 0, 0}, {CfgB::RT, CfgB::CT, CfgB::NT, 0, 1.28f, 0, 0},
    {CfgC::RT, CfgC::CT, CfgC::NT, 0, 1.62f, 1, 0}, {CfgD::RT, CfgD::CT, CfgD::NT, 1, 2.30f, 0, 0},
    {CfgE::RT, CfgE::CT, CfgE::NT, 0, 1.65f, 1, 0}, {CfgF::RT, CfgF::CT, CfgF::NT, 0, 6.30f, 1, 0},
    {CfgG::RT, CfgG::CT, CfgG::NT, 1, 1.00f, 0, 0}, {CfgH::RT, CfgH::CT, CfgH::NT, 1, 1.33f, 0, 0},
    {CfgI::RT, CfgI::CT, CfgI::NT, 1, 1.10f, 0, 1},
};

// FP32D keeps 2 x WR x WC x 4 FP64 sums per lane beside the FP32 chains: only
// the tiles with at most 8 MFMA blocks per wave (64 x 256, 128 x 128 / 8
// waves, 16 x 256) fit the register file without spilling.
template <int SCHEME, class T>
constexpr bool tile_built() {
    if constexpr (std::is_same_v<T, CfgI>) return SCHEME != SCH_FP32 && SCHEME != SCH_FP32D;   // split schemes only
    return SCHEME != SCH_FP32D || T::WR * T::WC <= 8;
}

int gemm_num_tiles() { return (int)(sizeof(kTiles) / sizeof(kTiles[0])); }
TileInfo gemm_tile_info(int t) { return kTiles[t]; }

template <int SCHEME, class T>
static void launch_one(const AtomRec *rec, const int *slice_beg, float2 *partial, const float *sq, const int *slice_sq,
                       const GemmGeom &g, int slice0, int nslice, int nwg, hipStream_t st) {
    const long long items = (long long)((g.C + T::CT - 1) / T::CT) * ((g.R + T::RT - 1) / T::RT) * nslice;
    const int grid = (int)(nwg > 0 && nwg < items ? nwg : items);
    if (sq)
        phase_gemm_kernel<SCHEME, T, true><<<grid, dim3(T::NT), 0, st>>>(rec, slice_beg, partial, sq, slice_sq, g, slice0, (int)items);
    else
        phase_gemm_kernel<SCHEME, T, false><<<grid, dim3(T::NT), 0, st>>>(rec, slice_beg, partial, sq, slice_sq, g, slice0, (int)items);
}

template <int SCHEME, class T>
static hipError_t launch_if_built(const AtomRec *rec, const int *slice_beg, float2 *partial, const float *sq,
                                  const int *slice_sq, const GemmGeom &g, int slice0, int nslice, int nwg,
                                  hipStream_t st) {
    if constexpr (tile_built<SCHEME, T>()) {
        launch_one<SCHEME, T>(rec, slice_beg, partial, sq, slice_sq, g, slice0, nslice, nwg, st);
        return hipGetLastError();
    } else {
        return hipErrorInvalidValue;
    }
}

template <int SCHEME>
static hipError_t launch_scheme(int tile, const AtomRec *rec, const int *slice_beg, float2 *partial, const float *sq,
                                const int *slice_sq, const GemmGeom &g, int slice0, int nslice, int nwg,
                                hipStream_t st) {
    switch (tile) {
    case 0: return launch_if_built<SCHEME, CfgA>(rec, slice_beg, partial, sq, slice_sq, g, slice0, nslice, nwg, st);
    case 1: return launch_if_built<SCHEME, CfgB>(rec, slice_beg, partial, sq, slice_sq, g, slice0, nslice, nwg, st);
    case 2: return launch_if_built<SCHEME, CfgC>(rec, slice_beg, partial, sq, slice_sq, g, slice0, nslice, nwg, st);
    case 3: return launch_if_built<SCHEME, CfgD>(rec, slice_beg, partial, sq, slice_sq, g, slice0, nslice, nwg, st);
    case 4: return launch_if_built<SCHEME, CfgE>(rec, slice_beg, partial, sq, slice_sq, g, slice0, nslice, nwg, st);
    case 5: return launch_if_built<SCHEME, CfgF>(rec, slice_beg, partial, sq, slice_sq, g, slice0, nslice, nwg, st);
    case 6: return launch_if_built<SCHEME, CfgG>(rec, slice_beg, partial, sq, slice_sq, g, slice0, nslice, nwg, st);
    case 7: return launch_if_built<SCHEME, CfgH>(rec, slice_beg, partial, sq, slice_sq, g, slice0, nslice, nwg, st);
    case 8: return launch_if_built<SCHEME, CfgI>(rec, slice_beg, partial, sq, slice_sq, g, slice0, nslice, nwg, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_phase_gemm(int scheme, int tile, const AtomRec *rec, const int *slice_beg, float2 *partial,
                             const float *sq, const GemmGeom &g, int slice0, int nslice, hipStream_t st, int nwg,
                             const int *slice_sq) {
    // the two-level operand tables need wrap periods of at least 16
    if (g.R0 < 16 && g.R0 != g.R) return hipErrorInvalidValue;
    if (g.C0 < 16 && g.C0 != g.C) return hipErrorInvalidValue;
    if (nslice <= 0) return hipSuccess;
    if ((long long)((g.C + 15) / 16) * ((g.R + 15) / 16) * nslice > INT32_MAX) return hipErrorInvalidValue;
    switch (scheme) {
    case SCH_FP32: return launch_scheme<SCH_FP32>(tile, rec, slice_beg, partial, sq, slice_sq, g, slice0, nslice, nwg, st);
    case SCH_BF16X3: return launch_scheme<SCH_BF16X3>(tile, rec, slice_beg, partial, sq, slice_sq, g, slice0, nslice, nwg, st);
    case SCH_FP16X2: return launch_scheme<SCH_FP16X2>(tile, rec, slice_beg, partial, sq, slice_sq, g, slice0, nslice, nwg, st);
    case SCH_FP16X3: return launch_scheme<SCH_FP16X3>(tile, rec, slice_beg, partial, sq, slice_sq, g, slice0, nslice, nwg, st);
    case SCH_FP32D: return launch_scheme<SCH_FP32D>(tile, rec, slice_beg, partial, sq, slice_sq, g, slice0, nslice, nwg, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace dc
