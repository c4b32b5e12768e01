// dc_common.h -- shared device/host definitions of the phase-GEMM library.
//
// Phase representation.  A phase is a 64-bit unsigned fixed-point number of
// cycles: value/2^64 turns, arithmetic modulo 2^64 == modulo one turn.  This
// is DISCUS's integer-phase idea (four_strucf_OMP_mod.f90:143-170 keeps
// NINT(2^22 frac(v.r)) and indexes a 2^16 table) carried with 64 bits and
// evaluated exactly instead of through the table.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dc {

constexpr int CFPKT = 9999;
constexpr double CFINC = 0.001;

// Per-atom record of one GEMM problem.  The phase of output element
// (row r, col c) is  p0 + phi_row(r) + phi_col(c)  with
//   phi_row(r) = (r % R0) * ra + (r / R0) * ra1
//   phi_col(c) = (c % C0) * ca + (c / C0) * ca1
// (all modulo 2^64) and its amplitude is w.
struct alignas(16) AtomRec {
    uint64_t p0;
    uint64_t ra, ra1;
    uint64_t ca, ca1;
    float w;
    float pad;
};
static_assert(sizeof(AtomRec) == 48, "AtomRec layout");

// GEMM problem geometry.  R rows x C cols per slice, output element
// (r, c) stored at partial[slice * P + r + Rs * c]; the column stride Rs is R
// rounded up to 16 rows where that adds at most 1/16 to the plane, so that
// the 16-row blocks the epilogue writes are whole 128-byte lines.  P = Rs * C.
struct GemmGeom {
    int R, R0, C, C0;
    int Rs;
    long long P;
    uint32_t dither;   // per-atom phase rotation salt (0: none), see dither_turn
    const int *order;  // launch order of the slices (null: slice0, slice0 + 1, ...)
    int *queue;        // work-item counter, zero at launch (null: static assignment)
    int pad_hash;      // split schemes, undamped: AtomRec::pad holds dither_hash32 of the record (prep)
};

// Per-atom phase rotation of the phase GEMM: a pseudo-random turn fraction
// (splitmix64 finaliser of the atom index and a salt).  Row operands carry
// +psi and column operands -psi, so every product is unchanged in exact
// arithmetic, while the rounding of the two operands (sin/cos polynomial,
// table products, the split into hi/lo halves) differs from atom to atom.
// Atoms with equal phases (a crystal at its Bragg peaks) then no longer add
// the same rounding error N times: those errors grow as sqrt(N), not N.
__host__ __device__ inline uint64_t dither_turn(uint32_t k, uint32_t salt) {
    if (salt == 0) return 0;
    uint64_t z = ((uint64_t)salt << 32 | k) + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
// The same rotation from a 32-bit avalanche hash (the murmur3 finaliser) in
// the top 32 bits of the turn: about a third of the VALU work of the 64-bit
// one, which every thread of a GEMM stage evaluates for its atom (split
// schemes: headline GEMM -1.9%).  The FP32 schemes keep the 64-bit hash:
// their results stay as they were, one of their 1e-6 tests sitting within
// 10% of the bar.
__host__ __device__ inline uint32_t dither_hash32(uint32_t k, uint32_t salt) {
    uint32_t h = k * 0x9E3779B1u + salt;
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return h;
}
__host__ __device__ inline uint64_t dither_turn32(uint32_t k, uint32_t salt) {
    return salt == 0 ? 0 : (uint64_t)dither_hash32(k, salt) << 32;
}

// fraction of a FP64 value as 64-bit fixed point (exact for |t| < 2^11:
// every mantissa bit of t - floor(t) lands inside the 64 fraction bits).
__host__ __device__ inline uint64_t frac_to_u64(double t) {
    double f = t - floor(t);
    if (f >= 1.0) f = 0.0;                 // t a tiny negative number: one full turn
    double s = f * 4294967296.0;           // exact (power of two)
    double hi = floor(s);
    double lo = (s - hi) * 4294967296.0;   // exact (Sterbenz), < 2^32
    return ((uint64_t)(uint32_t)hi << 32) | (uint64_t)(uint32_t)lo;
}

// sin(2 pi x/2^32) and cos(2 pi x/2^32) for a 32-bit turn fraction x.
// Integer quadrant reduction (exact), then least-squares fitted odd/even
// polynomials on [-pi/4, pi/4] (coefficients fitted for this library on
// Chebyshev nodes; max error 4e-8 / 7e-8 in float arithmetic).
__device__ __forceinline__ void sincos_turn(uint32_t x, float &c, float &s) {
    uint32_t xs = x + 0x20000000u;                 // shift by 1/8 turn
    uint32_t q = xs >> 30;                          // quadrant 0..3
    int32_t r = (int32_t)(xs & 0x3FFFFFFFu) - 0x20000000;  // [-2^29, 2^29)
    float t = (float)r * 1.4629180792671596e-09f;   // 2 pi / 2^32
    float u = t * t;
    float ps = __builtin_fmaf(u, 2.716169092e-06f, -1.983906404e-04f);
    ps = __builtin_fmaf(u, ps, 8.333328180e-03f);
    ps = __builtin_fmaf(u, ps, -1.666666716e-01f);
    float sn = __builtin_fmaf(t * u, ps, t);
    float pc = __builtin_fmaf(u, 2.439410309e-05f, -1.388680888e-03f);
    pc = __builtin_fmaf(u, pc, 4.166662320e-02f);
    pc = __builtin_fmaf(u, pc, -5.000000000e-01f);
    float cs = __builtin_fmaf(u, pc, 1.0f);
    // rotate by q quarter turns: (c,s) -> q=1: (-s,c) q=2: (-c,-s) q=3: (s,-c)
    float c1 = (q & 1) ? sn : cs;
    float s1 = (q & 1) ? cs : sn;
    uint32_t negc = ((q + 1) & 2) << 30;   // q = 1, 2
    uint32_t negs = (q & 2) << 30;         // q = 2, 3
    c = __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, c1) ^ negc);
    s = __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, s1) ^ negs);
}

}  // namespace dc
