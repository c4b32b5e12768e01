// dc_internal.h -- kernel argument blocks and launchers shared between the
// kernel translation units and the C-ABI layer.
#pragma once
#include "dc_common.h"

namespace dc {

constexpr int PREP_CORNER = 0;
constexpr int PREP_LMN = 1;

struct PrepArgs {
    const double *src;      // positions: (3,N) rows, or xat(ld,3) when src_ld > 0
    long long src_ld;
    const int *idx;         // gathered atom index (null: identity)
    const int *oid;         // offset id per gathered atom (null: none)
    const int4 *runs;       // (first, src, oid, -) runs of gathered atoms: atom first + q takes
                            // idx[src + q] and offset oid (null: idx/oid per atom); by first
    int nruns;
    const double *offs;     // offset table (3 per id)
    const float *w;         // per-atom amplitude (null: 1)
    const float *bdw;       // per-atom isotropic B in AtomRec::pad (null: 0)
    const double *occ_src;  // caller-order occupancy read through idx (overrides w)
    const double *bdw_src;  // caller-order B read through idx (overrides bdw)
    uint32_t hash_salt;     // nonzero: AtomRec::pad = dither_hash32(record, salt) instead of B
                            // (records of undamped GEMMs only; the adjoints read pad as B)
    long long n;
    int mode;               // PREP_CORNER / PREP_LMN
    int mapping;            // 0: rows = abscissa ; 1: rows = (abscissa, ordinate)
    double vec[18];         // uin, vin, win, off_shift(:,1..3)
    double h0[3];           // corner (PREP_CORNER)
    int lmn[6];             // PREP_LMN
};

// Reciprocal grid: full size n (num), computed region f (fnum), corner and
// increments, metric; P = f0*f1*f2 pixels of the computed region.
struct Grid {
    int n[3];
    int f[3];
    double xm[3], uin[3], vin[3], win[3];
    double rten[9];
    long long P;           // pixels of the computed region (compact)
    long long ps[3];       // strides of (i, j, l) inside a GEMM slice plane
    long long plane;       // elements of one slice plane (R * C of the GEMM)
    int friedel;
};

// sin(theta)/lambda bin, four_stltab's expression (fourier_sup.f90:2663-2674,
// skalpro metric.f90:90-105): h(k) is summed in FP64 and rounded to FP32 by
// REAL() of default kind before the metric is applied in FP64; evaluated
// left to right with no contraction so the bins equal the host restatement
// bit for bit.
__host__ __device__ inline double stl_h(double xm, double u, double v, double w, int i, int j, int l) {
#pragma clang fp contract(off)
    return (double)(float)(xm + u * (double)i + v * (double)j + w * (double)l);
}
// the bin as four_stltab stores it (NINT, no clamp; values beyond 2^30 saturate)
__device__ inline int istl_raw(const Grid &g, int i, int j, int l) {
#pragma clang fp contract(off)
    double h[3], t[3];
    for (int k = 0; k < 3; ++k) h[k] = stl_h(g.xm[k], g.uin[k], g.vin[k], g.win[k], i, j, l);
    for (int r = 0; r < 3; ++r) t[r] = g.rten[r] * h[0] + g.rten[r + 3] * h[1] + g.rten[r + 6] * h[2];
    double q2 = (h[0] * t[0] + h[1] * t[1] + h[2] * t[2]) / 4.0;
    double v = __builtin_sqrt(q2) * (1.0 / CFINC);
    return (int)__builtin_fmin(__builtin_round(v), 1073741824.0);
}
// the bin used to index cfact (the run checks the grid's largest bin first)
__device__ inline int istl_point(const Grid &g, int i, int j, int l) {
    const int b = istl_raw(g, i, j, l);
    return b < 0 ? 0 : (b > CFPKT ? CFPKT : b);
}

struct TileInfo {
    int rt, ct, nt;
    int split_only;         // tile offered to the split schemes only (its FP32 variant spills)
    float eff;              // split schemes: time per padded output x atom, relative to 256 x 176
    int acc64;              // built for SCHEME_FP32_ACC64
    int no_fp32;            // no FP32 variant built (split schemes only)
};

// GEMM schemes: DC_PREC_FP32 .. DC_PREC_FP16X3 (0..3) and the FP32 scheme
// with FP64 partial sums, which writes two planes (hi, lo) per slice.
constexpr int SCHEME_FP32_ACC64 = 4;
inline int planes_per_slice(int scheme) { return scheme == SCHEME_FP32_ACC64 ? 2 : 1; }

struct ReduceArgs {
    Grid g;
    const float2 *partial;  // slices [slice0, ...) planes of g.P
    int slice0;
    const int *lot_slice;   // nlots+1 slice boundaries (absolute slice ids)
    const int *slice_type;  // 0-based type (cfact column) per slice
    int nlots;
    int lot0;               // lots already accumulated (Welford count)
    int lform;
    const double2 *cfact;   // (ncfact, nscat)
    int ncfact;
    int ntype;              // cfact columns (slice types < ntype); 0 = not given
    const int *istl;        // optional bins (full-grid layout), null = compute
    const double *acsf;     // compact region, null = zero
    double *dsi, *csf, *wm, *wv;
    int fresh;              // dsi, wm, wv start from zero (not read)
};

struct FinalArgs {
    Grid g;
    long long ntot;
    const double *csf, *acsf, *dsi, *wm, *wv;   // compact region
    double *csf_out, *acsf_out, *dsi_out, *fmean_out, *dvar_out;  // full grid
    double sq, dscale, vscale;
};

// sq (R + C floats: s^2 parts of rows and columns) selects the variant with a
// per-atom displacement factor exp(-B_n s^2), B_n in AtomRec::pad; null = none.
// nwg > 0 launches that many persistent work-groups that stride over the
// (slice, tile) work items; 0 = one work-group per item.  slice_sq (absolute
// slice ids) offsets sq per slice when slices use different metrics.
hipError_t launch_phase_gemm(int scheme, int tile, const AtomRec *rec, const int *slice_beg, float2 *partial,
                             const float *sq, const GemmGeom &g, int slice0, int nslice, hipStream_t st,
                             int nwg = 0, const int *slice_sq = nullptr);
int gemm_num_tiles();
TileInfo gemm_tile_info(int t);
hipError_t launch_prep(const PrepArgs &a, AtomRec *out, hipStream_t st);
hipError_t launch_reduce_lots(const ReduceArgs &a, hipStream_t st);
hipError_t launch_reduce_sum(const ReduceArgs &a, double *out, int accumulate, hipStream_t st);
hipError_t launch_aver_combine(const Grid &g, double *acsf, const double *lat, double norm, hipStream_t st);
hipError_t launch_finalize(const FinalArgs &a, hipStream_t st);
hipError_t launch_strucf_out(const ReduceArgs &a, double *tcsf, long long ntot, int s0, int s1, int accumulate,
                             hipStream_t st);
hipError_t launch_istl(const Grid &g, int *out, hipStream_t st);

// ---- refinement model (dc_adjoint.hip) --------------------------------------
// Adjoint of the phase GEMM.  With Q_t(p) = 2 (dL/dI_p) conj(F_p) f_t(p) and
// E_n(p) = exp(-B_n s_p^2) exp(2 pi i h_p.x_n), every atom n of type t needs
//   S0 = sum_p Q E,  S_r = sum_p (r - rc) Q E,  S_j, S_m (column indices),
//   S_s = sum_p s_p^2 Q E;
// the kernel writes their per-column-tile partial sums
// [Re S0, Im S0, Im S_r, Im S_j, Im S_m, Re S_s] to part[tile][atom][6].
struct AdjArgs {
    const AtomRec *rec;     // type-grouped atoms (mapping 0 records, B_n in pad)
    const int4 *tiles;      // atom tiles: (first, end, Q plane, member), each inside one type
    const float2 *q;        // Q planes: q[t * plane + c + C * r] (column fastest)
    const float *sq;        // s^2 parts: R rows, then C columns (per member: + tile.w * sq_stride)
    long long sq_stride;
    float *part;            // [partial set][natoms][6]: one set per column tile or column range
    const uint32_t *qs;     // split scheme: Q as BF16 hi/lo words, per plane [hi|lo][row group][C][4 rows]
    int Rg;                 // row groups of 4 rows in qs (R padded to a multiple of 16)
    int R, C, C0, f1;       // rows, padded columns, column period, valid columns per period
    long long plane;        // R * C
    int natoms;
    float rc, jc, mc;       // centres of the row / column index weights
    int ctpg;               // column tiles per work-group (set by adjoint_plan; 1 = one tile each)
};
struct AdjFinishArgs {
    const float *part;
    int ntile_c, natoms;
    const int *idx;         // sorted atom -> caller index
    const double *occ;      // caller-order occupancies
    double hc[3], v[9];     // h at the weight centres; grid increments vi(:,1..3)
    double *grad;           // (5, caller atoms): d/dx, d/dy, d/dz, d/docc, d/dB
};
struct QArgs {
    Grid g;
    int R, C, C0;
    long long plane;
    const double *f;        // F over the grid (complex, abscissa fastest)
    const double *gi;       // dL/dI per pixel
    const double2 *cfact;   // (ncfact, nscat)
    int ncfact;
    const int *qtype;       // cfact column of each Q plane
    int nq;
    float2 *q;
};
// Batched scoring of a population of models sharing one grid (member m owns
// GEMM slices [mslice[m], mslice[m+1]) and Q planes m*nscat .. m*nscat+nscat-1).
// Per pixel: F_m (FP64) = sum_s cfact_m(type_s, bin_m) partial_s, I = |F|^2,
// G = 2 w (I - T), chi^2 block sums, Q planes 2 G conj(F) cfact_m(t, bin_m).
struct BatchReduceArgs {
    Grid g;                 // shared grid (rten per member below), ps/plane of the GEMM
    int R, C, C0;
    long long plane;        // R * C (Q plane)
    const float2 *partial;
    const int *mslice;      // nmembers + 1
    const int *slice_type;  // 0-based type per slice
    const double2 *cfact;   // (ncfact, nscat, nmembers)
    int ncfact, nscat, nmembers;
    const double *rten;     // (9, nmembers)
    const double *target;   // (P, nmembers)
    const double *weight;   // (P, nmembers) or null
    double *icalc;          // (P, nmembers) or null
    double *chi_part;       // (nblocks, nmembers)
    float2 *q;              // (plane, nscat, nmembers) or null
};
hipError_t launch_batch_reduce(const BatchReduceArgs &a, int nblocks, hipStream_t st);
hipError_t launch_fill_zero(double *p, long long n, hipStream_t st);
// Forward-mode derivative (JVP) of I along ndir parameter directions: slice
// group 5k + q of the GEMM holds direction k weighted by q = (docc,
// occ dx, occ dy, occ dz, occ dB); per pixel
//   dF_k = sum_t f_t(bin) [G0 + 2 pi i h.(G1, G2, G3) - s^2 G4],  dI_k = 2 Re(conj(F) dF_k).
struct JvpArgs {
    Grid g;                 // grid with the GEMM plan's ps / plane
    const float2 *partial;
    const int *dslice;      // 5 ndir + 1 slice boundaries (relative to partial)
    const int *slice_type;  // per slice (relative)
    const double2 *cfact;
    int ncfact;
    const double *f;        // F over the grid (complex)
    int ndir;
    double *di;             // (P, ndir)
    double *df;             // (P, ndir) complex or null
};
hipError_t launch_jvp_reduce(const JvpArgs &a, hipStream_t st);
int adjoint_tile_atoms();
int adjoint_tile_cols();
// Chooses the adjoint kernel's column walk for a.R, a.C and ntiles atom
// tiles (sets a.ctpg) and returns the number of partial sets it writes
// (<= ceil(C / adjoint_tile_cols())), which adjoint_finish sums.  force > 0
// asks for that many column tiles per work-group where the walk applies.
int adjoint_plan(AdjArgs &a, int ntiles, bool split, int force = 0);
// Compute units of the current device (queried once per device; 256 if unknown).
int device_cu_count();
int adjoint_max_walk();
// split = true: BF16x3 MFMA on qs (made by launch_q_split), else FP32 on q.
hipError_t launch_adjoint(const AdjArgs &a, int ntiles, hipStream_t st, bool split = false);
// FP32 complex Q planes (c + C r) -> BF16 hi/lo words for the split adjoint.
hipError_t launch_q_split(const float2 *q, int nq, int R, int C, long long plane, int Rg, uint32_t *qs,
                          hipStream_t st);
hipError_t launch_adjoint_finish(const AdjFinishArgs &a, hipStream_t st);
hipError_t launch_model_q(const QArgs &a, hipStream_t st);
hipError_t launch_chi2_resid(const double *icalc, const double *tgt, const double *wt, long long n, double *gi,
                             double *blocksum, int nblocks, hipStream_t st);

}  // namespace dc
