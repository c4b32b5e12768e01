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
    const double *offs;     // offset table (3 per id)
    const float *w;         // per-atom amplitude (null: 1)
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

struct TileInfo {
    int rt, ct, nt;
};

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
    const int *istl;        // optional bins (full-grid layout), null = compute
    const double *acsf;     // compact region, null = zero
    double *dsi, *csf, *wm, *wv;
};

struct FinalArgs {
    Grid g;
    long long ntot;
    const double *csf, *acsf, *dsi, *wm, *wv;   // compact region
    double *csf_out, *acsf_out, *dsi_out, *fmean_out, *dvar_out;  // full grid
    double sq, dscale, vscale;
};

hipError_t launch_phase_gemm(int scheme, int tile, const AtomRec *rec, const int *slice_beg, float2 *partial,
                             const GemmGeom &g, int slice0, int nslice, hipStream_t st);
int gemm_num_tiles();
TileInfo gemm_tile_info(int t);
hipError_t launch_prep(const PrepArgs &a, AtomRec *out, hipStream_t st);
hipError_t launch_reduce_lots(const ReduceArgs &a, hipStream_t st);
hipError_t launch_reduce_sum(const ReduceArgs &a, double *out, int accumulate, hipStream_t st);
hipError_t launch_aver_combine(const Grid &g, double *acsf, const double *lat, double norm, hipStream_t st);
hipError_t launch_finalize(const FinalArgs &a, hipStream_t st);
hipError_t launch_strucf_out(const ReduceArgs &a, double *tcsf, long long ntot, hipStream_t st);
hipError_t launch_istl(const Grid &g, int *out, hipStream_t st);

}  // namespace dc
