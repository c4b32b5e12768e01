// dc_lot.hip -- HBM-bound passes around the phase GEMM:
//   * atom gather + exact phase reduction (per-atom FP64 -> 64-bit fixed point)
//   * per-pixel reduction of GEMM slices with the form factor cfact(istl(h)),
//     lot loop with Bragg subtraction, |F|^2 and FP64 Welford moments
//   * average-structure combination (four_aver) and the Friedel mirror
//   * scatter of the computed region into the Fortran-ordered full grid.
#include "dc_internal.h"

namespace dc {

// ---------------------------------------------------------------------------
// Record preparation.  Position of gathered atom k:
//   x = src[3*ia + c] - offs[o].c,  ia = idx[k], o = oid[k]  (idx/oid/offs may be null)
// or, with runs (the lot gather, four_getatm), ia = idx[run.src + k - run.first] and
// o = run.oid for the run holding k (binary search over the runs' first atoms);
// or, with src_ld > 0, the Fortran xat(ld,3) layout x = src[k + ld*c].
// Phases: p0 from the corner (mode CORNER) or from DISCUS's lmn/off_shift
// decomposition (mode LMN, four_strucf_OMP_mod.f90:135-149,163-165).
__global__ void prep_records_kernel(PrepArgs a, AtomRec *__restrict__ out) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.n) return;
    double x[3];
    long long src = k;     // caller-order index of the atom
    if (a.src_ld > 0) {
        for (int c = 0; c < 3; ++c) x[c] = a.src[k + a.src_ld * c];
    } else {
        long long o = -1;
        if (a.runs) {
            int lo = 0, hi = a.nruns - 1;   // last run with first <= k
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (a.runs[mid].x <= k) lo = mid; else hi = mid - 1;
            }
            const int4 r = a.runs[lo];
            src = a.idx[r.y + (k - r.x)];
            o = r.z;
        } else {
            src = a.idx ? (long long)a.idx[k] : k;
            if (a.oid) o = a.oid[k];
        }
        for (int c = 0; c < 3; ++c) x[c] = a.src[3 * src + c];
        if (o >= 0) {
            const double *of = a.offs + 3 * o;
            for (int c = 0; c < 3; ++c) x[c] -= of[c];
        }
    }
    uint64_t inc[3];
    for (int m = 0; m < 3; ++m)
        inc[m] = frac_to_u64(a.vec[3 * m] * x[0] + a.vec[3 * m + 1] * x[1] + a.vec[3 * m + 2] * x[2]);
    uint64_t p0;
    if (a.mode == PREP_LMN) {
        p0 = 0;
        for (int m = 0; m < 6; ++m) {
            const double *v = a.vec + 3 * m;
            uint64_t f = m < 3 ? inc[m] : frac_to_u64(v[0] * x[0] + v[1] * x[1] + v[2] * x[2]);
            p0 += (uint64_t)(int64_t)a.lmn[m] * f;
        }
    } else {
        p0 = frac_to_u64(a.h0[0] * x[0] + a.h0[1] * x[1] + a.h0[2] * x[2]);
    }
    AtomRec r;
    r.p0 = p0;
    if (a.mapping == 0) {          // rows = abscissa, cols = (ordinate, top)
        r.ra = inc[0]; r.ra1 = 0;
        r.ca = inc[1]; r.ca1 = inc[2];
    } else {                       // rows = (abscissa, ordinate), cols = top
        r.ra = inc[0]; r.ra1 = inc[1];
        r.ca = inc[2]; r.ca1 = 0;
    }
    r.w = a.occ_src ? (float)a.occ_src[src] : (a.w ? a.w[k] : 1.0f);
    r.pad = a.bdw_src ? (float)a.bdw_src[src] : (a.bdw ? a.bdw[k] : 0.0f);
    if (a.hash_salt) r.pad = __builtin_bit_cast(float, dither_hash32((uint32_t)k, a.hash_salt));
    out[k] = r;
}

__device__ __forceinline__ void decode(const Grid &g, long long p, int &i, int &j, int &l) {
    i = (int)(p % g.f[0]);
    long long q = p / g.f[0];
    j = (int)(q % g.f[1]);
    l = (int)(q / g.f[1]);
}

// offset of compact pixel p inside a GEMM slice plane
__device__ __forceinline__ long long plane_off(const Grid &g, long long p) {
    int i, j, l;
    decode(g, p, i, j, l);
    return i * g.ps[0] + j * g.ps[1] + l * g.ps[2];
}

// bin of compact pixel p, clamped into the cfact table (0..ncfact-1)
__device__ __forceinline__ int pixel_bin(const Grid &g, const int *istl, long long p, int ncfact) {
    int i, j, l;
    decode(g, p, i, j, l);
    int b = istl ? istl[i + (long long)g.n[0] * (j + (long long)g.n[1] * l)] : istl_point(g, i, j, l);
    return b < 0 ? 0 : (b >= ncfact ? ncfact - 1 : b);
}

// ---------------------------------------------------------------------------
// Lot reduction: for every pixel p of the computed region and every lot of
// the batch, F_l(p) = sum_{slices s of lot l} cfact(type_s, bin p) * partial_s(p)
// (FP64), then (four_run, fourier_sup.f90:224-254)
//   csf = F_l - acsf ; dsi += |csf|^2 ; Welford moments of F_l.
// fresh: the first batch, whose dsi and moments start from zero.
//
// One thread per pixel walks the batch's slices in order, RED_BATCH planes at
// a time, and issues the loads of the next RED_BATCH planes before it sums the
// current ones, so that a wave keeps loads in flight while it computes (with
// loads issued only between sums, a many-lot batch ran at under half the
// rate a plain read of its planes reaches).  CF_LDS: the pixel's form factor
// of every type (1..RED_LDS_TYPES types, every slice type < a.ntype) is read
// once into LDS, so that no global load in the sums waits behind the
// prefetched planes; otherwise the sums load it from cfact.  Slices of a lot
// come in runs of one type: a run's planes are summed, then that type's form
// factor is applied once.  A lot without slices adds F_l = 0.  Batches of 4
// planes keep the kernel at 64 VGPRs (8 waves per SIMD); batches of 8 ran a
// 64-lot batch no faster and the short walks of one-lot runs slower.
constexpr int RED_BATCH = 4;
constexpr int RED_BLOCK = 256;
constexpr int RED_LDS_TYPES = 8;
template <bool CF_LDS>
__global__ __launch_bounds__(RED_BLOCK) void reduce_lots_kernel(ReduceArgs a) {
    extern __shared__ double2 cf_lds[];   // CF_LDS: [type][thread]
    const long long p = (long long)blockIdx.x * RED_BLOCK + threadIdx.x;
    if (p >= a.g.P) return;
    const int bin = a.lform ? pixel_bin(a.g, a.istl, p, a.ncfact) : 0;
    if constexpr (CF_LDS)
        for (int t = 0; t < a.ntype; ++t) cf_lds[t * RED_BLOCK + threadIdx.x] = a.cfact[(long long)t * a.ncfact + bin];
    double ar = 0.0, ai = 0.0;
    if (a.acsf) { ar = a.acsf[2 * p]; ai = a.acsf[2 * p + 1]; }
    const bool rd = !a.fresh;
    double dsi = rd ? a.dsi[p] : 0.0;
    double mr = a.wm && rd ? a.wm[2 * p] : 0.0, mi = a.wm && rd ? a.wm[2 * p + 1] : 0.0;
    double m2 = a.wv && rd ? a.wv[p] : 0.0;
    double cr = 0.0, ci = 0.0;
    const int s0 = a.lot_slice[0], s1 = a.lot_slice[a.nlots];
    int l = 0;
    double fr = 0.0, fi = 0.0, sr = 0.0, si = 0.0;
    auto finish_lots = [&](int pos) {   // every lot ending at slice position pos
        while (l < a.nlots && a.lot_slice[l + 1] == pos) {
            if (a.wm) {
                const double inv = 1.0 / (double)(a.lot0 + l + 1);
                const double dr = fr - mr, di = fi - mi;
                mr += dr * inv;
                mi += di * inv;
                m2 += dr * (fr - mr) + di * (fi - mi);
            }
            cr = fr - ar;
            ci = fi - ai;
            dsi += cr * cr + ci * ci;
            fr = fi = 0.0;
            ++l;
        }
    };
    finish_lots(s0);
    if (s0 < s1) {
        const long long pl = a.g.plane;
        const float2 *base = a.partial + (long long)(s0 - a.slice0) * pl + plane_off(a.g, p);
        // planes past the last slice repeat it: every load is unconditional
        // and in bounds, and the sums stop at s1
        float2 nxt[RED_BATCH];
#pragma unroll
        for (int u = 0; u < RED_BATCH; ++u) nxt[u] = base[(long long)min(u, s1 - 1 - s0) * pl];
        for (int sb = s0; sb < s1; sb += RED_BATCH) {
            float2 v[RED_BATCH];
#pragma unroll
            for (int u = 0; u < RED_BATCH; ++u) v[u] = nxt[u];
            if (sb + RED_BATCH < s1) {
#pragma unroll
                for (int u = 0; u < RED_BATCH; ++u)
                    nxt[u] = base[(long long)min(sb + RED_BATCH + u - s0, s1 - 1 - s0) * pl];
            }
#pragma unroll
            for (int u = 0; u < RED_BATCH; ++u) {
                const int s = sb + u;
                if (s >= s1) break;
                sr += (double)v[u].x;
                si += (double)v[u].y;
                const bool lot_end = a.lot_slice[l + 1] == s + 1;
                const int t = a.slice_type ? a.slice_type[s] : 0;
                if (lot_end || (a.slice_type ? a.slice_type[s + 1] : 0) != t) {
                    if (a.lform) {
                        const double2 cf = CF_LDS ? cf_lds[t * RED_BLOCK + threadIdx.x]
                                                  : a.cfact[(long long)t * a.ncfact + bin];
                        fr += cf.x * sr - cf.y * si;
                        fi += cf.x * si + cf.y * sr;
                    } else {
                        fr += sr;
                        fi += si;
                    }
                    sr = si = 0.0;
                }
                if (lot_end) finish_lots(s + 1);
            }
        }
    }
    a.dsi[p] = dsi;
    a.csf[2 * p] = cr;
    a.csf[2 * p + 1] = ci;
    if (a.wm) {
        a.wm[2 * p] = mr;
        a.wm[2 * p + 1] = mi;
        a.wv[p] = m2;
    }
}

// Sum of the slices of one batch into a complex FP64 plane, with or without
// the form factor: used for <F> per type and for the lot lattice.
__global__ void reduce_sum_kernel(ReduceArgs a, double *__restrict__ out, int accumulate) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.g.P) return;
    const int bin = a.lform ? pixel_bin(a.g, a.istl, p, a.ncfact) : 0;
    double fr = accumulate ? out[2 * p] : 0.0, fi = accumulate ? out[2 * p + 1] : 0.0;
    const long long po = plane_off(a.g, p);
    for (int s = a.lot_slice[0]; s < a.lot_slice[a.nlots]; ++s) {
        const float2 v = a.partial[(long long)(s - a.slice0) * a.g.plane + po];
        if (a.lform) {
            const double2 cf = a.cfact[(long long)a.slice_type[s] * a.ncfact + bin];
            fr += cf.x * (double)v.x - cf.y * (double)v.y;
            fi += cf.x * (double)v.y + cf.y * (double)v.x;
        } else {
            fr += (double)v.x;
            fi += (double)v.y;
        }
    }
    out[2 * p] = fr;
    out[2 * p + 1] = fi;
}

// acsf = (sum_types F_t) * F_lattice * norm  (four_aver, fourier_sup.f90:1337-1343).
// The lattice plane is computed over the full grid (lines 1315-1318) but only
// the computed region is needed here; lat is indexed by the full grid.
__global__ void aver_combine_kernel(Grid g, double *__restrict__ acsf, const double *__restrict__ lat, double norm) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.P) return;
    int i, j, l;
    decode(g, p, i, j, l);
    const long long q = i + (long long)g.n[0] * (j + (long long)g.n[1] * l);
    const double ar = acsf[2 * p], ai = acsf[2 * p + 1];
    const double tr = lat[2 * q], ti = lat[2 * q + 1];
    acsf[2 * p] = (ar * tr - ai * ti) * norm;
    acsf[2 * p + 1] = (ar * ti + ai * tr) * norm;
}

// ---------------------------------------------------------------------------
// Full-grid outputs.  q runs over the num grid; values come from the
// computed region (compact, f-grid).  With the Friedel reduction on, the
// result equals DISCUS's sequential in-place mirror four_apply_friedel
// (fourier_sup.f90:3748-3758): a pixel whose mirror M(q) lies in the computed
// region and precedes q in loop order (or that lies outside the region) takes
// the mirror's value, conjugated for complex arrays; the centre pixel
// M(q) = q is conjugated; every other pixel keeps its own value.
__device__ __forceinline__ bool region_src(const Grid &g, long long q, long long &src, bool &conj) {
    const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
    const int i = (int)(q % n0);
    const long long qq = q / n0;
    const int j = (int)(qq % n1), l = (int)(qq / n1);
    const bool in = i < g.f[0] && j < g.f[1] && l < g.f[2];
    conj = false;
    if (g.friedel) {
        const int mi = n0 - 1 - i, mj = n1 - 1 - j, ml = n2 - 1 - l;
        const bool min_ = mi < g.f[0] && mj < g.f[1] && ml < g.f[2];
        bool take_mirror = false;
        if (mi == i && mj == j && ml == l) {
            conj = true;
        } else if (min_) {
            if (!in) take_mirror = true;
            else {
                // loop order: i outer, l inner
                take_mirror = (mi < i) || (mi == i && (mj < j || (mj == j && ml < l)));
            }
        }
        if (take_mirror) {
            src = mi + (long long)g.f[0] * (mj + (long long)g.f[1] * ml);
            conj = true;
            return true;
        }
    }
    if (!in) return false;
    src = i + (long long)g.f[0] * (j + (long long)g.f[1] * l);
    return true;
}

__global__ void finalize_kernel(FinalArgs a) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= a.ntot) return;
    long long s;
    bool conj;
    const bool has = region_src(a.g, q, s, conj);
    const double sg = conj ? -1.0 : 1.0;
    if (a.csf_out) {
        a.csf_out[2 * q] = has ? a.csf[2 * s] * a.sq : 0.0;
        a.csf_out[2 * q + 1] = has ? sg * a.csf[2 * s + 1] * a.sq : 0.0;
    }
    if (a.acsf_out) {
        a.acsf_out[2 * q] = (has && a.acsf) ? a.acsf[2 * s] : 0.0;
        a.acsf_out[2 * q + 1] = (has && a.acsf) ? sg * a.acsf[2 * s + 1] : 0.0;
    }
    if (a.dsi_out) a.dsi_out[q] = has ? a.dsi[s] * a.dscale : 0.0;
    if (a.fmean_out) {
        a.fmean_out[2 * q] = (has && a.wm) ? a.wm[2 * s] : 0.0;
        a.fmean_out[2 * q + 1] = (has && a.wm) ? sg * a.wm[2 * s + 1] : 0.0;
    }
    if (a.dvar_out) a.dvar_out[q] = (has && a.wv) ? a.wv[s] * a.vscale : 0.0;
}

// Layer call output: tcsf over the num grid, zero outside the fnum region
// (four_strucf_omp zeroes tcsf first, four_strucf_OMP_mod.f90:129).
// tcsf over the full num grid from the slices [s0, s1) (planes relative to
// a.slice0): cfact(istl) x their FP64 sum on the computed region, 0 outside;
// accumulate != 0 adds to tcsf instead (the computed region only).
__global__ void strucf_out_kernel(ReduceArgs a, double *__restrict__ tcsf, long long ntot, int s0, int s1,
                                  int accumulate) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ntot) return;
    const int i = (int)(q % a.g.n[0]);
    const long long qq = q / a.g.n[0];
    const int j = (int)(qq % a.g.n[1]), l = (int)(qq / a.g.n[1]);
    if (i >= a.g.f[0] || j >= a.g.f[1] || l >= a.g.f[2]) {
        if (!accumulate) {
            tcsf[2 * q] = 0.0;
            tcsf[2 * q + 1] = 0.0;
        }
        return;
    }
    const long long po = i * a.g.ps[0] + j * a.g.ps[1] + l * a.g.ps[2];
    double fr = 0.0, fi = 0.0;
    for (int s = s0; s < s1; ++s) {
        const float2 v = a.partial[(long long)(s - a.slice0) * a.g.plane + po];
        fr += (double)v.x;
        fi += (double)v.y;
    }
    if (a.lform) {
        int b = a.istl ? a.istl[q] : istl_point(a.g, i, j, l);
        b = b < 0 ? 0 : (b >= a.ncfact ? a.ncfact - 1 : b);
        const double2 cf = a.cfact[b];
        const double r = fr * cf.x - fi * cf.y, im = fr * cf.y + fi * cf.x;
        fr = r;
        fi = im;
    }
    if (accumulate) {
        tcsf[2 * q] += fr;
        tcsf[2 * q + 1] += fi;
    } else {
        tcsf[2 * q] = fr;
        tcsf[2 * q + 1] = fi;
    }
}

// four_stltab over the full grid (fourier_sup.f90:2660-2685), abscissa fastest:
// the unclamped bins, which the caller checks against CFPKT.
__global__ void istl_kernel(Grid g, int *__restrict__ out) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long ntot = (long long)g.n[0] * g.n[1] * g.n[2];
    if (q >= ntot) return;
    const int i = (int)(q % g.n[0]);
    const long long qq = q / g.n[0];
    out[q] = istl_raw(g, i, (int)(qq % g.n[1]), (int)(qq / g.n[1]));
}

static inline unsigned nblk(long long n, int b) { return (unsigned)((n + b - 1) / b); }

hipError_t launch_prep(const PrepArgs &a, AtomRec *out, hipStream_t st) {
    if (a.n <= 0) return hipSuccess;
    prep_records_kernel<<<nblk(a.n, 256), 256, 0, st>>>(a, out);
    return hipGetLastError();
}
hipError_t launch_reduce_lots(const ReduceArgs &a, hipStream_t st) {
    if (a.lform && a.ntype >= 1 && a.ntype <= RED_LDS_TYPES)
        reduce_lots_kernel<true><<<nblk(a.g.P, RED_BLOCK), RED_BLOCK, sizeof(double2) * RED_BLOCK * a.ntype, st>>>(a);
    else
        reduce_lots_kernel<false><<<nblk(a.g.P, RED_BLOCK), RED_BLOCK, 0, st>>>(a);
    return hipGetLastError();
}
hipError_t launch_reduce_sum(const ReduceArgs &a, double *out, int accumulate, hipStream_t st) {
    reduce_sum_kernel<<<nblk(a.g.P, 256), 256, 0, st>>>(a, out, accumulate);
    return hipGetLastError();
}
hipError_t launch_aver_combine(const Grid &g, double *acsf, const double *lat, double norm, hipStream_t st) {
    aver_combine_kernel<<<nblk(g.P, 256), 256, 0, st>>>(g, acsf, lat, norm);
    return hipGetLastError();
}
hipError_t launch_finalize(const FinalArgs &a, hipStream_t st) {
    finalize_kernel<<<nblk(a.ntot, 256), 256, 0, st>>>(a);
    return hipGetLastError();
}
hipError_t launch_strucf_out(const ReduceArgs &a, double *tcsf, long long ntot, int s0, int s1, int accumulate,
                             hipStream_t st) {
    strucf_out_kernel<<<nblk(ntot, 256), 256, 0, st>>>(a, tcsf, ntot, s0, s1, accumulate);
    return hipGetLastError();
}
hipError_t launch_istl(const Grid &g, int *out, hipStream_t st) {
    istl_kernel<<<nblk((long long)g.n[0] * g.n[1] * g.n[2], 256), 256, 0, st>>>(g, out);
    return hipGetLastError();
}

}  // namespace dc
