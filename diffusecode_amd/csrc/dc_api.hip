// dc_api.hip -- C ABI of libdiffuse_turbo.so (declared in include/diffuse_turbo.h).
//
// Host side of the MI355X structure-factor path.  It owns device buffers
// inside a context, restates the host bookkeeping of DISCUS's four_run
// (grid set-up, Friedel test, lot gathering, average structure, lot loop;
// discus/prog/fourier_sup.f90:11-323) and drives the kernels of dc_gemm.hip
// and dc_lot.hip on one HIP stream.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <functional>
#include <vector>

#include "../../include/diffuse_turbo.h"
#include "dc_host.h"

using namespace dc;

namespace dc {

constexpr int GEMM_KS = 16;
constexpr int KCHUNK_MIN = 512;
constexpr int KCHUNK_MAX = 16384;
// FP32 scheme: atoms per slice.  The FP32 kernel adds groups of 64 atoms
// into a second set of FP32 sums; with 1,024-atom slices summed in FP64 the
// outer sum has 16 terms, keeping max|dI|/max(I) near 1e-7 on coherent sums
// (Bragg peaks of displaced crystals), where one FP32 accumulator over
// thousands of atoms reaches 1e-6.  The split schemes keep long slices: with
// the per-atom phase rotation of the GEMM (dither_turn) their accumulator and
// operand rounding stays incoherent (shifted fcc 40^3 crystal, all 256,000
// atoms in phase at the Bragg peaks: FP16X3 4.0e-7 over 4,096-atom slices,
// 4.9e-5 without the rotation).
constexpr int KCHUNK_FP32 = 1024;
// Longest slice the planner's search (kchunk_for) may pick for the split and
// FP64-sum schemes: the headline 161^3 block's 85,334-atom type segments
// then split 5 ways (15 slices x 51 tiles = 765 items, 3 rounds of 256 CUs
// with 3 idle slots) instead of 10 ways (30 x 51 = 1,530 items, 6 rounds):
// half the planes to write and reduce (reduction 0.122 -> 0.061 ms), the
// same GEMM time, and on the coherent sums of the shifted crystal FP16X3
// 8.7e-7, BF16X3 3.3e-6, FP16X2 1.3e-6 at 17,067 atoms against 1.7e-6,
// 5.0e-6, 7.5e-7 at 8,534 (max|dI|/max(I) vs FP64).
constexpr int KCHUNK_SEARCH_MAX = 32768;
constexpr double PARTIAL_BUDGET = 6.0e9;   // bytes of slice planes per batch

}  // namespace dc

namespace {

// four_test_friedel (fourier_sup.f90:3645-3728): reduced sizes; returns the
// halved axis (1..3) or 0.
int test_friedel(const double *eck, const double *vi, const int *inc, int ano, int four_friedel, int fnum[3]) {
    fnum[0] = inc[0];
    fnum[1] = inc[1];
    fnum[2] = inc[2];
    if (!four_friedel || ano) return 0;
    double point[3];
    for (int r = 0; r < 3; ++r) {
        double rut = eck[r] + vi[r] * (fnum[0] - 1) + vi[r + 3] * (fnum[1] - 1) + vi[r + 6] * (fnum[2] - 1);
        point[r] = eck[r] + rut;
    }
    const double eps = (double)1.E-6f;
    if ((fnum[0] % 2) && (fnum[1] % 2) && (fnum[2] % 2)) {
        if (std::fabs(point[0]) < eps && std::fabs(point[2]) < eps && std::fabs(point[1]) < eps) {
            if (fnum[0] > 1) { fnum[0] = (fnum[0] + 1) / 2; return 1; }
            if (fnum[1] > 1) { fnum[1] = (fnum[1] + 1) / 2; return 2; }
            if (fnum[2] > 1) { fnum[2] = (fnum[2] + 1) / 2; return 3; }
        }
    }
    return 0;
}

// Zero increment vector test of fourier_lmn (fourier_lmn_mod.f90:74-93),
// with its rounding of vi to 1e-7.
bool zero_increment(const double *vi, const int *inc) {
    for (int i = 0; i < 3; ++i) {
        if (inc[i] <= 1) continue;
        double s = 0.0;
        for (int r = 0; r < 3; ++r) {
            double v = (double)(long long)std::llround(vi[r + 3 * i] * 1.0e7) / 1.0e7;
            s += v * v;
        }
        if (!(std::sqrt(s) > 1.0e-7)) return true;
    }
    return false;
}

}  // namespace

namespace dc {

// sin(theta)/lambda bin with the expression of four_stltab
// (fourier_sup.f90:2663-2674, h rounded to FP32); mirrors istl_point of
// dc_internal.h.
#pragma clang fp contract(off)
int host_istl(const double *xm, const double *u, const double *v, const double *w, int i, int j, int l,
              const double *rten) {
    double h[3], t[3];
    for (int k = 0; k < 3; ++k) h[k] = stl_h(xm[k], u[k], v[k], w[k], i, j, l);
    for (int r = 0; r < 3; ++r) t[r] = rten[r] * h[0] + rten[r + 3] * h[1] + rten[r + 6] * h[2];
    double q2 = (h[0] * t[0] + h[1] * t[1] + h[2] * t[2]) / 4.0;
    return (int)std::llround(std::sqrt(q2) * (1.0 / CFINC));
}
#pragma clang fp contract(on)

// s is a norm of h, so its maximum over the parallelepiped grid sits at a corner.
int max_corner_bin(const Grid &g) {
    int m = 0;
    for (int ci = 0; ci < 8; ++ci) {
        int i = (ci & 1) ? g.n[0] - 1 : 0, j = (ci & 2) ? g.n[1] - 1 : 0, l = (ci & 4) ? g.n[2] - 1 : 0;
        m = std::max(m, host_istl(g.xm, g.uin, g.vin, g.win, i, j, l, g.rten));
    }
    return m;
}

long long padded(long long n, long long t) { return (n + t - 1) / t * t; }

// Atoms per slice of the split schemes: enough slices that a launch has
// about 1024 work items (wgs = work-groups per slice), within
// [KCHUNK_MIN, KCHUNK_MAX]; FP32 uses a fixed slice length that bounds its
// accumulators' rounding on coherent sums.
static int kchunk_atoms(long long wgs, long long natoms, int prec) {
    if (prec == DC_PREC_FP32) return KCHUNK_FP32;   // (FP32 with FP64 sums takes long slices)
    long long want = std::max<long long>(1, (1024 + wgs - 1) / std::max<long long>(wgs, 1));
    long long k = (natoms + want - 1) / want;
    k = std::min<long long>(std::max<long long>(k, KCHUNK_MIN), KCHUNK_MAX);
    return (int)padded(k, GEMM_KS);
}

// GEMM plan for the computed region f of grid g.
//   mapping 0: rows = abscissa (no wrap), cols = (ordinate, top), fast period C0
//   mapping 1: rows = (abscissa, ordinate), fast period R0, cols = top (no wrap)
// Fast periods are padded to >= 16 when the side wraps (the two-level operand
// tables of dc_gemm.hip need at most one wrap per 16-group); padded pixels
// are computed and never read.  forced >= 0 fixes the tile (the mapping is
// then the one with the least padded work), only_mapping >= 0 restricts the
// mapping, and the split-only tiles are skipped for the FP32 scheme.
// Cost of a (mapping, tile) candidate:
//   FP32: padded work x (1 + (16/RT + 16/CT) / 2);
//   split schemes: the launch's rounds of work-groups over the CUs x the
//   work of one work-group (RT x CT x slice atoms) x the tile's measured
//   cost factor (TileInfo::eff), with the slices cut as kchunk_atoms cuts
//   them for natoms atoms.  Small problems fill the device with few work
//   items, so the rounds term favours smaller tiles there; large problems
//   reduce to padded work x eff.  natoms <= 0: padded work x eff.
Plan plan_gemm(Grid &g, int forced, int only_mapping, int prec, long long natoms) {
    Plan best;
    double best_cost = 1e300;
    const int nt = gemm_num_tiles();
    const int ncu = device_cu_count();
    for (int mapping = 0; mapping < 2; ++mapping) {
        if (mapping == 1 && g.f[2] <= 1) continue;
        if (only_mapping >= 0 && mapping != only_mapping) continue;
        GemmGeom gg{};
        long long ps[3];
        if (mapping == 0) {
            gg.R = g.f[0];
            gg.R0 = g.f[0];
            gg.C0 = g.f[2] > 1 ? std::max(16, g.f[1]) : g.f[1];
            gg.C = gg.C0 * g.f[2];
        } else {
            gg.R0 = g.f[1] > 1 ? std::max(16, g.f[0]) : g.f[0];
            gg.R = gg.R0 * g.f[1];
            gg.C = g.f[2];
            gg.C0 = g.f[2];
        }
        // plane column stride: whole 16-row blocks (128-byte lines) unless
        // that grows the plane by more than 1/16 (short columns)
        gg.Rs = (padded(gg.R, 16) - gg.R) * 16 <= gg.R ? (int)padded(gg.R, 16) : gg.R;
        if (mapping == 0) {
            ps[0] = 1; ps[1] = gg.Rs; ps[2] = (long long)gg.Rs * gg.C0;
        } else {
            ps[0] = 1; ps[1] = gg.R0; ps[2] = gg.Rs;
        }
        gg.P = (long long)gg.Rs * gg.C;
        const bool acc64 = prec == SCHEME_FP32_ACC64;
        // FP32 with FP64 sums is built for a few tiles only; a forced tile
        // outside them is ignored for it
        const bool use_forced = forced >= 0 && (!acc64 || gemm_tile_info(forced).acc64) &&
                                !(prec == DC_PREC_FP32 && gemm_tile_info(forced).no_fp32);
        for (int t = 0; t < nt; ++t) {
            if (use_forced && t != forced) continue;
            TileInfo ti = gemm_tile_info(t);
            if (!use_forced && ti.split_only && prec == DC_PREC_FP32) continue;
            if (acc64 && !ti.acc64) continue;
            if ((acc64 || prec == DC_PREC_FP32) && ti.no_fp32) continue;
            const double work = (double)padded(gg.R, ti.rt) * (double)padded(gg.C, ti.ct);
            double cost;
            if (prec == DC_PREC_FP32 || acc64 || use_forced) {
                cost = work * (1.0 + 0.5 * (16.0 / ti.rt + 16.0 / ti.ct));
            } else if (natoms <= 0) {
                cost = work * ti.eff;
            } else {
                const long long wgs = (long long)((gg.R + ti.rt - 1) / ti.rt) * ((gg.C + ti.ct - 1) / ti.ct);
                const long long k = kchunk_atoms(wgs, natoms, prec);
                const long long nsl = (natoms + k - 1) / k;
                const long long rounds = (wgs * nsl + ncu - 1) / ncu;
                cost = (double)rounds * ti.rt * ti.ct * (double)std::min(k, natoms) * ti.eff;
            }
            if (cost < best_cost) {
                best_cost = cost;
                best.gg = gg;
                best.mapping = mapping;
                best.tile = t;
                best.scheme = prec;
                g.ps[0] = ps[0]; g.ps[1] = ps[1]; g.ps[2] = ps[2];
                g.plane = gg.P;
            }
        }
    }
    return best;
}

int kchunk_for(const Plan &pl, long long natoms_batch, int prec) {
    TileInfo ti = gemm_tile_info(pl.tile);
    long long wgs = (long long)((pl.gg.R + ti.rt - 1) / ti.rt) * ((pl.gg.C + ti.ct - 1) / ti.ct);
    return kchunk_atoms(wgs, natoms_batch, prec);
}

// Slice length for the (type, count) segments of one
// launch.  The launch has one work-group per (slice, tile), and every tile's
// LDS footprint allows one work-group per CU, so it runs in rounds of ncu
// work items; a last round that is mostly empty costs as much as a full one
// (the headline 161^3 block had 21 slices x 51 tiles = 1,071 items, 4.2
// rounds run as 5).  A candidate length k splits every segment into
// ceil(n / k) near-equal slices (add_segments); its predicted time is
// ceil(W / ncu) x (L + 64), with L the longest slice, 64 atoms standing for
// an item's set-up, and W the items weighted by (length + 64) / (L + 64):
// exact when the slices are equal, close to the fluid limit when short
// slices fill the gaps.  The candidates are the lengths in [KCHUNK_MIN,
// KCHUNK_SEARCH_MAX] at which some segment's slice count changes; the shortest
// predicted time wins, ties going to the longer slices (fewer planes to
// reduce).  FP32 searches between half its fixed slice cap and the cap.
// Launches that already have 16 rounds of items or more, and
// populations (many segments), keep the rule of kchunk_atoms.
int kchunk_for(const Plan &pl, const std::vector<std::pair<int, int>> &segs, int prec) {
    long long natoms = 0;
    for (auto &sg : segs) natoms += std::max(sg.second, 0);
    const int k0 = kchunk_for(pl, natoms, prec);
    if (natoms <= 0) return k0;
    // FP32 caps its slices (accumulator rounding on coherent sums); shorter
    // slices are allowed, down to half the cap
    const bool capped = prec == DC_PREC_FP32;
    const int kmax = capped ? k0 : KCHUNK_SEARCH_MAX;
    const int kmin = capped ? k0 / 2 : KCHUNK_MIN;
    const TileInfo ti = gemm_tile_info(pl.tile);
    const long long wgs = (long long)((pl.gg.R + ti.rt - 1) / ti.rt) * ((pl.gg.C + ti.ct - 1) / ti.ct);
    const int ncu = device_cu_count();
    long long items0 = 0;
    for (auto &sg : segs)
        if (sg.second > 0) items0 += (sg.second + k0 - 1) / k0;
    items0 *= wgs;
    // 16 rounds or more: the last one costs at most 1/16 -- keep the rule
    if (items0 >= 16LL * ncu || segs.size() > 512) return k0;
    // candidates: the lengths at which a segment's slice count changes
    std::vector<int> cand;
    for (auto &sg : segs) {
        if (sg.second <= 0) continue;
        const int n = sg.second;
        for (int ns = (n + kmax - 1) / kmax; ns <= (n + kmin - 1) / kmin; ++ns) {
            const int k = (n + ns - 1) / ns;
            if (k >= kmin && k <= kmax) cand.push_back(k);
        }
        if ((long long)cand.size() > 4096) break;
    }
    cand.push_back(k0);
    std::sort(cand.begin(), cand.end(), std::greater<int>());
    cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
    int best = k0;
    double best_t = 1e300;
    for (int k : cand) {
        // add_segments splits a segment of n atoms into ns = ceil(n / k)
        // slices of ceil(n / ns) atoms at most
        double L = 0.0;
        for (auto &sg : segs) {
            if (sg.second <= 0) continue;
            const long long ns = (sg.second + k - 1) / k;
            L = std::max(L, (double)((sg.second + ns - 1) / ns));
        }
        L += 64.0;
        double w = 0.0;
        for (auto &sg : segs) {
            if (sg.second <= 0) continue;
            const long long ns = (sg.second + k - 1) / k;
            w += (double)ns * ((double)((sg.second + ns - 1) / ns) + 64.0) / L;
        }
        const double t = std::ceil(w * (double)wgs / ncu - 1e-9) * L;
        if (t < best_t * (1.0 - 1e-9)) {
            best_t = t;
            best = k;
        }
    }
    return best;
}


// Each segment of n atoms becomes ceil(n / kchunk) slices of near-equal
// length (they differ by at most one atom, none longer than kchunk), so the
// work items of a launch are uniform.
void add_segments(SliceSet &s, const std::vector<std::pair<int, int>> &segs, int kchunk, int &atom_off) {
    for (auto &sg : segs) {
        const int n = sg.second;
        s.seg_first.push_back((int)s.type.size());
        const int ns = (n + kchunk - 1) / kchunk;
        int b = 0;
        for (int q = 0; q < ns; ++q) {
            const int len = n / ns + (q < n % ns ? 1 : 0);
            if (s.beg.empty()) s.beg.push_back(atom_off);
            s.type.push_back(sg.first);
            b += len;
            s.beg.push_back(atom_off + b);
        }
        atom_off += n;
    }
}

int run_gemm(dc_ctx *c, const Plan &pl, int nslice, float2 *partial, const float *sq) {
    GemmGeom gg = launch_geom(c, pl.gg);
    gg.pad_hash = c->rec_hash && !sq;
    if (nslice <= 0) return 0;
    int st = 0;
    hipEvent_t e0, e1;
    if ((st = mark(c, e0))) return st;
    // launch in chunks of slices so that the cancel flag is polled
    const int per = 4096;
    for (int s0 = 0; s0 < nslice; s0 += per) {
        if (cancelled(c)) return set_err(c, DC_ERR_CANCEL, DC_ER_COMM, "cancelled");
        int ns = std::min(per, nslice - s0);
        HIPCHK(launch_phase_gemm(pl.scheme, pl.tile, ptr<AtomRec>(c->rec), ptr<int>(c->slice_beg),
                                 partial + (long long)s0 * gg.P * planes_per_slice(pl.scheme), sq, gg, s0, ns,
                                 c->st));
        c->stats.gemm_launches++;
    }
    if ((st = mark(c, e1))) return st;
    span(c, e0, e1, &c->stats.gemm_ms);
    return 0;
}

PrepArgs base_prep(const Grid &g, int mapping) {
    // (mapping selects which increments feed the row and column sides)
    PrepArgs a{};
    for (int k = 0; k < 3; ++k) {
        a.vec[k] = g.uin[k];
        a.vec[3 + k] = g.vin[k];
        a.vec[6 + k] = g.win[k];
        a.h0[k] = g.xm[k];
    }
    a.mode = PREP_CORNER;
    a.mapping = mapping;
    return a;
}

}  // namespace dc

namespace {

// Compute sum over atom segments (idx list on device or explicit positions)
// into planes; used by the average structure and the lattice.
struct GatherList {
    std::vector<int> idx;
    std::vector<int> oid;
    std::vector<int4> runs;                  // lot gather: (first, src in type_idx, oid, -)
    std::vector<double> offs;
    std::vector<std::pair<int, int>> segs;   // (type 0-based, count)
};

}  // namespace

// ============================================================================
extern "C" {

const char *dc_version(void) { return "diffuse_turbo 0.1 (gfx950)"; }

int dc_create(dc_ctx **out, int device) {
    if (!out) return DC_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return DC_ERR_NODEV;
    if (device < 0) {
        if (hipGetDevice(&device) != hipSuccess) return DC_ERR_NODEV;
    }
    if (device >= ndev) return DC_ERR_NODEV;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return DC_ERR_NODEV;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return DC_ERR_NODEV;
    dc_ctx *c = new dc_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return DC_ERR_DEVICE;
    }
    for (auto &e : c->ev)
        if (hipEventCreate(&e) != hipSuccess) {
            delete c;
            return DC_ERR_DEVICE;
        }
    if (noalloc_attach(c) != 0) {
        dc_destroy(c);
        return DC_ERR_DEVICE;
    }
    *out = c;
    return DC_OK;
}

void dc_destroy(dc_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->st);
    batch_release(c);
    post_release(c);
    nufft_release(c);
    for (DevBuf *b : c->buffers())
        if (b->p) (void)hipFree(b->p);
    for (auto &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : c->tev) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(c->st);
    delete c;
}

const char *dc_last_error(const dc_ctx *c) { return c ? c->err.c_str() : "null context"; }
int dc_last_error_type(const dc_ctx *c) { return c ? c->err_typ : DC_ER_LIB; }

int dc_set_precision(dc_ctx *c, int p) {
    if (!c) return DC_ERR_ARG;
    if (p != DC_PREC_FP32 && p != DC_PREC_BF16X3 && p != DC_PREC_FP16X2 && p != DC_PREC_FP16X3)
        return set_err(c, DC_ERR_ARG, DC_ER_LIB, "unknown precision");
    c->prec = p;
    return DC_OK;
}
int dc_get_precision(const dc_ctx *c) { return c ? c->prec : -1; }

int dc_set_gemm_tile(dc_ctx *c, int tile) {
    if (!c) return DC_ERR_ARG;
    if (tile < -1 || tile >= gemm_num_tiles()) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "unknown tile configuration");
    c->tile = tile;
    return DC_OK;
}

int dc_gemm_num_tiles(void) { return gemm_num_tiles(); }

int dc_set_phase_dither(dc_ctx *c, int on) {
    if (!c) return DC_ERR_ARG;
    c->dither = on ? 1 : 0;
    return DC_OK;
}

int dc_set_slice_atoms(dc_ctx *c, int n) {
    if (!c) return DC_ERR_ARG;
    if (n < 0 || n > (1 << 24)) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "slice length out of range");
    c->slice_atoms = n;
    return DC_OK;
}

int dc_set_adjoint_walk(dc_ctx *c, int walk) {
    if (!c) return DC_ERR_ARG;
    if (walk < 0 || walk > adjoint_max_walk())
        return set_err(c, DC_ERR_ARG, DC_ER_LIB, "adjoint walk out of range");
    c->adj_walk = walk;
    return DC_OK;
}

int dc_gemm_tile_shape(int tile, int *rows, int *cols, int *threads) {
    if (tile < 0 || tile >= gemm_num_tiles()) return DC_ERR_ARG;
    TileInfo t = gemm_tile_info(tile);
    if (rows) *rows = t.rt;
    if (cols) *cols = t.ct;
    if (threads) *threads = t.nt;
    return DC_OK;
}

int dc_set_cancel_flag(dc_ctx *c, const volatile int *flag) {
    if (!c) return DC_ERR_ARG;
    c->cancel = flag;
    return DC_OK;
}

int dc_device_info(const dc_ctx *c, char *name, int len, int *num_cu) {
    if (!c) return DC_ERR_ARG;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c->device) != hipSuccess) return DC_ERR_DEVICE;
    if (name && len > 0) {
        std::strncpy(name, prop.gcnArchName, len - 1);
        name[len - 1] = 0;
    }
    if (num_cu) *num_cu = prop.multiProcessorCount;
    return DC_OK;
}

int dc_last_stats(const dc_ctx *cc, dc_stats *out) {
    if (!cc || !out) return DC_ERR_ARG;
    // timings still in flight are part of the last call's statistics
    dc_ctx *c = const_cast<dc_ctx *>(cc);
    int st = resolve_spans(c);
    if (st) return st;
    *out = c->stats;
    return DC_OK;
}

int dc_synchronize(dc_ctx *c) {
    if (!c) return DC_ERR_ARG;
    HIPCHK(hipStreamSynchronize(c->st));
    return DC_OK;
}

}  // extern "C"

// ----------------------------------------------------------------------------
namespace dc {

// Validation and grid of the layer call (four_strucf's inputs).  Returns 0 or
// a status; fills g and ntot.
int strucf_setup(dc_ctx *c, const void *xat, int64_t nxat, int64_t ld_xat, const double uin[3], const double vin[3],
                 const double win[3], const double off_shift[9], const int lmn[6], const int num[3],
                 const int fnum[3], const void *cfact_col, int64_t ncfact, const void *istl, const double xm[3],
                 const double rten[9], int lform, const void *tcsf, Grid &g, long long &ntot) {
    c->err.clear();
    c->err_typ = DC_ER_NONE;
    if (!uin || !vin || !win || !off_shift || !lmn || !num || !fnum || !tcsf || nxat < 0 || (nxat > 0 && (!xat || ld_xat < nxat)))
        return set_err(c, DC_ERR_ARG, DC_ER_LIB, "dc_strucf: null pointer or bad size");
    for (int k = 0; k < 3; ++k)
        if (num[k] < 1 || fnum[k] < 1 || fnum[k] > num[k]) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "dc_strucf: bad num/fnum");
    if (lform && (!cfact_col || ncfact < 1)) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "dc_strucf: lform needs cfact");
    if (lform && !istl && (!xm || !rten)) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "dc_strucf: need istl or xm+rten");
    if (nxat > INT32_MAX) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "dc_strucf: too many atoms");
    if (cancelled(c)) return set_err(c, DC_ERR_CANCEL, DC_ER_COMM, "cancelled");
    g = Grid{};
    for (int k = 0; k < 3; ++k) {
        g.n[k] = num[k];
        g.f[k] = fnum[k];
        g.uin[k] = uin[k];
        g.vin[k] = vin[k];
        g.win[k] = win[k];
        g.xm[k] = xm ? xm[k] : 0.0;
    }
    for (int k = 0; k < 9; ++k) g.rten[k] = rten ? rten[k] : 0.0;
    g.P = (long long)fnum[0] * fnum[1] * fnum[2];
    ntot = (long long)num[0] * num[1] * num[2];
    if (lform && !istl && max_corner_bin(g) >= ncfact)
        return set_err(c, DC_ERR_ISTL, DC_ER_FOUR, "sin(theta)/lambda beyond table");
    return 0;
}

// The layer call on device buffers: records of the nxat atoms (d_xat in
// xat(ld,3) layout) into rec, the slices of the plan, and the phase GEMM in
// batches of slices whose planes fit plane_bytes of partial, each batch
// reduced into d_tcsf (cfact(istl) x FP64 sum; the first batch writes the
// whole num grid, the next ones add).  slice_beg holds nslice + 1 ints on the
// device.  Allocates nothing.
int strucf_core(dc_ctx *c, const Grid &g0, long long ntot, const double *d_xat, int64_t nxat, int64_t ld_xat,
                const double off_shift[9], const int lmn[6], const double2 *d_cfact, int64_t ncfact,
                const int *d_istl, int lform, AtomRec *rec, int *slice_beg, float2 *partial, double plane_bytes,
                double *d_tcsf) {
    Grid g = g0;
    const Plan pl = plan_gemm(g, c->tile, -1, c->prec, nxat);
    const double one_plane = (double)sizeof(float2) * (double)pl.gg.P * planes_per_slice(pl.scheme);
    if (one_plane > plane_bytes) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "dc_strucf: one slice plane exceeds the plane budget");
    PrepArgs pa = base_prep(g, pl.mapping);
    pa.mode = PREP_LMN;
    for (int k = 0; k < 9; ++k) pa.vec[9 + k] = off_shift[k];
    for (int k = 0; k < 6; ++k) pa.lmn[k] = lmn[k];
    pa.src = d_xat;
    pa.src_ld = std::max<int64_t>(ld_xat, 1);
    pa.n = nxat;
    pa.hash_salt = launch_geom(c, pl.gg).dither;   // undamped launches below
    HIPCHK(launch_prep(pa, rec, c->st));
    SliceSet ss;
    int off = 0;
    add_segments(ss, {{0, (int)nxat}}, slice_len(c, kchunk_for(pl, nxat, pl.scheme)), off);
    if (ss.beg.empty()) ss.beg.push_back(0);
    const int nslice = (int)ss.type.size();
    HIPCHK(hipMemcpyAsync(slice_beg, ss.beg.data(), sizeof(int) * ss.beg.size(), hipMemcpyHostToDevice, c->st));
    ReduceArgs ra{};
    ra.g = g;
    ra.partial = partial;
    ra.lform = lform;
    ra.cfact = d_cfact;
    ra.ncfact = (int)ncfact;
    ra.istl = d_istl;
    const int per = std::max(1, (int)std::min<double>(4096.0, std::floor(plane_bytes / one_plane)));
    GemmGeom gg = launch_geom(c, pl.gg);
    gg.pad_hash = 1;
    HIPCHK(hipEventRecord(c->ev[2], c->st));
    for (int s0 = 0; s0 < std::max(nslice, 1); s0 += per) {
        if (cancelled(c)) return set_err(c, DC_ERR_CANCEL, DC_ER_COMM, "cancelled");
        const int s1 = std::min(nslice, s0 + per);
        if (s1 > s0) {
            HIPCHK(launch_phase_gemm(pl.scheme, pl.tile, rec, slice_beg, partial, nullptr, gg, s0, s1 - s0, c->st));
            c->stats.gemm_launches++;
        }
        ra.slice0 = s0;
        HIPCHK(launch_strucf_out(ra, d_tcsf, ntot, s0, s1, s0 > 0, c->st));
    }
    HIPCHK(hipEventRecord(c->ev[3], c->st));
    c->stats.slices = nslice;
    c->stats.gemm_scheme = pl.scheme;
    c->stats.pairs = (double)nxat * (double)ntot;
    return 0;
}

}  // namespace dc

extern "C" {

int dc_strucf(dc_ctx *c, const double *xat, int64_t nxat, int64_t ld_xat, const double uin[3],
              const double vin[3], const double win[3], const double off_shift[9], const int lmn[6],
              const int num[3], const int fnum[3], const double *cfact_col, int64_t ncfact, const int *istl,
              const double xm[3], const double rten[9], int lform, double *tcsf) {
    if (!c) return DC_ERR_ARG;
    Grid g{};
    long long ntot = 0;
    int st = strucf_setup(c, xat, nxat, ld_xat, uin, vin, win, off_shift, lmn, num, fnum, cfact_col, ncfact, istl, xm,
                          rten, lform, tcsf, g, ntot);
    if (st) return st;
    if (lform && istl) {
        for (int l = 0; l < fnum[2]; ++l)
            for (int j = 0; j < fnum[1]; ++j)
                for (int i = 0; i < fnum[0]; ++i) {
                    int b = istl[i + (long long)num[0] * (j + (long long)num[1] * l)];
                    if (b < 0 || b >= ncfact) return set_err(c, DC_ERR_ISTL, DC_ER_FOUR, "istl outside cfact");
                }
    }
    HIPCHK(hipSetDevice(c->device));
    stats_begin(c);
    HIPCHK(hipEventRecord(c->ev[0], c->st));
    // positions: xat(ld,3) columns packed into (n,3) Fortran layout on device
    if ((st = ensure(c, c->xat, sizeof(double) * 3 * std::max<int64_t>(nxat, 1)))) return st;
    for (int col = 0; col < 3 && nxat > 0; ++col)
        HIPCHK(hipMemcpyAsync(ptr<double>(c->xat) + col * nxat, xat + col * ld_xat, sizeof(double) * nxat,
                              hipMemcpyHostToDevice, c->st));
    if (lform) {
        if ((st = upload(c, c->cfact, cfact_col, 2 * (size_t)ncfact))) return st;
        if (istl && (st = upload(c, c->istl, istl, (size_t)ntot))) return st;
    }
    // buffers for the records, the slice boundaries and the planes of all
    // slices, or of as many as PARTIAL_BUDGET holds (the GEMM then runs in
    // batches of slices)
    Grid gp = g;
    const Plan pl = plan_gemm(gp, c->tile, -1, c->prec, nxat);
    const int klen = slice_len(c, kchunk_for(pl, nxat, pl.scheme));
    const long long nslice = std::max<long long>(1, (nxat + klen - 1) / klen);
    const double plane = (double)sizeof(float2) * (double)pl.gg.P;
    const double planes = std::max(plane, std::min((double)PARTIAL_BUDGET, plane * (double)nslice));
    if ((st = ensure(c, c->rec, sizeof(AtomRec) * std::max<int64_t>(nxat, 1)))) return st;
    if ((st = ensure(c, c->slice_beg, sizeof(int) * (size_t)(nslice + 1)))) return st;
    if ((st = ensure(c, c->partial, (size_t)planes))) return st;
    if ((st = ensure(c, c->out_tcsf, sizeof(double) * 2 * ntot))) return st;
    if ((st = strucf_core(c, g, ntot, ptr<double>(c->xat), nxat, std::max<int64_t>(nxat, 1), off_shift, lmn,
                          lform ? ptr<double2>(c->cfact) : nullptr, ncfact, lform && istl ? ptr<int>(c->istl) : nullptr,
                          lform, ptr<AtomRec>(c->rec), ptr<int>(c->slice_beg), ptr<float2>(c->partial), planes,
                          ptr<double>(c->out_tcsf))))
        return st;
    HIPCHK(hipMemcpyAsync(tcsf, c->out_tcsf.p, sizeof(double) * 2 * ntot, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipEventRecord(c->ev[1], c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    c->stats.total_ms = ms;
    HIPCHK(hipEventElapsedTime(&ms, c->ev[2], c->ev[3]));
    c->stats.gemm_ms = ms;
    return DC_OK;
}

// ----------------------------------------------------------------------------
int dc_four_stltab(dc_ctx *c, const double xm[3], const double uin[3], const double vin[3], const double win[3],
                   const int num[3], const double rten[9], int *istl) {
    if (!c) return DC_ERR_ARG;
    c->err.clear();
    c->err_typ = DC_ER_NONE;
    if (!xm || !uin || !vin || !win || !num || !rten || !istl)
        return set_err(c, DC_ERR_ARG, DC_ER_LIB, "dc_four_stltab: null pointer");
    for (int k = 0; k < 3; ++k)
        if (num[k] < 1) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "dc_four_stltab: num < 1");
    const long long ntot = (long long)num[0] * num[1] * num[2];
    if (ntot > INT32_MAX) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "dc_four_stltab: grid too large");
    HIPCHK(hipSetDevice(c->device));
    Grid g{};
    for (int k = 0; k < 3; ++k) {
        g.n[k] = g.f[k] = num[k];
        g.xm[k] = xm[k];
        g.uin[k] = uin[k];
        g.vin[k] = vin[k];
        g.win[k] = win[k];
    }
    for (int k = 0; k < 9; ++k) g.rten[k] = rten[k];
    g.P = ntot;
    int st = 0;
    if ((st = ensure(c, c->istl, sizeof(int) * (size_t)ntot))) return st;
    HIPCHK(launch_istl(g, ptr<int>(c->istl), c->st));
    HIPCHK(hipMemcpyAsync(istl, c->istl.p, sizeof(int) * (size_t)ntot, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    for (long long q = 0; q < ntot; ++q)
        if (istl[q] > CFPKT) return set_err(c, DC_ERR_ISTL, DC_ER_FOUR, "sin(theta)/lambda beyond the form factor table");
    return DC_OK;
}

// ----------------------------------------------------------------------------
int dc_crystal_set(dc_ctx *c, const dc_crystal_desc *d) {
    if (!c) return DC_ERR_ARG;
    c->err.clear();
    if (!d || d->natoms < 0 || (d->natoms > 0 && (!d->cr_pos || !d->cr_iscat)) || d->nscat < 1 || d->ncatoms < 1 ||
        d->icc[0] < 1 || d->icc[1] < 1 || d->icc[2] < 1 || d->natoms > INT32_MAX)
        return set_err(c, DC_ERR_ARG, DC_ER_LIB, "dc_crystal_set: bad descriptor");
    HIPCHK(hipSetDevice(c->device));
    c->natoms = d->natoms;
    c->ncatoms = d->ncatoms;
    c->nscat = d->nscat;
    for (int k = 0; k < 3; ++k) {
        c->icc[k] = d->icc[k];
        c->dim0[k] = d->dim0[k];
        c->dim_lo[k] = d->dim_lo[k];
    }
    c->h_pos.assign(d->cr_pos, d->cr_pos + 3 * d->natoms);
    c->h_iscat.assign(d->cr_iscat, d->cr_iscat + d->natoms);
    int st = 0;
    if ((st = upload(c, c->pos, c->h_pos.data(), c->h_pos.size()))) return st;
    // per-type atom lists in crystal order (LOT_OFF gather, four_getatm :1839-1848);
    // for a DISCUS-ordered crystal (atom = cell * ncatoms + site) crystal order
    // is (cell, site) order, so the same lists serve the lot gather, with the
    // first position of every cell per type in tc_start
    std::vector<std::vector<int>> lists(c->nscat);
    for (int64_t i = 0; i < d->natoms; ++i) {
        int t = d->cr_iscat[i];
        if (t >= 1 && t <= c->nscat) lists[t - 1].push_back((int)i);
    }
    std::vector<int> flat;
    c->type_beg.assign(c->nscat + 1, 0);
    for (int t = 0; t < c->nscat; ++t) {
        c->type_beg[t] = (int)flat.size();
        flat.insert(flat.end(), lists[t].begin(), lists[t].end());
    }
    c->type_beg[c->nscat] = (int)flat.size();
    if ((st = upload(c, c->type_idx, flat.data(), flat.size()))) return st;
    const int64_t ncell = (int64_t)c->icc[0] * c->icc[1] * c->icc[2];
    c->tc_start.clear();
    if (d->natoms == ncell * c->ncatoms && ncell * c->nscat < (int64_t)1 << 28) {
        c->tc_start.assign((size_t)c->nscat * (ncell + 1), 0);
        for (int t = 0; t < c->nscat; ++t) {
            int *cs = c->tc_start.data() + (size_t)t * (ncell + 1);
            int q = 0;
            const std::vector<int> &L = lists[t];
            for (int64_t cell = 0; cell <= ncell; ++cell) {
                while (q < (int)L.size() && L[q] < cell * c->ncatoms) ++q;
                cs[cell] = q;
            }
        }
    }
    HIPCHK(hipStreamSynchronize(c->st));
    c->have_crystal = true;
    return DC_OK;
}

namespace {

// four_getatm for box/ellipsoid lots (fourier_sup.f90:1852-1968): the cells
// of lot `lot` in DISCUS's order (kk outer, ii inner) with periodic wrap, as
// runs of consecutive cells (ii direction) that share a wrap pattern; appends
// the lot's 8 offsets (one per wrap pattern, cr_dim0 + lbeg - 1 - icc where
// wrapped) to offs and returns the number of cells.  The atoms of type t in
// a run of cells are one contiguous range of the type's (cell, site)-ordered
// list (tc_start).  four_getatm's scan of the atoms beyond cr_end
// (:1886-1899) is absent: DISCUS's 'run' refuses lots unless cr_natoms equals
// cells x sites (fourier.f90:644-651, DC_ERR_LOTS here), so that list is
// always empty on this path.
struct CellRun {
    long long cell;   // first cell (0-based crystal cell index)
    int n;            // consecutive cells along ii
    int wrap;         // wrap bits of the run (1: a, 2: b, 4: c)
};

long long lot_cells(const dc_ctx *c, const dc_run_desc *r, int lot, std::vector<CellRun> &cruns,
                    std::vector<double> &offs) {
    const int *lbeg = r->lbeg + 3 * lot;
    double x0[3];
    for (int k = 0; k < 3; ++k) x0[k] = (double)(float)r->ls[k] / 2.0;
    for (int wb = 0; wb < 8; ++wb)
        for (int k = 0; k < 3; ++k)
            offs.push_back(c->dim0[k] + lbeg[k] - 1 - (((wb >> k) & 1) ? (double)c->icc[k] : 0.0));
    long long ncell = 0;
    cruns.clear();
    for (int kk = 0; kk < r->ls[2]; ++kk) {
        int ic2 = kk + lbeg[2], w2 = 0;
        double xt2 = ((double)kk - x0[2] + 0.5) * ((double)kk - x0[2] + 0.5) / (x0[2] * x0[2]);
        if (ic2 > c->icc[2]) { ic2 -= c->icc[2]; w2 = 4; }
        for (int jj = 0; jj < r->ls[1]; ++jj) {
            int ic1 = jj + lbeg[1], w1 = 0;
            double xt1 = ((double)jj - x0[1] + 0.5) * ((double)jj - x0[1] + 0.5) / (x0[1] * x0[1]);
            if (ic1 > c->icc[1]) { ic1 -= c->icc[1]; w1 = 2; }
            const long long row = (long long)(ic2 - 1) * c->icc[0] * c->icc[1] + (long long)(ic1 - 1) * c->icc[0];
            bool open = false;
            for (int ii = 0; ii < r->ls[0]; ++ii) {
                int ic0 = ii + lbeg[0], w0 = 0;
                double xt0 = ((double)ii - x0[0] + 0.5) * ((double)ii - x0[0] + 0.5) / (x0[0] * x0[0]);
                if (ic0 > c->icc[0]) { ic0 -= c->icc[0]; w0 = 1; }
                if (r->ilots == DC_LOT_ELI && !(xt0 + xt1 + xt2 <= 1.0)) {
                    open = false;
                    continue;
                }
                ++ncell;
                const long long cell = row + (ic0 - 1);
                const int wrap = w0 | w1 | w2;
                if (open && cruns.back().wrap == wrap && cruns.back().cell + cruns.back().n == cell) {
                    cruns.back().n++;
                } else {
                    cruns.push_back({cell, 1, wrap});
                    open = true;
                }
            }
        }
    }
    return ncell;
}

// The runs of gathered atoms of type t1 (1-based) over the cell runs of a
// lot; dst is the position of the lot's next atom in the launch's list.
// Returns the number of atoms.
long long lot_type_runs(const dc_ctx *c, const std::vector<CellRun> &cruns, int t1, int obase, long long dst,
                        std::vector<int4> &runs) {
    const long long ncell = (long long)c->icc[0] * c->icc[1] * c->icc[2];
    const int *cs = c->tc_start.data() + (size_t)(t1 - 1) * (ncell + 1);
    long long n = 0;
    for (const CellRun &cr : cruns) {
        const int b = cs[cr.cell], e = cs[cr.cell + cr.n];
        if (e > b) {
            runs.push_back(make_int4((int)(dst + n), c->type_beg[t1 - 1] + b, obase + cr.wrap, 0));
            n += e - b;
        }
    }
    return n;
}

// Run the phase GEMM over gathered segments and leave the slice planes in
// c->partial.  Positions come from the resident crystal (idx/oid/offs) or,
// when explicit != null, from an explicit (n,3) list.  group_seg (ascending
// segment indices, the last one = gl.segs.size()) marks groups of segments
// (the lots of a batch): their slice boundaries go to c->lot_slice, uploaded
// with the slice tables before the launch.
int gemm_segments(dc_ctx *c, const Grid &g, const Plan &pl, GatherList &gl,
                  const double *explicit_pos, SliceSet &ss, int &nslice, bool use_type_lists,
                  const std::vector<int> &group_seg) {
    const GemmGeom &gg = pl.gg;
    const int mapping = pl.mapping;
    int st = 0;
    long long n = 0;
    for (auto &s : gl.segs) n += s.second;
    PrepArgs pa = base_prep(g, mapping);
    pa.n = n;
    if (explicit_pos) {
        if ((st = upload(c, c->xat, explicit_pos, 3 * (size_t)std::max<long long>(n, 1)))) return st;
        pa.src = ptr<double>(c->xat);
        pa.src_ld = 0;
    } else {
        pa.src = ptr<double>(c->pos);
        if (use_type_lists) {
            pa.idx = ptr<int>(c->type_idx);
        } else if (!gl.runs.empty()) {
            if ((st = upload(c, c->runs, gl.runs.data(), gl.runs.size()))) return st;
            if ((st = upload(c, c->offs, gl.offs.data(), gl.offs.size()))) return st;
            pa.idx = ptr<int>(c->type_idx);
            pa.runs = ptr<int4>(c->runs);
            pa.nruns = (int)gl.runs.size();
            pa.offs = ptr<double>(c->offs);
        } else {
            if ((st = upload(c, c->idx, gl.idx.data(), gl.idx.size()))) return st;
            pa.idx = ptr<int>(c->idx);
            if (!gl.oid.empty()) {
                if ((st = upload(c, c->oid, gl.oid.data(), gl.oid.size()))) return st;
                if ((st = upload(c, c->offs, gl.offs.data(), gl.offs.size()))) return st;
                pa.oid = ptr<int>(c->oid);
                pa.offs = ptr<double>(c->offs);
            }
        }
    }
    if ((st = ensure(c, c->rec, sizeof(AtomRec) * std::max<long long>(n, 1)))) return st;
    pa.hash_salt = launch_geom(c, gg).dither;   // run_gemm below launches undamped
    HIPCHK(launch_prep(pa, ptr<AtomRec>(c->rec), c->st));
    c->rec_hash = true;
    int off = 0;
    ss = SliceSet();
    add_segments(ss, gl.segs, slice_len(c, kchunk_for(pl, gl.segs, pl.scheme)), off);
    if (ss.beg.empty()) ss.beg.push_back(0);
    const int ngemm = (int)ss.type.size();
    // schemes with two planes per slice (hi, lo): the reductions see each
    // plane as a slice of the same type
    const int pps = planes_per_slice(pl.scheme);
    if (pps > 1) {
        std::vector<int> t2;
        for (int v : ss.type)
            for (int q = 0; q < pps; ++q) t2.push_back(v);
        ss.type.swap(t2);
        for (int &v : ss.seg_first) v *= pps;
    }
    nslice = ngemm * pps;
    if ((st = upload(c, c->slice_beg, ss.beg.data(), ss.beg.size()))) return st;
    if ((st = upload(c, c->slice_type, ss.type.data(), std::max<size_t>(ss.type.size(), 1)))) return st;
    ss.seg_first.push_back(nslice);
    std::vector<int> gsl;
    for (int v : group_seg) gsl.push_back(ss.seg_first[v]);
    if ((st = upload(c, c->lot_slice, gsl.data(), gsl.size()))) return st;
    if ((st = ensure(c, c->partial, sizeof(float2) * gg.P * std::max(nslice, 1)))) return st;
    return run_gemm(c, pl, ngemm, ptr<float2>(c->partial));
}

}  // namespace

int dc_four_run_dev(dc_ctx *c, const dc_run_desc *r, double *d_csf, double *d_acsf, double *d_dsi,
                    double *d_fmean, double *d_dvar) {
    if (!c) return DC_ERR_ARG;
    c->err.clear();
    c->err_typ = DC_ER_NONE;
    if (!r || !r->cfact || r->ncfact < 1) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "dc_four_run: bad descriptor");
    if (!c->have_crystal) return set_err(c, DC_ERR_STATE, DC_ER_LIB, "dc_four_run: no crystal (dc_crystal_set)");
    if (c->natoms < 1) return set_err(c, DC_ERR_NOATOMS, DC_ER_APPL, "No fourier will be calculated");
    for (int k = 0; k < 3; ++k)
        if (r->inc[k] < 1) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "dc_four_run: inc < 1");
    if (r->ilots != DC_LOT_OFF && r->ilots != DC_LOT_BOX && r->ilots != DC_LOT_ELI)
        return set_err(c, DC_ERR_ARG, DC_ER_LIB, "dc_four_run: bad lot shape");
    const bool lots = r->ilots != DC_LOT_OFF;
    if (lots) {
        if (c->natoms != (int64_t)c->icc[0] * c->icc[1] * c->icc[2] * c->ncatoms)
            return set_err(c, DC_ERR_LOTS, DC_ER_APPL, "Lots are on but the number of atoms differs from the crystal size");
        if (c->tc_start.empty()) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "dc_four_run: crystal too large for lots");
        if (r->nlots < 1 || !r->lbeg) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "dc_four_run: lots need lbeg");
        for (int k = 0; k < 3; ++k)
            if (r->ls[k] < 1 || r->ls[k] > c->icc[k]) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "dc_four_run: lot size outside crystal");
    }
    if (zero_increment(r->vi, r->inc)) return set_err(c, DC_ERR_ZEROVEC, DC_ER_APPL, "Zero increment vector");
    int csize[3];
    for (int k = 0; k < 3; ++k) csize[k] = r->lperiod ? c->icc[k] : std::max(1, c->icc[k] - (lots ? r->ls[k] : 1) + 1);
    if (lots)
        for (int l = 0; l < r->nlots; ++l)
            for (int k = 0; k < 3; ++k)
                if (r->lbeg[3 * l + k] < 1 || r->lbeg[3 * l + k] > csize[k])
                    return set_err(c, DC_ERR_ARG, DC_ER_LIB, "dc_four_run: lot origin outside crystal");
    if (r->fave != 0.0 && !r->sel_cell) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "dc_four_run: fave needs sel_cell");
    if (cancelled(c)) return set_err(c, DC_ERR_CANCEL, DC_ER_COMM, "cancelled");
    HIPCHK(hipSetDevice(c->device));
    stats_begin(c);
    HIPCHK(hipEventRecord(c->ev[0], c->st));

    Grid g{};
    int fnum[3];
    int fried = test_friedel(r->eck, r->vi, r->inc, r->ano, r->four_friedel, fnum);
    for (int k = 0; k < 3; ++k) {
        g.n[k] = r->inc[k];
        g.f[k] = fnum[k];
        g.xm[k] = r->eck[k];
        g.uin[k] = r->vi[k];
        g.vin[k] = r->vi[3 + k];
        g.win[k] = r->vi[6 + k];
    }
    for (int k = 0; k < 9; ++k) g.rten[k] = r->rten[k];
    g.P = (long long)fnum[0] * fnum[1] * fnum[2];
    g.friedel = fried != 0;
    const long long ntot = (long long)g.n[0] * g.n[1] * g.n[2];
    if (max_corner_bin(g) > std::min(CFPKT, r->ncfact - 1))
        return set_err(c, DC_ERR_ISTL, DC_ER_FOUR, "sin(theta)/lambda beyond the form factor table");

    // atoms of one GEMM launch: the crystal, or the lots of a batch
    long long natoms_launch = c->natoms;
    if (lots)
        natoms_launch = std::min<long long>(natoms_launch, (long long)std::max(r->nlots, 1) * r->ls[0] * r->ls[1] *
                                                               r->ls[2] * c->ncatoms);
    // Runs whose result is a catastrophic difference -- the <F> subtraction
    // (fave != 0) or the cross-lot variance -- amplify the relative error of
    // F_l by |F_l| / |F_l - <F>| at the Bragg pixels.  FP32 keeps the GEMM's
    // partial sums in FP64 there (SCHEME_FP32_ACC64): an FP32 running sum's
    // rounding, relative to |F_l|, would stay in the difference.  BF16X3
    // runs the FP16X3 products there (the same MFMA rate): BF16's 8-bit low
    // halves leave about 2^-17 of every term, which the difference amplifies
    // past 1e-5; FP16's 11-bit halves leave 2^-22.
    const bool want_var = d_fmean || d_dvar;
    const bool diff = r->fave != 0.0 || want_var;
    const int scheme = !diff ? c->prec
                       : c->prec == DC_PREC_FP32 ? SCHEME_FP32_ACC64
                       : c->prec == DC_PREC_BF16X3 ? DC_PREC_FP16X3 : c->prec;
    const Plan pl = plan_gemm(g, c->tile, -1, scheme, natoms_launch);
    int st = 0;
    if ((st = upload(c, c->cfact, r->cfact, 2 * (size_t)r->ncfact * c->nscat))) return st;

    // compact-region state: every element is written before it is read (<F>
    // by its reduction, dsi and the moments by the first lot batch's
    // reduction, which starts them from zero), so nothing is cleared
    if ((st = ensure(c, c->acsf, sizeof(double) * 2 * g.P))) return st;
    if ((st = ensure(c, c->dsi, sizeof(double) * g.P))) return st;
    if ((st = ensure(c, c->csf, sizeof(double) * 2 * g.P))) return st;
    if (want_var) {
        if ((st = ensure(c, c->wm, sizeof(double) * 2 * g.P))) return st;
        if ((st = ensure(c, c->wv, sizeof(double) * g.P))) return st;
    }
    const double *acsf = r->fave != 0.0 ? ptr<double>(c->acsf) : nullptr;   // null: <F> = 0

    double pairs = 0.0;
    long long nslices_total = 0;
    int warn = 0;
    // ---- average structure <F> (four_aver, fourier_sup.f90:1194-1357) ----
    if (r->fave != 0.0) {
        GatherList gl;
        // selected cells in four_aver's loop order (ii outer, kk inner, :1250-1252)
        std::vector<long long> cells;
        for (int ii = 1; ii <= csize[0]; ++ii)
            for (int jj = 1; jj <= csize[1]; ++jj)
                for (int kk = 1; kk <= csize[2]; ++kk) {
                    if (!r->sel_cell[(ii - 1) + (long long)csize[0] * ((jj - 1) + (long long)csize[1] * (kk - 1))]) continue;
                    cells.push_back(((long long)(kk - 1) * c->icc[0] * c->icc[1] + (long long)(jj - 1) * c->icc[0] + (ii - 1)));
                    gl.offs.push_back((double)(ii - 1) + c->dim0[0]);
                    gl.offs.push_back((double)(jj - 1) + c->dim0[1]);
                    gl.offs.push_back((double)(kk - 1) + c->dim0[2]);
                }
        const long long nsel = (long long)cells.size();
        for (int t = 1; t <= c->nscat; ++t) {
            long long cnt = 0;
            for (long long q = 0; q < nsel; ++q)
                for (int is = 0; is < c->ncatoms; ++is) {
                    long long ia = cells[q] * c->ncatoms + is;
                    if (ia < c->natoms && c->h_iscat[ia] == t) {
                        gl.idx.push_back((int)ia);
                        gl.oid.push_back((int)q);
                        ++cnt;
                    }
                }
            gl.segs.push_back({t - 1, (int)cnt});
        }
        SliceSet ss;
        int nslice = 0;
        if ((st = gemm_segments(c, g, pl, gl, nullptr, ss, nslice, false, {0, (int)gl.segs.size()}))) return st;
        pairs += (double)gl.idx.size() * (double)ntot;
        nslices_total += nslice;
        ReduceArgs ra{};
        ra.g = g;
        ra.partial = ptr<float2>(c->partial);
        ra.lot_slice = ptr<int>(c->lot_slice);
        ra.slice_type = ptr<int>(c->slice_type);
        ra.nlots = 1;
        ra.lform = 1;
        ra.cfact = ptr<double2>(c->cfact);
        ra.ncfact = r->ncfact;
        HIPCHK(launch_reduce_sum(ra, ptr<double>(c->acsf), 0, c->st));
        // lattice of the lot shape over the full grid (four_getav :2440-2510)
        std::vector<double> lp;
        if (!lots) {
            for (int ii = 0; ii < c->icc[0]; ++ii)
                for (int jj = 0; jj < c->icc[1]; ++jj)
                    for (int kk = 0; kk < c->icc[2]; ++kk) {
                        lp.push_back(ii + c->dim0[0]);
                        lp.push_back(jj + c->dim0[1]);
                        lp.push_back(kk + c->dim0[2]);
                    }
        } else {
            double x0[3];
            for (int k = 0; k < 3; ++k) x0[k] = (double)(float)r->ls[k] / 2.0;
            for (int kk = 0; kk < r->ls[2]; ++kk) {
                double xt2 = ((double)kk - x0[2] + 0.5) * ((double)kk - x0[2] + 0.5) / (x0[2] * x0[2]);
                for (int jj = 0; jj < r->ls[1]; ++jj) {
                    double xt1 = ((double)jj - x0[1] + 0.5) * ((double)jj - x0[1] + 0.5) / (x0[1] * x0[1]);
                    for (int ii = 0; ii < r->ls[0]; ++ii) {
                        double xt0 = ((double)ii - x0[0] + 0.5) * ((double)ii - x0[0] + 0.5) / (x0[0] * x0[0]);
                        if (r->ilots == DC_LOT_ELI && !(xt0 + xt1 + xt2 <= 1.0)) continue;
                        lp.push_back(ii);
                        lp.push_back(jj);
                        lp.push_back(kk);
                    }
                }
            }
        }
        Grid gl_full = g;
        for (int k = 0; k < 3; ++k) gl_full.f[k] = g.n[k];
        gl_full.P = ntot;
        gl_full.friedel = 0;
        const Plan pll = plan_gemm(gl_full, c->tile, -1, scheme, (long long)(lp.size() / 3));
        GatherList gll;
        gll.segs.push_back({0, (int)(lp.size() / 3)});
        if ((st = gemm_segments(c, gl_full, pll, gll, lp.data(), ss, nslice, false, {0, 1}))) return st;
        pairs += (double)(lp.size() / 3) * (double)ntot;
        nslices_total += nslice;
        if ((st = ensure(c, c->lat, sizeof(double) * 2 * ntot))) return st;
        ReduceArgs rl{};
        rl.g = gl_full;
        rl.partial = ptr<float2>(c->partial);
        rl.lot_slice = ptr<int>(c->lot_slice);
        rl.nlots = 1;
        rl.lform = 0;
        HIPCHK(launch_reduce_sum(rl, ptr<double>(c->lat), 0, c->st));
        double norm = nsel > 0 ? 1.0 / (double)nsel : 0.0;
        if (nsel == 0) warn = DC_WARN_NOCELL;
        HIPCHK(launch_aver_combine(g, ptr<double>(c->acsf), ptr<double>(c->lat), norm, c->st));
    }

    // ---- lot loop (fourier_sup.f90:140-257) ----
    long long ncell = (long long)c->icc[0] * c->icc[1] * c->icc[2];
    const int nlots = lots ? r->nlots : 1;
    const double plane_bytes = (double)sizeof(float2) * (double)g.plane * planes_per_slice(scheme);
    int lot = 0;
    while (lot < nlots) {
        if (cancelled(c)) return set_err(c, DC_ERR_CANCEL, DC_ER_COMM, "cancelled");
        GatherList gl;
        std::vector<int> lot_first_seg;
        int l1 = lot;
        long long n_batch = 0;
        std::vector<CellRun> cruns;
        while (l1 < nlots) {
            std::vector<std::pair<int, int>> segs_one;
            std::vector<int4> runs_one;
            std::vector<double> offs8;
            long long nc = ncell;
            if (!lots) {
                for (int t = 0; t < c->nscat; ++t) segs_one.push_back({t, c->type_beg[t + 1] - c->type_beg[t]});
            } else {
                const int obase = (int)(gl.offs.size() / 3);
                nc = lot_cells(c, r, l1, cruns, offs8);
                long long dst = n_batch;
                for (int t = 1; t <= c->nscat; ++t) {
                    const long long n = lot_type_runs(c, cruns, t, obase, dst, runs_one);
                    segs_one.push_back({t - 1, (int)n});
                    dst += n;
                }
            }
            long long n_one = 0;
            for (auto &sg : segs_one) n_one += sg.second;
            const double est_slices = (double)(gl.segs.size() + segs_one.size()) +
                                      (double)(n_batch + n_one) / (double)KCHUNK_MIN;
            if (l1 > lot && est_slices * plane_bytes > PARTIAL_BUDGET) break;
            lot_first_seg.push_back((int)gl.segs.size());
            gl.offs.insert(gl.offs.end(), offs8.begin(), offs8.end());
            gl.runs.insert(gl.runs.end(), runs_one.begin(), runs_one.end());
            gl.segs.insert(gl.segs.end(), segs_one.begin(), segs_one.end());
            n_batch += n_one;
            ncell = nc;
            ++l1;
            if (!lots) break;
        }
        lot_first_seg.push_back((int)gl.segs.size());
        SliceSet ss;
        int nslice = 0;
        if ((st = gemm_segments(c, g, pl, gl, nullptr, ss, nslice, !lots, lot_first_seg))) return st;
        pairs += (double)n_batch * (double)ntot;
        nslices_total += nslice;
        ReduceArgs ra{};
        ra.g = g;
        ra.partial = ptr<float2>(c->partial);
        ra.slice0 = 0;
        ra.lot_slice = ptr<int>(c->lot_slice);
        ra.slice_type = ptr<int>(c->slice_type);
        ra.nlots = l1 - lot;
        ra.lot0 = lot;
        ra.lform = 1;
        ra.cfact = ptr<double2>(c->cfact);
        ra.ncfact = r->ncfact;
        ra.ntype = c->nscat;
        ra.acsf = acsf;
        ra.dsi = ptr<double>(c->dsi);
        ra.csf = ptr<double>(c->csf);
        ra.wm = want_var ? ptr<double>(c->wm) : nullptr;
        ra.wv = want_var ? ptr<double>(c->wv) : nullptr;
        ra.fresh = lot == 0;
        hipEvent_t r0, r1;
        if ((st = mark(c, r0))) return st;
        HIPCHK(launch_reduce_lots(ra, c->st));
        if ((st = mark(c, r1))) return st;
        span(c, r0, r1, &c->stats.reduce_ms);
        lot = l1;
    }

    // ---- normalisation, Friedel mirror, outputs (fourier_sup.f90:256-282) ----
    const double dnorm = nlots != 1 ? (double)((long long)c->icc[0] * c->icc[1] * c->icc[2]) / (double)(ncell * nlots) : 1.0;
    FinalArgs fa{};
    fa.g = g;
    fa.ntot = ntot;
    fa.csf = ptr<double>(c->csf);
    fa.acsf = acsf;
    fa.dsi = ptr<double>(c->dsi);
    fa.wm = want_var ? ptr<double>(c->wm) : nullptr;
    fa.wv = want_var ? ptr<double>(c->wv) : nullptr;
    fa.csf_out = d_csf;
    fa.acsf_out = d_acsf;
    fa.dsi_out = d_dsi;
    fa.fmean_out = d_fmean;
    fa.dvar_out = d_dvar;
    fa.sq = std::sqrt(r->diff_scale);
    fa.dscale = dnorm * r->diff_scale;
    fa.vscale = dnorm * r->diff_scale;
    HIPCHK(launch_finalize(fa, c->st));
    HIPCHK(hipEventRecord(c->ev[1], c->st));
    HIPCHK(hipEventSynchronize(c->ev[1]));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    c->stats.total_ms = ms;
    if ((st = resolve_spans(c))) return st;
    c->stats.pairs = pairs;
    c->stats.slices = nslices_total;
    c->stats.gemm_scheme = scheme;
    return warn;
}

int dc_four_run(dc_ctx *c, const dc_run_desc *r, double *csf, double *acsf, double *dsi, double *fmean,
                double *dvar) {
    if (!c || !r) return DC_ERR_ARG;
    const long long ntot = (long long)std::max(r->inc[0], 1) * std::max(r->inc[1], 1) * std::max(r->inc[2], 1);
    double *host[5] = {csf, acsf, dsi, fmean, dvar};
    const int width[5] = {2, 2, 1, 2, 1};
    double *dev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    int st = 0;
    for (int k = 0; k < 5; ++k)
        if (host[k]) {
            if ((st = ensure(c, c->out[k], sizeof(double) * width[k] * ntot))) return st;
            dev[k] = ptr<double>(c->out[k]);
        }
    st = dc_four_run_dev(c, r, dev[0], dev[1], dev[2], dev[3], dev[4]);
    if (st < 0) return st;
    for (int k = 0; k < 5; ++k)
        if (host[k]) HIPCHK(hipMemcpyAsync(host[k], dev[k], sizeof(double) * width[k] * ntot, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return st;
}

}  // extern "C"
