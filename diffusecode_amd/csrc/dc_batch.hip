// dc_batch.hip -- population scoring: chi^2 (and its gradient) of many trial
// structures against their own target grids in one pass.
//
// DIFFEV evaluates a generation by running one DISCUS macro per member over
// its MPI farm (diffev/prog/diffev_kdo.f90:676, lib_f90/with_mpi.f90:281-609).
// Here every member's atoms become GEMM slices of ONE persistent phase-GEMM
// launch (uneven member sizes are balanced by slicing), one fused reduction
// forms F, I, chi^2 and the adjoint's Q planes for all members, and one
// adjoint launch gives the gradients.  The whole step is captured once per
// batch shape as a hipGraph and replayed for every later generation; the
// GEMM tile shape and persistent grid size are chosen by timing the
// candidates on the batch itself (autotuner).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "dc_host.h"

using namespace dc;

struct BatchState {
    // key: the shape of the batch
    int nmembers = 0, nscat = 0, ncfact = 0, prec = -1, tile_forced = -2, autotune = -1, graphs = -1;
    int slice_atoms = -1, dither = -1;
    bool want_grad = false, want_icalc = false, have_weight = false, any_b = false;
    std::vector<int64_t> atom_beg;
    std::vector<int> type;
    double eck[12] = {}, vi[9] = {};
    int inc[3] = {0, 0, 0};
    // plan
    Grid g{};
    Plan pl;
    int nwg = 0;
    std::vector<int> order;
    std::vector<std::pair<int, int>> segs;   // (type, atoms) per member and type, member-major
    std::vector<int> seg_member;
    int kchunk = 0;                          // atoms per GEMM slice
    int nsorted = 0, nslice = 0, ntiles = 0, ntc = 0, nblk = 0;
    int walk = 1, nset = 0;              // adjoint column tiles per work-group, partial sets
    int walk_req = 0;                    // the context's adjoint walk setting at planning
    int64_t natoms = 0;
    DevBuf pos, occ, bdw, idx, rec, sbeg, stype, ssq, msl, sorder, partial, cfact, rten, sq, target, weight, icalc,
        chi, q, qs, atiles, part, grad, wq;
    bool queue = false;                  // GEMM items from a device counter (autotuner's choice)
    bool primed = false;                 // records built and autotuned once
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    float tune_best_ms = 0.f;

    std::vector<DevBuf *> buffers() {
        return {&pos, &occ, &bdw, &idx, &rec, &sbeg, &stype, &ssq, &msl, &sorder, &partial, &cfact, &rten, &sq,
                &target, &weight, &icalc, &chi, &q, &qs, &atiles, &part, &grad, &wq};
    }
};

namespace dc {

void batch_release(dc_ctx *c) {
    BatchState *b = c->batch;
    if (!b) return;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (DevBuf *d : b->buffers())
        if (d->p) (void)hipFree(d->p);
    delete b;
    c->batch = nullptr;
}

}  // namespace dc

namespace {

double dotg(const double *a, const double *g, const double *b) {
    double s = 0.0;
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) s += a[r] * g[r + 3 * k] * b[k];
    return s;
}

bool same_shape(const BatchState *b, const dc_ctx *c, const dc_batch_desc *d, bool want_grad, bool want_icalc,
                bool have_weight, bool any_b) {
    if (!b) return false;
    if (b->nmembers != d->nmembers || b->nscat != d->nscat || b->ncfact != d->ncfact || b->prec != c->prec ||
        b->tile_forced != c->tile || b->autotune != c->autotune || b->graphs != c->graphs ||
        b->want_grad != want_grad || b->want_icalc != want_icalc || b->have_weight != have_weight ||
        b->any_b != any_b || b->walk_req != c->adj_walk || b->slice_atoms != c->slice_atoms || b->dither != c->dither)
        return false;
    if (std::memcmp(b->eck, d->eck, sizeof(b->eck)) || std::memcmp(b->vi, d->vi, sizeof(b->vi)) ||
        std::memcmp(b->inc, d->inc, sizeof(b->inc)))
        return false;
    if (!std::equal(b->atom_beg.begin(), b->atom_beg.end(), d->atom_beg)) return false;
    return std::equal(b->type.begin(), b->type.end(), d->type);
}

int check_desc(dc_ctx *c, const dc_batch_desc *d) {
    if (!d || d->nmembers < 1 || !d->atom_beg || !d->pos || !d->type || !d->occ || !d->cfact || !d->rten ||
        d->nscat < 1 || d->ncfact < 1)
        return set_err(c, DC_ERR_ARG, DC_ER_LIB, "batch: null pointer or bad size");
    if (d->atom_beg[0] != 0) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "batch: atom_beg[0] must be 0");
    for (int m = 0; m < d->nmembers; ++m)
        if (d->atom_beg[m + 1] < d->atom_beg[m]) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "batch: atom_beg decreasing");
    const int64_t n = d->atom_beg[d->nmembers];
    if (n > INT32_MAX / 8) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "batch: too many atoms");
    for (int64_t k = 0; k < n; ++k)
        if (d->type[k] < 0 || d->type[k] > d->nscat) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "batch: type outside 0..nscat");
    for (int k = 0; k < 3; ++k)
        if (d->inc[k] < 1) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "batch: bad inc");
    return 0;
}

// s^2 row / column parts for every member's metric; errors on grids that are
// not metric-orthogonal between rows and columns when B or gradients need it.
int member_sq(dc_ctx *c, const BatchState &b, const dc_batch_desc *d, std::vector<float> &sq, bool need_sep) {
    const GemmGeom &gg = b.pl.gg;
    const int stride = gg.R + gg.C;
    sq.assign((size_t)stride * d->nmembers, 0.0f);
    const double *v1 = d->vi, *v2 = d->vi + 3, *v3 = d->vi + 6;
    for (int m = 0; m < d->nmembers; ++m) {
        const double *G = d->rten + 9 * m;
        if (need_sep) {
            const double n1 = std::sqrt(std::fabs(dotg(v1, G, v1)));
            for (int ax = 1; ax < 3; ++ax) {
                if (d->inc[ax] <= 1) continue;
                const double *v = d->vi + 3 * ax;
                const double nv = std::sqrt(std::fabs(dotg(v, G, v)));
                if (std::fabs(dotg(v1, G, v)) > 1e-12 * std::max(1.0, n1 * nv))
                    return set_err(c, DC_ERR_ARG, DC_ER_LIB, "batch: vi(:,1) not metric-orthogonal to vi(:,2), vi(:,3)");
            }
        }
        float *s = sq.data() + (size_t)stride * m;
        for (int i = 0; i < gg.R; ++i) {
            double h[3];
            for (int k = 0; k < 3; ++k) h[k] = d->eck[k] + i * v1[k];
            s[i] = (float)(dotg(h, G, h) / 4.0);
        }
        for (int col = 0; col < gg.C; ++col) {
            const int j = col % gg.C0, l = col / gg.C0;
            if (j >= b.g.f[1]) continue;
            double v[3];
            for (int k = 0; k < 3; ++k) v[k] = j * v2[k] + l * v3[k];
            s[gg.R + col] = (float)((dotg(v, G, v) + 2.0 * dotg(v, G, d->eck)) / 4.0);
        }
        Grid gm = b.g;
        for (int k = 0; k < 9; ++k) gm.rten[k] = G[k];
        if (max_corner_bin(gm) >= d->ncfact) return set_err(c, DC_ERR_ISTL, DC_ER_FOUR, "sin(theta)/lambda beyond table");
    }
    return 0;
}

// GEMM slices of at most kchunk atoms over the member-major, type-sorted
// segments: slice tables, each member's slice range and metric offset, the
// launch order and the plane buffer.
int make_slices(dc_ctx *c, BatchState &b, int kchunk) {
    const GemmGeom &gg = b.pl.gg;
    SliceSet ss;
    std::vector<int> mslice, slice_sq;
    int off = 0;
    size_t sg = 0;
    for (int m = 0; m < b.nmembers; ++m) {
        mslice.push_back((int)ss.type.size());
        for (; sg < b.segs.size() && b.seg_member[sg] == m; ++sg) {
            add_segments(ss, {b.segs[sg]}, kchunk, off);
            while ((int)slice_sq.size() < (int)ss.type.size()) slice_sq.push_back(m * (gg.R + gg.C));
        }
    }
    mslice.push_back((int)ss.type.size());
    if (ss.beg.empty()) ss.beg.push_back(0);
    b.nslice = (int)ss.type.size();
    b.kchunk = kchunk;
    int st;
    if ((st = upload(c, b.sbeg, ss.beg.data(), ss.beg.size()))) return st;
    if ((st = upload(c, b.stype, ss.type.data(), ss.type.size()))) return st;
    if ((st = upload(c, b.ssq, slice_sq.data(), slice_sq.size()))) return st;
    if ((st = upload(c, b.msl, mslice.data(), mslice.size()))) return st;
    // GEMM launch order of the slices: longest first.  Members of uneven size
    // give slices of uneven length, and in member order a long slice can
    // start in the launch's last round of work-groups and run alone
    std::vector<int> sorder(std::max(b.nslice, 1), 0);
    for (int s = 0; s < b.nslice; ++s) sorder[s] = s;
    std::stable_sort(sorder.begin(), sorder.begin() + b.nslice,
                     [&](int x, int y) { return ss.beg[x + 1] - ss.beg[x] > ss.beg[y + 1] - ss.beg[y]; });
    if ((st = upload(c, b.sorder, sorder.data(), sorder.size()))) return st;
    return ensure(c, b.partial, sizeof(float2) * gg.P * std::max(b.nslice, 1));
}

int build_state(dc_ctx *c, const dc_batch_desc *d, BatchState &b) {
    b.nmembers = d->nmembers;
    b.nscat = d->nscat;
    b.ncfact = d->ncfact;
    b.prec = c->prec;
    b.walk_req = c->adj_walk;
    b.tile_forced = c->tile;
    b.slice_atoms = c->slice_atoms;
    b.dither = c->dither;
    b.autotune = c->autotune;
    b.graphs = c->graphs;
    b.natoms = d->atom_beg[d->nmembers];
    b.atom_beg.assign(d->atom_beg, d->atom_beg + d->nmembers + 1);
    b.type.assign(d->type, d->type + b.natoms);
    std::memcpy(b.eck, d->eck, sizeof(b.eck));
    std::memcpy(b.vi, d->vi, sizeof(b.vi));
    std::memcpy(b.inc, d->inc, sizeof(b.inc));
    Grid &g = b.g;
    for (int k = 0; k < 3; ++k) {
        g.n[k] = g.f[k] = d->inc[k];
        g.xm[k] = d->eck[k];
        g.uin[k] = d->vi[k];
        g.vin[k] = d->vi[3 + k];
        g.win[k] = d->vi[6 + k];
    }
    g.P = (long long)g.f[0] * g.f[1] * g.f[2];
    b.pl = plan_gemm(g, c->tile, 0, c->prec, d->atom_beg[d->nmembers]);
    const GemmGeom &gg = b.pl.gg;

    // member-major, type-sorted atom order; adjoint tiles; slices
    std::vector<int> tiles;
    std::vector<std::pair<int, int>> &segs_all = b.segs;
    std::vector<int> &seg_member = b.seg_member;
    segs_all.clear();
    seg_member.clear();
    b.order.clear();
    for (int m = 0; m < d->nmembers; ++m) {
        std::vector<int> cnt(d->nscat + 1, 0);
        for (int64_t k = d->atom_beg[m]; k < d->atom_beg[m + 1]; ++k) cnt[d->type[k]]++;
        for (int t = 1; t <= d->nscat; ++t) {
            if (!cnt[t]) continue;
            for (int64_t k = d->atom_beg[m]; k < d->atom_beg[m + 1]; ++k)
                if (d->type[k] == t) b.order.push_back((int)k);
            segs_all.push_back({t - 1, cnt[t]});
            seg_member.push_back(m);
        }
    }
    b.nsorted = (int)b.order.size();
    int off = 0;
    const int na = adjoint_tile_atoms();
    for (size_t sg = 0; sg < segs_all.size(); ++sg) {
        const int m = seg_member[sg];
        for (int k = 0; k < segs_all[sg].second; k += na)
            tiles.insert(tiles.end(), {off + k, off + std::min(segs_all[sg].second, k + na),
                                       m * d->nscat + segs_all[sg].first, m});
        off += segs_all[sg].second;
    }
    b.ntiles = (int)tiles.size() / 4;
    b.ntc = (gg.C + adjoint_tile_cols() - 1) / adjoint_tile_cols();
    {
        AdjArgs pa{};
        pa.R = gg.R;
        pa.C = gg.C;
        b.nset = adjoint_plan(pa, b.ntiles, b.prec != DC_PREC_FP32, c->adj_walk);
        b.walk = pa.ctpg;
    }
    b.nblk = (int)std::min<long long>((g.P + 255) / 256, 256);

    int st;
    const size_t P = (size_t)g.P, M = (size_t)d->nmembers, T = (size_t)std::max<int64_t>(b.natoms, 1);
    if ((st = upload(c, b.idx, b.order.data(), b.order.size()))) return st;
    if ((st = make_slices(c, b, slice_len(c, kchunk_for(b.pl, b.nsorted, b.prec))))) return st;
    if ((st = upload(c, b.atiles, tiles.data(), tiles.size()))) return st;
    if ((st = ensure(c, b.pos, sizeof(double) * 3 * T))) return st;
    if ((st = ensure(c, b.occ, sizeof(double) * T))) return st;
    if ((st = ensure(c, b.bdw, sizeof(double) * T))) return st;
    if ((st = ensure(c, b.rec, sizeof(AtomRec) * std::max(b.nsorted, 1)))) return st;
    if ((st = ensure(c, b.cfact, sizeof(double) * 2 * (size_t)d->ncfact * d->nscat * M))) return st;
    if ((st = ensure(c, b.rten, sizeof(double) * 9 * M))) return st;
    if ((st = ensure(c, b.sq, sizeof(float) * (size_t)(gg.R + gg.C) * M))) return st;
    if ((st = ensure(c, b.target, sizeof(double) * P * M))) return st;
    if (b.have_weight && (st = ensure(c, b.weight, sizeof(double) * P * M))) return st;
    if (b.want_icalc && (st = ensure(c, b.icalc, sizeof(double) * P * M))) return st;
    if ((st = ensure(c, b.chi, sizeof(double) * (size_t)b.nblk * M))) return st;
    if (b.want_grad) {
        if ((st = ensure(c, b.q, sizeof(float2) * gg.P * d->nscat * M))) return st;
        // padded columns of the Q planes are never written: zero them once
        HIPCHK(hipMemsetAsync(b.q.p, 0, sizeof(float2) * gg.P * d->nscat * M, c->st));
        if ((st = ensure(c, b.part, sizeof(float) * 6 * (size_t)b.ntc * std::max(b.nsorted, 1)))) return st;
        if (c->prec != DC_PREC_FP32) {
            const size_t rg = (size_t)((gg.R + 15) / 16) * 4;
            if ((st = ensure(c, b.qs, sizeof(uint32_t) * 8 * rg * gg.C * d->nscat * M))) return st;
        }
        if ((st = ensure(c, b.grad, sizeof(double) * 5 * T))) return st;
    }
    return 0;
}

// The reduction of the slice planes: every member's I, its chi^2 partial sums
// and, when gradients are wanted, the Q planes.
int reduce_planes(dc_ctx *c, BatchState &b) {
    const GemmGeom &gg = b.pl.gg;
    BatchReduceArgs ra{};
    ra.g = b.g;
    ra.R = gg.R;
    ra.C = gg.C;
    ra.C0 = gg.C0;
    ra.plane = gg.P;
    ra.partial = ptr<float2>(b.partial);
    ra.mslice = ptr<int>(b.msl);
    ra.slice_type = ptr<int>(b.stype);
    ra.cfact = ptr<double2>(b.cfact);
    ra.ncfact = b.ncfact;
    ra.nscat = b.nscat;
    ra.nmembers = b.nmembers;
    ra.rten = ptr<double>(b.rten);
    ra.target = ptr<double>(b.target);
    ra.weight = b.have_weight ? ptr<double>(b.weight) : nullptr;
    ra.icalc = b.want_icalc ? ptr<double>(b.icalc) : nullptr;
    ra.chi_part = ptr<double>(b.chi);
    ra.q = b.want_grad ? ptr<float2>(b.q) : nullptr;
    HIPCHK(launch_batch_reduce(ra, b.nblk, c->st));
    return 0;
}

// Kernels of one population step, in stream order (captured into the graph).
// rest: 0 = records (prep) and GEMM only, 1 = and the reduction (I, chi^2,
// Q planes), 2 = and the adjoint when gradients are wanted.
__global__ void zero_counter_kernel(int *p) {
    if (threadIdx.x == 0) *p = 0;
}

int enqueue(dc_ctx *c, BatchState &b, int tile, int nwg, bool prep, int rest, bool queue) {
    const GemmGeom &gg = b.pl.gg;
    if (prep) {
        PrepArgs pa = base_prep(b.g, 0);
        pa.src = ptr<double>(b.pos);
        pa.src_ld = 0;
        pa.idx = ptr<int>(b.idx);
        pa.occ_src = ptr<double>(b.occ);
        pa.bdw_src = b.any_b ? ptr<double>(b.bdw) : nullptr;
        pa.n = b.nsorted;
        if (b.nsorted > 0) HIPCHK(launch_prep(pa, ptr<AtomRec>(b.rec), c->st));
    }
    GemmGeom lg = launch_geom(c, gg);
    lg.order = ptr<int>(b.sorder);
    if (queue) {
        int st = ensure(c, b.wq, 256);
        if (st) return st;
        zero_counter_kernel<<<1, 64, 0, c->st>>>(ptr<int>(b.wq));
        HIPCHK(hipGetLastError());
        lg.queue = ptr<int>(b.wq);
    }
    HIPCHK(launch_phase_gemm(c->prec, tile, ptr<AtomRec>(b.rec), ptr<int>(b.sbeg), ptr<float2>(b.partial),
                             b.any_b ? ptr<float>(b.sq) : nullptr, lg, 0, b.nslice, c->st, nwg, ptr<int>(b.ssq)));
    if (rest < 1) return 0;
    int st = reduce_planes(c, b);
    if (st) return st;
    if (rest >= 2 && b.want_grad) {
        HIPCHK(launch_fill_zero(ptr<double>(b.grad), 5 * b.natoms, c->st));
        AdjArgs aa{};
        aa.rec = ptr<AtomRec>(b.rec);
        aa.tiles = ptr<int4>(b.atiles);
        aa.q = ptr<float2>(b.q);
        aa.sq = ptr<float>(b.sq);
        aa.sq_stride = gg.R + gg.C;
        aa.part = ptr<float>(b.part);
        aa.R = gg.R;
        aa.C = gg.C;
        aa.C0 = gg.C0;
        aa.f1 = b.g.f[1];
        aa.plane = gg.P;
        aa.natoms = b.nsorted;
        aa.rc = 0.5f * (float)(b.g.f[0] - 1);
        aa.jc = 0.5f * (float)(b.g.f[1] - 1);
        aa.mc = 0.5f * (float)(b.g.f[2] - 1);
        const bool split = b.prec != DC_PREC_FP32;
        if (split) {
            aa.Rg = (gg.R + 15) / 16 * 4;
            aa.qs = ptr<uint32_t>(b.qs);
            HIPCHK(launch_q_split(ptr<float2>(b.q), b.nmembers * b.nscat, gg.R, gg.C, gg.P, aa.Rg, ptr<uint32_t>(b.qs), c->st));
        }
        aa.ctpg = b.walk;
        HIPCHK(launch_adjoint(aa, b.ntiles, c->st, split));
        AdjFinishArgs fa{};
        fa.part = ptr<float>(b.part);
        fa.ntile_c = b.nset;
        fa.natoms = b.nsorted;
        fa.idx = ptr<int>(b.idx);
        fa.occ = ptr<double>(b.occ);
        for (int k = 0; k < 3; ++k)
            fa.hc[k] = b.eck[k] + (double)aa.rc * b.vi[k] + (double)aa.jc * b.vi[3 + k] + (double)aa.mc * b.vi[6 + k];
        for (int k = 0; k < 9; ++k) fa.v[k] = b.vi[k];
        fa.grad = ptr<double>(b.grad);
        HIPCHK(launch_adjoint_finish(fa, c->st));
    }
    return 0;
}

// Time every tile configuration on this batch at the planned slice length,
// each with one work-group per item, #CU or 2 #CU persistent work-groups
// (static item assignment), or #CU work-groups taking items from a device
// counter (longest first, each as it finishes one: the cfg 5 population's
// GEMM 2.49 -> 2.43 ms and a steadier step); then the best tile at half and
// at twice that length, with the same four launch forms.  A candidate costs its GEMM plus the
// reduction of its slice planes, whose number the slice length sets: longer
// slices mean fewer planes to write and reduce, shorter ones more, smaller
// work items to balance over the CUs.  Another length must beat the planned
// one by 2%, more than the spread of these short timings: the slice length
// changes the FP32 summation order, so a choice made on noise would change
// the results' last bits from one context to the next for nothing (on the
// 96-member population the planned length wins).
int autotune(dc_ctx *c, BatchState &b) {
    const bool fp32_only = b.prec == DC_PREC_FP32;    // tiles built without an FP32 variant are skipped
    if (c->tile >= 0 && !(fp32_only && gemm_tile_info(c->tile).no_fp32)) b.pl.tile = c->tile;
    b.nwg = 0;
    if (c->tile >= 0 || !c->autotune || b.nslice == 0) return 0;
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, c->device));
    const int ncu = prop.multiProcessorCount;
    // mean of 2 launches after a warm-up: the GEMM of (tile, nwg), or the reduction
    auto time_ms = [&](int t, int nwg, bool reduce, float &ms) -> int {
        const bool qu = nwg < 0;
        const int n = qu ? -nwg : nwg;
        int st = reduce ? reduce_planes(c, b) : enqueue(c, b, t, n, false, 0, qu);
        if (st) return st;
        HIPCHK(hipEventRecord(c->ev[2], c->st));
        for (int r = 0; r < 2; ++r)
            if ((st = reduce ? reduce_planes(c, b) : enqueue(c, b, t, n, false, 0, qu))) return st;
        HIPCHK(hipEventRecord(c->ev[3], c->st));
        HIPCHK(hipEventSynchronize(c->ev[3]));
        HIPCHK(hipEventElapsedTime(&ms, c->ev[2], c->ev[3]));
        ms *= 0.5f;
        return 0;
    };
    float best = 1e30f, best_gemm = 0.f, red = 0.f;
    int bt = b.pl.tile, bn = 0, bk = b.kchunk;
    int st = time_ms(0, 0, true, red);
    if (st) return st;
    // candidates: tiles whose padded work is within 1.5x of the least padded one
    double wmin = 1e300;
    for (int t = 0; t < gemm_num_tiles(); ++t) {
        const TileInfo ti = gemm_tile_info(t);
        if (fp32_only && ti.no_fp32) continue;
        wmin = std::min(wmin, (double)padded(b.pl.gg.R, ti.rt) * (double)padded(b.pl.gg.C, ti.ct));
    }
    for (int t = 0; t < gemm_num_tiles(); ++t) {
        const TileInfo ti = gemm_tile_info(t);
        if (fp32_only && ti.no_fp32) continue;
        if ((double)padded(b.pl.gg.R, ti.rt) * (double)padded(b.pl.gg.C, ti.ct) > 1.5 * wmin) continue;
        for (int nwg : {0, ncu, 2 * ncu, -ncu}) {
            float g = 0.f;
            if ((st = time_ms(t, nwg, false, g))) return st;
            if (g + red < best) {
                best = g + red;
                best_gemm = g;
                bt = t;
                bn = nwg;
            }
        }
    }
    // slice length, on the best tile (a forced length is kept)
    if (c->slice_atoms <= 0) {
        const int k0 = b.kchunk;
        for (int kc : {(k0 / 2 + 15) / 16 * 16, 2 * k0}) {
            if (kc < 256 || kc == k0) continue;
            if ((st = make_slices(c, b, kc))) return st;
            if ((st = time_ms(0, 0, true, red))) return st;
            for (int nwg : {0, ncu, 2 * ncu, -ncu}) {
                float g = 0.f;
                if ((st = time_ms(bt, nwg, false, g))) return st;
                if (g + red < (bk == k0 ? 0.98f : 1.0f) * best) {
                    best = g + red;
                    best_gemm = g;
                    bn = nwg;
                    bk = kc;
                }
            }
        }
        if (b.kchunk != bk && (st = make_slices(c, b, bk))) return st;
    }
    b.pl.tile = bt;
    b.queue = bn < 0;
    b.nwg = bn < 0 ? -bn : bn;
    b.tune_best_ms = best_gemm;
    return 0;
}

}  // namespace

extern "C" {

int dc_set_autotune(dc_ctx *c, int on) {
    if (!c) return DC_ERR_ARG;
    c->autotune = on ? 1 : 0;
    return DC_OK;
}

int dc_set_graphs(dc_ctx *c, int on) {
    if (!c) return DC_ERR_ARG;
    c->graphs = on ? 1 : 0;
    return DC_OK;
}

int dc_batch_info(const dc_ctx *c, int *tile, int *nwg, int *graph, double *tune_ms) {
    if (!c || !c->batch) return DC_ERR_STATE;
    if (tile) *tile = c->batch->pl.tile;
    if (nwg) *nwg = c->batch->queue ? -c->batch->nwg : c->batch->nwg;
    if (graph) *graph = c->batch->exec != nullptr;
    if (tune_ms) *tune_ms = c->batch->tune_best_ms;
    return DC_OK;
}

}  // extern "C"

namespace {

// One population step.  dev = false: pos/occ/bdw of d, target, weight, icalc
// and grad are host arrays; dev = true: they are device arrays (the inputs
// are copied on the device into the buffers the graph reads, the outputs
// out of them), and a non-null bdw counts as per-atom B.
int batch_chi2(dc_ctx *c, const dc_batch_desc *d, const double *target, const double *weight, double *chi2,
               double *icalc, double *grad, bool dev) {
    if (!c) return DC_ERR_ARG;
    c->err.clear();
    c->err_typ = DC_ER_NONE;
    int st;
    if ((st = check_desc(c, d))) return st;
    if (!target || !chi2) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "batch: null target or chi2");
    if (cancelled(c)) return set_err(c, DC_ERR_CANCEL, DC_ER_COMM, "cancelled");
    HIPCHK(hipSetDevice(c->device));
    stats_begin(c);
    const int64_t natoms = d->atom_beg[d->nmembers];
    const hipMemcpyKind in = dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    const hipMemcpyKind out = dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    bool any_b = dev && d->bdw != nullptr;
    if (d->bdw && !dev)
        for (int64_t k = 0; k < natoms && !any_b; ++k) any_b = d->type[k] > 0 && d->bdw[k] != 0.0;
    const bool want_grad = grad != nullptr, want_icalc = icalc != nullptr, have_weight = weight != nullptr;
    if (!same_shape(c->batch, c, d, want_grad, want_icalc, have_weight, any_b)) {
        batch_release(c);
        c->batch = new BatchState();
        c->batch->want_grad = want_grad;
        c->batch->want_icalc = want_icalc;
        c->batch->have_weight = have_weight;
        c->batch->any_b = any_b;
        if ((st = build_state(c, d, *c->batch))) {
            batch_release(c);
            return st;
        }
    }
    BatchState &b = *c->batch;
    const size_t P = (size_t)b.g.P, M = (size_t)d->nmembers;
    std::vector<float> sq;
    if ((st = member_sq(c, b, d, sq, any_b || want_grad))) return st;
    HIPCHK(hipEventRecord(c->ev[0], c->st));
    HIPCHK(hipMemcpyAsync(b.pos.p, d->pos, sizeof(double) * 3 * natoms, in, c->st));
    HIPCHK(hipMemcpyAsync(b.occ.p, d->occ, sizeof(double) * natoms, in, c->st));
    if (any_b) HIPCHK(hipMemcpyAsync(b.bdw.p, d->bdw, sizeof(double) * natoms, in, c->st));
    HIPCHK(hipMemcpyAsync(b.cfact.p, d->cfact, sizeof(double) * 2 * d->ncfact * d->nscat * M, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(b.rten.p, d->rten, sizeof(double) * 9 * M, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(b.sq.p, sq.data(), sizeof(float) * sq.size(), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(b.target.p, target, sizeof(double) * P * M, in, c->st));
    if (weight) HIPCHK(hipMemcpyAsync(b.weight.p, weight, sizeof(double) * P * M, in, c->st));

    if (!b.primed) {
        // records of this call, then the tile / grid-size choice, then the graph
        if ((st = enqueue(c, b, b.pl.tile, 0, true, 0, false))) return st;
        if ((st = autotune(c, b))) return st;
        b.primed = true;
        if (c->graphs) {
            HIPCHK(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
            st = enqueue(c, b, b.pl.tile, b.nwg, true, 2, b.queue);
            hipGraph_t graph = nullptr;
            hipError_t e = hipStreamEndCapture(c->st, &graph);
            if (st) {
                if (graph) (void)hipGraphDestroy(graph);
                return st;
            }
            HIPCHK(e);
            b.graph = graph;
            HIPCHK(hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0));
        }
    }
    HIPCHK(hipEventRecord(c->ev[4], c->st));
    if (b.exec) {
        HIPCHK(hipGraphLaunch(b.exec, c->st));
    } else {
        if ((st = enqueue(c, b, b.pl.tile, b.nwg, true, 2, b.queue))) return st;
    }
    HIPCHK(hipEventRecord(c->ev[5], c->st));
    std::vector<double> part((size_t)b.nblk * M);
    HIPCHK(hipMemcpyAsync(part.data(), b.chi.p, sizeof(double) * part.size(), hipMemcpyDeviceToHost, c->st));
    if (icalc) HIPCHK(hipMemcpyAsync(icalc, b.icalc.p, sizeof(double) * P * M, out, c->st));
    if (grad) HIPCHK(hipMemcpyAsync(grad, b.grad.p, sizeof(double) * 5 * natoms, out, c->st));
    HIPCHK(hipEventRecord(c->ev[1], c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    for (size_t m = 0; m < M; ++m) {
        double s = 0.0;
        for (int k = 0; k < b.nblk; ++k) s += part[m * b.nblk + k];
        chi2[m] = s;
    }
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev[4], c->ev[5]));
    c->stats.forward_ms = ms;          // the whole population step (forward + adjoint when requested)
    HIPCHK(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    c->stats.total_ms = ms;
    c->stats.pairs = (double)b.nsorted * (double)P;
    c->stats.slices = b.nslice;
    c->stats.gemm_launches = 1;
    c->stats.gemm_scheme = b.prec;
    c->stats.gemm_ms = b.tune_best_ms;
    c->stats.adjoint_walk = want_grad ? b.walk : 0;
    return DC_OK;
}

}  // namespace

extern "C" {

int dc_batch_chi2(dc_ctx *c, const dc_batch_desc *d, const double *target, const double *weight, double *chi2,
                  double *icalc, double *grad) {
    return batch_chi2(c, d, target, weight, chi2, icalc, grad, false);
}

int dc_batch_chi2_dev(dc_ctx *c, const dc_batch_desc *d, const double *target, const double *weight, double *chi2,
                      double *icalc, double *grad) {
    return batch_chi2(c, d, target, weight, chi2, icalc, grad, true);
}

}  // extern "C"
