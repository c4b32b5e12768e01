// dc_refine.hip -- C ABI of the refinement model: intensity of a structure
// with per-atom occupancies and displacement parameters, and the reverse-mode
// derivative of a loss of that intensity (include/diffuse_turbo.h).
//
// REFINE (refine/prog/refine_run.f90:491-600) differentiates the Fourier
// calculation by re-running it once per parameter; the adjoint here costs a
// small multiple of one forward pass for all 5N parameters together.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "dc_host.h"

using namespace dc;

namespace {

constexpr int CHI2_BLOCKS = 1024;
constexpr double JVP_PARTIAL_BUDGET = 4.0e9;   // bytes of slice planes per batch of directions

struct ModelPlan {
    Grid g{};
    Plan pl;
    std::vector<int> order;                   // sorted atom -> caller index
    std::vector<std::pair<int, int>> segs;    // (type 0-based, count) in sorted order
    std::vector<float> sq;                    // s^2 row parts (R) then column parts (C)
    bool separable = true;
    bool any_b = false;
    int nsorted = 0;
};

double dotg(const double *a, const double *g, const double *b) {
    double s = 0.0;
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) s += a[r] * g[r + 3 * k] * b[k];
    return s;
}

int check_model(dc_ctx *c, const dc_model_desc *m) {
    if (!m || !m->pos || !m->type || !m->occ || !m->cfact || m->natoms < 0 || m->nscat < 1 || m->ncfact < 1)
        return set_err(c, DC_ERR_ARG, DC_ER_LIB, "model: null pointer or bad size");
    if (m->natoms > INT32_MAX / 8) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "model: too many atoms");
    for (int k = 0; k < 3; ++k)
        if (m->inc[k] < 1) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "model: bad inc");
    for (int64_t n = 0; n < m->natoms; ++n)
        if (m->type[n] < 0 || m->type[n] > m->nscat) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "model: type outside 0..nscat");
    return 0;
}

// Grid, GEMM plan (rows = abscissa), type-sorted atom order and the s^2 parts.
int plan_model(dc_ctx *c, const dc_model_desc *m, ModelPlan &mp) {
    Grid &g = mp.g;
    for (int k = 0; k < 3; ++k) {
        g.n[k] = g.f[k] = m->inc[k];
        g.xm[k] = m->eck[k];
        g.uin[k] = m->vi[k];
        g.vin[k] = m->vi[3 + k];
        g.win[k] = m->vi[6 + k];
    }
    for (int k = 0; k < 9; ++k) g.rten[k] = m->rten[k];
    g.P = (long long)g.f[0] * g.f[1] * g.f[2];
    mp.pl = plan_gemm(g, c->tile, 0, c->prec, m->natoms);
    const GemmGeom &gg = mp.pl.gg;

    // metric orthogonality of the abscissa against the other axes
    const double *v1 = m->vi, *v2 = m->vi + 3, *v3 = m->vi + 6;
    const double n1 = std::sqrt(std::fabs(dotg(v1, m->rten, v1)));
    for (int ax = 1; ax < 3; ++ax) {
        if (m->inc[ax] <= 1) continue;
        const double *v = m->vi + 3 * ax;
        const double nv = std::sqrt(std::fabs(dotg(v, m->rten, v)));
        if (std::fabs(dotg(v1, m->rten, v)) > 1e-12 * std::max(1.0, n1 * nv)) mp.separable = false;
    }
    mp.sq.assign((size_t)gg.R + gg.C, 0.0f);
    for (int i = 0; i < gg.R; ++i) {
        double h[3];
        for (int k = 0; k < 3; ++k) h[k] = m->eck[k] + i * v1[k];
        mp.sq[i] = (float)(dotg(h, m->rten, h) / 4.0);
    }
    for (int col = 0; col < gg.C; ++col) {
        const int j = col % gg.C0, l = col / gg.C0;
        if (j >= g.f[1]) continue;
        double d[3];
        for (int k = 0; k < 3; ++k) d[k] = j * v2[k] + l * v3[k];
        mp.sq[gg.R + col] = (float)((dotg(d, m->rten, d) + 2.0 * dotg(d, m->rten, m->eck)) / 4.0);
    }

    // counting sort by type, vacancies dropped
    std::vector<int> cnt(m->nscat + 1, 0);
    for (int64_t n = 0; n < m->natoms; ++n) cnt[m->type[n]]++;
    std::vector<int> beg(m->nscat + 2, 0);
    for (int t = 1; t <= m->nscat; ++t) beg[t + 1] = beg[t] + cnt[t];
    mp.nsorted = beg[m->nscat + 1];
    mp.order.assign(mp.nsorted, 0);
    std::vector<int> fill(beg.begin(), beg.end());
    for (int64_t n = 0; n < m->natoms; ++n) {
        const int t = m->type[n];
        if (t > 0) mp.order[fill[t]++] = (int)n;
        if (t > 0 && m->bdw && m->bdw[n] != 0.0) mp.any_b = true;
    }
    for (int t = 1; t <= m->nscat; ++t)
        if (cnt[t] > 0) mp.segs.push_back({t - 1, cnt[t]});
    if (mp.any_b && !mp.separable)
        return set_err(c, DC_ERR_ARG, DC_ER_LIB, "model: per-atom B needs vi(:,1) metric-orthogonal to vi(:,2), vi(:,3)");
    if (max_corner_bin(g) >= m->ncfact)
        return set_err(c, DC_ERR_ISTL, DC_ER_FOUR, "sin(theta)/lambda beyond table");
    return 0;
}

// Sorted per-atom inputs on the device and the GEMM records (corner phases).
int upload_model(dc_ctx *c, const dc_model_desc *m, const ModelPlan &mp) {
    int st;
    const int ns = mp.nsorted;
    std::vector<float> w(ns), b(ns);
    for (int k = 0; k < ns; ++k) {
        const int n = mp.order[k];
        w[k] = (float)m->occ[n];
        b[k] = m->bdw ? (float)m->bdw[n] : 0.0f;
    }
    if ((st = upload(c, c->m_pos, m->pos, 3 * (size_t)m->natoms))) return st;
    if ((st = upload(c, c->idx, mp.order.data(), mp.order.size()))) return st;
    if ((st = upload(c, c->m_w, w.data(), w.size()))) return st;
    if ((st = upload(c, c->m_b, b.data(), b.size()))) return st;
    if ((st = upload(c, c->m_sq, mp.sq.data(), mp.sq.size()))) return st;
    if ((st = ensure(c, c->rec, sizeof(AtomRec) * std::max(ns, 1)))) return st;
    PrepArgs pa = base_prep(mp.g, 0);
    pa.src = ptr<double>(c->m_pos);
    pa.src_ld = 0;
    pa.idx = ptr<int>(c->idx);
    pa.w = ptr<float>(c->m_w);
    pa.bdw = ptr<float>(c->m_b);
    pa.n = ns;
    HIPCHK(launch_prep(pa, ptr<AtomRec>(c->rec), c->st));
    c->rec_hash = false;
    return 0;
}

// F (complex, device m_f) and I (device dsi) over the grid.
int model_forward(dc_ctx *c, const dc_model_desc *m, ModelPlan &mp) {
    int st;
    const GemmGeom &gg = mp.pl.gg;
    SliceSet ss;
    int off = 0;
    add_segments(ss, mp.segs, slice_len(c, kchunk_for(mp.pl, mp.segs, c->prec)), off);
    if (ss.beg.empty()) ss.beg.push_back(0);
    const int nslice = (int)ss.type.size();
    ss.lot = {0, nslice};
    if ((st = upload(c, c->slice_beg, ss.beg.data(), ss.beg.size()))) return st;
    if ((st = upload(c, c->lot_slice, ss.lot.data(), ss.lot.size()))) return st;
    if (nslice > 0 && (st = upload(c, c->slice_type, ss.type.data(), ss.type.size()))) return st;
    if ((st = upload(c, c->cfact, m->cfact, 2 * (size_t)m->ncfact * m->nscat))) return st;
    if ((st = ensure(c, c->partial, sizeof(float2) * gg.P * std::max(nslice, 1)))) return st;
    if ((st = ensure(c, c->m_f, sizeof(double) * 2 * mp.g.P))) return st;
    if ((st = ensure(c, c->dsi, sizeof(double) * mp.g.P))) return st;
    HIPCHK(hipEventRecord(c->ev[4], c->st));
    if ((st = run_gemm(c, mp.pl, nslice, ptr<float2>(c->partial), mp.any_b ? ptr<float>(c->m_sq) : nullptr))) return st;
    ReduceArgs ra{};
    ra.g = mp.g;
    ra.partial = ptr<float2>(c->partial);
    ra.slice0 = 0;
    ra.lot_slice = ptr<int>(c->lot_slice);
    ra.slice_type = ptr<int>(c->slice_type);
    ra.nlots = nslice > 0 ? 1 : 0;
    ra.lform = 1;
    ra.cfact = ptr<double2>(c->cfact);
    ra.ncfact = m->ncfact;
    ra.ntype = m->nscat;
    ra.dsi = ptr<double>(c->dsi);
    ra.csf = ptr<double>(c->m_f);
    ra.fresh = 1;
    if (nslice > 0) {
        HIPCHK(launch_reduce_lots(ra, c->st));
    } else {
        HIPCHK(hipMemsetAsync(c->m_f.p, 0, sizeof(double) * 2 * mp.g.P, c->st));
        HIPCHK(hipMemsetAsync(c->dsi.p, 0, sizeof(double) * mp.g.P, c->st));
    }
    HIPCHK(hipEventRecord(c->ev[5], c->st));
    HIPCHK(hipEventSynchronize(c->ev[5]));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev[4], c->ev[5]));
    c->stats.forward_ms += ms;
    c->stats.pairs += (double)mp.nsorted * (double)mp.g.P;
    c->stats.slices += nslice;
    return 0;
}

// Gradient (5, natoms) on the device (m_grad) from F (m_f) and G (m_gi).
int model_adjoint(dc_ctx *c, const dc_model_desc *m, ModelPlan &mp) {
    int st;
    if (!mp.separable)
        return set_err(c, DC_ERR_ARG, DC_ER_LIB, "model gradient: needs vi(:,1) metric-orthogonal to vi(:,2), vi(:,3)");
    const GemmGeom &gg = mp.pl.gg;
    const int nq = (int)mp.segs.size();
    std::vector<int> qtype(std::max(nq, 1), 0);
    for (int t = 0; t < nq; ++t) qtype[t] = mp.segs[t].first;
    // atom tiles inside one type
    const int na = adjoint_tile_atoms();
    std::vector<int> tiles;
    int a0 = 0;
    for (int t = 0; t < nq; ++t) {
        const int n = mp.segs[t].second;
        for (int b = 0; b < n; b += na) {
            tiles.insert(tiles.end(), {a0 + b, a0 + std::min(n, b + na), t, 0});
        }
        a0 += n;
    }
    const int ntiles = (int)tiles.size() / 4;
    const int ntc = (gg.C + adjoint_tile_cols() - 1) / adjoint_tile_cols();

    HIPCHK(hipEventRecord(c->ev[4], c->st));
    if ((st = upload(c, c->m_qt, qtype.data(), qtype.size()))) return st;
    if (ntiles > 0 && (st = upload(c, c->m_tiles, tiles.data(), tiles.size()))) return st;
    if ((st = upload(c, c->m_tgt, m->occ, (size_t)m->natoms))) return st;
    if ((st = ensure(c, c->m_q, sizeof(float2) * gg.P * std::max(nq, 1)))) return st;
    if ((st = ensure(c, c->m_part, sizeof(float) * 6 * (size_t)ntc * std::max(mp.nsorted, 1)))) return st;
    if ((st = ensure(c, c->m_grad, sizeof(double) * 5 * std::max<int64_t>(m->natoms, 1)))) return st;
    HIPCHK(hipMemsetAsync(c->m_grad.p, 0, sizeof(double) * 5 * std::max<int64_t>(m->natoms, 1), c->st));
    if (ntiles > 0) {
        HIPCHK(hipMemsetAsync(c->m_q.p, 0, sizeof(float2) * gg.P * nq, c->st));
        QArgs qa{};
        qa.g = mp.g;
        qa.R = gg.R;
        qa.C = gg.C;
        qa.C0 = gg.C0;
        qa.plane = gg.P;
        qa.f = ptr<double>(c->m_f);
        qa.gi = ptr<double>(c->m_gi);
        qa.cfact = ptr<double2>(c->cfact);
        qa.ncfact = m->ncfact;
        qa.qtype = ptr<int>(c->m_qt);
        qa.nq = nq;
        qa.q = ptr<float2>(c->m_q);
        HIPCHK(launch_model_q(qa, c->st));

        AdjArgs aa{};
        aa.rec = ptr<AtomRec>(c->rec);
        aa.tiles = ptr<int4>(c->m_tiles);
        aa.q = ptr<float2>(c->m_q);
        aa.sq = ptr<float>(c->m_sq);
        aa.part = ptr<float>(c->m_part);
        aa.R = gg.R;
        aa.C = gg.C;
        aa.C0 = gg.C0;
        aa.f1 = mp.g.f[1];
        aa.plane = gg.P;
        aa.natoms = mp.nsorted;
        aa.rc = 0.5f * (float)(mp.g.f[0] - 1);
        aa.jc = 0.5f * (float)(mp.g.f[1] - 1);
        aa.mc = 0.5f * (float)(mp.g.f[2] - 1);
        const bool split = c->prec != DC_PREC_FP32;
        if (split) {
            aa.Rg = (int)((gg.R + 15) / 16) * 4;
            if ((st = ensure(c, c->m_qs, sizeof(uint32_t) * 8 * (size_t)aa.Rg * gg.C * nq))) return st;
            aa.qs = ptr<uint32_t>(c->m_qs);
            HIPCHK(launch_q_split(ptr<float2>(c->m_q), nq, gg.R, gg.C, gg.P, aa.Rg, ptr<uint32_t>(c->m_qs), c->st));
        }
        const int nset = adjoint_plan(aa, ntiles, split, c->adj_walk);
        c->stats.adjoint_walk = aa.ctpg;
        HIPCHK(launch_adjoint(aa, ntiles, c->st, split));

        AdjFinishArgs fa{};
        fa.part = ptr<float>(c->m_part);
        fa.ntile_c = nset;
        fa.natoms = mp.nsorted;
        fa.idx = ptr<int>(c->idx);
        fa.occ = ptr<double>(c->m_tgt);
        for (int k = 0; k < 3; ++k)
            fa.hc[k] = m->eck[k] + (double)aa.rc * m->vi[k] + (double)aa.jc * m->vi[3 + k] + (double)aa.mc * m->vi[6 + k];
        for (int k = 0; k < 9; ++k) fa.v[k] = m->vi[k];
        fa.grad = ptr<double>(c->m_grad);
        HIPCHK(launch_adjoint_finish(fa, c->st));
    }
    HIPCHK(hipEventRecord(c->ev[5], c->st));
    HIPCHK(hipEventSynchronize(c->ev[5]));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev[4], c->ev[5]));
    c->stats.adjoint_ms += ms;
    return 0;
}

int model_begin(dc_ctx *c, const dc_model_desc *m, ModelPlan &mp) {
    c->err.clear();
    c->err_typ = DC_ER_NONE;
    int st;
    if ((st = check_model(c, m))) return st;
    if (cancelled(c)) return set_err(c, DC_ERR_CANCEL, DC_ER_COMM, "cancelled");
    HIPCHK(hipSetDevice(c->device));
    stats_begin(c);
    HIPCHK(hipEventRecord(c->ev[0], c->st));
    if ((st = plan_model(c, m, mp))) return st;
    return upload_model(c, m, mp);
}

int model_end(dc_ctx *c) {
    HIPCHK(hipEventRecord(c->ev[1], c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    c->stats.total_ms = ms;
    return DC_OK;
}

}  // namespace

extern "C" {

int dc_model_intensity(dc_ctx *c, const dc_model_desc *m, double *icalc, double *fcalc) {
    if (!c) return DC_ERR_ARG;
    ModelPlan mp;
    int st;
    if ((st = model_begin(c, m, mp))) return st;
    if ((st = model_forward(c, m, mp))) return st;
    if (icalc) HIPCHK(hipMemcpyAsync(icalc, c->dsi.p, sizeof(double) * mp.g.P, hipMemcpyDeviceToHost, c->st));
    if (fcalc) HIPCHK(hipMemcpyAsync(fcalc, c->m_f.p, sizeof(double) * 2 * mp.g.P, hipMemcpyDeviceToHost, c->st));
    return model_end(c);
}

int dc_model_vjp(dc_ctx *c, const dc_model_desc *m, const double *fcalc, const double *gi, double *grad) {
    if (!c) return DC_ERR_ARG;
    if (!fcalc || !gi || !grad) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "dc_model_vjp: null pointer");
    ModelPlan mp;
    int st;
    if ((st = model_begin(c, m, mp))) return st;
    if ((st = upload(c, c->m_f, fcalc, 2 * (size_t)mp.g.P))) return st;
    if ((st = upload(c, c->m_gi, gi, (size_t)mp.g.P))) return st;
    if ((st = upload(c, c->cfact, m->cfact, 2 * (size_t)m->ncfact * m->nscat))) return st;
    if ((st = model_adjoint(c, m, mp))) return st;
    HIPCHK(hipMemcpyAsync(grad, c->m_grad.p, sizeof(double) * 5 * m->natoms, hipMemcpyDeviceToHost, c->st));
    return model_end(c);
}

int dc_model_jvp(dc_ctx *c, const dc_model_desc *m, const double *fcalc, int ndir, const double *dparam,
                 double *di, double *df) {
    if (!c) return DC_ERR_ARG;
    if (!fcalc || !dparam || !di || ndir < 0) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "dc_model_jvp: null pointer or bad ndir");
    ModelPlan mp;
    int st;
    if ((st = model_begin(c, m, mp))) return st;
    const GemmGeom &gg = mp.pl.gg;
    const long long P = mp.g.P;
    const int ns = mp.nsorted;
    if ((st = upload(c, c->m_f, fcalc, 2 * (size_t)P))) return st;
    if ((st = upload(c, c->cfact, m->cfact, 2 * (size_t)m->ncfact * m->nscat))) return st;
    if (m->bdw && (st = upload(c, c->m_bd, m->bdw, (size_t)m->natoms))) return st;
    if ((st = ensure(c, c->m_di, sizeof(double) * (size_t)P * std::max(ndir, 1) * (df ? 3 : 1)))) return st;
    // directions per batch so that the slice planes fit the budget
    const int kchunk = slice_len(c, kchunk_for(mp.pl, (long long)ns * 5, c->prec));
    long long slices_per_dir = 0;
    for (auto &sg : mp.segs) slices_per_dir += 5 * ((sg.second + kchunk - 1) / kchunk);
    const double dir_bytes = (double)std::max<long long>(slices_per_dir, 1) * gg.P * sizeof(float2);
    const int per = std::max(1, std::min(ndir, (int)(JVP_PARTIAL_BUDGET / dir_bytes)));
    HIPCHK(hipEventRecord(c->ev[4], c->st));
    for (int d0 = 0; d0 < ndir; d0 += per) {
        if (cancelled(c)) return set_err(c, DC_ERR_CANCEL, DC_ER_COMM, "cancelled");
        const int nd = std::min(per, ndir - d0);
        // weights per (direction, component) in sorted order, records, slices
        std::vector<float> w((size_t)nd * 5 * ns);
        std::vector<int> idx((size_t)nd * 5 * ns);
        for (int k = 0; k < nd; ++k)
            for (int q = 0; q < 5; ++q)
                for (int t = 0; t < ns; ++t) {
                    const int n = mp.order[t];
                    const double *dp = dparam + 5 * ((size_t)(d0 + k) * m->natoms + n);
                    const double o = m->occ[n];
                    const double v = q == 0 ? dp[3] : (q < 4 ? o * dp[q - 1] : o * dp[4]);
                    w[((size_t)k * 5 + q) * ns + t] = (float)v;
                    idx[((size_t)k * 5 + q) * ns + t] = n;
                }
        SliceSet ss;
        std::vector<int> dsl;
        int off = 0;
        for (int g5 = 0; g5 < 5 * nd; ++g5) {
            dsl.push_back((int)ss.type.size());
            add_segments(ss, mp.segs, kchunk, off);
        }
        dsl.push_back((int)ss.type.size());
        if (ss.beg.empty()) ss.beg.push_back(0);
        const int nslice = (int)ss.type.size();
        if ((st = upload(c, c->m_jw, w.data(), w.size()))) return st;
        if ((st = upload(c, c->m_jidx, idx.data(), idx.size()))) return st;
        if ((st = upload(c, c->m_dsl, dsl.data(), dsl.size()))) return st;
        if ((st = upload(c, c->slice_beg, ss.beg.data(), ss.beg.size()))) return st;
        if (nslice > 0 && (st = upload(c, c->slice_type, ss.type.data(), ss.type.size()))) return st;
        if ((st = ensure(c, c->rec, sizeof(AtomRec) * std::max<size_t>(w.size(), 1)))) return st;
        if ((st = ensure(c, c->partial, sizeof(float2) * gg.P * std::max(nslice, 1)))) return st;
        PrepArgs pa = base_prep(mp.g, 0);
        pa.src = ptr<double>(c->m_pos);
        pa.src_ld = 0;
        pa.idx = ptr<int>(c->m_jidx);
        pa.w = ptr<float>(c->m_jw);
        pa.bdw_src = m->bdw ? ptr<double>(c->m_bd) : nullptr;
        pa.n = (long long)w.size();
        if (pa.n > 0) HIPCHK(launch_prep(pa, ptr<AtomRec>(c->rec), c->st));
        c->rec_hash = false;
        if ((st = run_gemm(c, mp.pl, nslice, ptr<float2>(c->partial), mp.any_b ? ptr<float>(c->m_sq) : nullptr))) return st;
        JvpArgs ja{};
        ja.g = mp.g;
        ja.partial = ptr<float2>(c->partial);
        ja.dslice = ptr<int>(c->m_dsl);
        ja.slice_type = ptr<int>(c->slice_type);
        ja.cfact = ptr<double2>(c->cfact);
        ja.ncfact = m->ncfact;
        ja.f = ptr<double>(c->m_f);
        ja.ndir = nd;
        ja.di = ptr<double>(c->m_di) + (size_t)d0 * P;
        ja.df = df ? ptr<double>(c->m_di) + (size_t)ndir * P + 2 * (size_t)d0 * P : nullptr;
        HIPCHK(launch_jvp_reduce(ja, c->st));
        c->stats.pairs += (double)w.size() * (double)P;
        c->stats.slices += nslice;
    }
    HIPCHK(hipEventRecord(c->ev[5], c->st));
    if (ndir > 0) {
        HIPCHK(hipMemcpyAsync(di, c->m_di.p, sizeof(double) * (size_t)P * ndir, hipMemcpyDeviceToHost, c->st));
        if (df)
            HIPCHK(hipMemcpyAsync(df, ptr<double>(c->m_di) + (size_t)ndir * P, sizeof(double) * 2 * (size_t)P * ndir,
                                  hipMemcpyDeviceToHost, c->st));
    }
    if ((st = model_end(c))) return st;
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev[4], c->ev[5]));
    c->stats.forward_ms = ms;
    return DC_OK;
}

int dc_model_chi2(dc_ctx *c, const dc_model_desc *m, const double *target, const double *weight, double *chi2,
                  double *icalc, double *grad) {
    if (!c) return DC_ERR_ARG;
    if (!target || !chi2) return set_err(c, DC_ERR_ARG, DC_ER_LIB, "dc_model_chi2: null pointer");
    ModelPlan mp;
    int st;
    if ((st = model_begin(c, m, mp))) return st;
    if ((st = model_forward(c, m, mp))) return st;
    const long long P = mp.g.P;
    if ((st = ensure(c, c->m_gi, sizeof(double) * (P + CHI2_BLOCKS)))) return st;
    if ((st = ensure(c, c->w, sizeof(double) * 2 * P))) return st;
    double *d_t = ptr<double>(c->w);
    double *d_w = weight ? d_t + P : nullptr;
    HIPCHK(hipMemcpyAsync(d_t, target, sizeof(double) * P, hipMemcpyHostToDevice, c->st));
    if (weight) HIPCHK(hipMemcpyAsync(d_w, weight, sizeof(double) * P, hipMemcpyHostToDevice, c->st));
    double *d_gi = ptr<double>(c->m_gi);
    HIPCHK(launch_chi2_resid(ptr<double>(c->dsi), d_t, d_w, P, d_gi, d_gi + P, CHI2_BLOCKS, c->st));
    std::vector<double> bs(CHI2_BLOCKS);
    HIPCHK(hipMemcpyAsync(bs.data(), d_gi + P, sizeof(double) * CHI2_BLOCKS, hipMemcpyDeviceToHost, c->st));
    if (icalc) HIPCHK(hipMemcpyAsync(icalc, c->dsi.p, sizeof(double) * P, hipMemcpyDeviceToHost, c->st));
    if (grad) {
        if ((st = model_adjoint(c, m, mp))) return st;
        HIPCHK(hipMemcpyAsync(grad, c->m_grad.p, sizeof(double) * 5 * m->natoms, hipMemcpyDeviceToHost, c->st));
    }
    if ((st = model_end(c))) return st;
    double s = 0.0;
    for (double v : bs) s += v;
    *chi2 = s;
    return DC_OK;
}

}  // extern "C"
