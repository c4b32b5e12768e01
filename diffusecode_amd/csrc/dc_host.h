// dc_host.h -- host-side state and helpers shared by the C-ABI translation
// units (dc_api.hip: the Fourier run, dc_refine.hip: the refinement model).
#pragma once
#include <string>
#include <utility>
#include <vector>

#include "../../include/diffuse_turbo.h"
#include "dc_internal.h"

struct BatchState;   // dc_batch.hip
struct PostState;    // dc_post.hip
struct NufftState;   // dc_nufft.hip

struct DevBuf {
    void *p = nullptr;
    size_t n = 0;
};

struct dc_ctx {
    int device = 0;
    hipStream_t st = nullptr;
    int prec = DC_PREC_FP32;
    int tile = -1;                  // forced GEMM tile configuration, -1 = planned
    int autotune = 1;               // batch: time tile shapes on the first call of a shape
    int graphs = 1;                 // batch: capture the population step as a hipGraph
    int adj_walk = 0;               // split adjoint column walk: 0 planned, else forced (short grids)
    int dither = 1;                 // per-atom phase rotation of the phase GEMM (dither_turn)
    int slice_atoms = 0;            // forced atoms per GEMM slice, 0 = planned
    BatchState *batch = nullptr;    // cached plan / buffers / graph of the last batch shape
    PostState *post = nullptr;      // FFT plans and buffers of the post-processing calls
    NufftState *nufft = nullptr;    // fine grid, plans and buffers of the NUFFT technique
    std::string err;
    int err_typ = DC_ER_NONE;
    const volatile int *cancel = nullptr;
    hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    dc_stats stats{};
    // timed spans of the current call whose events the host has not waited
    // for: (start, end, stats field the elapsed ms are added to); tev is the
    // pool their events come from (tev_used in use)
    struct Span {
        hipEvent_t a, b;
        double *dst;
    };
    std::vector<Span> spans;
    std::vector<hipEvent_t> tev;
    size_t tev_used = 0;
    // device buffers of the Fourier run
    DevBuf rec, slice_beg, slice_type, lot_slice, partial, cfact, istl, idx, oid, offs, xat, w;
    DevBuf acsf, dsi, csf, wm, wv, lat, out_tcsf, out[5];
    // device buffers of the refinement model
    DevBuf m_pos, m_w, m_b, m_sq, m_q, m_part, m_grad, m_tiles, m_gi, m_f, m_tgt, m_qt, m_qs, m_bd, m_jw, m_jidx, m_dsl, m_di;
    // crystal
    bool have_crystal = false;
    DevBuf pos, iscat;
    std::vector<double> h_pos;
    std::vector<int> h_iscat;
    int64_t natoms = 0;
    int icc[3] = {1, 1, 1};
    int ncatoms = 1, nscat = 0;
    double dim0[3] = {0, 0, 0}, dim_lo[3] = {0, 0, 0};
    std::vector<int> type_beg;      // per type offsets into d type list
    DevBuf type_idx;
    // lot gather (DISCUS-ordered crystals, natoms == cells x sites): per type,
    // the atoms in (cell, site) order on the device (tc_idx, type t from
    // type_beg[t]) and per type the first position of every cell in that list
    // (tc_start[t * (ncell + 1) + cell], host), so that a run of consecutive
    // cells of a lot is one contiguous range
    std::vector<int> tc_start;
    DevBuf runs;
    // this device's dc_scratch array (dc_noalloc.hip), looked up at creation
    void *scratch = nullptr;
    // rec's records carry the rotation hash in AtomRec::pad (PrepArgs::hash_salt):
    // set by the Fourier run's prep, cleared by the refinement's, whose
    // adjoints read pad as B
    bool rec_hash = false;

    std::vector<DevBuf *> buffers() {
        return {&rec, &slice_beg, &slice_type, &lot_slice, &partial, &cfact, &istl, &idx, &oid, &offs, &xat, &w,
                &acsf, &dsi, &csf, &wm, &wv, &lat, &out_tcsf, &out[0], &out[1], &out[2], &out[3], &out[4],
                &m_pos, &m_w, &m_b, &m_sq, &m_q, &m_part, &m_grad, &m_tiles, &m_gi, &m_f, &m_tgt, &m_qt, &m_qs, &m_bd, &m_jw, &m_jidx, &m_dsl, &m_di,
                &pos, &iscat, &type_idx, &runs};
    }
};

namespace dc {

inline int set_err(dc_ctx *c, int code, int typ, const std::string &msg) {
    if (c) {
        c->err = msg;
        c->err_typ = typ;
    }
    return code;
}

inline int hip_err(dc_ctx *c, hipError_t e, const char *where) {
    return set_err(c, e == hipErrorOutOfMemory ? DC_ERR_NOMEM : DC_ERR_DEVICE, DC_ER_LIB,
                   std::string(where) + ": " + hipGetErrorString(e));
}

#define HIPCHK(call)                                        \
    do {                                                    \
        hipError_t e_ = (call);                             \
        if (e_ != hipSuccess) return hip_err(c, e_, #call); \
    } while (0)

// Grow-only device buffer (reallocation synchronises the context stream).
inline int ensure(dc_ctx *c, DevBuf &b, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (b.n >= bytes) return 0;
    if (b.p) {
        hipError_t e = hipStreamSynchronize(c->st);
        if (e != hipSuccess) return hip_err(c, e, "sync before realloc");
        (void)hipFree(b.p);
        b.p = nullptr;
        b.n = 0;
    }
    size_t n = bytes + bytes / 8;
    hipError_t e = hipMalloc(&b.p, n);
    if (e != hipSuccess) {
        b.p = nullptr;
        return hip_err(c, e, "hipMalloc");
    }
    b.n = n;
    return 0;
}

template <class T>
T *ptr(DevBuf &b) {
    return reinterpret_cast<T *>(b.p);
}

template <class T>
int upload(dc_ctx *c, DevBuf &b, const T *src, size_t count) {
    int st = ensure(c, b, sizeof(T) * count);
    if (st) return st;
    if (count) HIPCHK(hipMemcpyAsync(b.p, src, sizeof(T) * count, hipMemcpyHostToDevice, c->st));
    return 0;
}

// Deferred timing: mark() records a pool event on the context stream;
// span() files (start, end) to be added to a stats field once the host has
// waited for the stream, so that timing a kernel does not stall the host
// between launches.  stats_begin() starts a call's statistics (dropping the
// spans of an earlier call), resolve_spans() waits for the last span and
// adds every span's elapsed time.
inline int mark(dc_ctx *c, hipEvent_t &e) {
    if (c->tev_used == c->tev.size()) {
        hipEvent_t n = nullptr;
        HIPCHK(hipEventCreate(&n));
        c->tev.push_back(n);
    }
    e = c->tev[c->tev_used++];
    HIPCHK(hipEventRecord(e, c->st));
    return 0;
}
inline void span(dc_ctx *c, hipEvent_t a, hipEvent_t b, double *dst) { c->spans.push_back({a, b, dst}); }
inline void stats_begin(dc_ctx *c) {
    c->stats = dc_stats{};
    c->spans.clear();
    c->tev_used = 0;
}
inline int resolve_spans(dc_ctx *c) {
    if (c->spans.empty()) return 0;
    HIPCHK(hipEventSynchronize(c->spans.back().b));
    for (const auto &s : c->spans) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, s.a, s.b));
        *s.dst += ms;
    }
    c->spans.clear();
    c->tev_used = 0;
    return 0;
}

inline bool cancelled(dc_ctx *c) { return c->cancel && *c->cancel; }

// GEMM plan of a grid region (dc_api.hip).
struct Plan {
    GemmGeom gg{};
    int mapping = 0;
    int tile = 0;
    int scheme = 0;              // GEMM scheme the plan is for (DC_PREC_* or SCHEME_FP32_ACC64)
};

// Slices over a list of (type, count) segments laid out consecutively.
struct SliceSet {
    std::vector<int> beg;        // nslice + 1 atom offsets
    std::vector<int> type;       // per slice
    std::vector<int> lot;        // lot boundaries (slice ids), nlots + 1
    std::vector<int> seg_first;  // first slice of each segment
};

long long padded(long long n, long long t);
int max_corner_bin(const Grid &g);   // largest sin(theta)/lambda bin over the grid corners
Plan plan_gemm(Grid &g, int forced, int only_mapping = -1, int prec = 1, long long natoms = 0);
int kchunk_for(const Plan &pl, long long natoms_batch, int prec);
int kchunk_for(const Plan &pl, const std::vector<std::pair<int, int>> &segs, int prec);
void add_segments(SliceSet &s, const std::vector<std::pair<int, int>> &segs, int kchunk, int &atom_off);
int run_gemm(dc_ctx *c, const Plan &pl, int nslice, float2 *partial, const float *sq = nullptr);
// atoms per slice: the context's forced length (dc_set_slice_atoms) or the planned one
inline int slice_len(const dc_ctx *c, int planned) { return c->slice_atoms > 0 ? c->slice_atoms : planned; }
// the plan's geometry with the context's dither salt
inline GemmGeom launch_geom(const dc_ctx *c, const GemmGeom &g) {
    GemmGeom r = g;
    r.dither = c->dither ? 0x2545F491u : 0u;
    return r;
}
void batch_release(dc_ctx *c);   // frees the cached batch plan, buffers and graph
void post_release(dc_ctx *c);    // frees the FFT plans and buffers of the post-processing calls
void nufft_release(dc_ctx *c);   // frees the NUFFT fine grid, plans and buffers
int noalloc_attach(dc_ctx *c);   // device address of this device's dc_scratch (dc_noalloc.hip)
PrepArgs base_prep(const Grid &g, int mapping);

}  // namespace dc
