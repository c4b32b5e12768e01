// dc_noalloc.hip -- the layer call with no allocation (dc_strucf_dev_noalloc).
//
// Callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc) get the layer call
// with every scratch buffer -- atom records, slice boundaries, slice planes
// -- carved out of one file-scope device array, dc_scratch.  Each GPU holds
// its own copy of the array until the process ends.  Its device address is
// looked up once per device (hipGetSymbolAddress) when a context is created
// on that device; host code uses only that address, never the array's name.
#include <cmath>
#include <cstring>
#include <mutex>

#include "../../include/diffuse_turbo.h"
#include "dc_host.h"

using namespace dc;

namespace {

constexpr size_t kScratchBytes = (size_t)DC_NOALLOC_SCRATCH_BYTES;
__device__ __align__(256) unsigned char dc_scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void *g_scratch[kMaxDevices] = {};

inline size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

}  // namespace

namespace dc {

// Device address of this device's dc_scratch (the current device must be
// `device`); looked up on the first call per device.  0 on success.
int noalloc_attach(dc_ctx *c) {
    if (c->device < 0 || c->device >= kMaxDevices) return set_err(c, DC_ERR_DEVICE, DC_ER_LIB, "device index");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[c->device]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(dc_scratch));
        if (e != hipSuccess || !p) return hip_err(c, e == hipSuccess ? hipErrorInvalidSymbol : e, "hipGetSymbolAddress(dc_scratch)");
        g_scratch[c->device] = p;
    }
    c->scratch = g_scratch[c->device];
    return 0;
}

int strucf_setup(dc_ctx *c, const void *xat, int64_t nxat, int64_t ld_xat, const double uin[3], const double vin[3],
                 const double win[3], const double off_shift[9], const int lmn[6], const int num[3],
                 const int fnum[3], const void *cfact_col, int64_t ncfact, const void *istl, const double xm[3],
                 const double rten[9], int lform, const void *tcsf, Grid &g, long long &ntot);
int strucf_core(dc_ctx *c, const Grid &g0, long long ntot, const double *d_xat, int64_t nxat, int64_t ld_xat,
                const double off_shift[9], const int lmn[6], const double2 *d_cfact, int64_t ncfact,
                const int *d_istl, int lform, AtomRec *rec, int *slice_beg, float2 *partial, double plane_bytes,
                double *d_tcsf);

}  // namespace dc

extern "C" {

int dc_strucf_dev_noalloc(dc_ctx *c, const double *d_xat, int64_t nxat, int64_t ld_xat, const double uin[3],
                          const double vin[3], const double win[3], const double off_shift[9], const int lmn[6],
                          const int num[3], const int fnum[3], const double *d_cfact_col, int64_t ncfact,
                          const int *d_istl, const double xm[3], const double rten[9], int lform, double *d_tcsf) {
    if (!c) return DC_ERR_ARG;
    Grid g{};
    long long ntot = 0;
    int st = strucf_setup(c, d_xat, nxat, ld_xat, uin, vin, win, off_shift, lmn, num, fnum, d_cfact_col, ncfact,
                          d_istl, xm, rten, lform, d_tcsf, g, ntot);
    if (st) return st;
    if (!c->scratch) return set_err(c, DC_ERR_STATE, DC_ER_LIB, "dc_strucf_dev_noalloc: no scratch on this device");
    HIPCHK(hipSetDevice(c->device));
    // scratch layout: records | slice boundaries | slice planes
    Grid gp = g;
    const Plan pl = plan_gemm(gp, c->tile, -1, c->prec, nxat);
    const int klen = slice_len(c, kchunk_for(pl, nxat, pl.scheme));
    const long long nslice = std::max<long long>(1, (nxat + klen - 1) / klen);
    const size_t rec_bytes = align256(sizeof(AtomRec) * (size_t)std::max<int64_t>(nxat, 1));
    const size_t beg_bytes = align256(sizeof(int) * (size_t)(nslice + 1));
    const size_t plane = sizeof(float2) * (size_t)pl.gg.P;
    if (rec_bytes + beg_bytes + plane > kScratchBytes) {
        char msg[200];
        std::snprintf(msg, sizeof(msg),
                      "dc_strucf_dev_noalloc: needs %zu bytes of scratch (records and one slice plane), holds %zu",
                      rec_bytes + beg_bytes + plane, kScratchBytes);
        return set_err(c, DC_ERR_ARG, DC_ER_LIB, msg);
    }
    unsigned char *base = static_cast<unsigned char *>(c->scratch);
    stats_begin(c);
    HIPCHK(hipEventRecord(c->ev[0], c->st));
    st = strucf_core(c, g, ntot, d_xat, nxat, std::max<int64_t>(ld_xat, 1), off_shift, lmn,
                     lform ? reinterpret_cast<const double2 *>(d_cfact_col) : nullptr, ncfact, lform ? d_istl : nullptr,
                     lform, reinterpret_cast<AtomRec *>(base), reinterpret_cast<int *>(base + rec_bytes),
                     reinterpret_cast<float2 *>(base + rec_bytes + beg_bytes),
                     (double)(kScratchBytes - rec_bytes - beg_bytes), d_tcsf);
    if (st) return st;
    HIPCHK(hipEventRecord(c->ev[1], c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    c->stats.total_ms = ms;
    HIPCHK(hipEventElapsedTime(&ms, c->ev[2], c->ev[3]));
    c->stats.gemm_ms = ms;
    return DC_OK;
}

}  // extern "C"
