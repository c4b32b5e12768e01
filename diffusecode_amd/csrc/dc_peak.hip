// dc_peak.hip -- the device's ceilings for the instructions the path relies
// on, measured with plain streams of them:
//   * dense MFMA chains: v_mfma_f32_16x16x32_bf16 (BF16x3), _f16 (FP16x2,
//     FP16x3) and v_mfma_f32_16x16x4_f32 (FP32), 8 independent accumulators
//     per wave, operands in registers, 8 waves per SIMD, about 1 ms per launch;
//   * HBM streams: a 16-byte-per-lane read of a buffer much larger than the
//     L2 and MALL caches, and a copy of it.
// The roofline fractions of the kernels are priced against the published
// peaks; these figures say how much of a published peak a plain stream of
// the same instruction reaches on the part at hand (clocks under load).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "dc_host.h"

namespace dc {
namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

constexpr int PK_THREADS = 256;   // 4 waves per work-group
constexpr int PK_CHAINS = 8;      // independent accumulators per wave
constexpr int PK_BF16 = 0, PK_F16 = 1, PK_F32 = 2;

// flops of one MFMA of each kind (2 per multiply-add)
constexpr double pk_flops(int kind) { return kind == PK_F32 ? 2.0 * 16 * 16 * 4 : 2.0 * 16 * 16 * 32; }

template <int KIND>
__global__ void __launch_bounds__(PK_THREADS) mfma_peak_kernel(float *__restrict__ out, int iters, float seed) {
    const int lane = threadIdx.x & 63;
    // small lane-dependent operands and distinct starting sums: the sums stay
    // finite and the compiler cannot fold or merge the chains
    const float x = seed * (float)(1 + (lane & 7));
    f32x4 acc[PK_CHAINS];
#pragma unroll
    for (int k = 0; k < PK_CHAINS; ++k) acc[k] = f32x4{x * (float)(k + 1), x, 2.f * x, 3.f * x};
    if constexpr (KIND == PK_F32) {
        for (int i = 0; i < iters; ++i)
#pragma unroll
            for (int k = 0; k < PK_CHAINS; ++k) acc[k] = __builtin_amdgcn_mfma_f32_16x16x4f32(x, seed, acc[k], 0, 0, 0);
    } else if constexpr (KIND == PK_BF16) {
        bf16x8 a, b;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            a[e] = (__bf16)(x * (float)(e + 1));
            b[e] = (__bf16)seed;
        }
        for (int i = 0; i < iters; ++i)
#pragma unroll
            for (int k = 0; k < PK_CHAINS; ++k) acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[k], 0, 0, 0);
    } else {
        f16x8 a, b;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            a[e] = (_Float16)(x * (float)(e + 1));
            b[e] = (_Float16)seed;
        }
        for (int i = 0; i < iters; ++i)
#pragma unroll
            for (int k = 0; k < PK_CHAINS; ++k) acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, acc[k], 0, 0, 0);
    }
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < PK_CHAINS; ++k) s += acc[k][0] + acc[k][1] + acc[k][2] + acc[k][3];
    out[(long long)blockIdx.x * PK_THREADS + threadIdx.x] = s;
}

// Streaming read: each lane sums 16-byte words with a grid stride, 8 loads in
// flight per iteration (non-temporal: the data are used once).
constexpr int RD_UNROLL = 8;
__global__ void __launch_bounds__(PK_THREADS) hbm_read_kernel(const f32x4 *__restrict__ src, long long n,
                                                               float *__restrict__ out) {
    const long long stride = (long long)gridDim.x * PK_THREADS;
    long long i = (long long)blockIdx.x * PK_THREADS + threadIdx.x;
    f32x4 s = f32x4{0.f, 0.f, 0.f, 0.f};
    for (; i + (RD_UNROLL - 1) * stride < n; i += RD_UNROLL * stride) {
        f32x4 v[RD_UNROLL];
#pragma unroll
        for (int u = 0; u < RD_UNROLL; ++u) v[u] = __builtin_nontemporal_load(src + i + u * stride);
#pragma unroll
        for (int u = 0; u < RD_UNROLL; ++u) s += v[u];
    }
    for (; i < n; i += stride) s += __builtin_nontemporal_load(src + i);
    out[(long long)blockIdx.x * PK_THREADS + threadIdx.x] = s[0] + s[1] + s[2] + s[3];
}

// Streaming copy, 4 words in flight per lane and iteration.
constexpr int CP_UNROLL = 4;
__global__ void __launch_bounds__(PK_THREADS) hbm_copy_kernel(const f32x4 *__restrict__ src, f32x4 *__restrict__ dst,
                                                               long long n) {
    const long long stride = (long long)gridDim.x * PK_THREADS;
    long long i = (long long)blockIdx.x * PK_THREADS + threadIdx.x;
    for (; i + (CP_UNROLL - 1) * stride < n; i += CP_UNROLL * stride) {
        f32x4 v[CP_UNROLL];
#pragma unroll
        for (int u = 0; u < CP_UNROLL; ++u) v[u] = __builtin_nontemporal_load(src + i + u * stride);
#pragma unroll
        for (int u = 0; u < CP_UNROLL; ++u) __builtin_nontemporal_store(v[u], dst + i + u * stride);
    }
    for (; i < n; i += stride) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    ~Timer() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

struct Mem {
    void *p = nullptr;
    ~Mem() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace
}  // namespace dc

using namespace dc;

// Best of 3 timed launches after one untimed launch (ms).
template <class F>
static int best_ms(dc_ctx *c, Timer &t, F launch, double &ms) {
    HIPCHK(hipSetDevice(c->device));
    launch();
    HIPCHK(hipGetLastError());
    ms = 1e30;
    for (int r = 0; r < 3; ++r) {
        HIPCHK(hipEventRecord(t.a, c->st));
        launch();
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(t.b, c->st));
        HIPCHK(hipEventSynchronize(t.b));
        float m = 0.f;
        HIPCHK(hipEventElapsedTime(&m, t.a, t.b));
        ms = std::min(ms, (double)m);
    }
    return DC_OK;
}

extern "C" int dc_measure_peaks(dc_ctx *c, int64_t hbm_bytes, dc_peaks *out) {
    if (!c || !out) return DC_ERR_ARG;
    c->err.clear();
    if (hbm_bytes == 0) hbm_bytes = (int64_t)1 << 30;
    if (hbm_bytes < ((int64_t)1 << 24) || hbm_bytes > ((int64_t)1 << 36) || hbm_bytes % 16)
        return set_err(c, DC_ERR_ARG, DC_ER_LIB, "dc_measure_peaks: hbm_bytes must be a multiple of 16 in [16 MiB, 64 GiB]");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->st));
    int ncu = 0;
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device));
    if (ncu <= 0) return set_err(c, DC_ERR_DEVICE, DC_ER_LIB, "dc_measure_peaks: no compute units reported");

    Timer t;
    HIPCHK(hipEventCreate(&t.a));
    HIPCHK(hipEventCreate(&t.b));
    // MFMA: 8 waves per SIMD (8 work-groups of 4 waves per CU)
    const int grid = ncu * 8;
    const long long nbuf = std::max<long long>((long long)grid * PK_THREADS, (long long)ncu * 64 * PK_THREADS);
    Mem sink;
    HIPCHK(hipMalloc(&sink.p, sizeof(float) * nbuf));
    float *d_out = static_cast<float *>(sink.p);
    const long long waves = (long long)grid * (PK_THREADS / 64);
    double ms = 0;
    // iteration counts: about 1 ms per launch at the published peaks (short
    // launches under-read the ceiling: 0.13 ms launches reached 2.06-2.10
    // PFLOP/s BF16 where 1 ms launches reach 2.4)
    const int it16 = 2048, it32 = 1024;
    int st = best_ms(c, t, [&] { mfma_peak_kernel<PK_BF16><<<grid, PK_THREADS, 0, c->st>>>(d_out, it16, 1.0f / 4096); }, ms);
    if (st) return st;
    out->mfma_bf16_tflops = waves * (double)it16 * PK_CHAINS * pk_flops(PK_BF16) / (ms * 1e-3) / 1e12;
    st = best_ms(c, t, [&] { mfma_peak_kernel<PK_F16><<<grid, PK_THREADS, 0, c->st>>>(d_out, it16, 1.0f / 4096); }, ms);
    if (st) return st;
    out->mfma_f16_tflops = waves * (double)it16 * PK_CHAINS * pk_flops(PK_F16) / (ms * 1e-3) / 1e12;
    st = best_ms(c, t, [&] { mfma_peak_kernel<PK_F32><<<grid, PK_THREADS, 0, c->st>>>(d_out, it32, 1.0f / 4096); }, ms);
    if (st) return st;
    out->mfma_f32_tflops = waves * (double)it32 * PK_CHAINS * pk_flops(PK_F32) / (ms * 1e-3) / 1e12;

    // HBM: two buffers of hbm_bytes, zero-filled so the sums stay finite
    Mem src, dst;
    HIPCHK(hipMalloc(&src.p, (size_t)hbm_bytes));
    HIPCHK(hipMalloc(&dst.p, (size_t)hbm_bytes));
    HIPCHK(hipMemsetAsync(src.p, 0, (size_t)hbm_bytes, c->st));
    const long long n16 = hbm_bytes / 16;
    const int sgrid = ncu * 64;   // 16 waves per CU, outputs fit the sink (ncu * 64 * 256 floats)
    st = best_ms(c, t, [&] { hbm_read_kernel<<<sgrid, PK_THREADS, 0, c->st>>>(static_cast<const f32x4 *>(src.p), n16, d_out); }, ms);
    if (st) return st;
    out->hbm_read_gbs = (double)hbm_bytes / (ms * 1e-3) / 1e9;
    st = best_ms(c, t, [&] {
        hbm_copy_kernel<<<sgrid, PK_THREADS, 0, c->st>>>(static_cast<const f32x4 *>(src.p), static_cast<f32x4 *>(dst.p), n16);
    }, ms);
    if (st) return st;
    out->hbm_copy_gbs = 2.0 * (double)hbm_bytes / (ms * 1e-3) / 1e9;
    out->hbm_bytes = hbm_bytes;
    out->num_cu = ncu;
    HIPCHK(hipStreamSynchronize(c->st));
    return DC_OK;
}
