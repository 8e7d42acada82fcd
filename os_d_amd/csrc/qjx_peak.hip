// qjx_peak.hip -- measured throughput ceilings of the device the engine runs on:
// matrix-core FP32 (32x32x2), FP64 (16x16x4) and BF16 (32x32x16) MFMA issue rates,
// the FP64 vector FMA rate and the HBM copy bandwidth.  The roofline fractions of
// the trajectory kernels are quoted against these as well as the published peaks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>

#include "../../include/qjx.h"

namespace {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef double doublex4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kChains = 4;   // independent accumulators per wave (cover the MFMA latency)
constexpr int kChains64 = 8; // the FP64 MFMA has the longer latency per flop

__global__ __launch_bounds__(256) void peak_f32_kernel(float *out, int iters, float seed)
{
    floatx16 acc[kChains];
    for (int c = 0; c < kChains; c++)
        for (int i = 0; i < 16; i++) acc[c][i] = 0.0f;
    const float a = seed + threadIdx.x * 1e-7f, b = seed - threadIdx.x * 1e-7f;
    for (int it = 0; it < iters; it++)
#pragma unroll
        for (int c = 0; c < kChains; c++) acc[c] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc[c], 0, 0, 0);
    float s = 0.0f;
    for (int c = 0; c < kChains; c++)
        for (int i = 0; i < 16; i++) s += acc[c][i];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void peak_f64_kernel(double *out, int iters, double seed)
{
    doublex4 acc[kChains64];
    for (int c = 0; c < kChains64; c++)
        for (int i = 0; i < 4; i++) acc[c][i] = 0.0;
    const double a = seed + threadIdx.x * 1e-9, b = seed - threadIdx.x * 1e-9;
    for (int it = 0; it < iters; it++)
#pragma unroll
        for (int c = 0; c < kChains64; c++) acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[c], 0, 0, 0);
    double s = 0.0;
    for (int c = 0; c < kChains64; c++)
        for (int i = 0; i < 4; i++) s += acc[c][i];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void peak_bf16_kernel(float *out, int iters, float seed)
{
    floatx16 acc[kChains];
    for (int c = 0; c < kChains; c++)
        for (int i = 0; i < 16; i++) acc[c][i] = 0.0f;
    bf16x8 a, b;
    for (int i = 0; i < 8; i++) {
        a[i] = (__bf16)(seed + i * 1e-3f);
        b[i] = (__bf16)(seed - i * 1e-3f);
    }
    for (int it = 0; it < iters; it++)
#pragma unroll
        for (int c = 0; c < kChains; c++) acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[c], 0, 0, 0);
    float s = 0.0f;
    for (int c = 0; c < kChains; c++)
        for (int i = 0; i < 16; i++) s += acc[c][i];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// FP64 vector FMA: 8 independent chains per lane
__global__ __launch_bounds__(256) void peak_f64_valu_kernel(double *out, int iters, double seed)
{
    double v[8];
    for (int c = 0; c < 8; c++) v[c] = seed + c * 1e-3 + threadIdx.x * 1e-9;
    const double m = 1.0 - 1e-12, k = 1e-13;
    for (int it = 0; it < iters; it++)
#pragma unroll
        for (int c = 0; c < 8; c++) v[c] = __builtin_fma(v[c], m, k);
    double s = 0.0;
    for (int c = 0; c < 8; c++) s += v[c];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// four independent 16-byte loads in flight per thread; n is a multiple of 4 * grid size
__global__ __launch_bounds__(256) void copy_kernel(const float4 *__restrict__ src, float4 *__restrict__ dst, size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i + 3 * stride < n; i += 4 * stride) {
        const float4 v0 = src[i], v1 = src[i + stride], v2 = src[i + 2 * stride], v3 = src[i + 3 * stride];
        dst[i] = v0;
        dst[i + stride] = v1;
        dst[i + 2 * stride] = v2;
        dst[i + 3 * stride] = v3;
    }
}

template <typename F>
hipError_t timed(F launch, hipStream_t s, float *ms)
{
    hipEvent_t a, b;
    hipError_t e = hipEventCreate(&a);
    if (e != hipSuccess) return e;
    if ((e = hipEventCreate(&b)) != hipSuccess) {
        (void)hipEventDestroy(a);
        return e;
    }
    launch();                                   // warm-up
    e = hipEventRecord(a, s);
    if (e == hipSuccess) launch();
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(b, s);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return e;
}

}  // namespace

namespace qjx {
void set_last_error(const std::string &msg);
}

extern "C" int qjx_peak_probe(double *f32_mfma_tflops, double *f64_mfma_tflops, double *bf16_mfma_tflops,
                              double *f64_valu_tflops, double *hbm_gbps)
{
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) {
        qjx::set_last_error("no HIP device available");
        return QJX_ERR_HIP;
    }
    const int blocks = cus * 8, threads = 256;   // 32 waves per CU (8 per SIMD)
    const size_t nout = (size_t)blocks * threads;
    const size_t nbytes = (size_t)2 << 30;       // 2 GiB copied per launch
    void *out = nullptr, *src = nullptr, *dst = nullptr;
    hipStream_t s = nullptr;
    auto release = [&]() {
        if (out) (void)hipFree(out);
        if (src) (void)hipFree(src);
        if (dst) (void)hipFree(dst);
        if (s) (void)hipStreamDestroy(s);
    };
    if (hipStreamCreate(&s) != hipSuccess || hipMalloc(&out, nout * sizeof(double)) != hipSuccess ||
        hipMalloc(&src, nbytes) != hipSuccess || hipMalloc(&dst, nbytes) != hipSuccess ||
        hipMemsetAsync(src, 0, nbytes, s) != hipSuccess) {
        release();
        qjx::set_last_error("peak probe allocation failed");
        return QJX_ERR_NOMEM;
    }
    float ms = 0.0f;
    hipError_t e = hipSuccess;
    const double waves = double(blocks) * threads / 64.0;
    const int it32 = 4096, it64 = 4096, itbf = 4096, itv = 32768;
    if (e == hipSuccess && f32_mfma_tflops) {
        e = timed([&] { peak_f32_kernel<<<blocks, threads, 0, s>>>((float *)out, it32, 1.0f); }, s, &ms);
        *f32_mfma_tflops = waves * it32 * kChains * (2.0 * 32 * 32 * 2) / (ms * 1e-3) / 1e12;
    }
    if (e == hipSuccess && f64_mfma_tflops) {
        e = timed([&] { peak_f64_kernel<<<blocks, threads, 0, s>>>((double *)out, it64, 1.0); }, s, &ms);
        *f64_mfma_tflops = waves * it64 * kChains64 * (2.0 * 16 * 16 * 4) / (ms * 1e-3) / 1e12;
    }
    if (e == hipSuccess && bf16_mfma_tflops) {
        e = timed([&] { peak_bf16_kernel<<<blocks, threads, 0, s>>>((float *)out, itbf, 1.0f); }, s, &ms);
        *bf16_mfma_tflops = waves * itbf * kChains * (2.0 * 32 * 32 * 16) / (ms * 1e-3) / 1e12;
    }
    if (e == hipSuccess && f64_valu_tflops) {
        e = timed([&] { peak_f64_valu_kernel<<<blocks, threads, 0, s>>>((double *)out, itv, 1.0); }, s, &ms);
        *f64_valu_tflops = double(blocks) * threads * itv * 8 * 2.0 / (ms * 1e-3) / 1e12;
    }
    if (e == hipSuccess && hbm_gbps) {
        const size_t n4 = nbytes / sizeof(float4);
        e = timed([&] { copy_kernel<<<cus * 32, 256, 0, s>>>((const float4 *)src, (float4 *)dst, n4); }, s, &ms);
        *hbm_gbps = 2.0 * double(nbytes) / (ms * 1e-3) / 1e9;
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    release();
    if (e != hipSuccess) {
        qjx::set_last_error(std::string("peak probe: ") + hipGetErrorString(e));
        return QJX_ERR_HIP;
    }
    return QJX_OK;
}
