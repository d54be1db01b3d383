// peak.hip -- register-only throughput probes of the units the distance
// kernels run on (the measured peaks the roofline fractions divide by):
//   int8 MFMA  v_mfma_i32_32x32x32_i8            (binary product, int8 form)
//   fp4 MFMA   v_mfma_scale_f32_32x32x64_f8f6f4  (binary product, fp4 form)
//   VALU       v_and_b32 + v_bcnt_u32_b32        (AND + popcount form)
//   HBM        16-byte streaming copy
// Same role as the reference's tensorcore/peak.cu:8-54 (1-bit mma loop and a
// cuBLAS int8 GEMM), measured here with kernels of this library.

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>

#include "cgmlst_hamming.h"

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v8i __attribute__((ext_vector_type(8)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef float v16f __attribute__((ext_vector_type(16)));

constexpr int CHAINS = 8;  // independent accumulators per wavefront

__global__ void __launch_bounds__(256) peak_i8(int iters, int* sink) {
    v4i a, b;
    for (int t = 0; t < 4; t++) {
        a[t] = (int)(threadIdx.x * 2654435761u + t);
        b[t] = (int)(threadIdx.x * 40503u + 7 * t);
    }
    v16i acc[CHAINS];
    for (int c = 0; c < CHAINS; c++)
        for (int e = 0; e < 16; e++) acc[c][e] = 0;
    for (int i = 0; i < iters; i++)
#pragma unroll
        for (int c = 0; c < CHAINS; c++) acc[c] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, acc[c], 0, 0, 0);
    int s = 0;
    for (int c = 0; c < CHAINS; c++)
        for (int e = 0; e < 16; e++) s += acc[c][e];
    if (s == 0x7fffffff) sink[0] = s;
}

__global__ void __launch_bounds__(256) peak_fp4(int iters, int* sink) {
    v8i a, b;
    for (int t = 0; t < 8; t++) {
        a[t] = t < 4 ? (int)((threadIdx.x * 2654435761u + t) & 0x22222222u) : 0;
        b[t] = t < 4 ? (int)((threadIdx.x * 40503u + 7 * t) & 0x22222222u) : 0;
    }
    v16f acc[CHAINS];
    for (int c = 0; c < CHAINS; c++)
        for (int e = 0; e < 16; e++) acc[c][e] = 0.f;
    for (int i = 0; i < iters; i++)
#pragma unroll
        for (int c = 0; c < CHAINS; c++)
            acc[c] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b, acc[c], 4, 4, 0, 127, 0, 127);
    float s = 0.f;
    for (int c = 0; c < CHAINS; c++)
        for (int e = 0; e < 16; e++) s += acc[c][e];
    if (s == 1234.5f) sink[0] = (int)s;
}

// 32 bit-pairs per (AND, BCNT) pair of VALU instructions.
__global__ void __launch_bounds__(256) peak_popc(int iters, int* sink) {
    uint32_t a[CHAINS], acc[CHAINS];
    uint32_t b = threadIdx.x * 2654435761u;
    for (int c = 0; c < CHAINS; c++) {
        a[c] = threadIdx.x * 40503u + 977u * c;
        acc[c] = 0;
    }
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int c = 0; c < CHAINS; c++) acc[c] = (uint32_t)__builtin_popcount(a[c] & b) + acc[c];
        b += 0x9e3779b9u;
    }
    uint32_t s = 0;
    for (int c = 0; c < CHAINS; c++) s += acc[c];
    if (s == 0xffffffffu) sink[0] = (int)s;
}

__global__ void __launch_bounds__(256) stream_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
        dst[k] = src[k];
}

void set_err(char* err, int errlen, const char* fmt, ...) __attribute__((format(printf, 3, 4)));
void set_err(char* err, int errlen, const char* fmt, ...) {
    if (!err || errlen <= 0) return;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(err, (size_t)errlen, fmt, ap);
    va_end(ap);
}

#define PK(x)                                                                  \
    do {                                                                       \
        hipError_t e_ = (x);                                                   \
        if (e_ != hipSuccess) {                                                \
            set_err(err, errlen, "%s: %s", #x, hipGetErrorString(e_));         \
            goto fail;                                                         \
        }                                                                      \
    } while (0)

}  // namespace

extern "C" int cgh_measure_peaks(int device, double* i8_tops, double* fp4_tflops, double* popc_tops,
                                 double* hbm_gbs, char* err, int errlen) {
    int* sink = NULL;
    uint4 *src = NULL, *dst = NULL;
    hipEvent_t t0 = NULL, t1 = NULL;
    hipDeviceProp_t prop;
    float ms = 0.f;
    int rc = 1;
    const int iters = 4000;
    const size_t bytes = (size_t)2 << 30;
    PK(hipSetDevice(device));
    PK(hipGetDeviceProperties(&prop, device));
    PK(hipMalloc(&sink, 4));
    PK(hipEventCreate(&t0));
    PK(hipEventCreate(&t1));
    {
        // 4 wavefronts per workgroup (one per SIMD), 2 workgroups per CU
        const dim3 grid((unsigned)prop.multiProcessorCount * 2), block(256);
        const double waves = (double)grid.x * 4.0;
        hipLaunchKernelGGL(peak_i8, grid, block, 0, 0, 16, sink);
        PK(hipEventRecord(t0));
        hipLaunchKernelGGL(peak_i8, grid, block, 0, 0, iters, sink);
        PK(hipEventRecord(t1));
        PK(hipEventSynchronize(t1));
        PK(hipEventElapsedTime(&ms, t0, t1));
        if (i8_tops) *i8_tops = waves * iters * CHAINS * 2.0 * 32 * 32 * 32 / (ms * 1e-3) / 1e12;
        hipLaunchKernelGGL(peak_fp4, grid, block, 0, 0, 16, sink);
        PK(hipEventRecord(t0));
        hipLaunchKernelGGL(peak_fp4, grid, block, 0, 0, iters, sink);
        PK(hipEventRecord(t1));
        PK(hipEventSynchronize(t1));
        PK(hipEventElapsedTime(&ms, t0, t1));
        if (fp4_tflops) *fp4_tflops = waves * iters * CHAINS * 2.0 * 32 * 32 * 64 / (ms * 1e-3) / 1e12;
        // popcount: 8 waves per CU-SIMD pair of launches keeps the VALU fed
        const dim3 pgrid((unsigned)prop.multiProcessorCount * 8);
        hipLaunchKernelGGL(peak_popc, pgrid, block, 0, 0, 16, sink);
        PK(hipEventRecord(t0));
        hipLaunchKernelGGL(peak_popc, pgrid, block, 0, 0, iters * 4, sink);
        PK(hipEventRecord(t1));
        PK(hipEventSynchronize(t1));
        PK(hipEventElapsedTime(&ms, t0, t1));
        // one AND + one BCNT per 32 bit-pairs; counted as 2 ops per bit pair
        if (popc_tops)
            *popc_tops = (double)pgrid.x * 256.0 * iters * 4.0 * CHAINS * 32.0 * 2.0 / (ms * 1e-3) / 1e12;
    }
    PK(hipMalloc(&src, bytes));
    PK(hipMalloc(&dst, bytes));
    PK(hipMemset(src, 1, bytes));
    {
        const dim3 grid((unsigned)prop.multiProcessorCount * 16), block(256);
        hipLaunchKernelGGL(stream_copy, grid, block, 0, 0, src, dst, bytes / 16);
        PK(hipEventRecord(t0));
        for (int r = 0; r < 5; r++) hipLaunchKernelGGL(stream_copy, grid, block, 0, 0, src, dst, bytes / 16);
        PK(hipEventRecord(t1));
        PK(hipEventSynchronize(t1));
        PK(hipEventElapsedTime(&ms, t0, t1));
        if (hbm_gbs) *hbm_gbs = 5.0 * 2.0 * (double)bytes / (ms * 1e-3) / 1e9;
    }
    PK(hipGetLastError());
    rc = 0;
fail:
    if (sink) (void)hipFree(sink);
    if (src) (void)hipFree(src);
    if (dst) (void)hipFree(dst);
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    return rc;
}
