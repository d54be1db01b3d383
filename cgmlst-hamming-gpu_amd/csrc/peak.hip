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

#include <algorithm>

#include "cgmlst_hamming.h"
#include "internal.h"

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

// Streaming copy: U independent 16-byte loads in flight per lane, each
// word read and written once, non-temporal (no reuse to keep in the caches).
template <int U>
__global__ void __launch_bounds__(256) stream_copy(const v4i* __restrict__ src, v4i* __restrict__ dst, size_t n) {
    const size_t stride = (size_t)gridDim.x * 256;
    size_t k = (size_t)blockIdx.x * 256 + threadIdx.x;
    for (; k + (U - 1) * stride < n; k += U * stride) {
        v4i v[U];
#pragma unroll
        for (int u = 0; u < U; u++) v[u] = __builtin_nontemporal_load(&src[k + u * stride]);
#pragma unroll
        for (int u = 0; u < U; u++) __builtin_nontemporal_store(v[u], &dst[k + u * stride]);
    }
    for (; k < n; k += stride) dst[k] = src[k];
}

// Read-only stream: `passes` sweeps over n 16-byte words, U loads in flight
// per lane; a word-sum that is never the sentinel keeps the loads alive.
template <int U>
__global__ void __launch_bounds__(256) stream_read(const v4i* __restrict__ src, size_t n, int passes, int* sink) {
    const size_t stride = (size_t)gridDim.x * 256;
    int acc = 0;
    for (int p = 0; p < passes; p++) {
        // each pass reads every word, rotated by half the array on odd
        // passes, so no pass can reuse another's loads
        const size_t rot = (p & 1) ? n / 2 : 0;
        for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < n; k += U * stride) {
            v4i v[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                size_t w = k + u * stride;
                w = w < n ? w + rot : k + rot;  // past the end: reload this lane's first word
                if (w >= n) w -= n;
                v[u] = src[w];
            }
#pragma unroll
            for (int u = 0; u < U; u++) acc ^= v[u][0] + v[u][1] + v[u][2] + v[u][3];
        }
        asm volatile("" : : : "memory");  // every pass loads again (no hoisting across passes)
    }
    if (acc == 0x7fffffff) sink[0] = acc;
}



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
    HIP_TRY(hipSetDevice(device));
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    HIP_TRY(hipMalloc(&sink, 4));
    HIP_TRY(hipEventCreate(&t0));
    HIP_TRY(hipEventCreate(&t1));
    {
        // 4 wavefronts per workgroup (one per SIMD), 2 workgroups per CU
        const dim3 grid((unsigned)prop.multiProcessorCount * 2), block(256);
        const double waves = (double)grid.x * 4.0;
        hipLaunchKernelGGL(peak_i8, grid, block, 0, 0, 16, sink);
        HIP_TRY(hipEventRecord(t0));
        hipLaunchKernelGGL(peak_i8, grid, block, 0, 0, iters, sink);
        HIP_TRY(hipEventRecord(t1));
        HIP_TRY(hipEventSynchronize(t1));
        HIP_TRY(hipEventElapsedTime(&ms, t0, t1));
        if (i8_tops) *i8_tops = waves * iters * CHAINS * 2.0 * 32 * 32 * 32 / (ms * 1e-3) / 1e12;
        hipLaunchKernelGGL(peak_fp4, grid, block, 0, 0, 16, sink);
        HIP_TRY(hipEventRecord(t0));
        hipLaunchKernelGGL(peak_fp4, grid, block, 0, 0, iters, sink);
        HIP_TRY(hipEventRecord(t1));
        HIP_TRY(hipEventSynchronize(t1));
        HIP_TRY(hipEventElapsedTime(&ms, t0, t1));
        if (fp4_tflops) *fp4_tflops = waves * iters * CHAINS * 2.0 * 32 * 32 * 64 / (ms * 1e-3) / 1e12;
        // popcount: 8 waves per CU-SIMD pair of launches keeps the VALU fed
        const dim3 pgrid((unsigned)prop.multiProcessorCount * 8);
        hipLaunchKernelGGL(peak_popc, pgrid, block, 0, 0, 16, sink);
        HIP_TRY(hipEventRecord(t0));
        hipLaunchKernelGGL(peak_popc, pgrid, block, 0, 0, iters * 4, sink);
        HIP_TRY(hipEventRecord(t1));
        HIP_TRY(hipEventSynchronize(t1));
        HIP_TRY(hipEventElapsedTime(&ms, t0, t1));
        // one AND + one BCNT per 32 bit-pairs; counted as 2 ops per bit pair
        if (popc_tops)
            *popc_tops = (double)pgrid.x * 256.0 * iters * 4.0 * CHAINS * 32.0 * 2.0 / (ms * 1e-3) / 1e12;
    }
    HIP_TRY(hipMalloc(&src, bytes));
    HIP_TRY(hipMalloc(&dst, bytes));
    HIP_TRY(hipMemset(src, 1, bytes));
    {
        // the best of a few grid sizes of the 4-loads-in-flight copy
        double best = 0;
        for (unsigned per_cu : {4u, 8u, 16u}) {
            const dim3 grid((unsigned)prop.multiProcessorCount * per_cu), block(256);
            const v4i* s4 = reinterpret_cast<const v4i*>(src);
            v4i* d4 = reinterpret_cast<v4i*>(dst);
            hipLaunchKernelGGL(stream_copy<4>, grid, block, 0, 0, s4, d4, bytes / 16);
            HIP_TRY(hipEventRecord(t0));
            for (int r = 0; r < 5; r++) hipLaunchKernelGGL(stream_copy<4>, grid, block, 0, 0, s4, d4, bytes / 16);
            HIP_TRY(hipEventRecord(t1));
            HIP_TRY(hipEventSynchronize(t1));
            HIP_TRY(hipEventElapsedTime(&ms, t0, t1));
            best = std::max(best, 5.0 * 2.0 * (double)bytes / (ms * 1e-3) / 1e9);
        }
        if (hbm_gbs) *hbm_gbs = best;
    }
    HIP_TRY(hipGetLastError());
    rc = 0;
fail:
    if (sink) (void)hipFree(sink);
    if (src) (void)hipFree(src);
    if (dst) (void)hipFree(dst);
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    return rc;
}

extern "C" int cgh_measure_memory(int device, double* copy_gbs, double* read_gbs, double* cache_read_gbs, char* err,
                                  int errlen) {
    int* sink = NULL;
    v4i *src = NULL, *dst = NULL;
    hipEvent_t t0 = NULL, t1 = NULL;
    hipDeviceProp_t prop;
    float ms = 0.f;
    int rc = 1;
    const size_t bytes = (size_t)4 << 30, cached = (size_t)64 << 20;
    HIP_TRY(hipSetDevice(device));
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    HIP_TRY(hipMalloc(&sink, 4));
    HIP_TRY(hipEventCreate(&t0));
    HIP_TRY(hipEventCreate(&t1));
    HIP_TRY(hipMalloc(&src, bytes));
    HIP_TRY(hipMalloc(&dst, bytes));
    HIP_TRY(hipMemset(src, 3, bytes));
    {
        const dim3 block(256);
        double best_copy = 0, best_read = 0, best_cache = 0;
        for (unsigned per_cu : {4u, 8u, 16u, 32u}) {
            const dim3 grid((unsigned)prop.multiProcessorCount * per_cu);
            // copy: 4 GiB read + 4 GiB written per launch
            hipLaunchKernelGGL(stream_copy<4>, grid, block, 0, 0, src, dst, bytes / 16);
            HIP_TRY(hipEventRecord(t0));
            for (int r = 0; r < 3; r++) hipLaunchKernelGGL(stream_copy<4>, grid, block, 0, 0, src, dst, bytes / 16);
            HIP_TRY(hipEventRecord(t1));
            HIP_TRY(hipEventSynchronize(t1));
            HIP_TRY(hipEventElapsedTime(&ms, t0, t1));
            best_copy = std::max(best_copy, 3.0 * 2.0 * (double)bytes / (ms * 1e-3) / 1e9);
            // read: 4 GiB once per launch (far beyond every cache)
            hipLaunchKernelGGL(stream_read<4>, grid, block, 0, 0, src, bytes / 16, 1, sink);
            HIP_TRY(hipEventRecord(t0));
            for (int r = 0; r < 3; r++) hipLaunchKernelGGL(stream_read<4>, grid, block, 0, 0, src, bytes / 16, 1, sink);
            HIP_TRY(hipEventRecord(t1));
            HIP_TRY(hipEventSynchronize(t1));
            HIP_TRY(hipEventElapsedTime(&ms, t0, t1));
            best_read = std::max(best_read, 3.0 * (double)bytes / (ms * 1e-3) / 1e9);
            // 64 MiB read 40 times: resident in the Infinity Cache, not in the L2s
            hipLaunchKernelGGL(stream_read<4>, grid, block, 0, 0, src, cached / 16, 2, sink);
            HIP_TRY(hipEventRecord(t0));
            hipLaunchKernelGGL(stream_read<4>, grid, block, 0, 0, src, cached / 16, 40, sink);
            HIP_TRY(hipEventRecord(t1));
            HIP_TRY(hipEventSynchronize(t1));
            HIP_TRY(hipEventElapsedTime(&ms, t0, t1));
            best_cache = std::max(best_cache, 40.0 * (double)cached / (ms * 1e-3) / 1e9);
        }
        if (copy_gbs) *copy_gbs = best_copy;
        if (read_gbs) *read_gbs = best_read;
        if (cache_read_gbs) *cache_read_gbs = best_cache;
    }
    HIP_TRY(hipGetLastError());
    rc = 0;
fail:
    if (sink) (void)hipFree(sink);
    if (src) (void)hipFree(src);
    if (dst) (void)hipFree(dst);
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    return rc;
}
