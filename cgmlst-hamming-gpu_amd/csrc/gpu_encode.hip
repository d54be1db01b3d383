// gpu_encode.hip -- the encoder on the device: u64 allele ids straight to a
// device-resident dataset.
//
// Per locus the calls are sorted by id with their row index (one segmented
// radix sort over the locus-major transpose; the sort is stable, so equal ids
// keep ascending rows).  In sorted order, the dense rank of a call is the
// number of distinct non-zero ids up to it in its locus (a flag scan), and
// every (locus, rank) is one contiguous run whose length is the allele count
// and whose row indices, ascending, are the allele's member list.  Only the
// per-allele counts go to the host, where tau is chosen and the column and
// group tables are laid out (a few bytes per distinct allele); then the
// device scatters the ranks into X, the rare runs into the member array, the
// row index of member ranges (a counting sort by row) and the O/P planes.
//
// Restates on the device: tensorcore/rs/src/main.rs:134-165 (dense_ranks),
// 167-198 (choose_tau, via cgh_choose_tau), 366-403 (column map, bit pack),
// 404-417 (rare groups; the reference then expands them into a CSR).

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "cgmlst_hamming.h"
#include "internal.h"

namespace {

void set_err(char* err, int errlen, const char* fmt, ...) __attribute__((format(printf, 3, 4)));
void set_err(char* err, int errlen, const char* fmt, ...) {
    if (!err || errlen <= 0) return;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(err, (size_t)errlen, fmt, ap);
    va_end(ap);
}

size_t up(size_t x, size_t m) { return (x + m - 1) / m * m; }


// keys[k n + i] = calls[i L + k], vals[k n + i] = i, through a 32 x 32 LDS
// tile so that both the reads and the writes are coalesced.
__global__ void __launch_bounds__(256)
transpose_calls(const uint64_t* __restrict__ calls, int n, int L, uint64_t* __restrict__ keys,
                uint32_t* __restrict__ vals) {
    __shared__ uint64_t t[32][33];
    const int k0 = blockIdx.x * 32, i0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int r = ty; r < 32; r += 8) {
        const int i = i0 + r, k = k0 + tx;
        if (i < n && k < L) t[r][tx] = calls[(size_t)i * L + k];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int k = k0 + r, i = i0 + tx;
        if (i < n && k < L) {
            keys[(size_t)k * n + i] = t[tx][r];
            vals[(size_t)k * n + i] = (uint32_t)i;
        }
    }
}

// flag[p] = 1 where a new distinct non-zero id starts inside its locus.
__global__ void __launch_bounds__(256)
distinct_flags(const uint64_t* __restrict__ keys, int n, size_t total, uint32_t* __restrict__ flag) {
    const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= total) return;
    const uint64_t v = keys[p];
    flag[p] = v != 0 && (p % (size_t)n == 0 || keys[p - 1] != v);
}

// distinct[k] = distinct non-zero ids of locus k (from the inclusive scan).
__global__ void __launch_bounds__(256)
locus_distinct(const uint32_t* __restrict__ scan, int n, int L, uint32_t* __restrict__ distinct) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= L) return;
    const uint32_t before = k ? scan[(size_t)k * n - 1] : 0u;
    distinct[k] = scan[(size_t)(k + 1) * n - 1] - before;
}

__device__ __forceinline__ uint32_t rank_at(const uint64_t* keys, const uint32_t* scan, int n, size_t p, int k) {
    const uint32_t before = k ? scan[(size_t)k * n - 1] : 0u;
    return keys[p] ? scan[p] - before : 0u;
}

// X[row, k] = rank; count of every (locus, rank) (rank 0 = missing).
__global__ void __launch_bounds__(256)
scatter_ranks(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals, const uint32_t* __restrict__ scan,
              int n, int L, const uint32_t* __restrict__ colstart, uint16_t* __restrict__ X,
              uint32_t* __restrict__ cnt) {
    const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= (size_t)n * L) return;
    const int k = (int)(p / (size_t)n);
    const uint32_t a = rank_at(keys, scan, n, p, k);
    X[(size_t)vals[p] * L + k] = (uint16_t)a;
    atomicAdd(&cnt[colstart[k] + a], 1u);
}

// Members of rare alleles: position in the allele's run -> its slot in the
// member array (group g's members at gstart[g] .., ascending rows), and the
// row index counts (member idx of a group of c has later partners iff
// idx + 1 < c).
template <typename M>
__global__ void __launch_bounds__(256)
rare_members(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals, const uint32_t* __restrict__ scan,
             int n, int L, const uint32_t* __restrict__ colstart, const int64_t* __restrict__ gid,
             const uint32_t* __restrict__ runstart, const uint32_t* __restrict__ gstart, M* __restrict__ gmem,
             uint32_t* __restrict__ rc) {
    const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= (size_t)n * L) return;
    const int k = (int)(p / (size_t)n);
    const uint32_t a = rank_at(keys, scan, n, p, k);
    const int64_t g = gid[colstart[k] + a];
    if (g < 0) return;
    const uint32_t idx = (uint32_t)(p % (size_t)n) - runstart[colstart[k] + a];
    const uint32_t row = vals[p];
    gmem[gstart[g] + idx] = (M)row;
    if (gstart[g] + idx + 1 < gstart[g + 1]) atomicAdd(&rc[row], 1u);
}

// Row index entries: row gmem[s] gets (end << 32) | (s + 1) for every member
// slot s of a group [begin, end) with s + 1 < end.
template <typename M>
__global__ void __launch_bounds__(256)
row_entries(const M* __restrict__ gmem, const uint32_t* __restrict__ gend_of, size_t members,
            uint32_t* __restrict__ cursor, uint64_t* __restrict__ rentry) {
    const size_t s = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (s >= members) return;
    const uint32_t end = gend_of[s];
    if (s + 1 >= end) return;
    const uint32_t slot = atomicAdd(&cursor[(uint32_t)gmem[s]], 1u);
    rentry[slot] = ((uint64_t)end << 32) | (uint64_t)(s + 1);
}

// gend_of[s] = end of the group holding member slot s.
__global__ void __launch_bounds__(256)
group_ends(const uint32_t* __restrict__ gstart, size_t groups, uint32_t* __restrict__ gend_of) {
    const size_t g = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= groups) return;
    for (uint32_t s = gstart[g]; s < gstart[g + 1]; s++) gend_of[s] = gstart[g + 1];
}

unsigned blocks(size_t items) { return (unsigned)((items + 255) / 256); }

}  // namespace

extern "C" int cgh_dataset_encode(int device, const uint64_t* calls, int n, int L, int tau, size_t out_bytes,
                                  void (*log)(void* ctx, const char* line), void* log_ctx, cgh_dataset** out,
                                  int* tau_used, uint16_t* X_out, char* err, int errlen) {
    const size_t total = (size_t)n * (size_t)L;
    uint64_t *dcalls = NULL, *dkeys = NULL, *dkeys2 = NULL;
    uint32_t *dvals = NULL, *dvals2 = NULL, *dflag = NULL, *dscan = NULL, *ddist = NULL, *dcs = NULL, *dcnt = NULL;
    uint32_t *dcm = NULL, *drun = NULL, *dgs = NULL, *drc = NULL, *dcur = NULL, *dgend = NULL;
    int64_t* dgid = NULL;
    uint16_t* dX = NULL;
    void* dtemp = NULL;
    size_t temp_bytes = 0;
    int* doff = NULL;
    cghi::DatasetParts ds;
    std::vector<uint32_t> distinct, colstart, cnt, colmap, runstart, gstart;
    std::vector<int64_t> gid;
    std::vector<int> off;
    size_t ncells = 0, D = 0, groups = 0, members = 0, rentries = 0;
    uint64_t rare = 0;
    *out = NULL;
    if (n < 1 || L < 1 || L > 65535 || tau < 0) {
        set_err(err, errlen, "cgh_dataset_encode: bad arguments (n=%d L=%d tau=%d)", n, L, tau);
        return 1;
    }
    if (total >= (1ull << 31)) {
        set_err(err, errlen, "cgh_dataset_encode: %zu calls exceed one sort (2^31)", total);
        return 1;
    }
    ds.device = device;
    ds.n = n;
    ds.L = L;
    HIP_TRY(hipSetDevice(device));
    // 1. locus-major keys with row indices, sorted per locus
    HIP_TRY(hipMalloc(&dcalls, total * 8));
    HIP_TRY(hipMemcpy(dcalls, calls, total * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&dkeys, total * 8));
    HIP_TRY(hipMalloc(&dvals, total * 4));
    hipLaunchKernelGGL(transpose_calls, dim3((unsigned)((L + 31) / 32), (unsigned)((n + 31) / 32)), dim3(256), 0, 0,
                       dcalls, n, L, dkeys, dvals);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipFree(dcalls));
    dcalls = NULL;
    HIP_TRY(hipMalloc(&dkeys2, total * 8));
    HIP_TRY(hipMalloc(&dvals2, total * 4));
    off.resize((size_t)L + 1);
    for (int k = 0; k <= L; k++) off[k] = k * n;
    HIP_TRY(hipMalloc(&doff, off.size() * sizeof(int)));
    HIP_TRY(hipMemcpy(doff, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipcub::DeviceSegmentedRadixSort::SortPairs(NULL, temp_bytes, dkeys, dkeys2, dvals, dvals2, (int)total, L, doff,
                                                   doff + 1));
    HIP_TRY(hipMalloc(&dtemp, std::max<size_t>(temp_bytes, 256)));
    HIP_TRY(hipcub::DeviceSegmentedRadixSort::SortPairs(dtemp, temp_bytes, dkeys, dkeys2, dvals, dvals2, (int)total, L,
                                                   doff, doff + 1));
    HIP_TRY(hipFree(dkeys));
    dkeys = NULL;
    HIP_TRY(hipFree(dvals));
    dvals = NULL;
    HIP_TRY(hipFree(dtemp));
    dtemp = NULL;
    // 2. dense ranks: scan of "new distinct id" flags
    HIP_TRY(hipMalloc(&dflag, total * 4));
    HIP_TRY(hipMalloc(&dscan, total * 4));
    hipLaunchKernelGGL(distinct_flags, dim3(blocks(total)), dim3(256), 0, 0, dkeys2, n, total, dflag);
    HIP_TRY(hipGetLastError());
    temp_bytes = 0;
    HIP_TRY(hipcub::DeviceScan::InclusiveSum(NULL, temp_bytes, dflag, dscan, (int)total));
    HIP_TRY(hipMalloc(&dtemp, std::max<size_t>(temp_bytes, 256)));
    HIP_TRY(hipcub::DeviceScan::InclusiveSum(dtemp, temp_bytes, dflag, dscan, (int)total));
    HIP_TRY(hipFree(dtemp));
    dtemp = NULL;
    HIP_TRY(hipFree(dflag));
    dflag = NULL;
    HIP_TRY(hipMalloc(&ddist, (size_t)L * 4));
    hipLaunchKernelGGL(locus_distinct, dim3(blocks((size_t)L)), dim3(256), 0, 0, dscan, n, L, ddist);
    HIP_TRY(hipGetLastError());
    distinct.resize((size_t)L);
    HIP_TRY(hipMemcpy(distinct.data(), ddist, (size_t)L * 4, hipMemcpyDeviceToHost));
    for (int k = 0; k < L; k++)
        if (distinct[k] > 65535) {
            set_err(err, errlen, "locus %d has %u distinct alleles (limit 65535)", k, distinct[k]);
            goto fail;
        }
    // 3. X and the allele counts
    colstart.resize((size_t)L + 1);
    for (int k = 0; k < L; k++) {
        colstart[k] = (uint32_t)ncells;
        ncells += (size_t)distinct[k] + 1;
    }
    colstart[L] = (uint32_t)ncells;
    HIP_TRY(hipMalloc(&dcs, colstart.size() * 4));
    HIP_TRY(hipMemcpy(dcs, colstart.data(), colstart.size() * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&dcnt, ncells * 4));
    HIP_TRY(hipMemset(dcnt, 0, ncells * 4));
    HIP_TRY(hipMalloc(&dX, total * 2));
    hipLaunchKernelGGL(scatter_ranks, dim3(blocks(total)), dim3(256), 0, 0, dkeys2, dvals2, dscan, n, L, dcs, dX,
                       dcnt);
    HIP_TRY(hipGetLastError());
    cnt.resize(ncells);
    HIP_TRY(hipMemcpy(cnt.data(), dcnt, ncells * 4, hipMemcpyDeviceToHost));
    // 4. tau and the tables (host: a few bytes per distinct allele)
    if (tau == 0) {
        std::vector<uint32_t> c1;
        c1.reserve(ncells);
        for (int k = 0; k < L; k++)
            for (uint32_t a = 1; a <= distinct[k]; a++) c1.push_back(cnt[colstart[k] + a]);
        tau = cgh_choose_tau(c1.data(), c1.size(), n, L, out_bytes, tch_free_vram(device), log, log_ctx);
        if (!tau) {
            set_err(err, errlen, "no tau fits in %.2f GB of free VRAM", tch_free_vram(device) / 1e9);
            goto fail;
        }
    }
    tau = std::min(tau, n + 1);
    if (tau_used) *tau_used = tau;
    ds.tau = tau;
    colmap.assign(ncells, 0xffffffffu);
    runstart.assign(ncells, 0);
    gid.assign(ncells, -1);
    gstart.push_back(0);
    for (int k = 0; k < L; k++) {
        uint32_t pos = 0;
        for (uint32_t a = 0; a <= distinct[k]; a++) {
            const size_t t = colstart[k] + a;
            const uint32_t c = cnt[t];
            runstart[t] = pos;
            pos += c;
            if (!a) continue;
            if (c >= (uint32_t)tau) {
                colmap[t] = (uint32_t)D++;
            } else if (c >= 2) {
                gid[t] = (int64_t)groups++;
                members += c;
                rentries += c - 1;
                rare += (uint64_t)c * (c - 1) / 2;
                gstart.push_back((uint32_t)members);
            }
        }
    }
    if (members >= 0xffffffffull) {
        set_err(err, errlen, "too many rare-group members: raise --tau");
        goto fail;
    }
    // 5. O/P planes from X
    ds.D = D;
    ds.n_pad = up((size_t)n, 256);
    ds.kw1 = (size_t)cghi::device_words(up(std::max<size_t>(D, 1), 256) / 32);
    ds.kw2 = (size_t)cghi::device_words(up((size_t)L, 256) / 32);
    HIP_TRY(hipMalloc(&dcm, ncells * 4));
    HIP_TRY(hipMemcpy(dcm, colmap.data(), ncells * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&ds.dO, ds.n_pad * ds.kw1 * 4));
    HIP_TRY(hipMalloc(&ds.dP, ds.n_pad * ds.kw2 * 4));
    HIP_TRY(hipMemset(ds.dO, 0, ds.n_pad * ds.kw1 * 4));
    HIP_TRY(hipMemset(ds.dP, 0, ds.n_pad * ds.kw2 * 4));
    HIP_TRY(cghi::launch_build_planes(dX, n, L, dcs, dcm, ds.dO, ds.dP, (int)ds.n_pad, 0));
    // 6. rare member lists and their row index
    ds.n_groups = groups;
    ds.n_members = members;
    ds.n_rentries = rentries;
    ds.n_rare = rare;
    ds.m16 = n <= 65536;
    if (groups) {
        HIP_TRY(hipMalloc(&drun, ncells * 4));
        HIP_TRY(hipMemcpy(drun, runstart.data(), ncells * 4, hipMemcpyHostToDevice));
        HIP_TRY(hipMalloc(&dgid, ncells * 8));
        HIP_TRY(hipMemcpy(dgid, gid.data(), ncells * 8, hipMemcpyHostToDevice));
        HIP_TRY(hipMalloc(&dgs, gstart.size() * 4));
        HIP_TRY(hipMemcpy(dgs, gstart.data(), gstart.size() * 4, hipMemcpyHostToDevice));
        HIP_TRY(hipMalloc(&ds.dgmem, cghi::member_array_bytes(members, ds.m16)));
        HIP_TRY(hipMalloc(&drc, ((size_t)n + 1) * 4));
        HIP_TRY(hipMemset(drc, 0, ((size_t)n + 1) * 4));
        if (ds.m16)
            hipLaunchKernelGGL(rare_members<uint16_t>, dim3(blocks(total)), dim3(256), 0, 0, dkeys2, dvals2, dscan, n,
                               L, dcs, dgid, drun, dgs, (uint16_t*)ds.dgmem, drc);
        else
            hipLaunchKernelGGL(rare_members<uint32_t>, dim3(blocks(total)), dim3(256), 0, 0, dkeys2, dvals2, dscan, n,
                               L, dcs, dgid, drun, dgs, (uint32_t*)ds.dgmem, drc);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMalloc(&ds.drstart, ((size_t)n + 1) * 4));
        temp_bytes = 0;
        HIP_TRY(hipcub::DeviceScan::ExclusiveSum(NULL, temp_bytes, drc, ds.drstart, n + 1));
        HIP_TRY(hipMalloc(&dtemp, std::max<size_t>(temp_bytes, 256)));
        HIP_TRY(hipcub::DeviceScan::ExclusiveSum(dtemp, temp_bytes, drc, ds.drstart, n + 1));
        HIP_TRY(hipMalloc(&dcur, ((size_t)n + 1) * 4));
        HIP_TRY(hipMemcpy(dcur, ds.drstart, ((size_t)n + 1) * 4, hipMemcpyDeviceToDevice));
        HIP_TRY(hipMalloc(&dgend, members * 4));
        hipLaunchKernelGGL(group_ends, dim3(blocks(groups)), dim3(256), 0, 0, dgs, groups, dgend);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMalloc(&ds.drentry, std::max<size_t>(rentries, 1) * 8));
        if (ds.m16)
            hipLaunchKernelGGL(row_entries<uint16_t>, dim3(blocks(members)), dim3(256), 0, 0,
                               (const uint16_t*)ds.dgmem, dgend, members, dcur, ds.drentry);
        else
            hipLaunchKernelGGL(row_entries<uint32_t>, dim3(blocks(members)), dim3(256), 0, 0,
                               (const uint32_t*)ds.dgmem, dgend, members, dcur, ds.drentry);
        HIP_TRY(hipGetLastError());
    }
    if (X_out) HIP_TRY(hipMemcpy(X_out, dX, total * 2, hipMemcpyDeviceToHost));
    HIP_TRY(hipDeviceSynchronize());
    *out = cghi::dataset_adopt(ds);
    if (!*out) {
        set_err(err, errlen, "out of host memory");
        goto fail;
    }
    ds = cghi::DatasetParts();  // owned by the dataset now
    {
        void* tmp[] = {dkeys2, dvals2, dscan, ddist, dcs, dcnt, dcm, drun, dgs, drc, dcur, dgend, dgid, dX, doff, dtemp};
        for (void* q : tmp)
            if (q) (void)hipFree(q);
    }
    return 0;
fail: {
    void* tmp[] = {dcalls, dkeys, dkeys2, dvals, dvals2, dflag, dscan, ddist, dcs, dcnt, dcm, drun, dgs, drc, dcur,
                   dgend, dgid, dX, doff, dtemp, ds.dO, ds.dP, ds.dgmem, ds.drstart, ds.drentry};
    for (void* q : tmp)
        if (q) (void)hipFree(q);
}
    return 1;
}
