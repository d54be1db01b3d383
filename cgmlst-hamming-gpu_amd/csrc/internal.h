// internal.h -- what the device encoder (gpu_encode.hip) needs from the
// dataset code in hamming.hip.  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "cgmlst_hamming.h"

namespace cghi {

// Device buffers and sizes of a dataset; cgh_dataset takes ownership.
struct DatasetParts {
    int device = 0, n = 0, L = 0, tau = 0;
    size_t D = 0;                   // one-hot columns
    size_t n_pad = 0;
    size_t kw1 = 0, kw2 = 0;        // device words per row of O and P (multiples of the product slab)
    uint32_t* dO = nullptr;         // planes of n_pad rows
    uint32_t* dP = nullptr;
    size_t n_groups = 0, n_members = 0, n_rentries = 0;
    uint64_t n_rare = 0;
    void* dgmem = nullptr;          // members, u16 when m16
    int m16 = 0;
    uint32_t* drstart = nullptr;    // n + 1
    uint64_t* drentry = nullptr;
};

cgh_dataset* dataset_adopt(const DatasetParts& p);

// True when e is an error; then "HIP error '<text>' in <call>" goes to err
// (may be NULL, truncated to errlen bytes).
bool hip_failed(hipError_t e, const char* call, char* err, int errlen);

// Inside an entry point that has `err`, `errlen` and a `fail:` cleanup
// label: report the first failing HIP call and leave through the label.
#define HIP_TRY(x)                                                     \
    do {                                                               \
        if (cghi::hip_failed((x), #x, err, errlen)) goto fail;         \
    } while (0)

// Bytes of a device array of `count` group members (u16 when m16, else u32):
// whole 16-byte words, at least one, since the rare pass reads members 16
// bytes at a time and its last word of a list may extend past the list.
inline size_t member_array_bytes(size_t count, bool m16) {
    return ((count * (m16 ? 2 : 4) + 15) / 16 + 1) * 16;
}

// Device words per row for host words kw (rounded up to the product slab).
int device_words(size_t kw);

// O/P planes (zeroed by the caller) from dense ranks dX (n x L u16, device)
// and the column table (colstart[L + 1], colmap), rows per plane ld.
hipError_t launch_build_planes(const uint16_t* dX, int n, int L, const uint32_t* dcolstart, const uint32_t* dcolmap,
                               uint32_t* dO, uint32_t* dP, int ld, hipStream_t st);

}  // namespace cghi
