// hamming.hip -- gfx950 kernels and the device side of the C ABI.
//
//   H = P P^T - O O^T - R        (exact, integer arithmetic throughout)
//
// Binary products.  O (frequent-allele one-hot bits) and P (presence bits)
// live in HBM as bit planes: plane g holds 128 bits (words 4g..4g+3) of every
// row, 16 bytes per row, rows contiguous, so the 16-byte pieces of 64
// consecutive rows are one contiguous 1 KiB block.  A workgroup computes a
// 256 x 256 (or 128 x 128) output tile; each step stages a KC-word slab of its
// A rows and B rows in an LDS ring straight from global memory
// (global_load_lds_dwordx4: one 1 KiB block per wavefront load, no register
// round trip), NST-1 slabs ahead of the one being multiplied.  Every
// wavefront expands the words of its rows into MFMA operands in registers:
// fp4 (e2m1 0 / +1 / -1, v_mfma_scale_f32_32x32x64_f8f6f4, 64 bits of k per
// instruction, fp32 accumulation of integers <= L: exact) or int8 (0/1,
// v_mfma_i32_32x32x32_i8).  The same bit -> operand position map is used for
// A and B, so every bit is counted exactly once.  One accumulator carries
// O O^T - P P^T (the P slabs multiply a negated B operand) and the epilogue
// stores its negation as u16, zeroing cells with j <= i: no int32 stripe and
// no pack pass.
//
// Rare equal pairs (alleles with count < tau) are subtracted afterwards, row
// by row: a row's output cells are staged in LDS, every partner j listed for
// the row is subtracted there with an LDS atomic, and the row is written back
// once (HBM sees each cell read and written once).
//
// Reference structure this replaces: tensorcore/rs/cuda/tch.cu:22-66 (CUTLASS
// 1-bit GEMM, two launches with beta = -1), 36-43 (sub_rare), 45-50
// (pack_u16), 82-170 (tch_run stripe pipeline).

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>
#include <utility>
#include <unordered_map>
#include <vector>

#include "cgmlst_hamming.h"
#include "internal.h"

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef int v8i __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));

constexpr int KC = 16;  // 32-bit words per row staged per pipeline step (512 bits)

// One launch covers `ns` consecutive stripes of S rows starting at row r0.
// Stripe s (rows r0 + s*S ..) is stored at out + stripe_off(n - r0, S, s) as
// rows x w u16 with w = n - r0 - s*S, column c <-> j = r0 + s*S + c.  Its
// tiles are BM x BM, row tiles I = 0..ceil(rows/BM)-1 by column tiles
// J = 0..ceil(w/BM)-1, numbered stripe by stripe, row-major.  With ns > 1, S
// is a multiple of BM, so a full stripe has q = S/BM row tiles and nbc - s*q
// column tiles, and the tile count before stripe s has a closed form.
struct TileJob {
    const uint32_t* O;  // kw1 / 4 planes of ldA rows (frequent-allele one-hot bits)
    const uint32_t* P;  // kw2 / 4 planes of ldA rows (presence bits)
    int kw1, kw2;       // words per row, multiples of KC
    int n;              // valid rows
    int r0;             // first row of the layout
    int S;              // stripe height
    uint16_t* out;      // stripe layout base
    // Rectangular form (rect != 0): rows of (O, P) against the nB rows of
    // (OB, PB) in the same column space; out is n x nB row-major, every cell
    // kept.  The symmetric form has OB = O, PB = P, nB = n.
    const uint32_t* OB;
    const uint32_t* PB;
    int nB, rect;
    int ldA, ldB;       // rows per plane of (O, P) and of (OB, PB)
    // Cap (main.rs:251-253: a value above cap is written as cap; 0 = none).
    // Rows with rare partners are capped by the rare pass, after the
    // subtraction; the epilogue caps the others (rstart[i + 1] == rstart[i]).
    const uint32_t* rstart;
    int cap;
    // Words of O and P that can hold a set bit (ceil(D / 32), ceil(L / 32));
    // the padding after them is zero, and the VALU form skips its planes.
    // 0 = unknown (every word).
    int kv1, kv2;
    // MFMA forms, symmetric: add the tile's value to the stored cells j > i
    // instead of storing it (a later pass of a one-hot split along K; cells
    // below the diagonal are left alone).  With kw2 = 0 (no P planes) the
    // value is -O O^T.
    int accum;
};

// Cell value of row i (global) in the epilogue, capped unless the rare pass
// still has to subtract from that row.
__device__ __forceinline__ uint16_t epilogue_cap(const TileJob& a, int i, uint32_t v) {
    if (a.cap > 0 && (!a.rstart || a.rstart[i + 1] == a.rstart[i])) v = min(v, (uint32_t)a.cap);
    return (uint16_t)v;
}

struct TileArgs {
    TileJob job;
    int ns;             // stripes in the launch
    int q;              // row tiles of a full stripe (ns > 1)
    int qc;             // column tiles the stripe height spans (S / tile width)
    int nbc;            // column tiles of the first stripe
    long long T;        // tiles in the launch
    long long chunk;    // tiles per XCD: consecutive tiles share a row panel in one L2
};

// Offset (u16 elements) of stripe s in the concatenated upper-triangle layout.
__host__ __device__ inline size_t stripe_off(size_t n, size_t S, size_t s) {
    return S * n * s - S * S * (s * (s - 1) / 2);
}

// Tiles of the stripes before stripe s: q row tiles per stripe, nbc column
// tiles in stripe 0, qc fewer in each later stripe (S / tile width).
__host__ __device__ inline long long tiles_before(long long s, long long q, long long qc, long long nbc) {
    return q * (s * nbc - qc * (s * (s - 1) / 2));
}
__host__ __device__ inline long long tiles_before(long long s, long long q, long long nbc) {
    return tiles_before(s, q, q, nbc);
}

// 32 bits of one row -> 16 int8 operand bytes of lane half g (0 or 1):
// byte b of dword t is bit 8b + 4g + t.
__device__ __forceinline__ v4i expand_bits(uint32_t w, int g) {
    const uint32_t s = w >> (4 * g);
    v4i r;
    r[0] = (int)(s & 0x01010101u);
    r[1] = (int)((s >> 1) & 0x01010101u);
    r[2] = (int)((s >> 2) & 0x01010101u);
    r[3] = (int)((s >> 3) & 0x01010101u);
    return r;
}

// The same bytes negated: 0x00 / 0xFF (int8 -1).  v_perm_b32 returns 0x00 for
// selector byte 12 and 0xFF for 13, so the selector is 12 + bit per byte.
__device__ __forceinline__ v4i expand_bits_neg(uint32_t w, int g) {
    const uint32_t s = w >> (4 * g);
    v4i r;
#pragma unroll
    for (int t = 0; t < 4; t++)
        r[t] = (int)__builtin_amdgcn_perm(0u, 0u, ((s >> t) & 0x01010101u) | 0x0c0c0c0cu);
    return r;
}

// The MFMA helpers take the accumulator type of either kernel flavour (the
// other flavour's body is compiled but never reached).
template <typename Acc>
__device__ __forceinline__ Acc mfma_i8(v4i a, v4i b, Acc c) {
    if constexpr (std::is_same<Acc, v16i>::value) return __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c, 0, 0, 0);
    else return c;
}

// fp4 (e2m1) operands.  A 32-bit word of one row becomes the 32 nibbles of
// a lane's 64-deep operand half: nibble q of dword t holds bit 4q + t.  The
// nibble values differ by dword so that the A side needs no shift for
// t = 0, 1, 2 (the bit stays in place: 0x1 = 0.5, 0x2 = 1.0, 0x4 = 2.0) and
// the B side carries the reciprocals (2.0, 1.0, 0.5); t = 3 is 1.0 on both
// sides (bit 3 would be the sign).  Every product is 0 or exactly 1 and every
// partial sum an integer of magnitude <= L <= 65,535, so the fp32
// accumulation is exact.
__device__ __forceinline__ v8i expand_fp4_a(uint32_t w) {
    v8i r;
    r[0] = (int)(w & 0x11111111u);
    r[1] = (int)(w & 0x22222222u);
    r[2] = (int)(w & 0x44444444u);
    r[3] = (int)((w >> 2) & 0x22222222u);
    r[4] = r[5] = r[6] = r[7] = 0;
    return r;
}

__device__ __forceinline__ v8i expand_fp4_b(uint32_t w) {
    v8i r;
    r[0] = (int)((w << 2) & 0x44444444u);
    r[1] = (int)(w & 0x22222222u);
    r[2] = (int)((w >> 2) & 0x11111111u);
    r[3] = (int)((w >> 2) & 0x22222222u);
    r[4] = r[5] = r[6] = r[7] = 0;
    return r;
}

// v_mfma_scale_f32_32x32x64_f8f6f4 with both operands fp4 (format 4) and
// unit scales (E8M0 127 = 2^0).
template <typename Acc>
__device__ __forceinline__ Acc mfma_fp4(v8i a, v8i b, Acc c) {
    if constexpr (std::is_same<Acc, v16f>::value)
        return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b, c, 4, 4, 0, 127, 0, 127);
    else return c;
}

// Square BM x BM output tile per workgroup, WM x WN per wavefront; FP4
// selects the fp4 MFMA (64 bits of k per instruction) over the int8 one (32).
constexpr int KG = KC / 4;   // 4-word groups per slab
constexpr int NST_DEFAULT = 3;  // LDS ring stages (slabs in flight); the autotuner also tries 2 and 4

// ring.a[stage][group][row] (BM tile rows) and ring.b[stage][group][row] (BN
// tile columns) of 4 consecutive words: one wavefront-wide direct-to-LDS load
// fills 64 consecutive rows of a group (1 KiB, contiguous), and a wavefront's
// 16-byte reads of 32 consecutive rows are contiguous, hence conflict-free.
template <int BM, int BN, int NST>
struct Ring {
    uint4 a[NST][KG][BM];
    uint4 b[NST][KG][BN];
};

// One BM x BN tile (I, J) of stripe s of job a, WM x WN per wavefront.
template <int BM, int BN, int WM, int WN, bool FP4, int NST>
__device__ __forceinline__ void tile_body(const TileJob& a, int s, int I, int J, Ring<BM, BN, NST>& ring) {
    constexpr int WAVES_N = BN / WN;
    constexpr int NT = (BM / WM) * WAVES_N * 64;
    constexpr int TM = WM / 32, TN = WN / 32;
    constexpr int WLOADS = KG * ((BM + BN) / 64);  // wavefront-wide 16-byte loads per slab
    constexpr int LPW = WLOADS / (NT / 64);     // of which each wavefront issues
    static_assert(WLOADS % (NT / 64) == 0, "slab loads must divide evenly");

    const int tid = threadIdx.x;
    const int sr0 = a.r0 + s * a.S;              // first row of the stripe
    const int rows = min(a.S, a.n - sr0), w = a.rect ? a.nB : a.n - sr0;
    uint16_t* __restrict__ out = a.rect ? a.out + (size_t)(sr0 - a.r0) * a.nB
                                        : a.out + stripe_off((size_t)(a.n - a.r0), (size_t)a.S, (size_t)s);
    const int rt0 = I * BM, ct0 = J * BN;  // tile origin, stripe-relative

    if (!a.rect && ct0 + BN <= rt0) {  // entirely below the diagonal: every cell has j < i
        if (a.accum) return;           // written as 0 by the first pass
        const int rend = min(BM, rows - rt0), cend = min(BN, w - ct0);
        for (int e = tid; e < rend * BN; e += NT) {
            const int r = e / BN, c = e % BN;
            if (c < cend) out[(size_t)(rt0 + r) * w + ct0 + c] = 0;
        }
        return;
    }

    const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm0 = (wave / WAVES_N) * WM, wn0 = (wave % WAVES_N) * WN;
    const int l32 = lane & 31, g = lane >> 5;

    // int8: the O slabs (kw1 / KC), then the P slabs.  fp4: one K sequence of
    // 128-bit planes, the npO planes of O that can hold a set bit followed by
    // the npP such planes of P, KG planes per slab (so no slab is padding
    // between the two); a plane past the end reads some P plane and is not
    // multiplied.
    const int nc1 = a.kw1 / KC;
    const int npO = ((a.kv1 ? a.kv1 : a.kw1) + 3) / 4, npP = ((a.kv2 ? a.kv2 : a.kw2) + 3) / 4;
    const int npl = npO + npP;
    const int nc = FP4 ? (npl + KG - 1) / KG : nc1 + a.kw2 / KC;
    const int growA = sr0 + rt0, growB = a.rect ? ct0 : sr0 + ct0;

    // Slab c (KC words of every A and B row of the tile) -> ring stage st,
    // straight from global memory into LDS (global_load_lds_dwordx4: lane l
    // of the wavefront lands at the wavefront's LDS base + 16 l).  Rows past
    // n read row n-1: their products only reach output cells that are never
    // stored.
    // What does not change from slab to slab is set up once per tile: for
    // each of the wavefront's loads u, its 4-word group, the O and P bases of
    // its side, the bytes per plane, the lane's row offset in a plane and the
    // LDS address of its 64-row block in stage 0.  A slab's load then costs a
    // plane index, a base select and a 64-bit add in scalar registers (the
    // bases selected per slab from the kernel arguments were reloaded from
    // them every time, with a scalar-memory wait each).
    const char* base_o[LPW];
    const char* base_p[LPW];
    uint32_t plane_bytes[LPW], row_off[LPW], lds_blk[LPW], lds_stage[LPW];
    int grp_of[LPW];
    {
        using Lds = __attribute__((address_space(3))) char;
        const uint32_t lds_ring = __builtin_amdgcn_readfirstlane((uint32_t)(uintptr_t)(Lds*)&ring);
#pragma unroll
        for (int u = 0; u < LPW; u++) {
            const int v = wave * LPW + u;
            const int side = v >= KG * (BM / 64);
            const int rem = side ? v - KG * (BM / 64) : v;
            const int nrb = side ? BN / 64 : BM / 64;
            const int grp = rem / nrb, rb = rem % nrb;
            grp_of[u] = grp;
            base_o[u] = reinterpret_cast<const char*>(side ? a.OB : a.O);
            base_p[u] = reinterpret_cast<const char*>(side ? a.PB : a.P);
            plane_bytes[u] = (uint32_t)(side ? a.ldB : a.ldA) * 16u;
            const int row = side ? min(growB + rb * 64 + lane, a.nB - 1) : min(growA + rb * 64 + lane, a.n - 1);
            row_off[u] = (uint32_t)row * 16u;
            // ring.a[st][grp][r] / ring.b[st][grp][r] as byte offsets from the ring
            lds_blk[u] = lds_ring + (side ? (uint32_t)(NST * KG * BM * 16 + (grp * BN + rb * 64) * 16)
                                          : (uint32_t)((grp * BM + rb * 64) * 16));
            lds_stage[u] = side ? KG * BN * 16 : KG * BM * 16;
        }
    }
    // The global address of load u for slab c.
    auto slab_addr = [&](int c, int u) -> const char* {
        bool isO;
        int plane;
        if (FP4) {
            const int p = c * KG + grp_of[u];
            isO = p < npO || npP == 0;  // with no P planes a plane past the end reads O plane 0
            plane = p < npO ? p : npP ? min(p - npO, npP - 1) : 0;
        } else {
            isO = c < nc1;
            plane = (isO ? c : c - nc1) * KG + grp_of[u];
        }
        return (isO ? base_o[u] : base_p[u]) + (size_t)plane * plane_bytes[u];
    };
    // Running addresses of the last slab issued (cur): inside the O planes
    // and inside the P planes, slab c + 1 of a load is slab c plus KG planes,
    // so only the slabs where some load crosses from O to P or reaches the
    // end (and the first slab) recompute them.
    const char* addr[LPW];
    int cur = -2;
    auto issue_slab = [&](int c, int st) {
        const bool step1 = c == cur + 1 &&
                           (FP4 ? !((c * KG <= npO + KG && (c + 1) * KG >= npO) || (c + 1) * KG > npl) : c != nc1);
        if (step1) {
#pragma unroll
            for (int u = 0; u < LPW; u++) addr[u] += (size_t)KG * plane_bytes[u];
        } else if (c != cur) {
#pragma unroll
            for (int u = 0; u < LPW; u++) addr[u] = slab_addr(c, u);
        }
        cur = c;
#pragma unroll
        for (int u = 0; u < LPW; u++) {
            const char* base = addr[u];
            const uint32_t lds = lds_blk[u] + (uint32_t)st * lds_stage[u];
            // The LDS base of this wavefront's 64 rows goes to M0; the global
            // address is the scalar base plus the lane's 32-bit row offset.
            // Written as an instruction rather than the builtin so that the
            // compiler does not wait for every LDS-DMA transfer before each
            // LDS read (it cannot tell the ring stages apart); the waits are
            // explicit below.  The load has no register result.
            asm volatile("s_mov_b32 m0, %2\n\tglobal_load_lds_dwordx4 %0, %1"
                         :
                         : "v"(row_off[u]), "s"(base), "s"(lds)
                         : "memory", "m0");
        }
    };
    // Every step issues exactly one slab (past the last slab it reloads slab
    // nc-1 into the stage nobody reads again), so "slab c has landed" is
    // always "at most NST-2 slabs of this wavefront's loads pending": one
    // constant wait, no branch in the loop body.
    constexpr int AHEAD = (NST - 2) * LPW;
    static_assert(AHEAD < 64, "vmcnt field is 6 bits");
    constexpr int WAIT_AHEAD = (AHEAD & 15) | ((AHEAD >> 4) << 14) | (7 << 4) | (15 << 8);

    using Acc = typename std::conditional<FP4, v16f, v16i>::type;
    Acc acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++)
#pragma unroll
            for (int e = 0; e < 16; e++) acc[i][j][e] = 0;

#pragma unroll
    for (int p = 0; p < NST - 1; p++) issue_slab(min(p, nc - 1), p);

    // int8: acc accumulates O O^T - P P^T (O slabs with B = +O, P slabs with
    // B = -P, the short ones), negated in the epilogue.  fp4: acc holds
    // O O^T over the O planes, is negated once at the first P plane and then
    // adds P P^T.
    // Words of group gi+1 are read from LDS while group gi is multiplied.
    // fp4: each lane half g reads words 2g, 2g+1 of a 4-word group (one
    // 8-byte read), one word per 64-deep MFMA step.
    auto compute_slab_fp4 = [&](int buf, int c) {
        __builtin_amdgcn_iglp_opt(1);  // interleave the LDS reads and expansions with the MFMAs
        uint2 wa[TM], wb[TN];
#pragma unroll
        for (int i = 0; i < TM; i++) wa[i] = reinterpret_cast<const uint2*>(&ring.a[buf][0][wm0 + i * 32 + l32])[g];
#pragma unroll
        for (int j = 0; j < TN; j++) wb[j] = reinterpret_cast<const uint2*>(&ring.b[buf][0][wn0 + j * 32 + l32])[g];
#pragma unroll
        for (int gi = 0; gi < KG; gi++) {
            uint2 na[TM], nb[TN];
            if (gi + 1 < KG) {
#pragma unroll
                for (int i = 0; i < TM; i++)
                    na[i] = reinterpret_cast<const uint2*>(&ring.a[buf][gi + 1][wm0 + i * 32 + l32])[g];
#pragma unroll
                for (int j = 0; j < TN; j++)
                    nb[j] = reinterpret_cast<const uint2*>(&ring.b[buf][gi + 1][wn0 + j * 32 + l32])[g];
            }
            // plane c * KG + gi (uniform): the first P plane negates acc (a
            // pass with no P planes negates after the loop instead), a plane
            // past the end is skipped
            const int p = c * KG + gi;
            if (p == npO && npP > 0) {
#pragma unroll
                for (int i = 0; i < TM; i++)
#pragma unroll
                    for (int j = 0; j < TN; j++)
#pragma unroll
                        for (int e = 0; e < 16; e++) acc[i][j][e] = -acc[i][j][e];
            }
            if (p < npl) {
#pragma unroll
                for (int kk = 0; kk < 2; kk++) {
                    v8i av[TM], bv[TN];
#pragma unroll
                    for (int i = 0; i < TM; i++) av[i] = expand_fp4_a(kk ? wa[i].y : wa[i].x);
#pragma unroll
                    for (int j = 0; j < TN; j++) bv[j] = expand_fp4_b(kk ? wb[j].y : wb[j].x);
#pragma unroll
                    for (int i = 0; i < TM; i++)
#pragma unroll
                        for (int j = 0; j < TN; j++)
                            acc[i][j] = mfma_fp4(av[i], bv[j], acc[i][j]);
                }
            }
            if (gi + 1 < KG) {
#pragma unroll
                for (int i = 0; i < TM; i++) wa[i] = na[i];
#pragma unroll
                for (int j = 0; j < TN; j++) wb[j] = nb[j];
            }
        }
    };
    auto compute_slab_i8 = [&](int buf, auto neg) {
        uint4 wa[TM], wb[TN];
#pragma unroll
        for (int i = 0; i < TM; i++) wa[i] = ring.a[buf][0][wm0 + i * 32 + l32];
#pragma unroll
        for (int j = 0; j < TN; j++) wb[j] = ring.b[buf][0][wn0 + j * 32 + l32];
#pragma unroll
        for (int gi = 0; gi < KG; gi++) {
            uint4 na[TM], nb[TN];
            if (gi + 1 < KG) {
#pragma unroll
                for (int i = 0; i < TM; i++) na[i] = ring.a[buf][gi + 1][wm0 + i * 32 + l32];
#pragma unroll
                for (int j = 0; j < TN; j++) nb[j] = ring.b[buf][gi + 1][wn0 + j * 32 + l32];
            }
#pragma unroll
            for (int kk = 0; kk < 4; kk++) {
                v4i av[TM], bv[TN];
#pragma unroll
                for (int i = 0; i < TM; i++) av[i] = expand_bits(((const uint32_t*)&wa[i])[kk], g);
#pragma unroll
                for (int j = 0; j < TN; j++) {
                    const uint32_t x = ((const uint32_t*)&wb[j])[kk];
                    bv[j] = decltype(neg)::value ? expand_bits_neg(x, g) : expand_bits(x, g);
                }
#pragma unroll
                for (int i = 0; i < TM; i++)
#pragma unroll
                    for (int j = 0; j < TN; j++)
                        acc[i][j] = mfma_i8(av[i], bv[j], acc[i][j]);
            }
            if (gi + 1 < KG) {
#pragma unroll
                for (int i = 0; i < TM; i++) wa[i] = na[i];
#pragma unroll
                for (int j = 0; j < TN; j++) wb[j] = nb[j];
            }
        }
    };
    auto compute_slab = [&](int buf, int c, auto neg) {
        if constexpr (FP4) compute_slab_fp4(buf, c);
        else compute_slab_i8(buf, neg);
    };

    // Slab c: wait for its loads, barrier (it is visible to every wavefront
    // and stage (c-1) % NST is free), refill that stage with slab c+NST-1,
    // multiply slab c.
    auto step = [&](int c, auto neg) {
        __builtin_amdgcn_s_waitcnt(WAIT_AHEAD);
        // every LDS read of the previous slab has returned, then the
        // workgroup barrier (no memory fence: the DMA waits are explicit)
        asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" : : : "memory");
        issue_slab(min(c + NST - 1, nc - 1), (c + NST - 1) % NST);
        compute_slab(c % NST, c, neg);
    };
    if constexpr (FP4) {
        for (int c = 0; c < nc; c++) step(c, std::false_type{});
    } else {
        int c = 0;
        for (; c < nc1; c++) step(c, std::false_type{});
        for (; c < nc; c++) step(c, std::true_type{});
    }
    // the trailing reloads must land before the workgroup's LDS is released
    __builtin_amdgcn_s_waitcnt((7 << 4) | (15 << 8));
    if (FP4 && npP == 0) {  // O planes only: the value is -O O^T, as in the int8 form
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
            for (int j = 0; j < TN; j++)
#pragma unroll
                for (int e = 0; e < 16; e++) acc[i][j][e] = -acc[i][j][e];
    }

    // Epilogue: lane holds column l32 and rows 8*(e/4) + 4*g + (e%4) of the
    // distance (fp4) or of its negation (int8).  The capped form is a
    // separate loop so that the common one stays a plain store sequence.
    auto store_all = [&](auto capped) {
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
            for (int j = 0; j < TN; j++) {
                const int c = ct0 + wn0 + j * 32 + l32;
                if (c >= w) continue;
#pragma unroll
                for (int e = 0; e < 16; e++) {
                    const int r = rt0 + wm0 + i * 32 + (e >> 2) * 8 + g * 4 + (e & 3);
                    if (r >= rows) continue;
                    const uint32_t v = FP4 ? (uint32_t)(int)acc[i][j][e] : (uint32_t)(-(int)acc[i][j][e]);
                    uint16_t o;
                    if (a.rect) o = (uint16_t)v;
                    else if constexpr (decltype(capped)::value) o = c > r ? epilogue_cap(a, sr0 + r, v) : 0;
                    else o = c > r ? (uint16_t)v : 0;
                    out[(size_t)r * w + c] = o;
                }
            }
    };
    // A later pass of a one-hot split along K: the value is added to the
    // stored cell (it only ever lowers it towards the final distance); the
    // cap goes with the last pass.
    auto add_all = [&](auto capped) {
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
            for (int j = 0; j < TN; j++) {
                const int c = ct0 + wn0 + j * 32 + l32;
                if (c >= w) continue;
#pragma unroll
                for (int e = 0; e < 16; e++) {
                    const int r = rt0 + wm0 + i * 32 + (e >> 2) * 8 + g * 4 + (e & 3);
                    if (r >= rows || c <= r) continue;
                    const uint32_t v = FP4 ? (uint32_t)(int)acc[i][j][e] : (uint32_t)(-(int)acc[i][j][e]);
                    uint16_t* cell = out + (size_t)r * w + c;
                    const uint32_t nv = (uint32_t)*cell + v;
                    *cell = decltype(capped)::value ? epilogue_cap(a, sr0 + r, nv) : (uint16_t)nv;
                }
            }
    };
    if (a.accum) {
        if (a.cap > 0) add_all(std::true_type{});
        else add_all(std::false_type{});
        return;
    }
    // Interior tiles (every row < rows, every column < w and, in the
    // triangle, every column right of every row) need no per-cell test: one
    // convert and one store per cell, addressed as a wavefront-uniform base
    // plus a 32-bit lane offset; the row offsets are wavefront-uniform too.
    const bool interior = rt0 + BM <= rows && ct0 + BN <= w && (a.rect || ct0 >= rt0 + BM);
    if (interior && a.cap <= 0) {
        char* __restrict__ wb0 = reinterpret_cast<char*>(out + (size_t)(rt0 + wm0) * w + (ct0 + wn0));
        const uint32_t lo = 2u * ((uint32_t)l32 + (uint32_t)(g * 4) * (uint32_t)w);  // byte offsets below 2^32
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
            for (int e = 0; e < 16; e++) {
                const uint32_t ro = lo + 2u * (uint32_t)(i * 32 + (e >> 2) * 8 + (e & 3)) * (uint32_t)w;
#pragma unroll
                for (int j = 0; j < TN; j++)
                    *reinterpret_cast<uint16_t*>(wb0 + (ro + j * 64)) =
                        (uint16_t)(FP4 ? (uint32_t)(int)acc[i][j][e] : (uint32_t)(-(int)acc[i][j][e]));
            }
    } else if (a.cap > 0) {
        store_all(std::true_type{});
    } else {
        store_all(std::false_type{});
    }
}

// Tile shapes of the closed-form launch: 256 x 128 by 4 wavefronts of
// 128 x 64 with a 2-slab ring (48 KiB of LDS, at most 256 VGPRs), so that
// two workgroups share a CU and one's tile start, barriers and epilogue are
// covered by the other's MFMAs; and 128 x 128 by 4 wavefronts of 64 x 64.
// The batch queue keeps 256 x 256 by 8 wavefronts of 128 x 64 (two per
// SIMD).
//
// One launch over the closed-form tile numbering of TileArgs.  Workgroups
// are dealt round-robin to the 8 XCDs; each XCD gets a contiguous run of
// tiles so concurrently resident tiles share panels in that XCD's L2.
template <int BM, int BN, int WM, int WN, bool FP4, int NST>
__global__ void __launch_bounds__((BM / WM) * (BN / WN) * 64, BN < BM ? 2 : 1)
mfma_tile(TileArgs a) {
    __shared__ Ring<BM, BN, NST> ring;
    const long long t = (long long)(blockIdx.x & 7) * a.chunk + (blockIdx.x >> 3);
    if (t >= a.T) return;
    int s = 0;
    if (a.ns > 1) {
        int lo = 0, hi = a.ns - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (tiles_before(mid, a.q, a.qc, a.nbc) <= t) lo = mid; else hi = mid - 1;
        }
        s = lo;
    }
    const long long local = t - (s ? tiles_before(s, a.q, a.qc, a.nbc) : 0);
    const int tn = a.nbc - s * a.qc;
    tile_body<BM, BN, WM, WN, FP4, NST>(a.job, s, (int)(local / tn), (int)(local % tn), ring);
}

// A job read from memory by a uniform index is uniform: say so, so that its
// fields live in scalar registers.
template <typename T>
__device__ __forceinline__ T* uniform_ptr(T* p) {
    const uint64_t v = (uint64_t)(uintptr_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v), hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return (T*)(uintptr_t)(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ void scalarize(TileJob& j) {
    j.O = uniform_ptr(j.O);
    j.P = uniform_ptr(j.P);
    j.OB = uniform_ptr(j.OB);
    j.PB = uniform_ptr(j.PB);
    j.out = uniform_ptr(j.out);
    j.kw1 = __builtin_amdgcn_readfirstlane(j.kw1);
    j.kw2 = __builtin_amdgcn_readfirstlane(j.kw2);
    j.n = __builtin_amdgcn_readfirstlane(j.n);
    j.r0 = __builtin_amdgcn_readfirstlane(j.r0);
    j.S = __builtin_amdgcn_readfirstlane(j.S);
    j.nB = __builtin_amdgcn_readfirstlane(j.nB);
    j.rect = __builtin_amdgcn_readfirstlane(j.rect);
    j.ldA = __builtin_amdgcn_readfirstlane(j.ldA);
    j.ldB = __builtin_amdgcn_readfirstlane(j.ldB);
    j.rstart = uniform_ptr(j.rstart);
    j.cap = __builtin_amdgcn_readfirstlane(j.cap);
    j.kv1 = __builtin_amdgcn_readfirstlane(j.kv1);
    j.kv2 = __builtin_amdgcn_readfirstlane(j.kv2);
}

// Persistent form for many datasets: each workgroup pulls tiles (job,
// stripe, I, J) from a device work queue until the list is exhausted; the
// list is ordered longest-first, so uneven datasets finish together.
template <int BM, int WM, int WN, bool FP4>
__global__ void __launch_bounds__((BM / WM) * (BM / WN) * 64)
mfma_tile_queue(const TileJob* __restrict__ jobs, uint16_t* base, const size_t* __restrict__ outoff,
                const int4* __restrict__ tiles, int T, int* __restrict__ next) {
    __shared__ Ring<BM, BM, NST_DEFAULT> ring;
    __shared__ int slot;
    for (;;) {
        if (threadIdx.x == 0) slot = atomicAdd(next, 1);
        __syncthreads();
        const int t = __builtin_amdgcn_readfirstlane(slot);  // uniform: tile and job load as scalars
        __syncthreads();  // every wavefront has read slot, and left the previous tile
        if (t >= T) break;
        const int4 tl = tiles[t];
        TileJob jb = jobs[tl.x];
        jb.out = base + outoff[tl.x];
        scalarize(jb);
        tile_body<BM, BM, WM, WN, FP4, NST_DEFAULT>(jb, tl.y, tl.z, tl.w, ring);
    }
}

// ---------------------------------------------------------------------------
// VALU form of the binary product: AND + v_bcnt_u32_b32 on 32-bit words
// ---------------------------------------------------------------------------
//
// 128 x 128 sub-tiles by 256 threads, 8 x 8 cells per thread (rows ty + 16 i,
// columns tx + 16 j).  Each KC-word slab of the sub-tile's A and B rows is
// loaded with 16-byte reads (4 per thread, planes layout: consecutive threads
// read consecutive rows) into a double-buffered LDS stage as [row][4 x 16 B];
// a thread reads 4 words of each of its 16 rows with one ds_read_b128 and
// accumulates popcount(a & b) per cell.  The O slabs come first; the
// accumulator is negated before the first P slab, so it ends as
// P P^T - O O^T.  Tiles of 256 are done as 2 x 2 sub-tiles.
constexpr int VB = 128;
struct ValuStage {
    uint4 a[VB][KC / 4];
    uint4 b[VB][KC / 4];
};

__device__ __forceinline__ void valu_sub(const TileJob a, int sr0, int rt0, int ct0, int rows, int w,
                                         uint16_t* __restrict__ out, ValuStage (&st)[2]) {
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    const int nc1 = a.kw1 / KC, nc = nc1 + a.kw2 / KC;
    const int growA = sr0 + rt0, growB = a.rect ? ct0 : sr0 + ct0;
    uint32_t acc[8][8];
#pragma unroll
    for (int i = 0; i < 8; i++)
#pragma unroll
        for (int j = 0; j < 8; j++) acc[i][j] = 0;
    // piece q = tid + 256 u: side q / 512, plane q / 128 % 4, row q % 128.
    // Rows past the valid ones read the last valid row (never stored).
    size_t off[4], pitch[4];
    int dst_side[4], dst_pl[4], dst_r[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int q = tid + 256 * u;
        dst_side[u] = q >> 9;
        dst_pl[u] = (q >> 7) & 3;
        dst_r[u] = q & 127;
        const int grow = dst_side[u] ? min(growB + dst_r[u], a.nB - 1) : min(growA + dst_r[u], a.n - 1);
        pitch[u] = (size_t)(dst_side[u] ? a.ldB : a.ldA);
        off[u] = ((size_t)dst_pl[u] * pitch[u] + grow) * 4;  // words, from the slab's first plane
    }
    // slab c of the O or P planes: base pointer and first plane
#define VALU_PIECE(c, u)                                                                                       \
    (*reinterpret_cast<const uint4*>(                                                                         \
        ((c) < nc1 ? (dst_side[u] ? a.OB : a.O) : (dst_side[u] ? a.PB : a.P)) +                                 \
        ((size_t)(((c) < nc1 ? (c) : (c) - nc1) * (KC / 4)) * pitch[u]) * 4 + off[u]))
    uint4 v0 = VALU_PIECE(0, 0), v1 = VALU_PIECE(0, 1), v2 = VALU_PIECE(0, 2), v3 = VALU_PIECE(0, 3);
#define VALU_STORE(buf, u, v)                                      \
    do {                                                           \
        if (dst_side[u]) st[buf].b[dst_r[u]][dst_pl[u]] = (v);     \
        else st[buf].a[dst_r[u]][dst_pl[u]] = (v);                 \
    } while (0)
    VALU_STORE(0, 0, v0);
    VALU_STORE(0, 1, v1);
    VALU_STORE(0, 2, v2);
    VALU_STORE(0, 3, v3);
    __syncthreads();
    for (int c = 0; c < nc; c++) {
        const int buf = c & 1;
        if (c + 1 < nc) {  // next slab in flight during this one
            v0 = VALU_PIECE(c + 1, 0);
            v1 = VALU_PIECE(c + 1, 1);
            v2 = VALU_PIECE(c + 1, 2);
            v3 = VALU_PIECE(c + 1, 3);
        }
        if (c == nc1) {
#pragma unroll
            for (int i = 0; i < 8; i++)
#pragma unroll
                for (int j = 0; j < 8; j++) acc[i][j] = 0u - acc[i][j];
        }
        // planes of this slab that hold data (the rest is zero padding)
        const int rem = c < nc1 ? (a.kv1 ? a.kv1 : a.kw1) - c * KC : (a.kv2 ? a.kv2 : a.kw2) - (c - nc1) * KC;
        const int npl = min(KC / 4, max(0, (rem + 3) / 4));
#pragma unroll 1
        for (int pl = 0; pl < npl; pl++) {  // not unrolled: 16 rows x 4 words live at a time
            uint4 ar[8], br[8];
#pragma unroll
            for (int i = 0; i < 8; i++) ar[i] = st[buf].a[ty + 16 * i][pl];
#pragma unroll
            for (int j = 0; j < 8; j++) br[j] = st[buf].b[tx + 16 * j][pl];
#pragma unroll
            for (int i = 0; i < 8; i++)
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    uint32_t t = acc[i][j];
                    t += __builtin_popcount(ar[i].x & br[j].x);
                    t += __builtin_popcount(ar[i].y & br[j].y);
                    t += __builtin_popcount(ar[i].z & br[j].z);
                    t += __builtin_popcount(ar[i].w & br[j].w);
                    acc[i][j] = t;
                }
        }
        if (c + 1 < nc) {  // that stage was last read before the previous barrier
            VALU_STORE(buf ^ 1, 0, v0);
            VALU_STORE(buf ^ 1, 1, v1);
            VALU_STORE(buf ^ 1, 2, v2);
            VALU_STORE(buf ^ 1, 3, v3);
        }
        __syncthreads();
    }
#undef VALU_PIECE
#undef VALU_STORE
    const bool capped = a.cap > 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const int r = rt0 + ty + 16 * i;
        if (r >= rows) continue;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int cc = ct0 + tx + 16 * j;
            if (cc >= w) continue;
            const uint32_t v = acc[i][j];
            out[(size_t)r * w + cc] = a.rect ? (uint16_t)v
                                    : cc > r ? (capped ? epilogue_cap(a, sr0 + r, v) : (uint16_t)v) : (uint16_t)0;
        }
    }
}

// One BM x BM tile (I, J) of stripe s of job a (BM = 128 or 256).
template <int BM>
__device__ __forceinline__ void valu_body(const TileJob a, int s, int I, int J, ValuStage (&st)[2]) {
    const int sr0 = a.r0 + s * a.S;
    const int rows = min(a.S, a.n - sr0), w = a.rect ? a.nB : a.n - sr0;
    uint16_t* __restrict__ out = a.rect ? a.out + (size_t)(sr0 - a.r0) * a.nB
                                        : a.out + stripe_off((size_t)(a.n - a.r0), (size_t)a.S, (size_t)s);
    const int rt0 = I * BM, ct0 = J * BM;
    if (!a.rect && J < I) {  // entirely below the diagonal
        const int rend = min(BM, rows - rt0), cend = min(BM, w - ct0);
        for (int e = threadIdx.x; e < rend * BM; e += 256) {
            const int r = e / BM, c = e % BM;
            if (c < cend) out[(size_t)(rt0 + r) * w + ct0 + c] = 0;
        }
        return;
    }
    for (int di = 0; di < BM / VB; di++)
        for (int dj = 0; dj < BM / VB; dj++) {
            const int r0 = rt0 + di * VB, c0 = ct0 + dj * VB;
            if (r0 >= rows || c0 >= w) continue;
            valu_sub(a, sr0, r0, c0, rows, w, out, st);
            __syncthreads();  // the stages are reused by the next sub-tile
        }
}

template <int BM>
__global__ void __launch_bounds__(256) valu_tile(TileArgs a) {
    __shared__ ValuStage st[2];
    const long long t = (long long)(blockIdx.x & 7) * a.chunk + (blockIdx.x >> 3);
    if (t >= a.T) return;
    int s = 0;
    if (a.ns > 1) {
        int lo = 0, hi = a.ns - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (tiles_before(mid, a.q, a.nbc) <= t) lo = mid; else hi = mid - 1;
        }
        s = lo;
    }
    const long long local = t - (s ? tiles_before(s, a.q, a.nbc) : 0);
    const int tn = a.nbc - s * a.q;
    valu_body<BM>(a.job, s, (int)(local / tn), (int)(local % tn), st);
}

template <int BM>
__global__ void __launch_bounds__(256)
valu_tile_queue(const TileJob* __restrict__ jobs, uint16_t* base, const size_t* __restrict__ outoff,
                const int4* __restrict__ tiles, int T, int* __restrict__ next) {
    __shared__ ValuStage st[2];
    __shared__ int slot;
    for (;;) {
        if (threadIdx.x == 0) slot = atomicAdd(next, 1);
        __syncthreads();
        const int t = __builtin_amdgcn_readfirstlane(slot);
        __syncthreads();
        if (t >= T) break;
        const int4 tl = tiles[t];
        TileJob jb = jobs[tl.x];
        jb.out = base + outoff[tl.x];
        scalarize(jb);
        valu_body<BM>(jb, tl.y, tl.z, tl.w, st);
    }
}

__global__ void reset_counter(int* c) { *c = 0; }

// Counts the 16-byte words in which a and b differ (into *diff, cleared by
// the caller): the autotuner's check that every candidate agrees.
__global__ void count_diff(const uint4* __restrict__ a, const uint4* __restrict__ b, size_t words, int* diff) {
    int bad = 0;
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < words; k += (size_t)gridDim.x * blockDim.x) {
        const uint4 x = a[k], y = b[k];
        bad |= (x.x != y.x) | (x.y != y.y) | (x.z != y.z) | (x.w != y.w);
    }
    if (bad) atomicAdd(diff, 1);
}

// Rare pairs, row-local: a workgroup takes items (row i and a chunk of RCW
// columns, or a group of short rows) one after the other.  An item's cells
// are staged in LDS with 16-byte loads, every rare partner j of its rows is
// subtracted there with an LDS atomic, and the cells are written back: HBM
// sees each output cell read and written once and the scattered updates stay
// on-chip.  Row i's partners are ranges of the ascending group member lists
// (rentry).
constexpr int RCW = 65536;  // u16 cells per chunk (128 KiB of LDS: one workgroup per CU)
constexpr int RNT = 1024;   // threads per workgroup = ranges per batch
constexpr int RQ = (RCW / 8 + 1 + RNT - 1) / RNT;  // 16-byte words staged per thread
constexpr int MPL = 8;      // members per lane per unit of a range walk (one 16-byte word of u16 members)
constexpr int RFG = 1024;   // most rows of an item the flat walk takes (items of more take the subgroup walk)
// Lanes per range of the subgroup walk from the mean range length (partners
// per range): 4 below 20, 8 below 160, else 16.  Short ranges leave fewer
// lanes idle, long ones keep fewer ranges in flight.  Measured rare pass:
// mean 41 (12,000 rows, tau 256), 0.27 ms with 8 lanes against 0.32 ms with
// 16; mean 286 (30,000 rows, tau 2048), 2.57 ms with 16 against 2.66 ms
// with 8; the 64-set batch (means 4 to 170) 49.7 ms against 52.9 ms for the
// rule that gave 16 from a mean of 40 on.
int lanes_per_range(uint64_t pairs, size_t ranges) {
    const double mean = ranges ? (double)pairs / (double)ranges : 1.0;
    return mean >= 160.0 ? 16 : mean >= 20.0 ? 8 : 4;
}
struct RareJob {
    uint16_t* out;             // stripe layout of rows r0..
    int n, r0, S, chunks;      // chunks = column chunks per row
    const uint32_t* rstart;    // n + 1 offsets into rentry
    const uint64_t* rentry;    // partner ranges of gmem
    const void* gmem;          // group members: u16 if m16 (every row < 65536), else u32
    int rect, nB;              // rectangular form: rows against nB columns
    int m16;
    int cap;                   // > 0: staged cells are written back as min(value, cap)
    // Row groups: G > 1 consecutive rows of one stripe per workgroup (rows
    // short enough that G of them fit one chunk), gps groups per stripe;
    // G = 1: one row per workgroup, in `chunks` column chunks.
    int G, gps;
    int rsg;                   // lanes per range (a power of two, 4 .. 32)
    // Items of several rows take the flat walk rather than the subgroup walk
    // (set for long ranges, 16 lanes or more: a mean range length of 160 on).
    // Measured rare pass: 30,000 rows at tau 2048 (mean 286), 2.33 ms against
    // 2.55 ms; at a mean of 41 (12,000 rows, tau 256) 0.38 against 0.27 ms
    // and at 12 (4,000 rows, tau 64) 0.24 against 0.06 ms, where its two
    // edge words per range outweigh the short interiors.
    int mflat;
    // Members ascending within each range (group member lists).  A range
    // in arbitrary order (a caller's per-row partner list) is walked whole in
    // every column chunk, members outside the chunk masked out.
    int sorted;
};

// Rare job over the stripe layout `out` of rows r0.. (symmetric form).
RareJob rare_job(uint16_t* out, int n, int r0, int S, const uint32_t* rstart, const uint64_t* rentry,
                 const void* gmem, int m16, int cap, int rsg) {
    RareJob j{};
    j.out = out;
    j.n = n;
    j.r0 = r0;
    j.S = S;
    j.chunks = 1;
    j.rstart = rstart;
    j.rentry = rentry;
    j.gmem = gmem;
    j.m16 = m16;
    j.cap = cap;
    j.G = 1;
    j.rsg = rsg;
    j.mflat = rsg >= 16;
    j.sorted = 1;
    return j;
}

struct RareShared {
    uint4 cells4[RCW / 8 + 1];
    // Batch ranges: [ebeg, eend) of member indices (arbitrary-order ranges);
    // for sorted ranges, the first interior word of range r in the batch's
    // word list (ebeg) and the member-word index minus it (eend).
    uint32_t ebeg[RNT];
    uint32_t eend[RNT];
    uint32_t spare[64];  // one word per lane for the atomics of idle slots
    uint32_t wsum[RNT / 64];  // interior words per wavefront's ranges (scan)
    uint32_t next;       // next range of the batch to take
    // Flat walk of an item of several rows: LDS u16 offset of cell (i, 0)
    // of each batch range's row i, and the item's row starts in rentry.
    uint32_t wbo[RNT];
    uint32_t rs[RFG + 1];
};

// Work item b of job jb: with row groups (G > 1) rows i0 .. i1-1 of one
// stripe, every cell j > i of each (they fit one chunk); else one row i0 and
// column chunk ch of it.  The cells are staged as one contiguous range
// [g0, g1) of the layout (for several rows it includes their cells j <= i,
// which no partner touches and which are written back unchanged), as nq
// 16-byte words from a0, into LDS words lb .. lb + nq - 1.  False when the
// item has no rare partners.
struct RareItem {
    int i0, i1, j0, j1;
    size_t sbase, sr0, g0, g1, a0;
    int nq;
    int lb;  // first LDS word of the staged cells: 0, or RHALF (the upper half)
};

__device__ __forceinline__ bool rare_item(const RareJob& jb, long long b, RareItem& it) {
    int ch = 0;
    if (jb.G > 1) {
        const int s = (int)(b / jb.gps), g = (int)(b % jb.gps);
        const int sr0 = jb.r0 + s * jb.S;
        it.i0 = sr0 + g * jb.G;
        it.i1 = min(min(it.i0 + jb.G, sr0 + jb.S), jb.n);
        if (it.i0 >= it.i1) return false;
    } else {
        it.i0 = jb.r0 + (int)(b / jb.chunks);
        it.i1 = it.i0 + 1;
        ch = (int)(b % jb.chunks);
    }
    if (jb.rstart[it.i0] == jb.rstart[it.i1]) return false;  // no rare partners in these rows
    // symmetric: cells j in (i, n) of the stripe layout of rows r0..
    // (stripe s = rows r0+s*S ..); rectangular: cells j in [0, nB) of row
    // i - r0 of an (n - r0) x nB row-major matrix.
    const size_t s = jb.rect ? 0 : (size_t)((it.i0 - jb.r0) / jb.S);
    it.sr0 = (size_t)jb.r0 + s * jb.S;
    it.sbase = jb.rect ? 0 : stripe_off((size_t)(jb.n - jb.r0), (size_t)jb.S, s);
    const int jhi = jb.rect ? jb.nB : jb.n;
    it.j0 = (jb.rect ? 0 : it.i0 + 1) + ch * RCW;
    it.j1 = min(jhi, it.j0 + RCW);
    if (it.j0 >= jhi) return false;
    auto row_base = [&](int i) -> size_t {
        return jb.rect ? (size_t)(i - jb.r0) * (size_t)jb.nB
                       : it.sbase + ((size_t)i - it.sr0) * ((size_t)jb.n - it.sr0) - it.sr0;
    };
    it.g0 = row_base(it.i0) + it.j0;
    it.g1 = it.i1 - it.i0 == 1 ? row_base(it.i0) + it.j1 : row_base(it.i1 - 1) + jhi;
    it.a0 = it.g0 & ~(size_t)7;                    // first 16-byte word touched
    it.nq = (int)((it.g1 - it.a0 + 7) >> 3);       // 16-byte words staged
    it.lb = 0;
    return true;
}

// u16 index of cell (i, j) is rare_row_base(i) + j.
__device__ __forceinline__ size_t rare_row_base(const RareJob& jb, const RareItem& it, int i) {
    return jb.rect ? (size_t)(i - jb.r0) * (size_t)jb.nB
                   : it.sbase + ((size_t)i - it.sr0) * ((size_t)jb.n - it.sr0) - it.sr0;
}

// Staging loads of an item's words into registers (thread t holds words
// t, t + RNT, ...; indexes past nq reload word nq-1).  The indexes into v are
// template constants from the start, so v never becomes a stack array.
template <size_t... U>
__device__ __forceinline__ void rare_stage_issue(const RareJob& jb, const RareItem& it, v4i (&v)[RQ],
                                                 std::index_sequence<U...>) {
    const v4i* __restrict__ src = reinterpret_cast<const v4i*>(jb.out + it.a0);
    const int t = threadIdx.x, last = it.nq - 1;
    ((v[U] = src[min(t + (int)U * RNT, last)]), ...);
}

template <size_t... U>
__device__ __forceinline__ void rare_stage_land(const RareItem& it, const v4i (&v)[RQ], RareShared& sh,
                                                std::index_sequence<U...>) {
    v4i* cells = reinterpret_cast<v4i*>(sh.cells4 + it.lb);
    const int t = threadIdx.x;
    ((t + (int)U * RNT < it.nq ? (void)(cells[t + U * RNT] = v[U]) : (void)0), ...);
}

// Items of at most RHALF words are staged in one half of the LDS cells
// (words 0 .. RHALF-1, or RHALF .. 2 RHALF; the array holds 2 RHALF + 1), so
// the next such item can land in the other half while one is walked.
constexpr int RHALF = RCW / 16;

// An item's nq words straight from global memory into LDS words it.lb ..
// (global_load_lds_dwordx4: lane l of a wavefront lands at the wavefront's
// LDS base + 16 l), 64 words per wavefront instruction; lanes past the item
// reread its last word into the half's spare words (it.nq <= RHALF keeps all
// of them inside the half).  The loads count in the issuing wavefront's
// vmcnt: the caller waits for them (vmcnt 0) and then for every wavefront (a
// barrier) before the cells are used.  Written as an instruction rather than
// the builtin so that the compiler does not wait for these transfers before
// every LDS access of the walk running meanwhile in the other half.
__device__ __forceinline__ void rare_stage_dma(const RareJob& jb, const RareItem& it, RareShared& sh) {
    const v4i* __restrict__ src = reinterpret_cast<const v4i*>(jb.out + it.a0);
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int blocks = (it.nq + 63) / 64;
    for (int q = wave; q < blocks; q += RNT / 64) {
        const v4i* p = src + min(q * 64 + lane, it.nq - 1);
        const uint32_t lds = __builtin_amdgcn_readfirstlane(
            (uint32_t)(uintptr_t)(__attribute__((address_space(3))) void*)&sh.cells4[it.lb + q * 64]);
        asm volatile("s_mov_b32 m0, %1\n\tglobal_load_lds_dwordx4 %0, off" : : "v"(p), "s"(lds) : "memory", "m0");
    }
}

// Every rare partner of the item's rows is subtracted from its staged cell.
// Row i's partners are ranges of the ascending group member lists (rentry):
// a batch of up to RNT ranges is read (and clipped to the chunk) by all
// threads at once into LDS; then each rsg-lane subgroup takes ranges from
// LDS, its lanes reading consecutive members (MPL members in flight per lane)
// and applying them with LDS atomics.  The first barrier also orders the
// staging stores before the atomics.  MASK: ranges in arbitrary order (not
// clipped); a member outside the item's columns [j0, j1) is skipped.
template <typename M, bool MASK>
__device__ __forceinline__ void rare_ranges(const RareJob& jb, const M* __restrict__ gmem, const RareItem& it,
                                            RareShared& sh) {
    const uint32_t* __restrict__ rstart = jb.rstart;
    const uint64_t* __restrict__ rentry = jb.rentry;
    uint32_t* ebeg = sh.ebeg;
    uint32_t* eend = sh.eend;
    const int tid = threadIdx.x;
    // The rows' entries are contiguous in rentry, so the batch after the one
    // at eb of row i starts at eb + RNT inside the row, else at the row's end
    // (the start of the next row with entries).  Thread tid's entry of the
    // next batch is loaded while the current batch is walked.
    const uint32_t E1 = rstart[it.i1];
    uint64_t rn = 0;
    if (rstart[it.i0] + tid < E1) rn = rentry[rstart[it.i0] + tid];
    for (int i = it.i0; i < it.i1; i++) {
        const uint32_t e0 = rstart[i], e1 = rstart[i + 1];
        const size_t base = rare_row_base(jb, it, i);
        for (uint32_t eb = e0; eb < e1; eb += RNT) {
            const int nb = (int)min((uint32_t)RNT, e1 - eb);
            if (tid < nb) {
                const uint64_t r = rn;
                uint32_t b = (uint32_t)r, end = (uint32_t)(r >> 32);
                if (!MASK && jb.chunks > 1) {  // clip the ascending range to [j0, j1)
                    uint32_t lo = b, hi = end;
                    while (lo < hi) {
                        const uint32_t mid = (lo + hi) >> 1;
                        if ((int)gmem[mid] < it.j0) lo = mid + 1; else hi = mid;
                    }
                    b = lo;
                    hi = end;
                    while (lo < hi) {
                        const uint32_t mid = (lo + hi) >> 1;
                        if ((int)gmem[mid] < it.j1) lo = mid + 1; else hi = mid;
                    }
                    end = lo;
                }
                ebeg[tid] = b;
                eend[tid] = end;
            }
            if (tid == 0) sh.next = RNT / jb.rsg;  // ranges below this are taken without asking
            __syncthreads();
            {
                const uint32_t nbs = eb + RNT < e1 ? eb + RNT : e1;
                rn = nbs + tid < E1 ? rentry[nbs + tid] : 0;
            }
            // Subgroup sub starts with range sub and then takes the next free
            // one from sh.next, so a few long ranges do not leave the other
            // subgroups idle at the barrier.  A range is walked in units of
            // rsg x MPL members read as whole 16-byte words (lane sl takes
            // words sl, sl + rsg, ... of the unit, so each load instruction
            // of a subgroup reads consecutive words), the unit starting at
            // the word that holds the range's first member.  The next unit's
            // words are loaded while the current unit's atomics are issued,
            // so the member-load latency overlaps the LDS work.  Loads and
            // atomics are issued for every slot: a member outside the range
            // goes to a word of the lane's own (no branches, so the load
            // waits are fixed counts), and a word past the range's last one
            // reads that last word instead (never past the member array).
            //
            // A partner j of row i is the u16 cell at LDS offset o = bo + j:
            // the 32-bit word at byte 2o & ~3, low half when o is even.
            // Subtracting 1 from that cell adds -1 << (16 (o & 1)) to the word
            // (a cell never drops below its final value, so no borrow crosses
            // the halves); the shift amount is taken from o << 4, whose low
            // five bits are 16 (o & 1) (a 32-bit shift reads only those).
            constexpr int PER = 16 / (int)sizeof(M);  // members per 16-byte word
            constexpr int WPL = MPL / PER;            // words per lane per unit
            const v4i* __restrict__ words = reinterpret_cast<const v4i*>(gmem);
            const int rsg = jb.rsg, sub = tid / rsg, sl = tid % rsg, leader = (tid & 63) - sl;
            char* const lds = reinterpret_cast<char*>(sh.cells4);
            const uint32_t own = (uint32_t)(reinterpret_cast<char*>(&sh.spare[tid & 63]) - lds);
            const uint32_t bo = (uint32_t)(base - it.a0) + 8u * it.lb;  // LDS u16 offset of cell (i, 0), mod 2^32
            const uint32_t ustep = (uint32_t)(rsg * WPL);         // words per unit
            auto load_unit = [&](v4i (&w)[WPL], uint32_t uw, uint32_t end) {
                const uint32_t last = (max(end, 1u) - 1) / PER;
#pragma unroll
                for (int m = 0; m < WPL; m++) w[m] = words[min(uw + sl + m * rsg, last)];
            };
            // current unit: words [uw, uw + ustep) of the range [cp0, cend)
            uint32_t cp0 = 0, cend = 0;
            if (sub < nb) {
                cp0 = ebeg[sub];
                cend = eend[sub];
            }
            uint32_t uw = cp0 / PER;
            // One unit: find the next unit (the rest of this range, else the
            // next free range), load its words into `nxt`, then issue the
            // atomics of `cur`.  Returns false when no unit follows.
            auto unit = [&](const v4i (&cur)[WPL], v4i (&nxt)[WPL]) -> bool {
                uint32_t np0 = cp0, nend = cend, nuw = uw + ustep;
                bool more = true;
                if (nuw * PER >= cend) {  // this range ends in the current unit: take the next one
                    int nx = 0;
                    if (sl == 0) nx = (int)atomicAdd(&sh.next, 1u);
                    nx = __shfl(nx, leader, 64);
                    np0 = 0;
                    nend = 0;
                    if (nx < nb) {
                        np0 = ebeg[nx];
                        nend = eend[nx];
                    } else {
                        more = false;
                    }
                    nuw = np0 / PER;
                }
                load_unit(nxt, nuw, nend);
                __builtin_amdgcn_sched_barrier(0);  // the next unit's loads go out before this unit's atomics
                const uint32_t len = cend - cp0;
#pragma unroll
                for (int m = 0; m < WPL; m++) {
                    // position of the word's first member relative to the range start
                    const uint32_t q0 = (uw + sl + m * rsg) * PER - cp0;
#pragma unroll
                    for (int h = 0; h < PER; h++) {
                        const uint32_t x = (uint32_t)cur[m][PER == 8 ? h >> 1 : h];
                        const uint32_t j = PER == 8 ? __builtin_amdgcn_ubfe(x, (h & 1) * 16, 16) : x;
                        bool in = q0 + h < len;
                        if constexpr (MASK) in = in && j >= (uint32_t)it.j0 && j < (uint32_t)it.j1;
                        const uint32_t o = bo + j;
                        const uint32_t addr = in ? ((o << 1) & ~3u) : own;
                        atomicAdd(reinterpret_cast<uint32_t*>(lds + addr), ~0u << ((o << 4) & 31u));
                    }
                }
                cp0 = np0;
                cend = nend;
                uw = nuw;
                return more;
            };
            // two register sets in turn, so no unit's words are copied (a
            // copy would wait for the loads it was meant to overlap)
            v4i wa[WPL], wb[WPL];
            load_unit(wa, uw, cend);
            while (unit(wa, wb) && unit(wb, wa)) {
            }
            // both sets are used on the way out too, so the loads of a next
            // unit stay ahead of the atomics instead of sinking onto the
            // loop's back edge behind them
            asm volatile("" : : "v"(wa[0]), "v"(wb[0]));
            __syncthreads();  // ebeg / eend / next are rewritten by the next batch
        }
    }
}

// Subtracts 1 from the staged u16 cell at LDS u16 offset o: the 32-bit word
// at byte 2o & ~3, low half when o is even, i.e. -1 << (16 (o & 1)) added to
// the word (a cell never drops below its final value, so no borrow crosses
// the halves).  The shift amount is o << 4, whose low five bits are
// 16 (o & 1) (a 32-bit shift reads only those).
__device__ __forceinline__ void rare_dec(char* lds, uint32_t o) {
    atomicAdd(reinterpret_cast<uint32_t*>(lds + ((o << 1) & ~3u)), ~0u << ((o << 4) & 31u));
}

// The same walk as rare_ranges for ranges of ascending members, as one
// evenly split list of whole member words.  Range [b, end) of a batch has
// interior words wa = ceil(b / PER) .. wb = floor(end / PER) (every member of
// the word inside the range) and at most two partial edge words.  An
// exclusive scan over the batch's ranges lays the interior words end to end
// (ebeg = first position of range r, eend = its word index minus that
// position); wavefront w takes positions [W w / 16, W (w + 1) / 16) in
// chunks of 64, one word per lane, finding each lane's range from the range
// starts that fall inside the chunk.  Every member of an interior word is a
// partner: no per-member test.  The thread of range r applies its edge
// words' members itself, with a test per member.
// MR: an item of several rows (at most RFG).  Their entries are contiguous in
// rentry, so the batches run over all of them at once; each range's row is
// found by a binary search of the rows' starts (staged in LDS), its cell
// offset kept in wbo, and the walk carries each loaded word's offset with
// it through the register ring.
template <typename M, bool MR>
__device__ __forceinline__ void rare_ranges_flat(const RareJob& jb, const M* __restrict__ gmem, const RareItem& it,
                                                 RareShared& sh) {
    constexpr int PER = 16 / (int)sizeof(M);  // members per 16-byte word
    constexpr int NW = RNT / 64;              // wavefronts
    const v4i* __restrict__ words = reinterpret_cast<const v4i*>(gmem);
    const uint32_t* __restrict__ rstart = jb.rstart;
    const uint64_t* __restrict__ rentry = jb.rentry;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    char* const lds = reinterpret_cast<char*>(sh.cells4);
    const uint32_t own = (uint32_t)(reinterpret_cast<char*>(&sh.spare[lane]) - lds);
    uint32_t* const wstart = sh.ebeg;
    uint32_t* const wdelta = sh.eend;
    // members h of word x (word index wi) whose position lies in [lo, hi)
    auto edge = [&](const v4i& x, uint32_t wi, uint32_t lo, uint32_t hi, uint32_t bo) {
#pragma unroll
        for (int h = 0; h < PER; h++) {
            const uint32_t xv = (uint32_t)x[PER == 8 ? h >> 1 : h];
            const uint32_t j = PER == 8 ? __builtin_amdgcn_ubfe(xv, (h & 1) * 16, 16) : xv;
            const uint32_t o = bo + j;
            const bool in = wi * PER + h - lo < hi - lo;
            atomicAdd(reinterpret_cast<uint32_t*>(lds + (in ? ((o << 1) & ~3u) : own)), ~0u << ((o << 4) & 31u));
        }
    };
    // every member of word x
    auto whole = [&](const v4i& x, uint32_t bo) {
#pragma unroll
        for (int h = 0; h < PER; h++) {
            const uint32_t xv = (uint32_t)x[PER == 8 ? h >> 1 : h];
            rare_dec(lds, bo + (PER == 8 ? __builtin_amdgcn_ubfe(xv, (h & 1) * 16, 16) : xv));
        }
    };
    const uint32_t pitch = (uint32_t)((size_t)jb.n - it.sr0);  // MR: cell (i + 1, 0) - cell (i, 0)
    if constexpr (MR) {  // visible after the first batch's scan barrier, before any search
        int k = tid;
        asm volatile("" : "+v"(k));  // (formed here: a value kept from the kernel's start was spilled)
        for (; k <= it.i1 - it.i0; k += RNT) sh.rs[k] = rstart[it.i0 + k];
    }
    for (int i = it.i0; i < (MR ? it.i0 + 1 : it.i1); i++) {
        const uint32_t e0 = rstart[i], e1 = rstart[MR ? it.i1 : i + 1];
        // LDS u16 offset of cell (i, 0), mod 2^32
        const uint32_t bo = (uint32_t)(rare_row_base(jb, it, i) - it.a0) + 8u * it.lb;
        for (uint32_t eb = e0; eb < e1; eb += RNT) {
            const int nb = (int)min((uint32_t)RNT, e1 - eb);
            uint32_t b = 0, end = 0;
            if (tid < nb) {
                const uint64_t r = rentry[eb + tid];
                b = (uint32_t)r;
                end = (uint32_t)(r >> 32);
                if (jb.chunks > 1) {  // clip the ascending range to [j0, j1)
                    uint32_t lo = b, hi = end;
                    while (lo < hi) {
                        const uint32_t mid = (lo + hi) >> 1;
                        if ((int)gmem[mid] < it.j0) lo = mid + 1; else hi = mid;
                    }
                    b = lo;
                    hi = end;
                    while (lo < hi) {
                        const uint32_t mid = (lo + hi) >> 1;
                        if ((int)gmem[mid] < it.j1) lo = mid + 1; else hi = mid;
                    }
                    end = lo;
                }
            }
            const uint32_t wa = (b + PER - 1) / PER, wb = end / PER;
            const uint32_t nw = wb > wa ? wb - wa : 0;
            // edge words (read in any case: indexes stay inside the member array)
            const uint32_t fw = b / PER, lw = end > b ? (end - 1) / PER : fw;
            const v4i xf = words[fw], xl = words[lw];
            // exclusive scan of nw over the batch
            uint32_t incl = nw;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t y = (uint32_t)__shfl_up((int)incl, d, 64);
                if (lane >= d) incl += y;
            }
            if (lane == 63) sh.wsum[wave] = incl;
            __syncthreads();  // also orders the item's staging stores before the atomics
            uint32_t wpre = 0, W = 0;
#pragma unroll
            for (int w = 0; w < NW; w++) {
                const uint32_t s = sh.wsum[w];
                wpre += w < wave ? s : 0u;
                W += s;
            }
            const uint32_t ws = wpre + incl - nw;
            uint32_t bor = bo;  // the range's cell (row, 0)
            if (tid < nb) {
                wstart[tid] = ws;
                wdelta[tid] = wa - ws;
                if constexpr (MR) {
                    // the range's row: rs[lo] <= eb + tid < rs[hi] throughout
                    const uint32_t e = eb + (uint32_t)tid;
                    int lo = 0, hi = it.i1 - it.i0;
                    while (hi - lo > 1) {
                        const int mid = (lo + hi) >> 1;
                        if (sh.rs[mid] <= e) lo = mid; else hi = mid;
                    }
                    bor = bo + (uint32_t)lo * pitch;
                    sh.wbo[tid] = bor;
                }
            }
            if (tid < nb && end > b) {
                const bool ef = (b % PER) != 0 || nw == 0;
                const bool el = (end % PER) != 0 && lw != fw;
                edge(xf, fw, ef ? b : 0u, ef ? min(end, (fw + 1) * PER) : 0u, bor);
                edge(xl, lw, el ? lw * PER : 0u, el ? end : 0u, bor);
            }
            asm volatile("" : : "v"(xf[0]), "v"(xl[0]));
            __syncthreads();
            // wavefront-uniform (made explicit: the chunk loop's exit is then a
            // scalar branch)
            const uint32_t Wu = (uint32_t)__builtin_amdgcn_readfirstlane((int)W);
            const uint32_t wv = (uint32_t)__builtin_amdgcn_readfirstlane(wave);
            const uint32_t kbeg = (uint32_t)((uint64_t)Wu * wv / NW);
            const uint32_t kend = (uint32_t)((uint64_t)Wu * (wv + 1) / NW);
            if (kbeg < kend) {
                // r: the last range whose first position is <= kbeg
                // (positions ascend with r; range 0 starts at 0)
                int r;
                {
                    const uint32_t s = wstart[min(lane * 16, nb - 1)];
                    const int cb = __popcll(__ballot(lane * 16 < nb && s <= kbeg)) - 1;
                    const int c2 = cb * 16 + (lane & 15);
                    const uint32_t s2 = wstart[min(c2, nb - 1)];
                    r = cb * 16 + __popcll(__ballot(lane < 16 && c2 < nb && s2 <= kbeg)) - 1;
                }
                // Ranges of the chunk at k0 (lane: word k0 + lane) from the
                // starts s of ranges r + 1 + lane (~0 past the batch): the
                // last range starting at or before the lane's word.  r moves
                // to the last range starting inside the chunk; when all 64
                // ranges read start inside it, the next 64 are read here.
                auto starts_at = [&](int rr) -> uint32_t { return wstart[min(rr, nb - 1)]; };
                auto advance = [&](uint32_t k0, uint32_t s, int& rl) {
                    const uint32_t k = k0 + lane;
                    rl = r;
                    while (true) {
                        const int m = __popcll(__ballot(s <= k0 + 63));
                        for (int t = 0; t < m; t++) {
                            const uint32_t st = (uint32_t)__builtin_amdgcn_readlane((int)s, t);
                            if (k >= st) rl = r + 1 + t;
                        }
                        r += m;
                        if (m < 64) break;
                        const int rr = r + 1 + lane;
                        const uint32_t s0 = starts_at(rr);
                        s = rr < nb ? s0 : ~0u;
                    }
                };
                // Chunk c (at kbeg + 64 c) is loaded into set c % 4 of a
                // register ring and applied three steps later, so three
                // chunks' member loads are in flight while one is applied.
                // The sets are the quarters of one 16-register value and the
                // loop is unrolled by four with an atomic-issuing step each
                // time, so every load lands where the loop expects it and
                // every wait is a fixed count (a copy of a set, a branch
                // around the atomics or an exit inside the loop would cost
                // that).  In a step, the LDS reads for the next lookup are
                // issued before the atomics and used after them: LDS
                // operations complete in order, so a read issued after the
                // atomics would wait for them.  Chunks 0 .. nfull-1 are full;
                // chunk nfull, partial or empty, is applied at the end with
                // a lane test; loads past it read word 0.
                uint32_t k0 = kbeg;  // the next chunk to load
                v16i buf = v16i{};
                const uint32_t nfull = (kend - kbeg) / 64;
                auto set = [&](auto q, const v4i& x) {
                    constexpr int Q = decltype(q)::value;
                    if constexpr (Q == 0) buf.s0123 = x;
                    else if constexpr (Q == 1) buf.s4567 = x;
                    else if constexpr (Q == 2) buf.s89ab = x;
                    else buf.scdef = x;
                };
                auto get = [&](auto q) -> v4i {
                    constexpr int Q = decltype(q)::value;
                    if constexpr (Q == 0) return buf.s0123;
                    else if constexpr (Q == 1) return buf.s4567;
                    else if constexpr (Q == 2) return buf.s89ab;
                    else return buf.scdef;
                };
                // MR: the cell offset of each set's words (their ranges' rows)
                uint32_t bq0 = bo, bq1 = bo, bq2 = bo, bq3 = bo;
                auto setb = [&](auto q, uint32_t x) {
                    constexpr int Q = decltype(q)::value;
                    if constexpr (Q == 0) bq0 = x;
                    else if constexpr (Q == 1) bq1 = x;
                    else if constexpr (Q == 2) bq2 = x;
                    else bq3 = x;
                };
                auto getb = [&](auto q) -> uint32_t {
                    constexpr int Q = decltype(q)::value;
                    if constexpr (!MR) return bo;
                    else if constexpr (Q == 0) return bq0;
                    else if constexpr (Q == 1) return bq1;
                    else if constexpr (Q == 2) return bq2;
                    else return bq3;
                };
                int rl;  // lane's range in the chunk at k0
                {
                    const int rr = r + 1 + lane;
                    const uint32_t s0 = starts_at(rr);
                    advance(kbeg, rr < nb ? s0 : ~0u, rl);
                }
                using I0 = std::integral_constant<int, 0>;
                using I1 = std::integral_constant<int, 1>;
                using I2 = std::integral_constant<int, 2>;
                using I3 = std::integral_constant<int, 3>;
                // apply set A (when APPLY), then load the chunk at k0 into
                // set L and look up the next chunk's ranges
                auto step = [&](auto l, auto a, auto apply) {
                    const uint32_t d = wdelta[rl];
                    const uint32_t bl = MR ? sh.wbo[rl] : 0u;
                    const int rr = r + 1 + lane;
                    const uint32_t s0 = starts_at(rr);
                    __builtin_amdgcn_sched_barrier(0);
                    if constexpr (decltype(apply)::value) whole(get(a), getb(a));
                    __builtin_amdgcn_sched_barrier(0);
                    const uint32_t k = k0 + lane;
                    set(l, words[k < kend ? k + d : 0u]);
                    if constexpr (MR) setb(l, bl);
                    k0 += 64;
                    advance(k0, rr < nb ? s0 : ~0u, rl);
                };
                step(I0{}, I0{}, std::false_type{});  // chunks 0, 1, 2
                step(I1{}, I0{}, std::false_type{});
                step(I2{}, I0{}, std::false_type{});
                uint32_t c = 0;  // chunks applied
                for (; c + 4 <= nfull; c += 4) {
                    step(I3{}, I0{}, std::true_type{});
                    step(I0{}, I1{}, std::true_type{});
                    step(I1{}, I2{}, std::true_type{});
                    step(I2{}, I3{}, std::true_type{});
                }
                // the last nfull - c (< 4) full chunks; chunk nfull is then in
                // set nfull % 4 = nfull - c, with at most two loads after it
                if (c < nfull) step(I3{}, I0{}, std::true_type{});
                if (c + 1 < nfull) step(I0{}, I1{}, std::true_type{});
                if (c + 2 < nfull) step(I1{}, I2{}, std::true_type{});
                if (nfull * 64 < kend - kbeg) {
                    const uint32_t q = nfull - c;
                    const v4i t = q == 0 ? get(I0{}) : q == 1 ? get(I1{}) : q == 2 ? get(I2{}) : get(I3{});
                    const uint32_t tb = q == 0 ? getb(I0{}) : q == 1 ? getb(I1{}) : q == 2 ? getb(I2{}) : getb(I3{});
                    if (kbeg + nfull * 64 + lane < kend) whole(t, tb);
                }
                // the last loads may not be used: keep every set live so that
                // no load can land in a register that is already reused
                asm volatile("" : : "v"(buf[0]), "v"(buf[4]), "v"(buf[8]), "v"(buf[12]));
            }
            __syncthreads();  // wsum / ebeg / eend are rewritten by the next batch
        }
    }
}

// The staged words back to the layout (cap applied first when set): whole
// words inside [g0, g1) as 16-byte stores, the partial first and last words
// cell by cell (their other cells belong to other items).
__device__ __forceinline__ void rare_write_back(const RareJob& jb, const RareItem& it, RareShared& sh) {
    const int tid = threadIdx.x;
    const uint4* cells = sh.cells4 + it.lb;
    if (jb.cap > 0) {  // main.rs:251-253 on the final values
        uint16_t* c16 = reinterpret_cast<uint16_t*>(sh.cells4 + it.lb);
        const uint16_t cap = (uint16_t)min(jb.cap, 0xffff);
        for (int k = tid; k < it.nq * 8; k += RNT) c16[k] = min(c16[k], cap);
        __syncthreads();
    }
    uint16_t* __restrict__ out = jb.out;
    uint4* __restrict__ dst = reinterpret_cast<uint4*>(out + it.a0);
    for (int k = tid; k < it.nq; k += RNT) {
        const size_t lo = it.a0 + 8 * (size_t)k;  // u16 index of the word's first cell
        if (lo >= it.g0 && lo + 8 <= it.g1) {
            dst[k] = cells[k];
        } else {
            const uint16_t* c16 = reinterpret_cast<const uint16_t*>(&cells[k]);
            for (int h = 0; h < 8; h++)
                if (lo + h >= it.g0 && lo + h < it.g1) out[lo + h] = c16[h];
        }
    }
}

// Items b0, b0 + step, ... < nitems of job jb, one after the other in this
// workgroup.
// When this item and the next both fit half of the LDS cells (at most RHALF
// words), the next item's cells are copied straight into the other half
// before this item's walk starts (rare_stage_dma): the walk's first load wait
// covers them (loads complete in order), and the next item starts with its
// cells in place, with no staging round trip of its own.
// Otherwise the next item's staging loads go to registers before this item
// is written back, so their latency is spent under the write-back, and the
// next item lands at word 0 after it.
// (Measured and not kept: issuing the register staging loads, and the next
// item's first partner entries, as soon as the current item has landed: the
// registers they hold through the walk cost spills, 4.9 against 4.7 ms;
// counting partners in a cleared LDS buffer and subtracting the counts from
// the output with device-scope atomics instead of staging: 11.7 ms, the L2's
// atomic rate over 4 GB of output words.)
template <typename M, bool MASK>
__device__ __forceinline__ void rare_items(const RareJob& jb, const M* __restrict__ gmem, long long b0,
                                           long long nitems, long long step, RareShared& sh) {
    RareItem cur, nxt;
    long long b = b0;
    while (b < nitems && !rare_item(jb, b, cur)) b += step;
    if (b >= nitems) return;
    constexpr auto SEQ = std::make_index_sequence<RQ>();
    v4i v[RQ];
    rare_stage_issue(jb, cur, v, SEQ);
    bool landed = false;  // cur's cells are in LDS already (copied during the previous walk)
    while (true) {
        if (!landed) rare_stage_land(cur, v, sh, SEQ);
        long long nb = b + step;
        while (nb < nitems && !rare_item(jb, nb, nxt)) nb += step;
        const bool more = nb < nitems;
        const bool dma = more && cur.nq <= RHALF && nxt.nq <= RHALF;  // uniform
        if (dma) {
            nxt.lb = cur.lb ? 0 : RHALF;
            rare_stage_dma(jb, nxt, sh);
        }
        // Ascending ranges: one-row items take the flat walk (on the 63,005-row
        // long-tail set, one-row items only, 4.8 ms against 5.2 ms for the
        // subgroup walk); items of several rows take it too when the ranges
        // are long (jb.mflat), else the subgroup walk, whose cost per short
        // range is lower (the flat walk made the batch of 64 sets of 500 to
        // 20,000 rows 18% slower).
        if constexpr (MASK) rare_ranges<M, true>(jb, gmem, cur, sh);
        else if (cur.i1 - cur.i0 == 1) rare_ranges_flat<M, false>(jb, gmem, cur, sh);
        else if (jb.mflat && cur.i1 - cur.i0 <= RFG) rare_ranges_flat<M, true>(jb, gmem, cur, sh);
        else rare_ranges<M, false>(jb, gmem, cur, sh);
        if (more && !dma) rare_stage_issue(jb, nxt, v, SEQ);  // nxt.lb == 0
        rare_write_back(jb, cur, sh);
        if (!more) return;
        if (dma) __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0): this wavefront's copies have landed
        __syncthreads();  // the write-back's LDS reads before the next item lands; every copy landed
        landed = dma;
        b = nb;
        cur = nxt;
    }
}

// Member lists are read as u16 when every row index fits: half the bytes of
// the pass's dominant (L2-missing) reads.
template <bool MASK>
__device__ __forceinline__ void rare_items_any(const RareJob& jb, long long b0, long long nitems, long long step,
                                               RareShared& sh) {
    if (jb.m16) rare_items<uint16_t, MASK>(jb, static_cast<const uint16_t*>(jb.gmem), b0, nitems, step, sh);
    else rare_items<uint32_t, MASK>(jb, static_cast<const uint32_t*>(jb.gmem), b0, nitems, step, sh);
}

// One job, a persistent grid: workgroup w takes items w, w + grid, ...
// MASK: the job's ranges are unsorted (RareJob::sorted == 0).
template <bool MASK>
__global__ void __launch_bounds__(RNT) sub_rare_rows(RareJob jb, long long nitems) {
    __shared__ RareShared sh;
    rare_items_any<MASK>(jb, blockIdx.x, nitems, gridDim.x, sh);
}


size_t up(size_t x, size_t m) { return (x + m - 1) / m * m; }

// Workgroups of a persistent rare launch on the current device: one per CU
// (the LDS of one workgroup leaves no room for a second), at most one per item.
unsigned rare_grid(long long items) {
    static std::atomic<int> cus[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    int c = cus[dev].load(std::memory_order_relaxed);
    if (c <= 0) {
        if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c <= 0) c = 256;
        cus[dev].store(c, std::memory_order_relaxed);
    }
    return (unsigned)std::max<long long>(1, std::min<long long>(items, c));
}

// Work items of a rare job over `rows` rows from jb.r0: row groups when
// G >= 2 rows of the widest stripe fit one chunk, else one row per item in
// column chunks.  Returns the number of items.
long long rare_layout(RareJob& jb, int rows) {
    const int pitch = jb.rect ? jb.nB : jb.n - jb.r0;  // the widest row of the layout
    jb.G = jb.rect ? 1 : std::max(1, RCW / std::max(pitch, 1));
    if (jb.G > 1) {
        jb.chunks = 1;
        jb.gps = (jb.S + jb.G - 1) / jb.G;
        return (long long)((rows + jb.S - 1) / jb.S) * jb.gps;
    }
    const int width = jb.rect ? jb.nB : jb.n - jb.r0 - 1;
    jb.chunks = std::max(1, (width + RCW - 1) / RCW);
    jb.gps = 0;
    return (long long)rows * jb.chunks;
}

// One persistent rare launch of job jb over `items` work items.
hipError_t launch_rare(const RareJob& jb, long long items, hipStream_t st) {
    if (jb.sorted) hipLaunchKernelGGL(sub_rare_rows<false>, dim3(rare_grid(items)), dim3(RNT), 0, st, jb, items);
    else hipLaunchKernelGGL(sub_rare_rows<true>, dim3(rare_grid(items)), dim3(RNT), 0, st, jb, items);
    return hipGetLastError();
}

void set_err(char* err, int errlen, const char* fmt, ...) __attribute__((format(printf, 3, 4)));
void set_err(char* err, int errlen, const char* fmt, ...) {
    if (!err || errlen <= 0) return;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(err, (size_t)errlen, fmt, ap);
    va_end(ap);
}


// Tile shapes: 256 (4 wavefronts of 128 x 128) for large sets, 128 (4
// wavefronts of 64 x 64) when the 256 grid would leave most CUs idle.
int pick_tile(int n) { return n >= 8192 ? 256 : 128; }

// Stripes r0, r0+S, ... up to row n in one launch when S is a multiple of the
// tile (else one launch per stripe), output in the stripe layout at out.
// Row-major operand rows (rows x kw words) -> rows row0 .. row0+rows-1 of
// planes: plane g holds words 4g..4g+3 of every row (16 bytes per row, ld
// rows per plane).  Planes past kw / 4 read zero.  One thread per 16-byte
// piece, in plane order, so the stores are contiguous.
__global__ void __launch_bounds__(256)
to_planes(const uint32_t* __restrict__ src, int rows, int kw, uint4* __restrict__ dst, int ld, int planes, int row0) {
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (size_t)planes * (size_t)rows) return;
    const int g = (int)(idx / (size_t)rows), r = (int)(idx % (size_t)rows);
    uint4 v = make_uint4(0u, 0u, 0u, 0u);
    if (4 * g + 3 < kw) {
        const uint32_t* p = src + (size_t)r * kw + 4 * g;
        v = make_uint4(p[0], p[1], p[2], p[3]);
    }
    dst[(size_t)g * ld + row0 + r] = v;
}

hipError_t launch_planes(const uint32_t* src, int rows, int kw, uint32_t* dst, int ld, int kwd, hipStream_t st,
                         int row0 = 0) {
    const size_t pieces = (size_t)(kwd / 4) * (size_t)rows;
    hipLaunchKernelGGL(to_planes, dim3((unsigned)((pieces + 255) / 256)), dim3(256), 0, st, src, rows, kw,
                       reinterpret_cast<uint4*>(dst), ld, kwd / 4, row0);
    return hipGetLastError();
}

// Host row-major operand (ld x kw words) -> a new device buffer in planes of
// kwd words (kwd >= kw, a multiple of KC), converted through a staging
// buffer of at most 256 MiB of rows at a time.
hipError_t upload_planes(const void* host, int ld, int kw, int kwd, uint32_t** out) {
    uint32_t* tmp = NULL;
    hipError_t e;
    const size_t row_bytes = (size_t)kw * 4;
    const int block = (int)std::max<size_t>(256, std::min<size_t>((size_t)ld, ((256u << 20) / row_bytes) / 256 * 256));
    *out = NULL;
    if ((e = hipMalloc(&tmp, (size_t)block * row_bytes))) return e;
    if ((e = hipMalloc(out, (size_t)ld * kwd * 4))) goto done;
    for (int r0 = 0; r0 < ld; r0 += block) {
        const int rows = std::min(block, ld - r0);
        if ((e = hipMemcpy(tmp, (const char*)host + (size_t)r0 * row_bytes, (size_t)rows * row_bytes,
                           hipMemcpyHostToDevice)))
            goto done;
        if ((e = launch_planes(tmp, rows, kw, *out, ld, kwd, NULL, r0))) goto done;
        if ((e = hipDeviceSynchronize())) goto done;  // the staging buffer is reused
    }
done:
    (void)hipFree(tmp);
    if (e && *out) {
        (void)hipFree(*out);
        *out = NULL;
    }
    return e;
}

// Device operand build from dense ranks X (n x L u16, device): P bit k and
// O bit colmap[colstart[k] + a] of every called cell, ORed into zeroed planes
// of ld rows.  Workgroup b: row b / lb, loci 256 (b % lb) .. +255.
__global__ void __launch_bounds__(256)
build_planes(const uint16_t* __restrict__ X, int L, int lb, const uint32_t* __restrict__ colstart,
             const uint32_t* __restrict__ colmap, uint32_t* __restrict__ O, uint32_t* __restrict__ P, int ld) {
    const int i = (int)(blockIdx.x / (unsigned)lb);
    const int k = (int)(blockIdx.x % (unsigned)lb) * 256 + (int)threadIdx.x;
    if (k >= L) return;
    const int a = X[(size_t)i * L + k];
    if (!a) return;
    auto word = [&](uint32_t bit) { return (((size_t)(bit >> 7) * ld + i) << 2) + ((bit >> 5) & 3); };
    atomicOr(&P[word((uint32_t)k)], 1u << (k & 31));
    const uint32_t c = colmap[colstart[k] + a];
    if (c != 0xffffffffu) atomicOr(&O[word(c)], 1u << (c & 31));
}

int kc_words(size_t kw) { return (int)up(kw, (size_t)KC); }

// Member array (u32 row indices) -> a new device buffer, narrowed to u16
// when m16.
hipError_t upload_members(const uint32_t* mem, size_t count, bool m16, void** out) {
    hipError_t e;
    *out = NULL;
    if ((e = hipMalloc(out, cghi::member_array_bytes(count, m16)))) return e;
    if (!count) return hipSuccess;
    if (m16) {
        std::vector<uint16_t> h(count);
        for (size_t k = 0; k < count; k++) h[k] = (uint16_t)mem[k];
        e = hipMemcpy(*out, h.data(), count * 2, hipMemcpyHostToDevice);
    } else {
        e = hipMemcpy(*out, mem, count * 4, hipMemcpyHostToDevice);
    }
    if (e) {
        (void)hipFree(*out);
        *out = NULL;
    }
    return e;
}

// Column width of a product tile: the MFMA forms' 256-row tile is 256 x 128
// (two workgroups per CU), every other tile is square.
int tile_width(int kernel, int tile) { return tile == 256 && kernel != CGH_KERNEL_VALU_POPC ? 128 : tile; }

// One product launch over the tiles of `a` with the kernel form, tile edge
// (rows) and LDS ring depth given (2, 3 or 4 slabs for the 128-tile; the
// 256 x 128 tile always runs a 2-slab ring, which keeps two workgroups per
// CU); grid = a.chunk * 8 workgroups.
hipError_t launch_tile_args(const TileArgs& a, int kernel, int tile, int nst, hipStream_t st) {
    const dim3 grid((unsigned)(a.chunk * 8));
    const bool fp4 = kernel == CGH_KERNEL_MFMA_FP4;
    if (kernel == CGH_KERNEL_VALU_POPC) {
        if (tile == 256) hipLaunchKernelGGL((valu_tile<256>), grid, dim3(256), 0, st, a);
        else hipLaunchKernelGGL((valu_tile<128>), grid, dim3(256), 0, st, a);
        return hipGetLastError();
    }
#define CGH_MFMA_LAUNCH(F, K) hipLaunchKernelGGL((mfma_tile<128, 128, 64, 64, F, K>), grid, dim3(256), 0, st, a)
#define CGH_MFMA_STAGES(F)                         \
    switch (nst) {                                 \
        case 2: CGH_MFMA_LAUNCH(F, 2); break;      \
        case 4: CGH_MFMA_LAUNCH(F, 4); break;      \
        default: CGH_MFMA_LAUNCH(F, 3); break;     \
    }
    if (tile == 256 && fp4) hipLaunchKernelGGL((mfma_tile<256, 128, 128, 64, true, 2>), grid, dim3(256), 0, st, a);
    else if (tile == 256) hipLaunchKernelGGL((mfma_tile<256, 128, 128, 64, false, 2>), grid, dim3(256), 0, st, a);
    else if (fp4) { CGH_MFMA_STAGES(true) }
    else { CGH_MFMA_STAGES(false) }
#undef CGH_MFMA_STAGES
#undef CGH_MFMA_LAUNCH
    return hipGetLastError();
}

hipError_t launch_tiles(int kernel, int tile, const uint32_t* O, int kw1, const uint32_t* P, int kw2, int ld, int n,
                        int r0, int S, int ns, uint16_t* out, hipStream_t st, const uint32_t* rstart = NULL,
                        int cap = 0, int kv1 = 0, int kv2 = 0, int nst = NST_DEFAULT, int accum = 0) {
    const int tw = tile_width(kernel, tile);
    if (ns > 1 && S % tile) {
        for (int s = 0; s < ns; s++) {
            const hipError_t e = launch_tiles(kernel, tile, O, kw1, P, kw2, ld, n, r0 + s * S, S, 1,
                                              out + stripe_off((size_t)(n - r0), (size_t)S, (size_t)s), st, rstart,
                                              cap, kv1, kv2, nst, accum);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
    TileArgs a;
    a.job.O = O; a.job.P = P; a.job.kw1 = kw1; a.job.kw2 = kw2; a.job.n = n; a.job.r0 = r0; a.job.S = S;
    a.job.out = out;
    a.job.OB = O; a.job.PB = P; a.job.nB = n; a.job.rect = 0;
    a.job.ldA = ld; a.job.ldB = ld;
    a.job.rstart = rstart; a.job.cap = cap;
    a.job.kv1 = kv1; a.job.kv2 = kv2;
    a.job.accum = accum;
    a.ns = ns;
    a.nbc = (n - r0 + tw - 1) / tw;
    a.q = S / tile;
    a.qc = S / tw;
    const int last = r0 + (ns - 1) * S;
    const long long tm_last = (std::min(S, n - last) + tile - 1) / tile;
    a.T = (ns > 1 ? tiles_before(ns - 1, a.q, a.qc, a.nbc) : 0) + tm_last * (a.nbc - (long long)(ns - 1) * a.qc);
    a.chunk = (a.T + 7) / 8;
    return launch_tile_args(a, kernel, tile, nst, st);
}

// CGH_KERNEL_AUTO resolves to the fastest measured product kernel.
int resolve_kernel(int kernel) { return kernel == CGH_KERNEL_AUTO ? CGH_KERNEL_MFMA_FP4 : kernel; }
bool valid_kernel(int kernel) {
    return kernel == CGH_KERNEL_AUTO || kernel == CGH_KERNEL_MFMA_I8 || kernel == CGH_KERNEL_MFMA_FP4 ||
           kernel == CGH_KERNEL_VALU_POPC;
}

}  // namespace

// ---------------------------------------------------------------------------
// Reference-compatible boundary
// ---------------------------------------------------------------------------

extern "C" int tch_device_count(void) {
    int c = 0;
    return hipGetDeviceCount(&c) == hipSuccess ? c : 0;
}

extern "C" size_t tch_free_vram(int device) {
    const char* env = getenv("CGH_FREE_BYTES");
    if (env && *env) return (size_t)strtoull(env, NULL, 10);
    size_t f = 0, t = 0;
    if (hipSetDevice(device) != hipSuccess || hipMemGetInfo(&f, &t) != hipSuccess) return 0;
    return f;
}

// ---------------------------------------------------------------------------
// Device-resident dataset
// ---------------------------------------------------------------------------

namespace {
// f(lo, hi) over [0, count) split into contiguous ranges on the host's
// hardware threads.
template <class F>
void host_parallel(size_t count, F&& f) {
    const size_t threads = std::max<size_t>(1, std::min<size_t>(std::thread::hardware_concurrency(), 64));
    const size_t per = (count + threads - 1) / threads;
    if (threads == 1 || count < 2 * threads) {
        f((size_t)0, count);
        return;
    }
    std::vector<std::thread> pool;
    for (size_t lo = 0; lo < count; lo += per) pool.emplace_back([&f, lo, per, count] { f(lo, std::min(count, lo + per)); });
    for (auto& t : pool) t.join();
}
}  // namespace

struct cgh_dataset {
    int device;
    int n, L, tau;
    size_t D;                 // one-hot columns
    int cap;                  // > 0: outputs capped (cgh_dataset_set_cap)
    int tuned_kernel, tuned_tile, tuned_stages;  // from cgh_dataset_autotune (0 = not tuned)
    size_t n_pad, kw1, kw2;   // kw1, kw2: words per row on the device (multiples of KC)
    uint32_t *dO, *dP;        // planes of n_pad rows
    size_t n_groups, n_members;
    uint64_t n_rare;
    size_t n_rentries;
    void* dgmem;              // members, u16 when m16
    int m16;
    uint32_t* drstart;
    uint64_t* drentry;
    // One-hot kept in host memory (cgh_dataset_upload_split): hO holds the
    // kw1 / 4 O planes of n_pad rows in pinned memory, streamed through two
    // device buffers of chunk_planes planes each (dO is NULL).
    uint32_t* hO;
    uint32_t* dOc[2];
    size_t chunk_planes;
    hipStream_t ostream;
    hipEvent_t loaded[2], used[2];
};

extern "C" void cgh_dataset_free(cgh_dataset* ds) {
    if (!ds) return;
    (void)hipSetDevice(ds->device);
    if (ds->ostream) (void)hipStreamSynchronize(ds->ostream);
    for (int k = 0; k < 2; k++) {
        if (ds->used[k]) (void)hipEventSynchronize(ds->used[k]);
        if (ds->dOc[k]) (void)hipFree(ds->dOc[k]);
        if (ds->loaded[k]) (void)hipEventDestroy(ds->loaded[k]);
        if (ds->used[k]) (void)hipEventDestroy(ds->used[k]);
    }
    if (ds->ostream) (void)hipStreamDestroy(ds->ostream);
    if (ds->hO) (void)hipHostFree(ds->hO);
    if (ds->dO) (void)hipFree(ds->dO);
    if (ds->dP) (void)hipFree(ds->dP);
    if (ds->dgmem) (void)hipFree(ds->dgmem);
    if (ds->drstart) (void)hipFree(ds->drstart);
    if (ds->drentry) (void)hipFree(ds->drentry);
    free(ds);
}

namespace {

// A dataset's sizes and rare structures from an encoding (O/P not touched).
int dataset_meta(int device, const cgh_encoding* enc, cgh_dataset** out, char* err, int errlen) {
    cgh_dataset* ds = (cgh_dataset*)calloc(1, sizeof(cgh_dataset));
    if (!ds) {
        set_err(err, errlen, "out of host memory");
        return 1;
    }
    *out = NULL;
    ds->device = device;
    ds->n = enc->n;
    ds->L = enc->L;
    ds->tau = enc->tau;
    ds->D = enc->D;
    ds->n_pad = enc->n_pad;
    ds->kw1 = (size_t)kc_words(enc->kw1);
    ds->kw2 = (size_t)kc_words(enc->kw2);
    ds->n_groups = enc->n_groups;
    ds->n_members = enc->n_members;
    ds->n_rare = enc->n_rare;
    ds->n_rentries = enc->n_rentries;
    if (enc->kw1 % 4 || enc->kw2 % 4 || enc->kw1 == 0 || enc->n_pad % 256) {
        set_err(err, errlen, "cgh_dataset_upload: rows must be padded to 256 bits and 256 rows");
        free(ds);
        return 1;
    }
    HIP_TRY(hipSetDevice(device));
    if (enc->n_groups) {
        ds->m16 = enc->n <= 65536;
        HIP_TRY(upload_members(enc->gmem, enc->n_members, ds->m16, &ds->dgmem));
        HIP_TRY(hipMalloc(&ds->drstart, ((size_t)enc->n + 1) * 4));
        HIP_TRY(hipMalloc(&ds->drentry, std::max<size_t>(enc->n_rentries, 1) * 8));
        HIP_TRY(hipMemcpy(ds->drstart, enc->rstart, ((size_t)enc->n + 1) * 4, hipMemcpyHostToDevice));
        if (enc->n_rentries)
            HIP_TRY(hipMemcpy(ds->drentry, enc->rentry, enc->n_rentries * 8, hipMemcpyHostToDevice));
    }
    *out = ds;
    return 0;
fail:
    cgh_dataset_free(ds);
    return 1;
}

}  // namespace

extern "C" int cgh_dataset_upload(int device, const cgh_encoding* enc, cgh_dataset** out, char* err, int errlen) {
    cgh_dataset* ds = NULL;
    *out = NULL;
    if (dataset_meta(device, enc, &ds, err, errlen)) return 1;
    HIP_TRY(upload_planes(enc->O, (int)enc->n_pad, (int)enc->kw1, (int)ds->kw1, &ds->dO));
    HIP_TRY(upload_planes(enc->P, (int)enc->n_pad, (int)enc->kw2, (int)ds->kw2, &ds->dP));
    *out = ds;
    return 0;
fail:
    cgh_dataset_free(ds);
    return 1;
}

extern "C" int cgh_dataset_upload_split(int device, const cgh_encoding* enc, size_t o_bytes, cgh_dataset** out,
                                        char* err, int errlen) {
    cgh_dataset* ds = NULL;
    *out = NULL;
    if (dataset_meta(device, enc, &ds, err, errlen)) return 1;
    {
        const size_t n_pad = ds->n_pad, planes = ds->kw1 / 4, plane_bytes = n_pad * 16;
        // two chunk buffers, whole slabs (KG planes) each
        ds->chunk_planes = std::min(o_bytes / 2 / plane_bytes / KG * KG, planes);
        if (ds->chunk_planes < (size_t)KG) {
            set_err(err, errlen, "cgh_dataset_upload_split: %zu bytes hold less than two slabs of %zu one-hot planes",
                    o_bytes, planes);
            cgh_dataset_free(ds);
            return 1;
        }
        HIP_TRY(hipHostMalloc(&ds->hO, planes * plane_bytes, hipHostMallocDefault));
        // row-major words (kw1 per row) -> planes of 4 words; planes past the
        // encoding's words stay zero
        memset(ds->hO, 0, planes * plane_bytes);
        host_parallel((size_t)enc->n_pad, [&](size_t lo, size_t hi) {
            for (size_t r = lo; r < hi; r++)
                for (size_t w = 0; w < enc->kw1; w++) ds->hO[((w / 4) * n_pad + r) * 4 + w % 4] = enc->O[r * enc->kw1 + w];
        });
        for (int k = 0; k < 2; k++) {
            HIP_TRY(hipMalloc(&ds->dOc[k], ds->chunk_planes * plane_bytes));
            HIP_TRY(hipEventCreateWithFlags(&ds->loaded[k], hipEventDisableTiming));
            HIP_TRY(hipEventCreateWithFlags(&ds->used[k], hipEventDisableTiming));
        }
        HIP_TRY(hipStreamCreateWithFlags(&ds->ostream, hipStreamNonBlocking));
    }
    HIP_TRY(upload_planes(enc->P, (int)enc->n_pad, (int)enc->kw2, (int)ds->kw2, &ds->dP));
    *out = ds;
    return 0;
fail:
    cgh_dataset_free(ds);
    return 1;
}

extern "C" int cgh_dataset_build(int device, const uint16_t* X, int n, int L, int tau, cgh_dataset** out,
                                 char* err, int errlen) {
    cgh_encoding enc;
    cgh_dataset* ds = NULL;
    uint16_t* dX = NULL;
    uint32_t *dcs = NULL, *dcm = NULL;
    std::vector<uint32_t> colstart, colmap;
    size_t need = 0;
    *out = NULL;
    if (n < 1 || L < 1) {
        set_err(err, errlen, "cgh_dataset_build: bad arguments (n=%d L=%d)", n, L);
        return 1;
    }
    if (cgh_encode_meta(X, n, L, tau, 256, &enc, err, errlen)) return 1;
    if (dataset_meta(device, &enc, &ds, err, errlen)) {
        cgh_encoding_free(&enc);
        return 1;
    }
    need = cgh_column_table(X, n, L, tau, NULL, NULL, 0);
    colstart.resize((size_t)L + 1);
    colmap.resize(need);
    cgh_column_table(X, n, L, tau, colstart.data(), colmap.data(), need);
    HIP_TRY(hipMalloc(&ds->dO, ds->n_pad * ds->kw1 * 4));
    HIP_TRY(hipMalloc(&ds->dP, ds->n_pad * ds->kw2 * 4));
    HIP_TRY(hipMemset(ds->dO, 0, ds->n_pad * ds->kw1 * 4));
    HIP_TRY(hipMemset(ds->dP, 0, ds->n_pad * ds->kw2 * 4));
    HIP_TRY(hipMalloc(&dX, (size_t)n * L * 2));
    HIP_TRY(hipMalloc(&dcs, colstart.size() * 4));
    HIP_TRY(hipMalloc(&dcm, std::max<size_t>(need, 1) * 4));
    HIP_TRY(hipMemcpy(dX, X, (size_t)n * L * 2, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dcs, colstart.data(), colstart.size() * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dcm, colmap.data(), need * 4, hipMemcpyHostToDevice));
    {
        const int lb = (L + 255) / 256;
        hipLaunchKernelGGL(build_planes, dim3((unsigned)((size_t)n * lb)), dim3(256), 0, 0, dX, L, lb, dcs, dcm,
                           ds->dO, ds->dP, (int)ds->n_pad);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipDeviceSynchronize());
    }
    (void)hipFree(dX);
    (void)hipFree(dcs);
    (void)hipFree(dcm);
    cgh_encoding_free(&enc);
    *out = ds;
    return 0;
fail:
    if (dX) (void)hipFree(dX);
    if (dcs) (void)hipFree(dcs);
    if (dcm) (void)hipFree(dcm);
    cgh_encoding_free(&enc);
    cgh_dataset_free(ds);
    return 1;
}

namespace cghi {

bool hip_failed(hipError_t e, const char* call, char* err, int errlen) {
    if (e == hipSuccess) return false;
    set_err(err, errlen, "HIP error '%s' in %s", hipGetErrorString(e), call);
    return true;
}

cgh_dataset* dataset_adopt(const DatasetParts& p) {
    cgh_dataset* ds = (cgh_dataset*)calloc(1, sizeof(cgh_dataset));
    if (!ds) return NULL;
    ds->device = p.device;
    ds->n = p.n;
    ds->L = p.L;
    ds->tau = p.tau;
    ds->D = p.D;
    ds->n_pad = p.n_pad;
    ds->kw1 = p.kw1;
    ds->kw2 = p.kw2;
    ds->dO = p.dO;
    ds->dP = p.dP;
    ds->n_groups = p.n_groups;
    ds->n_members = p.n_members;
    ds->n_rare = p.n_rare;
    ds->n_rentries = p.n_rentries;
    ds->dgmem = p.dgmem;
    ds->m16 = p.m16;
    ds->drstart = p.drstart;
    ds->drentry = p.drentry;
    return ds;
}

int device_words(size_t kw) { return kc_words(kw); }

hipError_t launch_build_planes(const uint16_t* dX, int n, int L, const uint32_t* dcolstart, const uint32_t* dcolmap,
                               uint32_t* dO, uint32_t* dP, int ld, hipStream_t st) {
    const int lb = (L + 255) / 256;
    hipLaunchKernelGGL(build_planes, dim3((unsigned)((size_t)n * lb)), dim3(256), 0, st, dX, L, lb, dcolstart,
                       dcolmap, dO, dP, ld);
    return hipGetLastError();
}

}  // namespace cghi

extern "C" int cgh_dataset_set_cap(cgh_dataset* ds, int cap) {
    if (cap < 0) return 1;
    // every distance is <= L <= 65,535: a cap from 65,535 up changes nothing
    ds->cap = cap >= 65535 ? 0 : cap;
    return 0;
}

extern "C" void cgh_dataset_info(const cgh_dataset* ds, int* n, int* L, int* tau, size_t* D, uint64_t* n_rare) {
    if (n) *n = ds->n;
    if (L) *L = ds->L;
    if (tau) *tau = ds->tau;
    if (D) *D = ds->D;
    if (n_rare) *n_rare = ds->n_rare;
}

extern "C" size_t cgh_stripe_offset(int n, int S, int s) { return stripe_off((size_t)n, (size_t)S, (size_t)s); }

extern "C" size_t cgh_upper_elems(int n, int S) {
    if (n <= 0 || S <= 0) return 0;
    const int ns = (n + S - 1) / S;
    const size_t last = (size_t)(n - (ns - 1) * S);
    return up(stripe_off((size_t)n, (size_t)S, (size_t)(ns - 1)) + last * (size_t)(n - (ns - 1) * S), 8);
}

namespace {

// The dataset's product kernel and tile: the autotuned pair when there is
// one (for CGH_KERNEL_AUTO), else fp4 MFMA and 256-tiles from n >= 8192.
int ds_kernel(const cgh_dataset* ds, int kernel) {
    return kernel == CGH_KERNEL_AUTO && ds->tuned_kernel ? ds->tuned_kernel : resolve_kernel(kernel);
}
// Words of O and P that can hold a set bit: ceil(D / 32), ceil(L / 32).
int ds_kv1(const cgh_dataset* ds) { return (int)((ds->D + 31) / 32); }
int ds_kv2(const cgh_dataset* ds) { return (int)(((size_t)ds->L + 31) / 32); }
int ds_tile(const cgh_dataset* ds, int kernel) {
    return ds->tuned_tile && ds_kernel(ds, kernel) == ds->tuned_kernel ? ds->tuned_tile : pick_tile(ds->n);
}
// LDS ring depth of the dataset's product launches: the tuned one with the
// tuned kernel, else the default.
int ds_stages(const cgh_dataset* ds, int kernel) {
    return ds->tuned_stages && ds_kernel(ds, kernel) == ds->tuned_kernel ? ds->tuned_stages : NST_DEFAULT;
}

}  // namespace

namespace {
// The entry points that need O on the device refuse a split dataset.
bool split_unsupported(const cgh_dataset* ds, const char* what, char* err, int errlen) {
    if (!ds->hO) return false;
    set_err(err, errlen, "%s: not available for a dataset whose one-hot stays in host memory", what);
    return true;
}
}  // namespace

extern "C" int cgh_dataset_tile(const cgh_dataset* ds, int kernel) { return ds_tile(ds, kernel); }

extern "C" void cgh_dataset_tuning(const cgh_dataset* ds, int* kernel, int* tile, int* stages) {
    if (kernel) *kernel = ds_kernel(ds, CGH_KERNEL_AUTO);
    if (tile) *tile = ds_tile(ds, CGH_KERNEL_AUTO);
    if (stages) *stages = ds_stages(ds, CGH_KERNEL_AUTO);
}

extern "C" int cgh_dataset_autotune(cgh_dataset* ds, int S, int* kernel_out, int* tile_out, double* ms_out,
                                    char* err, int errlen) {
    // Candidates: fp4 and int8 MFMA x the 256 x 128 tile (its 2-slab ring)
    // and the 128-tile with LDS rings of 2, 3 and 4 slabs (the VALU form is
    // never faster), each timed on the first stripe (S rows, the widest)
    // after one warm-up launch; the fastest is kept.  Every candidate's stripe is compared with the first one's: a
    // candidate that disagrees is an error, not a choice.
    const int n = ds->n;
    const int ns = 1;
    const int kernels[2] = {CGH_KERNEL_MFMA_FP4, CGH_KERNEL_MFMA_I8}, tiles[2] = {256, 128}, stages[3] = {2, 3, 4};
    const size_t bytes = up(stripe_off((size_t)n, (size_t)S, (size_t)ns), 8) * 2 + 16;
    uint16_t *buf = NULL, *first = NULL;
    int* ddiff = NULL;
    hipStream_t st = NULL;
    hipEvent_t e0 = NULL, e1 = NULL;
    double best = 1e300;
    int bk = 0, bt = 0, bs = 0, cand = 0;
    if (S < 8 || S % 8) {
        set_err(err, errlen, "stripe height must be a positive multiple of 8 (got %d)", S);
        return 1;
    }
    if (split_unsupported(ds, "cgh_dataset_autotune", err, errlen)) return 1;
    HIP_TRY(hipSetDevice(ds->device));
    HIP_TRY(hipMalloc(&buf, bytes));
    HIP_TRY(hipMalloc(&first, bytes));
    HIP_TRY(hipMalloc(&ddiff, sizeof(int)));
    HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    HIP_TRY(hipMemsetAsync(buf, 0, bytes, st));
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    for (int k : kernels)
        for (int t : tiles)
            for (int g : stages) {
                if (t == 256 && g != 2) continue;  // one ring depth for the 256 x 128 tile
                float ms = 0.f;
                int diff = 0;
                HIP_TRY(launch_tiles(k, t, ds->dO, (int)ds->kw1, ds->dP, (int)ds->kw2, (int)ds->n_pad, n, 0, S, ns, buf,
                                     st, NULL, 0, ds_kv1(ds), ds_kv2(ds), g));
                HIP_TRY(hipEventRecord(e0, st));
                for (int rep = 0; rep < 3; rep++)
                    HIP_TRY(launch_tiles(k, t, ds->dO, (int)ds->kw1, ds->dP, (int)ds->kw2, (int)ds->n_pad, n, 0, S, ns,
                                         buf, st, NULL, 0, ds_kv1(ds), ds_kv2(ds), g));
                HIP_TRY(hipEventRecord(e1, st));
                if (cand++ == 0) {
                    HIP_TRY(hipMemcpyAsync(first, buf, bytes, hipMemcpyDeviceToDevice, st));
                } else {
                    HIP_TRY(hipMemsetAsync(ddiff, 0, sizeof(int), st));
                    hipLaunchKernelGGL(count_diff, dim3(1024), dim3(256), 0, st, (const uint4*)buf,
                                       (const uint4*)first, bytes / 16, ddiff);
                    HIP_TRY(hipGetLastError());
                    HIP_TRY(hipMemcpyAsync(&diff, ddiff, sizeof(int), hipMemcpyDeviceToHost, st));
                }
                HIP_TRY(hipStreamSynchronize(st));
                HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
                if (diff) {
                    set_err(err, errlen, "autotune: kernel %d, tile %d, ring %d disagrees with kernel %d, tile %d, ring %d",
                            k, t, g, kernels[0], tiles[0], stages[0]);
                    goto out;
                }
                if (ms / 3 < best) {
                    best = ms / 3;
                    bk = k;
                    bt = t;
                    bs = g;
                }
            }
    ds->tuned_kernel = bk;
    ds->tuned_tile = bt;
    ds->tuned_stages = bs;
    if (kernel_out) *kernel_out = bk;
    if (tile_out) *tile_out = bt;
    if (ms_out) *ms_out = best;
    cand = 0;
    goto out;
fail:
    cand = -1;
out:
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (ddiff) (void)hipFree(ddiff);
    if (first) (void)hipFree(first);
    if (buf) (void)hipFree(buf);
    return cand == 0 ? 0 : 1;
}

namespace {
// The product of a dataset whose one-hot stays in host memory: chunk c of
// chunk_planes O planes goes up on the dataset's upload stream into buffer
// c % 2 (after pass c - 2 has read that buffer), and pass c multiplies it on
// `st`.  Pass 0 also takes the P planes and stores P P^T - O_0 O_0^T; each
// later pass adds -O_c O_c^T to the stored cells, so every cell only falls
// towards P P^T - O O^T (never below it: no u16 underflow); the cap goes
// with the last pass.  The uploads of chunk c + 1 overlap pass c.
int split_passes(cgh_dataset* ds, int S, uint16_t* d_out, hipStream_t st, int kernel, int tile, char* err,
                 int errlen) {
    if (kernel != CGH_KERNEL_MFMA_FP4 && kernel != CGH_KERNEL_MFMA_I8) {
        set_err(err, errlen, "a one-hot in host memory needs an MFMA product kernel (got %d)", kernel);
        return 1;
    }
    const int n = ds->n, ns = (n + S - 1) / S;
    const size_t plane_bytes = ds->n_pad * 16;
    const size_t valid = ((size_t)ds_kv1(ds) + 3) / 4;                    // planes that can hold a bit
    const size_t cp = ds->chunk_planes, chunks = std::max<size_t>(1, (valid + cp - 1) / cp);
    for (size_t c = 0; c < chunks; c++) {
        const int k = (int)(c % 2);
        const size_t p0 = c * cp, np = std::min(cp, ds->kw1 / 4 - p0);  // whole slabs (kw1 is)
        const int kv1 = (int)std::min(np * 4, (size_t)ds_kv1(ds) - std::min((size_t)ds_kv1(ds), p0 * 4));
        if (c >= 2) HIP_TRY(hipStreamWaitEvent(ds->ostream, ds->used[k], 0));
        HIP_TRY(hipMemcpyAsync(ds->dOc[k], ds->hO + p0 * ds->n_pad * 4, np * plane_bytes, hipMemcpyHostToDevice,
                               ds->ostream));
        HIP_TRY(hipEventRecord(ds->loaded[k], ds->ostream));
        HIP_TRY(hipStreamWaitEvent(st, ds->loaded[k], 0));
        const bool first = c == 0, last = c + 1 == chunks;
        HIP_TRY(launch_tiles(kernel, tile, ds->dOc[k], (int)(np * 4), ds->dP, first ? (int)ds->kw2 : 0,
                             (int)ds->n_pad, n, 0, S, ns, d_out, st, ds->drstart, last ? ds->cap : 0, std::max(kv1, 1),
                             first ? ds_kv2(ds) : 0, ds_stages(ds, kernel), first ? 0 : 1));
        HIP_TRY(hipEventRecord(ds->used[k], st));
    }
    return 0;
fail:
    return 1;
}

}  // namespace

extern "C" int cgh_dataset_phase(cgh_dataset* ds, int S, uint16_t* d_out, void* stream, int kernel, int phases,
                                 char* err, int errlen) {
    hipStream_t st = (hipStream_t)stream;
    const int n = ds->n, tile = ds_tile(ds, kernel);
    if (S < 8 || S % 8) {
        set_err(err, errlen, "stripe height must be a positive multiple of 8 (got %d)", S);
        return 1;
    }
    if (!valid_kernel(kernel)) {
        set_err(err, errlen, "unknown kernel %d", kernel);
        return 1;
    }
    if (((uintptr_t)d_out) & 3) {
        set_err(err, errlen, "output buffer must be 4-byte aligned");
        return 1;
    }
    HIP_TRY(hipSetDevice(ds->device));
    if ((phases & CGH_PHASE_PRODUCT) && ds->hO) {
        if (split_passes(ds, S, d_out, st, ds_kernel(ds, kernel), tile, err, errlen)) return 1;
    } else if (phases & CGH_PHASE_PRODUCT) {
        HIP_TRY(launch_tiles(ds_kernel(ds, kernel), tile, ds->dO, (int)ds->kw1, ds->dP, (int)ds->kw2, (int)ds->n_pad, n, 0,
                             S, (n + S - 1) / S, d_out, st, ds->drstart, ds->cap, ds_kv1(ds), ds_kv2(ds),
                             ds_stages(ds, kernel)));
    }
    if ((phases & CGH_PHASE_RARE) && ds->n_rare) {
        RareJob jb = rare_job(d_out, n, 0, S, ds->drstart, ds->drentry, ds->dgmem, ds->m16, ds->cap,
                              lanes_per_range(ds->n_rare, ds->n_rentries));
        const long long items = rare_layout(jb, n);
        HIP_TRY(launch_rare(jb, items, st));
    }
    return 0;
fail:
    return 1;
}

extern "C" int cgh_dataset_upper(cgh_dataset* ds, int S, uint16_t* d_out, void* stream, int kernel, char* err,
                                 int errlen) {
    return cgh_dataset_phase(ds, S, d_out, stream, kernel, CGH_PHASE_PRODUCT | CGH_PHASE_RARE, err, errlen);
}

// ---------------------------------------------------------------------------
// Many independent datasets in one persistent launch
// ---------------------------------------------------------------------------

namespace {
// Tile edge of a dataset inside a batch: 256 from 4,096 rows (one persistent
// queue keeps the chip busy with fewer, larger tiles), else 128.
constexpr int BATCH_TILE_ROWS = 4096;
int batch_tile(int n) { return n >= BATCH_TILE_ROWS ? 256 : 128; }
}  // namespace

struct cgh_batch {
    int device, count, S;
    int cus = 0;                        // compute units: persistent grid size
    size_t total;                       // u16 elements of all outputs
    std::vector<size_t> offs;           // output offset of each dataset (16-byte aligned)
    std::vector<int> tile;              // tile edge chosen per dataset
    std::vector<const cgh_dataset*> sets;  // borrowed
    int fixed_tile = 0;                 // cgh_batch_create's tile (0 = by the threshold)
    int threshold = BATCH_TILE_ROWS;    // 256-tiles for datasets from this many rows
    int tuned_kernel = 0;               // cgh_batch_autotune's kernel for CGH_KERNEL_AUTO (0 = fp4)
    size_t tile_cap[2] = {0, 0};        // capacity of the device work lists
    TileJob* djobs = NULL;              // per dataset (out filled in-kernel)
    size_t* doffs = NULL;
    int4* dtiles[2] = {NULL, NULL};     // work lists of the 256 and 128 tile kernels
    int ntiles[2] = {0, 0};
    std::vector<RareJob> rare;          // one rare-pass launch per dataset with rare pairs
    std::vector<int> rare_ds;           // its dataset
    std::vector<long long> rare_items;  // its work items
    int* dcount = NULL;                 // two work-queue counters
    hipGraphExec_t exec = NULL;         // captured launch sequence
    uint16_t* graph_out = NULL;
    hipStream_t graph_stream = NULL;
    int graph_kernel = -1;
    // rare passes of different datasets run side by side on RSTREAMS streams
    // (the caller's plus side ones), balanced by work items
    static constexpr int RSTREAMS = 4;
    hipStream_t side[RSTREAMS - 1] = {};
    hipEvent_t fork = NULL, join[RSTREAMS - 1] = {};
    std::vector<int> rare_lane;  // stream of each rare job (0 = the caller's)
};

extern "C" void cgh_batch_free(cgh_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->djobs) (void)hipFree(b->djobs);
    if (b->doffs) (void)hipFree(b->doffs);
    for (int k = 0; k < 2; k++)
        if (b->dtiles[k]) (void)hipFree(b->dtiles[k]);
    if (b->dcount) (void)hipFree(b->dcount);
    for (int k = 0; k < cgh_batch::RSTREAMS - 1; k++) {
        if (b->side[k]) (void)hipStreamDestroy(b->side[k]);
        if (b->join[k]) (void)hipEventDestroy(b->join[k]);
    }
    if (b->fork) (void)hipEventDestroy(b->fork);
    delete b;
}

namespace {

// The two device work lists of (dataset, stripe, I, J) tiles for the current
// tile rule, longest first (below-diagonal zero tiles last); a captured graph
// refers to the old lists, so it is dropped.
hipError_t build_queues(cgh_batch* b) {
    std::vector<std::pair<long long, int4>> lists[2];  // (cost, tile) per kernel
    for (int d = 0; d < b->count; d++) {
        const cgh_dataset* ds = b->sets[(size_t)d];
        const int tl = b->fixed_tile ? b->fixed_tile : (ds->n >= b->threshold ? 256 : 128);
        b->tile[(size_t)d] = tl;
        const int k = tl == 256 ? 0 : 1;
        const long long cost = (long long)(ds->kw1 + ds->kw2);
        for (int s = 0, r0 = 0; r0 < ds->n; s++, r0 += b->S) {
            const int rows = std::min(b->S, ds->n - r0), w = ds->n - r0;
            for (int I = 0; I * tl < rows; I++)
                for (int J = 0; J * tl < w; J++) lists[k].push_back({J >= I ? cost : 0, make_int4(d, s, I, J)});
        }
    }
    if (b->exec) {
        const hipError_t e = hipGraphExecDestroy(b->exec);
        b->exec = NULL;
        if (e != hipSuccess) return e;
    }
    for (int k = 0; k < 2; k++) {
        std::stable_sort(lists[k].begin(), lists[k].end(),
                         [](const std::pair<long long, int4>& x, const std::pair<long long, int4>& y) {
                             return x.first > y.first;
                         });
        std::vector<int4> t;
        t.reserve(lists[k].size());
        for (auto& e : lists[k]) t.push_back(e.second);
        b->ntiles[k] = (int)t.size();
        if (t.empty()) continue;
        hipError_t e = hipSuccess;
        if (t.size() > b->tile_cap[k]) {
            if (b->dtiles[k]) e = hipFree(b->dtiles[k]);
            b->dtiles[k] = NULL;
            if (!e) e = hipMalloc(&b->dtiles[k], sizeof(int4) * t.size());
            if (e) return e;
            b->tile_cap[k] = t.size();
        }
        if ((e = hipMemcpy(b->dtiles[k], t.data(), sizeof(int4) * t.size(), hipMemcpyHostToDevice))) return e;
    }
    return hipSuccess;
}

}  // namespace

extern "C" int cgh_batch_create(int device, cgh_dataset* const* sets, int count, int S, int tile, cgh_batch** out,
                                char* err, int errlen) {
    cgh_batch* b = new cgh_batch();
    std::vector<TileJob> jobs((size_t)count);
    *out = NULL;
    b->device = device;
    b->count = count;
    b->S = S;
    b->total = 0;
    b->fixed_tile = tile;
    if (count < 1 || S < 8 || S % 8 || (tile != 0 && tile != 128 && tile != 256)) {
        set_err(err, errlen, "cgh_batch_create: bad arguments (count=%d S=%d tile=%d)", count, S, tile);
        delete b;
        return 1;
    }
    for (int d = 0; d < count; d++) {
        const cgh_dataset* ds = sets[d];
        if (split_unsupported(ds, "cgh_batch_create", err, errlen)) {
            delete b;
            return 1;
        }
        if (ds->device != device) {
            set_err(err, errlen, "dataset %d lives on device %d, not %d", d, ds->device, device);
            delete b;
            return 1;
        }
        b->sets.push_back(ds);
        b->tile.push_back(0);
        b->offs.push_back(b->total);
        b->total += cgh_upper_elems(ds->n, S);
        jobs[d] = {ds->dO, ds->dP, (int)ds->kw1, (int)ds->kw2, ds->n, 0, S, NULL, ds->dO, ds->dP, ds->n, 0,
                   (int)ds->n_pad, (int)ds->n_pad, ds->drstart, ds->cap, ds_kv1(ds), ds_kv2(ds)};
        if (ds->n_rare && ds->n > 1) {
            RareJob r = rare_job(NULL, ds->n, 0, S, ds->drstart, ds->drentry, ds->dgmem, ds->m16, ds->cap,
                                 lanes_per_range(ds->n_rare, ds->n_rentries));
            b->rare_items.push_back(rare_layout(r, ds->n));
            b->rare.push_back(r);
            b->rare_ds.push_back(d);
        }
    }
    HIP_TRY(hipSetDevice(device));
    HIP_TRY(hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, device));
    HIP_TRY(hipMalloc(&b->djobs, sizeof(TileJob) * (size_t)count));
    HIP_TRY(hipMemcpy(b->djobs, jobs.data(), sizeof(TileJob) * (size_t)count, hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&b->doffs, sizeof(size_t) * (size_t)count));
    HIP_TRY(hipMemcpy(b->doffs, b->offs.data(), sizeof(size_t) * (size_t)count, hipMemcpyHostToDevice));
    HIP_TRY(build_queues(b));
    HIP_TRY(hipMalloc(&b->dcount, 2 * sizeof(int)));
    if (!b->rare.empty()) {  // largest first, each to the least loaded stream
        std::vector<size_t> order(b->rare.size());
        for (size_t k = 0; k < order.size(); k++) order[k] = k;
        std::sort(order.begin(), order.end(), [&](size_t x, size_t y) { return b->rare_items[x] > b->rare_items[y]; });
        long long load[cgh_batch::RSTREAMS] = {};
        b->rare_lane.assign(b->rare.size(), 0);
        for (size_t k : order) {
            int best = 0;
            for (int l = 1; l < cgh_batch::RSTREAMS; l++)
                if (load[l] < load[best]) best = l;
            b->rare_lane[k] = best;
            load[best] += b->rare_items[k];
        }
        for (int k = 0; k < cgh_batch::RSTREAMS - 1; k++) {
            HIP_TRY(hipStreamCreateWithFlags(&b->side[k], hipStreamNonBlocking));
            HIP_TRY(hipEventCreateWithFlags(&b->join[k], hipEventDisableTiming));
        }
        HIP_TRY(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    }
    *out = b;
    return 0;
fail:
    cgh_batch_free(b);
    return 1;
}

extern "C" size_t cgh_batch_out_elems(const cgh_batch* b) { return b->total; }
extern "C" size_t cgh_batch_out_offset(const cgh_batch* b, int d) { return b->offs[(size_t)d]; }
extern "C" int cgh_batch_tile(const cgh_batch* b, int d) { return b->tile[(size_t)d]; }

namespace {

// The whole batch: counter resets, the persistent product kernel per tile
// size, then one rare-pair launch over every dataset's rows.
hipError_t batch_enqueue(cgh_batch* b, uint16_t* d_out, hipStream_t st, int kernel) {
    if (kernel == CGH_KERNEL_AUTO && b->tuned_kernel) kernel = b->tuned_kernel;
    const bool fp4 = resolve_kernel(kernel) == CGH_KERNEL_MFMA_FP4;
    const unsigned cus = (unsigned)b->cus;
    for (int k = 0; k < 2; k++) hipLaunchKernelGGL(reset_counter, dim3(1), dim3(1), 0, st, b->dcount + k);
    const bool valu = resolve_kernel(kernel) == CGH_KERNEL_VALU_POPC;
    if (b->ntiles[0]) {
        if (valu)
            hipLaunchKernelGGL((valu_tile_queue<256>), dim3(cus * 2), dim3(256), 0, st, b->djobs, d_out, b->doffs,
                               b->dtiles[0], b->ntiles[0], b->dcount);
        else if (fp4)
            hipLaunchKernelGGL((mfma_tile_queue<256, 128, 64, true>), dim3(cus), dim3(512), 0, st, b->djobs, d_out,
                               b->doffs, b->dtiles[0], b->ntiles[0], b->dcount);
        else
            hipLaunchKernelGGL((mfma_tile_queue<256, 128, 64, false>), dim3(cus), dim3(512), 0, st, b->djobs, d_out,
                               b->doffs, b->dtiles[0], b->ntiles[0], b->dcount);
    }
    if (b->ntiles[1]) {
        if (valu)
            hipLaunchKernelGGL((valu_tile_queue<128>), dim3(cus * 2), dim3(256), 0, st, b->djobs, d_out, b->doffs,
                               b->dtiles[1], b->ntiles[1], b->dcount + 1);
        else if (fp4)
            hipLaunchKernelGGL((mfma_tile_queue<128, 64, 64, true>), dim3(cus * 2), dim3(256), 0, st, b->djobs,
                               d_out, b->doffs, b->dtiles[1], b->ntiles[1], b->dcount + 1);
        else
            hipLaunchKernelGGL((mfma_tile_queue<128, 64, 64, false>), dim3(cus * 2), dim3(256), 0, st, b->djobs,
                               d_out, b->doffs, b->dtiles[1], b->ntiles[1], b->dcount + 1);
    }
    // rare pairs: one persistent launch per dataset, after every product,
    // spread over the caller's stream and the side streams (fork / join by
    // events, which a stream capture follows into the graph)
    if (b->rare.empty()) return hipGetLastError();
    hipError_t e = hipEventRecord(b->fork, st);
    for (int k = 0; k < cgh_batch::RSTREAMS - 1 && e == hipSuccess; k++) e = hipStreamWaitEvent(b->side[k], b->fork, 0);
    if (e != hipSuccess) return e;
    for (size_t k = 0; k < b->rare.size(); k++) {
        RareJob jb = b->rare[k];
        jb.out = d_out + b->offs[b->rare_ds[k]];
        const hipStream_t sk = b->rare_lane[k] ? b->side[b->rare_lane[k] - 1] : st;
        if ((e = launch_rare(jb, b->rare_items[k], sk)) != hipSuccess) return e;
    }
    for (int k = 0; k < cgh_batch::RSTREAMS - 1 && e == hipSuccess; k++) {
        e = hipEventRecord(b->join[k], b->side[k]);
        if (e == hipSuccess) e = hipStreamWaitEvent(st, b->join[k], 0);
    }
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

}  // namespace

extern "C" int cgh_batch_run(cgh_batch* b, uint16_t* d_out, void* stream, int kernel, int use_graph, char* err,
                             int errlen) {
    hipStream_t st = (hipStream_t)stream;
    hipGraph_t graph = NULL;
    if (!valid_kernel(kernel)) {
        set_err(err, errlen, "unknown kernel %d", kernel);
        return 1;
    }
    if (((uintptr_t)d_out) & 15) {
        set_err(err, errlen, "output buffer must be 16-byte aligned");
        return 1;
    }
    HIP_TRY(hipSetDevice(b->device));
    if (!use_graph || st == NULL) {
        HIP_TRY(batch_enqueue(b, d_out, st, kernel));
        return 0;
    }
    if (!b->exec || b->graph_out != d_out || b->graph_stream != st || b->graph_kernel != kernel) {
        if (b->exec) {
            HIP_TRY(hipGraphExecDestroy(b->exec));
            b->exec = NULL;
        }
        HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        {
            const hipError_t e = batch_enqueue(b, d_out, st, kernel);
            const hipError_t e2 = hipStreamEndCapture(st, &graph);
            HIP_TRY(e);
            HIP_TRY(e2);
        }
        HIP_TRY(hipGraphInstantiate(&b->exec, graph, NULL, NULL, 0));
        HIP_TRY(hipGraphDestroy(graph));
        graph = NULL;
        b->graph_out = d_out;
        b->graph_stream = st;
        b->graph_kernel = kernel;
    }
    HIP_TRY(hipGraphLaunch(b->exec, st));
    return 0;
fail:
    if (graph) (void)hipGraphDestroy(graph);
    return 1;
}

// Batch autotuner: the whole batch timed (eagerly, on `stream`, into d_out)
// under each candidate tile rule (256-tiles for every dataset, from 2,048,
// 4,096, 8,192 or 16,384 rows, or none) and product kernel (fp4, int8 MFMA);
// the fastest rule and kernel are kept for later runs with CGH_KERNEL_AUTO.
// A batch's best tile per dataset depends on what else shares the queue, so
// the batch is tuned as a whole rather than from each dataset's own timing.
extern "C" int cgh_batch_autotune(cgh_batch* b, uint16_t* d_out, void* stream, int* threshold_out, int* kernel_out,
                                  double* ms_out, char* err, int errlen) {
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t e0 = NULL, e1 = NULL;
    const int thresholds[6] = {0, 2048, 4096, 8192, 16384, INT32_MAX};
    const int kernels[2] = {CGH_KERNEL_MFMA_FP4, CGH_KERNEL_MFMA_I8};
    const size_t bytes = b->total * 2;
    uint16_t* first = NULL;
    int* ddiff = NULL;
    double best = 1e300;
    int bth = b->threshold, bk = CGH_KERNEL_MFMA_FP4, cand = 0;
    const int fixed = b->fixed_tile, threshold = b->threshold;
    if (((uintptr_t)d_out) & 15) {
        set_err(err, errlen, "output buffer must be 16-byte aligned");
        return 1;
    }
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(hipStreamSynchronize(st));  // earlier runs may still read the work lists
    HIP_TRY(hipMalloc(&first, bytes));
    HIP_TRY(hipMalloc(&ddiff, sizeof(int)));
    HIP_TRY(hipMemsetAsync(d_out, 0, bytes, st));
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    b->fixed_tile = 0;
    for (int th : thresholds) {
        b->threshold = th;
        HIP_TRY(build_queues(b));
        for (int k : kernels) {
            float ms = 0.f;
            int diff = 0;
            HIP_TRY(batch_enqueue(b, d_out, st, k));
            HIP_TRY(hipEventRecord(e0, st));
            for (int rep = 0; rep < 2; rep++) HIP_TRY(batch_enqueue(b, d_out, st, k));
            HIP_TRY(hipEventRecord(e1, st));
            if (cand++ == 0) {  // every candidate must write what the first one wrote
                HIP_TRY(hipMemcpyAsync(first, d_out, bytes, hipMemcpyDeviceToDevice, st));
            } else {
                HIP_TRY(hipMemsetAsync(ddiff, 0, sizeof(int), st));
                hipLaunchKernelGGL(count_diff, dim3(1024), dim3(256), 0, st, (const uint4*)d_out, (const uint4*)first,
                                   bytes / 16, ddiff);
                HIP_TRY(hipGetLastError());
                HIP_TRY(hipMemcpyAsync(&diff, ddiff, sizeof(int), hipMemcpyDeviceToHost, st));
            }
            HIP_TRY(hipStreamSynchronize(st));
            HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
            if (diff) {
                set_err(err, errlen, "autotune: threshold %d, kernel %d disagrees with threshold %d, kernel %d", th, k,
                        thresholds[0], kernels[0]);
                b->threshold = threshold;
                b->fixed_tile = fixed;
                (void)build_queues(b);
                goto out;
            }
            if (ms / 2 < best) {
                best = ms / 2;
                bth = th;
                bk = k;
            }
        }
    }
    b->threshold = bth;
    b->tuned_kernel = bk;
    HIP_TRY(build_queues(b));
    if (threshold_out) *threshold_out = bth;
    if (kernel_out) *kernel_out = bk;
    if (ms_out) *ms_out = best;
    cand = 0;
    goto out;
fail:
    b->fixed_tile = fixed;
    cand = -1;
out:
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (ddiff) (void)hipFree(ddiff);
    if (first) (void)hipFree(first);
    return cand == 0 ? 0 : 1;
}

// ---------------------------------------------------------------------------
// The batch without allocation (for callers that cannot allocate)
// ---------------------------------------------------------------------------

namespace {

// What the tile generator and the queue kernels read, at the head of the
// scratch array.  Per queue (0: 256-tiles, 1: 128-tiles) the datasets are
// listed longest K first; list entry p owns tiles first[p] .. first[p+1]-1.
struct NoAllocTable {
    int count, S;
    int ntiles[2];
    int lists[2];
    int order[2][CGH_NOALLOC_MAX_SETS];
    long long first[2][CGH_NOALLOC_MAX_SETS + 1];
    int tile[CGH_NOALLOC_MAX_SETS];
    size_t outoff[CGH_NOALLOC_MAX_SETS];
    TileJob jobs[CGH_NOALLOC_MAX_SETS];  // last: only `count` entries are stored
};

struct NoAllocScratch {
    int next[64];  // queue counters (two used), padded to 256 bytes
    NoAllocTable table;
    int4 tiles[CGH_NOALLOC_MAX_TILES];  // the 256-tile queue, then the 128-tile queue
};
static_assert(sizeof(NoAllocScratch) <= CGH_NOALLOC_SCRATCH_BYTES, "scratch size in the header");

}  // namespace

// Device scratch of cgh_batch_run_noalloc: one copy per GPU, held until the
// process ends.  Host code never uses this name as an address: the device's
// own address of its copy comes from hipGetSymbolAddress (noalloc_scratch).
static __device__ __align__(256) unsigned char scratch[CGH_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
NoAllocScratch* scratch_at[MAX_DEVICES] = {};

// The current device's address of `scratch`, looked up once per device.
hipError_t noalloc_scratch(int device, NoAllocScratch** out) {
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
    if (!scratch_at[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        scratch_at[device] = static_cast<NoAllocScratch*>(p);
    }
    *out = scratch_at[device];
    return hipSuccess;
}

// Bytes carried into device memory by value as a kernel argument (at most
// 4 KiB), so the table reaches the scratch in stream order with no host
// staging buffer and no copy engine.
constexpr unsigned PIECE = 2048;
struct TablePiece {
    unsigned offset, bytes;
    unsigned char data[PIECE];
};
__global__ void __launch_bounds__(256) store_piece(unsigned char* dst, TablePiece p) {
    for (unsigned k = threadIdx.x; k < p.bytes; k += 256) dst[p.offset + k] = p.data[k];
}

__global__ void zero_counters(int* next) {
    if (threadIdx.x < 2) next[threadIdx.x] = 0;
}

// Tile t of queue `list` -> (dataset, stripe, I, J): the dataset from the
// list's prefix sums, then the closed-form tile numbering of a one-dataset
// launch (stripe by stripe, row-major; S is a multiple of the tile).
__global__ void __launch_bounds__(256) noalloc_tiles(const NoAllocTable* __restrict__ tab, int list,
                                                     int4* __restrict__ out) {
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t >= tab->ntiles[list]) return;
    int lo = 0, hi = tab->lists[list] - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tab->first[list][mid] <= t) lo = mid; else hi = mid - 1;
    }
    const int d = tab->order[list][lo];
    const int tl = tab->tile[d], n = tab->jobs[d].n, S = tab->S;
    const long long local = t - tab->first[list][lo];
    const long long q = S / tl, nbc = (n + tl - 1) / tl;
    int slo = 0, shi = (n + S - 1) / S - 1;
    while (slo < shi) {
        const int mid = (slo + shi + 1) >> 1;
        if (tiles_before(mid, q, nbc) <= local) slo = mid; else shi = mid - 1;
    }
    const long long rem = local - tiles_before(slo, q, nbc), tn = nbc - slo * q;
    out[t] = make_int4(d, slo, (int)(rem / tn), (int)(rem % tn));
}

// Tiles of one dataset's upper-triangle layout in the closed-form numbering.
long long layout_tiles(int n, int S, int tl) {
    const int ns = (n + S - 1) / S;
    const long long q = S / tl, nbc = (n + tl - 1) / tl;
    const long long tm_last = (std::min(S, n - (ns - 1) * S) + tl - 1) / tl;
    return tiles_before(ns - 1, q, nbc) + tm_last * (nbc - (long long)(ns - 1) * q);
}

}  // namespace

extern "C" int cgh_batch_run_noalloc(cgh_dataset* const* sets, int count, int S, uint16_t* d_out, void* stream,
                                     int kernel, char* err, int errlen) {
    hipStream_t st = (hipStream_t)stream;
    NoAllocScratch* scr = nullptr;
    NoAllocTable tab{};  // host image (on the stack), stored into the scratch by kernel arguments
    if (count < 1 || count > CGH_NOALLOC_MAX_SETS) {
        set_err(err, errlen, "cgh_batch_run_noalloc: %d datasets (1..%d accepted)", count, CGH_NOALLOC_MAX_SETS);
        return 1;
    }
    if (S < 256 || S % 256) {
        set_err(err, errlen, "cgh_batch_run_noalloc: stripe height must be a positive multiple of 256 (got %d)", S);
        return 1;
    }
    if (!valid_kernel(kernel)) {
        set_err(err, errlen, "unknown kernel %d", kernel);
        return 1;
    }
    for (int d = 0; d < count; d++)
        if (split_unsupported(sets[d], "cgh_batch_run_noalloc", err, errlen)) return 1;
    if (((uintptr_t)d_out) & 15) {
        set_err(err, errlen, "output buffer must be 16-byte aligned");
        return 1;
    }
    const int device = sets[0]->device;
    tab.count = count;
    tab.S = S;
    {
        size_t off = 0;
        long long cost[CGH_NOALLOC_MAX_SETS];
        for (int d = 0; d < count; d++) {
            const cgh_dataset* ds = sets[d];
            if (ds->device != device) {
                set_err(err, errlen, "dataset %d lives on device %d, not %d", d, ds->device, device);
                return 1;
            }
            const int tl = batch_tile(ds->n), k = tl == 256 ? 0 : 1;
            tab.tile[d] = tl;
            tab.outoff[d] = off;
            off += cgh_upper_elems(ds->n, S);
            tab.jobs[d] = {ds->dO, ds->dP, (int)ds->kw1, (int)ds->kw2, ds->n, 0, S, NULL, ds->dO, ds->dP, ds->n, 0,
                           (int)ds->n_pad, (int)ds->n_pad, ds->drstart, ds->cap, ds_kv1(ds), ds_kv2(ds)};
            cost[d] = (long long)(ds->kw1 + ds->kw2);
            // insert into queue k's list, longest K first
            int p = tab.lists[k]++;
            while (p > 0 && cost[tab.order[k][p - 1]] < cost[d]) {
                tab.order[k][p] = tab.order[k][p - 1];
                p--;
            }
            tab.order[k][p] = d;
        }
        long long total = 0;
        for (int k = 0; k < 2; k++) {
            tab.first[k][0] = 0;
            for (int p = 0; p < tab.lists[k]; p++) {
                const int d = tab.order[k][p];
                tab.first[k][p + 1] = tab.first[k][p] + layout_tiles(sets[d]->n, S, tab.tile[d]);
            }
            total += tab.first[k][tab.lists[k]];
        }
        if (total > CGH_NOALLOC_MAX_TILES) {
            set_err(err, errlen, "cgh_batch_run_noalloc: %lld tiles (at most %d)", total, CGH_NOALLOC_MAX_TILES);
            return 1;
        }
        tab.ntiles[0] = (int)tab.first[0][tab.lists[0]];
        tab.ntiles[1] = (int)tab.first[1][tab.lists[1]];
    }
    HIP_TRY(hipSetDevice(device));
    HIP_TRY(noalloc_scratch(device, &scr));
    {
        // the table up to its last used job, then the counters, the tile
        // lists and the launches of cgh_batch_run (rare passes on `stream`)
        const size_t bytes = offsetof(NoAllocTable, jobs) + (size_t)count * sizeof(TileJob);
        unsigned char* dst = reinterpret_cast<unsigned char*>(&scr->table);
        for (size_t o = 0; o < bytes; o += PIECE) {
            TablePiece piece;
            piece.offset = (unsigned)o;
            piece.bytes = (unsigned)std::min<size_t>(PIECE, bytes - o);
            memcpy(piece.data, reinterpret_cast<const unsigned char*>(&tab) + o, piece.bytes);
            hipLaunchKernelGGL(store_piece, dim3(1), dim3(256), 0, st, dst, piece);
        }
        hipLaunchKernelGGL(zero_counters, dim3(1), dim3(64), 0, st, scr->next);
        int4* lists[2] = {scr->tiles, scr->tiles + tab.ntiles[0]};
        for (int k = 0; k < 2; k++)
            if (tab.ntiles[k])
                hipLaunchKernelGGL(noalloc_tiles, dim3((unsigned)((tab.ntiles[k] + 255) / 256)), dim3(256), 0, st,
                                   &scr->table, k, lists[k]);
        HIP_TRY(hipGetLastError());
        const int kern = resolve_kernel(kernel);
        const unsigned cus = (unsigned)rare_grid(1 << 30);
        const TileJob* jobs = scr->table.jobs;
        const size_t* outoff = scr->table.outoff;
        if (tab.ntiles[0]) {
            if (kern == CGH_KERNEL_VALU_POPC)
                hipLaunchKernelGGL((valu_tile_queue<256>), dim3(cus * 2), dim3(256), 0, st, jobs, d_out, outoff, lists[0],
                                   tab.ntiles[0], scr->next);
            else if (kern == CGH_KERNEL_MFMA_FP4)
                hipLaunchKernelGGL((mfma_tile_queue<256, 128, 64, true>), dim3(cus), dim3(512), 0, st, jobs, d_out,
                                   outoff, lists[0], tab.ntiles[0], scr->next);
            else
                hipLaunchKernelGGL((mfma_tile_queue<256, 128, 64, false>), dim3(cus), dim3(512), 0, st, jobs, d_out,
                                   outoff, lists[0], tab.ntiles[0], scr->next);
        }
        if (tab.ntiles[1]) {
            if (kern == CGH_KERNEL_VALU_POPC)
                hipLaunchKernelGGL((valu_tile_queue<128>), dim3(cus * 2), dim3(256), 0, st, jobs, d_out, outoff,
                                   lists[1], tab.ntiles[1], scr->next + 1);
            else if (kern == CGH_KERNEL_MFMA_FP4)
                hipLaunchKernelGGL((mfma_tile_queue<128, 64, 64, true>), dim3(cus * 2), dim3(256), 0, st, jobs, d_out,
                                   outoff, lists[1], tab.ntiles[1], scr->next + 1);
            else
                hipLaunchKernelGGL((mfma_tile_queue<128, 64, 64, false>), dim3(cus * 2), dim3(256), 0, st, jobs, d_out,
                                   outoff, lists[1], tab.ntiles[1], scr->next + 1);
        }
        HIP_TRY(hipGetLastError());
        for (int d = 0; d < count; d++) {
            const cgh_dataset* ds = sets[d];
            if (!ds->n_rare || ds->n < 2) continue;
            RareJob jb = rare_job(d_out + tab.outoff[d], ds->n, 0, S, ds->drstart, ds->drentry, ds->dgmem, ds->m16,
                                  ds->cap, lanes_per_range(ds->n_rare, ds->n_rentries));
            HIP_TRY(launch_rare(jb, rare_layout(jb, ds->n), st));
        }
    }
    return 0;
fail:
    return 1;
}

// ---------------------------------------------------------------------------
// Resident database and a continuous-batching query server
// ---------------------------------------------------------------------------

namespace {

constexpr int TOPK = 10;

// k nearest database rows of each query row of dist (rows x nB u16),
// ascending (distance, row index), as 64-bit keys dist << 32 | row.  Two
// levels so that every query row is spread over many workgroups: a part
// kernel reduces each TKC-column chunk of a row to its TOPK smallest keys,
// a final kernel reduces a row's chunk lists.  Inside a 256-thread
// workgroup every thread keeps a sorted list of its best TOPK, the lists
// meet in LDS and wavefront 0 draws the TOPK overall minima one by one.
constexpr int TKC = 8192;  // columns per part workgroup (32 per thread)

__device__ __forceinline__ void keep_best(uint64_t (&best)[TOPK], uint64_t key) {
    if (key < best[TOPK - 1]) {
        int k = TOPK - 1;
        while (k > 0 && best[k - 1] > key) {
            best[k] = best[k - 1];
            k--;
        }
        best[k] = key;
    }
}

// The TOPK smallest of the 256 threads' sorted lists -> out[0..TOPK) (thread 0
// writes; keys are unique, so a head advances only in the lane that held the
// minimum).
__device__ __forceinline__ void block_best(const uint64_t (&best)[TOPK], uint64_t* cand, uint64_t* out) {
    const int tid = threadIdx.x;
#pragma unroll
    for (int k = 0; k < TOPK; k++) cand[tid * TOPK + k] = best[k];
    __syncthreads();
    if (tid >= 64) return;
    int head[4] = {0, 0, 0, 0};  // lane l merges the lists of threads 4l .. 4l+3
    for (int o = 0; o < TOPK; o++) {
        uint64_t mine = ~0ull;
        int which = -1;
        for (int r = 0; r < 4; r++) {
            const int t = tid * 4 + r;
            if (head[r] < TOPK && cand[t * TOPK + head[r]] < mine) {
                mine = cand[t * TOPK + head[r]];
                which = r;
            }
        }
        uint64_t m = mine;
        for (int off = 32; off > 0; off >>= 1) {
            const uint64_t x = __shfl_xor(m, off, 64);
            m = x < m ? x : m;
        }
        if (which >= 0 && mine == m) head[which]++;
        if (tid == 0) out[o] = m;
    }
}

// Workgroup b: row b / chunks, columns [TKC * (b % chunks), +TKC).
__global__ void __launch_bounds__(256)
top_k_part(const uint16_t* __restrict__ dist, int nB, int chunks, uint64_t* __restrict__ part) {
    __shared__ uint64_t cand[256 * TOPK];
    const int row = blockIdx.x / chunks, ch = blockIdx.x % chunks;
    const uint16_t* d = dist + (size_t)row * nB;
    const int c1 = min(nB, (ch + 1) * TKC);
    uint64_t best[TOPK];
#pragma unroll
    for (int k = 0; k < TOPK; k++) best[k] = ~0ull;
    for (int j = ch * TKC + (int)threadIdx.x; j < c1; j += 256) keep_best(best, ((uint64_t)d[j] << 32) | (uint32_t)j);
    block_best(best, cand, part + (size_t)blockIdx.x * TOPK);
}

// Workgroup q: the chunks * TOPK keys of row q -> its TOPK nearest rows.
__global__ void __launch_bounds__(256)
top_k_final(const uint64_t* __restrict__ part, int chunks, uint16_t* __restrict__ top_d,
            uint32_t* __restrict__ top_i) {
    __shared__ uint64_t cand[256 * TOPK];
    __shared__ uint64_t res[TOPK];
    const int q = blockIdx.x;
    const uint64_t* p = part + (size_t)q * chunks * TOPK;
    uint64_t best[TOPK];
#pragma unroll
    for (int k = 0; k < TOPK; k++) best[k] = ~0ull;
    for (int j = threadIdx.x; j < chunks * TOPK; j += 256) keep_best(best, p[j]);
    block_best(best, cand, res);
    __syncthreads();
    if (threadIdx.x < TOPK) {
        const uint64_t m = res[threadIdx.x];
        top_d[(size_t)q * TOPK + threadIdx.x] = m == ~0ull ? 0xffff : (uint16_t)(m >> 32);
        top_i[(size_t)q * TOPK + threadIdx.x] = m == ~0ull ? 0xffffffffu : (uint32_t)m;
    }
}

}  // namespace

struct cgh_db {
    int device = 0, n = 0, L = 0, tau = 0;
    size_t kw1 = 0, kw2 = 0, D = 0;
    cgh_codec* codec = NULL;    // host dictionaries: query ids -> database columns and rare lists
    cgh_dataset* ds = NULL;     // O, P of the database on the device
    void* dqmem = NULL;         // the codec's member array on the device (u16 when m16)
    int m16 = 0;
};

extern "C" void cgh_db_free(cgh_db* db) {
    if (!db) return;
    (void)hipSetDevice(db->device);
    if (db->ds) cgh_dataset_free(db->ds);
    if (db->dqmem) (void)hipFree(db->dqmem);
    if (db->codec) cgh_codec_free(db->codec);
    delete db;
}

extern "C" int cgh_db_create(int device, const uint64_t* calls, int n, int L, int tau, cgh_db** out, char* err,
                             int errlen) {
    cgh_db* db = new cgh_db();
    cgh_encoding enc;
    size_t nmem = 0;
    memset(&enc, 0, sizeof enc);
    *out = NULL;
    db->device = device;
    if (cgh_codec_create(calls, n, L, tau, &enc, &db->codec, err, errlen)) goto fail;
    cgh_codec_info(db->codec, &db->n, &db->L, &db->tau, &db->kw1, &db->kw2, &db->D, &nmem);
    if (cgh_dataset_upload(device, &enc, &db->ds, err, errlen)) goto fail;
    cgh_encoding_free(&enc);
    HIP_TRY(hipSetDevice(device));
    db->m16 = db->n <= 65536;
    HIP_TRY(upload_members(cgh_codec_members(db->codec), nmem, db->m16, &db->dqmem));
    *out = db;
    return 0;
fail:
    cgh_encoding_free(&enc);
    cgh_db_free(db);
    return 1;
}

extern "C" int cgh_db_info(const cgh_db* db, int* n, int* L, int* tau, size_t* D) {
    if (n) *n = db->n;
    if (L) *L = db->L;
    if (tau) *tau = db->tau;
    if (D) *D = db->D;
    return 0;
}

struct cgh_server {
    struct Request {
        long id;
        int q;
        bool want_dist;
        std::vector<uint64_t> calls;
        double t_submit = 0, t_done = 0;
        std::vector<uint16_t> dist, top_d;
        std::vector<uint32_t> top_i;
        bool done = false;
        std::string err;
    };
    struct Slot {
        hipStream_t st = NULL;
        hipEvent_t ev = NULL;
        bool busy = false;
        std::vector<Request*> reqs;
        int q = 0;
        size_t bytes = 0;
        uint32_t *dO = NULL, *dP = NULL, *drs = NULL;  // dO, dP: query planes
        uint32_t *dOr = NULL, *dPr = NULL;               // the same rows as encoded (row-major)
        uint64_t* dre = NULL;
        uint16_t *ddist = NULL, *dtopd = NULL;
        uint64_t* dpart = NULL;  // per-chunk nearest-row keys
        uint32_t* dtopi = NULL;
        uint32_t *hO = NULL, *hP = NULL, *hrs = NULL;
        uint64_t* hre = NULL;
        uint16_t *hdist = NULL, *htopd = NULL;
        uint32_t* htopi = NULL;
        size_t re_cap = 0;
        std::vector<const uint64_t*> rows;  // the batch's query rows, request by request
    };
    cgh_db* db = NULL;
    int max_batch = 0, nslots = 0, kernel = 0;
    size_t budget = 0, inflight = 0;
    std::vector<Slot> slots;
    std::deque<Request*> pending;
    std::unordered_map<long, Request*> reqs;
    long next_id = 0;
    std::mutex mu;
    std::condition_variable cv;
    std::thread worker;
    bool stop = false;
    std::string fatal;
};

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

size_t batch_bytes(const cgh_db* db, int q) {
    return (size_t)q * (size_t)db->n * 2 + (size_t)up((size_t)q, 256) * (db->kw1 + db->kw2 + db->ds->kw1 + db->ds->kw2) * 4 +
           (size_t)q * (size_t)db->L * 8 + (size_t)q * TOPK * 6 +
           (size_t)q * (size_t)((db->n + TKC - 1) / TKC) * TOPK * 8;
}

// Enqueue one batch of queries on its slot: upload, product, rare pairs,
// top-k, download.
hipError_t launch_batch(cgh_server* sv, cgh_server::Slot& sl) {
    cgh_db* db = sv->db;
    const int q = sl.q, q_pad = (int)up((size_t)q, 256);
    // the batch's queries in the database's columns, all requests in one
    // call (one parallel region over every query) into the slot's pinned
    // buffers; rows q..q_pad-1 stay zero
    sl.rows.clear();
    for (auto* r : sl.reqs)
        for (int k = 0; k < r->q; k++) sl.rows.push_back(r->calls.data() + (size_t)k * db->L);
    const long m = cgh_codec_encode_rows(db->codec, sl.rows.data(), q, sl.hO, sl.hP, sl.hrs, sl.hre, sl.re_cap, NULL, 0);
    if (m < 0) return hipErrorInvalidValue;
    const size_t ne = (size_t)m;
    memset(sl.hO + (size_t)q * db->kw1, 0, (size_t)(q_pad - q) * db->kw1 * 4);
    memset(sl.hP + (size_t)q * db->kw2, 0, (size_t)(q_pad - q) * db->kw2 * 4);
    hipError_t e;
    if ((e = hipMemcpyAsync(sl.dOr, sl.hO, (size_t)q_pad * db->kw1 * 4, hipMemcpyHostToDevice, sl.st))) return e;
    if ((e = hipMemcpyAsync(sl.dPr, sl.hP, (size_t)q_pad * db->kw2 * 4, hipMemcpyHostToDevice, sl.st))) return e;
    if ((e = launch_planes(sl.dOr, q_pad, (int)db->kw1, sl.dO, q_pad, (int)db->ds->kw1, sl.st))) return e;
    if ((e = launch_planes(sl.dPr, q_pad, (int)db->kw2, sl.dP, q_pad, (int)db->ds->kw2, sl.st))) return e;
    if ((e = hipMemcpyAsync(sl.drs, sl.hrs, (size_t)(q + 1) * 4, hipMemcpyHostToDevice, sl.st))) return e;
    if (ne && (e = hipMemcpyAsync(sl.dre, sl.hre, ne * 8, hipMemcpyHostToDevice, sl.st))) return e;
    const int tile = q > 128 ? 256 : 128;
    TileArgs a;
    a.job = {sl.dO, sl.dP, (int)db->ds->kw1, (int)db->ds->kw2, q, 0, q, sl.ddist, db->ds->dO, db->ds->dP, db->n, 1,
             q_pad, (int)db->ds->n_pad, NULL, 0, ds_kv1(db->ds), ds_kv2(db->ds)};
    const int kernel = resolve_kernel(sv->kernel), tw = tile_width(kernel, tile);
    a.ns = 1;
    a.q = 0;
    a.qc = 0;
    a.nbc = (db->n + tw - 1) / tw;
    a.T = (long long)((q + tile - 1) / tile) * a.nbc;
    a.chunk = (a.T + 7) / 8;
    if ((e = launch_tile_args(a, kernel, tile, NST_DEFAULT, sl.st))) return e;
    if (ne) {
        RareJob jb = rare_job(sl.ddist, q, 0, q, sl.drs, sl.dre, db->dqmem, db->m16, 0, 32);
        jb.rect = 1;
        jb.nB = db->n;
        const long long items = rare_layout(jb, q);
        if ((e = launch_rare(jb, items, sl.st))) return e;
    }
    const int tchunks = (db->n + TKC - 1) / TKC;
    hipLaunchKernelGGL(top_k_part, dim3((unsigned)((size_t)q * tchunks)), dim3(256), 0, sl.st, sl.ddist, db->n,
                       tchunks, sl.dpart);
    hipLaunchKernelGGL(top_k_final, dim3((unsigned)q), dim3(256), 0, sl.st, sl.dpart, tchunks, sl.dtopd, sl.dtopi);
    if ((e = hipGetLastError())) return e;
    size_t row = 0;  // distances only of the requests that asked for them
    for (auto* r : sl.reqs) {
        const size_t off = row * (size_t)db->n, cnt = (size_t)r->q * db->n;
        if (r->want_dist && (e = hipMemcpyAsync(sl.hdist + off, sl.ddist + off, cnt * 2, hipMemcpyDeviceToHost, sl.st)))
            return e;
        row += (size_t)r->q;
    }
    if ((e = hipMemcpyAsync(sl.htopd, sl.dtopd, (size_t)q * TOPK * 2, hipMemcpyDeviceToHost, sl.st))) return e;
    if ((e = hipMemcpyAsync(sl.htopi, sl.dtopi, (size_t)q * TOPK * 4, hipMemcpyDeviceToHost, sl.st))) return e;
    return hipEventRecord(sl.ev, sl.st);
}

// Worker: retire finished batches, then give every free slot the longest
// run of waiting requests (arrival order) that fits max_batch queries and the
// HBM budget left by the batches in flight.  Only the worker touches slots
// and the requests it has taken from the queue, so result copies, query
// encoding and launches run with the lock released; a waiter sees a request
// only once `done` is set under the lock.
void serve_loop(cgh_server* sv) {
    (void)hipSetDevice(sv->db->device);
    const size_t n = (size_t)sv->db->n;
    std::unique_lock<std::mutex> lk(sv->mu);
    while (!sv->stop) {
        bool progressed = false;
        for (auto& sl : sv->slots) {
            if (!sl.busy) continue;
            const hipError_t qe = hipEventQuery(sl.ev);
            if (qe == hipErrorNotReady) continue;
            lk.unlock();
            size_t row = 0;
            for (auto* r : sl.reqs) {
                if (qe != hipSuccess) {
                    r->err = hipGetErrorString(qe);
                } else {
                    if (r->want_dist) r->dist.assign(sl.hdist + row * n, sl.hdist + (row + r->q) * n);
                    r->top_d.assign(sl.htopd + row * TOPK, sl.htopd + (row + r->q) * TOPK);
                    r->top_i.assign(sl.htopi + row * TOPK, sl.htopi + (row + r->q) * TOPK);
                }
                row += (size_t)r->q;
            }
            const double t = now_s();
            lk.lock();
            for (auto* r : sl.reqs) {
                r->t_done = t;
                r->done = true;
            }
            sl.reqs.clear();
            sl.busy = false;
            sv->inflight -= sl.bytes;
            progressed = true;
        }
        for (auto& sl : sv->slots) {
            if (sl.busy || sv->pending.empty()) continue;
            int q = 0;
            while (!sv->pending.empty()) {
                cgh_server::Request* r = sv->pending.front();
                const int nq = q + r->q;
                if (nq > sv->max_batch) break;
                if (sv->inflight + batch_bytes(sv->db, nq) > sv->budget) break;
                sl.reqs.push_back(r);
                sv->pending.pop_front();
                q = nq;
            }
            if (!q) break;  // the head waits for memory to free up
            sl.q = q;
            sl.bytes = batch_bytes(sv->db, q);
            sv->inflight += sl.bytes;
            sl.busy = true;
            lk.unlock();
            const hipError_t e = launch_batch(sv, sl);
            lk.lock();
            if (e != hipSuccess) {
                for (auto* r : sl.reqs) {
                    r->err = hipGetErrorString(e);
                    r->t_done = now_s();
                    r->done = true;
                }
                sl.reqs.clear();
                sl.busy = false;
                sv->inflight -= sl.bytes;
            }
            progressed = true;
        }
        if (progressed) sv->cv.notify_all();
        bool any_busy = false;
        for (auto& sl : sv->slots) any_busy |= sl.busy;
        if (any_busy) {
            lk.unlock();
            std::this_thread::sleep_for(std::chrono::microseconds(50));
            lk.lock();
        } else {
            sv->cv.wait(lk, [&] { return sv->stop || !sv->pending.empty(); });
        }
    }
}

}  // namespace

extern "C" void cgh_server_free(cgh_server* sv) {
    if (!sv) return;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;
    }
    sv->cv.notify_all();
    if (sv->worker.joinable()) sv->worker.join();
    (void)hipSetDevice(sv->db->device);
    for (auto& sl : sv->slots) {
        if (sl.st) (void)hipStreamSynchronize(sl.st);
        void* dev[] = {sl.dO, sl.dP, sl.dOr, sl.dPr, sl.drs, sl.dre, sl.ddist, sl.dtopd, sl.dtopi, sl.dpart};
        for (void* p : dev)
            if (p) (void)hipFree(p);
        void* host[] = {sl.hO, sl.hP, sl.hrs, sl.hre, sl.hdist, sl.htopd, sl.htopi};
        for (void* p : host)
            if (p) (void)hipHostFree(p);
        if (sl.ev) (void)hipEventDestroy(sl.ev);
        if (sl.st) (void)hipStreamDestroy(sl.st);
    }
    for (auto& kv : sv->reqs) delete kv.second;
    delete sv;
}

extern "C" int cgh_server_create(cgh_db* db, int slots, int max_batch, size_t hbm_budget, int kernel,
                                 cgh_server** out, char* err, int errlen) {
    cgh_server* sv = new cgh_server();
    size_t freeb = 0, totb = 0;
    *out = NULL;
    sv->db = db;
    sv->max_batch = max_batch;
    sv->kernel = kernel;
    if (slots < 1 || max_batch < 1 || !valid_kernel(kernel)) {
        set_err(err, errlen, "cgh_server_create: bad arguments (slots=%d max_batch=%d kernel=%d)", slots, max_batch,
                kernel);
        delete sv;
        return 1;
    }
    HIP_TRY(hipSetDevice(db->device));
    HIP_TRY(hipMemGetInfo(&freeb, &totb));
    {
        const char* env = getenv("CGH_FREE_BYTES");
        if (env && *env) freeb = std::min(freeb, (size_t)strtoull(env, NULL, 10));
    }
    // admission budget: what the caller allows, within the free HBM left
    // after the slot buffers (each slot holds one batch of up to max_batch)
    {
        const size_t per = batch_bytes(db, max_batch) + 4096;
        const int fit = (int)std::min<size_t>((size_t)slots, freeb / per);
        if (fit < 1) {
            set_err(err, errlen, "no query batch of %d fits in %.2f GB of free device memory", max_batch, freeb / 1e9);
            goto fail;
        }
        sv->nslots = fit;
        sv->budget = hbm_budget ? std::min(hbm_budget, per * (size_t)fit) : per * (size_t)fit;
    }
    sv->slots.resize((size_t)sv->nslots);
    for (auto& sl : sv->slots) {
        const size_t qp = up((size_t)max_batch, 256);
        sl.re_cap = (size_t)max_batch * db->L;
        HIP_TRY(hipStreamCreateWithFlags(&sl.st, hipStreamNonBlocking));
        HIP_TRY(hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming));
        HIP_TRY(hipMalloc(&sl.dOr, qp * db->kw1 * 4));
        HIP_TRY(hipMalloc(&sl.dPr, qp * db->kw2 * 4));
        HIP_TRY(hipMalloc(&sl.dO, qp * db->ds->kw1 * 4));
        HIP_TRY(hipMalloc(&sl.dP, qp * db->ds->kw2 * 4));
        HIP_TRY(hipMalloc(&sl.drs, ((size_t)max_batch + 1) * 4));
        HIP_TRY(hipMalloc(&sl.dre, std::max<size_t>(sl.re_cap, 1) * 8));
        HIP_TRY(hipMalloc(&sl.ddist, up((size_t)max_batch * db->n, 8) * 2));
        HIP_TRY(hipMalloc(&sl.dtopd, (size_t)max_batch * TOPK * 2));
        HIP_TRY(hipMalloc(&sl.dtopi, (size_t)max_batch * TOPK * 4));
        HIP_TRY(hipMalloc(&sl.dpart, (size_t)max_batch * ((db->n + TKC - 1) / TKC) * TOPK * 8));
        HIP_TRY(hipHostMalloc(&sl.hO, qp * db->kw1 * 4, hipHostMallocDefault));
        HIP_TRY(hipHostMalloc(&sl.hP, qp * db->kw2 * 4, hipHostMallocDefault));
        HIP_TRY(hipHostMalloc(&sl.hrs, ((size_t)max_batch + 1) * 4, hipHostMallocDefault));
        HIP_TRY(hipHostMalloc(&sl.hre, std::max<size_t>(sl.re_cap, 1) * 8, hipHostMallocDefault));
        HIP_TRY(hipHostMalloc(&sl.hdist, (size_t)max_batch * db->n * 2, hipHostMallocDefault));
        HIP_TRY(hipHostMalloc(&sl.htopd, (size_t)max_batch * TOPK * 2, hipHostMallocDefault));
        HIP_TRY(hipHostMalloc(&sl.htopi, (size_t)max_batch * TOPK * 4, hipHostMallocDefault));
    }
    sv->worker = std::thread(serve_loop, sv);
    *out = sv;
    return 0;
fail:
    cgh_server_free(sv);
    return 1;
}

extern "C" int cgh_server_slots(const cgh_server* sv) { return sv->nslots; }

extern "C" long cgh_server_submit(cgh_server* sv, const uint64_t* calls, int q, int flags, char* err, int errlen) {
    if (q < 1 || q > sv->max_batch) {
        set_err(err, errlen, "request of %d queries (1..%d accepted)", q, sv->max_batch);
        return -1;
    }
    if (batch_bytes(sv->db, q) > sv->budget) {
        set_err(err, errlen, "request of %d queries needs %.2f GB, over the %.2f GB admission budget", q,
                batch_bytes(sv->db, q) / 1e9, sv->budget / 1e9);
        return -1;
    }
    cgh_server::Request* r = new cgh_server::Request();
    r->q = q;
    r->want_dist = (flags & CGH_WANT_DIST) != 0;
    r->calls.assign(calls, calls + (size_t)q * sv->db->L);
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        r->id = sv->next_id++;
        r->t_submit = now_s();
        sv->reqs[r->id] = r;
        sv->pending.push_back(r);
    }
    sv->cv.notify_all();
    return r->id;
}

extern "C" int cgh_server_wait(cgh_server* sv, long ticket, uint16_t* dist, uint16_t* top_d, uint32_t* top_i,
                               double* latency_s, char* err, int errlen) {
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->reqs.find(ticket);
    if (it == sv->reqs.end()) {
        set_err(err, errlen, "unknown ticket %ld", ticket);
        return 1;
    }
    cgh_server::Request* r = it->second;
    if (dist && !r->want_dist) {
        set_err(err, errlen, "ticket %ld was submitted without CGH_WANT_DIST", ticket);
        return 1;
    }
    sv->cv.wait(lk, [&] { return r->done; });
    sv->reqs.erase(it);
    lk.unlock();
    const int rc = r->err.empty() ? 0 : 1;
    if (rc) {
        set_err(err, errlen, "GPU: %s", r->err.c_str());
    } else {
        if (dist) memcpy(dist, r->dist.data(), r->dist.size() * 2);
        if (top_d) memcpy(top_d, r->top_d.data(), r->top_d.size() * 2);
        if (top_i) memcpy(top_i, r->top_i.data(), r->top_i.size() * 4);
        if (latency_s) *latency_s = r->t_done - r->t_submit;
    }
    delete r;
    return rc;
}


// ---------------------------------------------------------------------------
// Stripes delivered to a host callback (tch_run and cgh_dataset_stream)
// ---------------------------------------------------------------------------

namespace {

// A ring of SLOTS stripe buffers between the device and a host consumer.
// Stripe s is computed into device slot s % SLOTS on the compute stream, its
// rows x w cells are copied to pinned host slot s % SLOTS on a separate copy
// stream (so the copy of stripe s runs under the computation of stripe
// s + 1), and the consumer takes the stripes in order on the calling thread.
// A slot is filled again only after the consumer has returned from the
// stripe it held, so the device works at most SLOTS - 1 stripes ahead and
// every buffer the consumer sees stays untouched while it reads it.
class StripeRing {
public:
    static constexpr int SLOTS = 3;

    StripeRing() = default;
    StripeRing(const StripeRing&) = delete;
    StripeRing& operator=(const StripeRing&) = delete;
    ~StripeRing() {
        // in-flight copies and kernels still reference the buffers
        if (compute_) (void)hipStreamSynchronize(compute_);
        if (copy_) (void)hipStreamSynchronize(copy_);
        if (host_stream_) (void)hipStreamSynchronize(host_stream_);
        for (int k = 0; k < SLOTS; k++) {
            if (dev_[k]) (void)hipFree(dev_[k]);
            if (host_[k]) (void)hipHostFree(host_[k]);
            if (computed_[k]) (void)hipEventDestroy(computed_[k]);
            if (landed_[k]) (void)hipEventDestroy(landed_[k]);
            if (consumed_[k]) (void)hipEventDestroy(consumed_[k]);
        }
        if (compute_) (void)hipStreamDestroy(compute_);
        if (copy_) (void)hipStreamDestroy(copy_);
        if (host_stream_) (void)hipStreamDestroy(host_stream_);
    }

    // Buffers of `cells` u16 per slot (a multiple of 8: the rare pass stages
    // whole 16-byte words); the streams are created at the default priority.
    hipError_t open(size_t cells) {
        hipError_t e = hipStreamCreateWithFlags(&compute_, hipStreamNonBlocking);
        if (!e) e = hipStreamCreateWithFlags(&copy_, hipStreamNonBlocking);
        if (!e) e = hipStreamCreateWithFlags(&host_stream_, hipStreamNonBlocking);
        for (int k = 0; k < SLOTS && !e; k++) {
            e = hipMalloc(&dev_[k], cells * 2);
            if (!e) e = hipHostMalloc(&host_[k], cells * 2, hipHostMallocDefault);
            if (!e) e = hipEventCreateWithFlags(&computed_[k], hipEventDisableTiming);
            if (!e) e = hipEventCreateWithFlags(&landed_[k], hipEventDisableTiming);
            if (!e) e = hipEventCreateWithFlags(&consumed_[k], hipEventDisableTiming);
        }
        return e;
    }

    // produce(r0, rows, dev, stream) enqueues stripe r0's device work into
    // dev (rows x (n - r0) u16 layout) and returns its status;
    // consume(r0, rows, w, host) is the host side of that stripe.
    template <class Produce, class Consume>
    hipError_t run(int n, int S, Produce&& produce, Consume&& consume) {
        const int stripes = (n + S - 1) / S;
        auto fill = [&](int s) -> hipError_t {
            const int k = s % SLOTS, r0 = s * S, rows = std::min(S, n - r0);
            const size_t cells = (size_t)rows * (size_t)(n - r0);
            hipError_t e = produce(r0, rows, dev_[k], compute_);
            if (!e) e = hipEventRecord(computed_[k], compute_);
            if (!e) e = hipStreamWaitEvent(copy_, computed_[k], 0);
            if (!e) e = hipMemcpyAsync(host_[k], dev_[k], cells * 2, hipMemcpyDeviceToHost, copy_);
            if (!e) e = hipEventRecord(landed_[k], copy_);
            return e;
        };
        hipError_t e = hipSuccess;
        for (int s = 0; s < std::min(stripes, SLOTS) && !e; s++) e = fill(s);
        for (int s = 0; s < stripes && !e; s++) {
            const int k = s % SLOTS, r0 = s * S;
            e = hipEventSynchronize(landed_[k]);
            if (e) break;
            consume(r0, std::min(S, n - r0), (size_t)(n - r0), host_[k]);
            if (s + SLOTS < stripes) e = fill(s + SLOTS);
        }
        return e;
    }

    // The same pipeline captured as one HIP graph and launched once: per
    // stripe a compute node sequence into device slot s % SLOTS, a copy node
    // to pinned host slot s % SLOTS, and a host node that calls
    // cb(ctx, r0, rows, w, cells); the host nodes run in stripe order, and
    // the compute of stripe s + SLOTS depends on the host node of stripe s,
    // so a slot is refilled only after its consumer has returned.  The
    // callback runs on a runtime thread while the caller waits, one stripe
    // at a time; it must not call HIP.  *build_ms gets the capture and
    // instantiation time.
    template <class Produce>
    hipError_t run_graph(int n, int S, Produce&& produce, tch_cb cb, void* ctx, double* build_ms) {
        struct HostCall {
            tch_cb cb;
            void* ctx;
            int r0, rows;
            size_t w;
            const uint16_t* cells;
        };
        const int stripes = (n + S - 1) / S;
        std::vector<HostCall> calls((size_t)stripes);
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        hipEvent_t fork = nullptr, copied = nullptr;
        const auto t0 = std::chrono::steady_clock::now();
        hipError_t e = hipEventCreateWithFlags(&fork, hipEventDisableTiming);
        if (!e) e = hipEventCreateWithFlags(&copied, hipEventDisableTiming);
        if (!e) e = hipStreamBeginCapture(compute_, hipStreamCaptureModeThreadLocal);
        if (!e) {
            // the copy and host streams join the capture through events
            hipError_t c = hipEventRecord(fork, compute_);
            if (!c) c = hipStreamWaitEvent(copy_, fork, 0);
            if (!c) c = hipStreamWaitEvent(host_stream_, fork, 0);
            for (int s = 0; s < stripes && !c; s++) {
                const int k = s % SLOTS, r0 = s * S, rows = std::min(S, n - r0);
                const size_t w = (size_t)(n - r0), cells = (size_t)rows * w;
                calls[(size_t)s] = HostCall{cb, ctx, r0, rows, w, host_[k]};
                if (s >= SLOTS) c = hipStreamWaitEvent(compute_, consumed_[k], 0);
                if (!c) c = produce(r0, rows, dev_[k], compute_);
                if (!c) c = hipEventRecord(computed_[k], compute_);
                if (!c) c = hipStreamWaitEvent(copy_, computed_[k], 0);
                if (!c) c = hipMemcpyAsync(host_[k], dev_[k], cells * 2, hipMemcpyDeviceToHost, copy_);
                if (!c) c = hipEventRecord(landed_[k], copy_);
                if (!c) c = hipStreamWaitEvent(host_stream_, landed_[k], 0);
                if (!c)
                    c = hipLaunchHostFunc(
                        host_stream_,
                        [](void* p) {
                            const HostCall* h = static_cast<const HostCall*>(p);
                            h->cb(h->ctx, h->r0, h->rows, h->w, h->cells);
                        },
                        &calls[(size_t)s]);
                if (!c) c = hipEventRecord(consumed_[k], host_stream_);
            }
            // join both side streams back before the capture ends
            if (!c) c = hipEventRecord(copied, copy_);
            if (!c) c = hipStreamWaitEvent(compute_, copied, 0);
            if (!c) c = hipStreamWaitEvent(compute_, consumed_[(stripes - 1) % SLOTS], 0);
            e = hipStreamEndCapture(compute_, &graph);
            if (c) e = c;
        }
        if (!e) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (build_ms) *build_ms = seconds(t0) * 1e3;
        if (!e) e = hipGraphLaunch(exec, compute_);
        if (!e) e = hipStreamSynchronize(compute_);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (fork) (void)hipEventDestroy(fork);
        if (copied) (void)hipEventDestroy(copied);
        return e;
    }

private:
    static double seconds(std::chrono::steady_clock::time_point t) {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t).count();
    }
    uint16_t* dev_[SLOTS] = {};
    uint16_t* host_[SLOTS] = {};
    hipEvent_t computed_[SLOTS] = {};
    hipEvent_t landed_[SLOTS] = {};
    hipEvent_t consumed_[SLOTS] = {};
    hipStream_t compute_ = nullptr;
    hipStream_t copy_ = nullptr;
    hipStream_t host_stream_ = nullptr;
};

double seconds_since(std::chrono::steady_clock::time_point t) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t).count();
}


// Device copy of a caller's rare CSR (start[n + 1], js[n_rare]: partners
// j > i of row i, in any order), read by the row-local rare pass as one
// member range per row that has partners: the partner array is the member
// array and row i's only range is [start[i], start[i + 1]).
struct CsrRanges {
    uint32_t* members = nullptr;  // js
    uint32_t* rstart = nullptr;   // n + 1 offsets into rentry
    uint64_t* rentry = nullptr;   // (end << 32) | begin
    size_t entries = 0;

    ~CsrRanges() {
        if (members) (void)hipFree(members);
        if (rstart) (void)hipFree(rstart);
        if (rentry) (void)hipFree(rentry);
    }
    static size_t device_bytes(int n, size_t n_rare) {
        return cghi::member_array_bytes(n_rare, false) + ((size_t)n + 1) * 4 + (size_t)n * 8;
    }
    hipError_t upload(int n, const uint32_t* start, const uint32_t* js, size_t n_rare) {
        std::vector<uint32_t> rs((size_t)n + 1);
        std::vector<uint64_t> re;
        re.reserve((size_t)n);
        for (int i = 0; i < n; i++) {
            rs[(size_t)i] = (uint32_t)re.size();
            if (start[i + 1] > start[i]) re.push_back(((uint64_t)start[i + 1] << 32) | start[i]);
        }
        rs[(size_t)n] = (uint32_t)re.size();
        entries = re.size();
        hipError_t e = hipMalloc(&members, cghi::member_array_bytes(n_rare, false));
        if (!e) e = hipMalloc(&rstart, rs.size() * 4);
        if (!e) e = hipMalloc(&rentry, std::max<size_t>(entries, 1) * 8);
        if (!e && n_rare) e = hipMemcpy(members, js, n_rare * 4, hipMemcpyHostToDevice);
        if (!e) e = hipMemcpy(rstart, rs.data(), rs.size() * 4, hipMemcpyHostToDevice);
        if (!e && entries) e = hipMemcpy(rentry, re.data(), entries * 8, hipMemcpyHostToDevice);
        return e;
    }
};

// Device operand planes of a caller's row-major bit matrices, freed with
// the object.
struct OwnedPlanes {
    uint32_t* O = nullptr;
    uint32_t* P = nullptr;
    ~OwnedPlanes() {
        if (O) (void)hipFree(O);
        if (P) (void)hipFree(P);
    }
};

bool stripe_height_ok(int S, char* err, int errlen) {
    if (S >= 8 && S % 8 == 0) return true;
    set_err(err, errlen, "stripe height must be a positive multiple of 8 (got %d)", S);
    return false;
}

}  // namespace

// The reference boundary as an adapter onto the library's own parts: the
// caller's bit matrices become device planes, its CSR becomes member ranges
// of the row-local rare pass (or, when it does not fit in device memory,
// is applied on the host to each stripe before the callback), and the
// stripes flow through the StripeRing.
extern "C" int tch_run(int device, int n, size_t n_pad, const uint8_t* O, size_t K1, const uint8_t* P,
                       size_t K2, const uint32_t* start, const uint32_t* js, size_t n_rare, int S, tch_cb cb,
                       void* ctx, double* gpu_secs, int* rare_on_gpu, char* err, int errlen) {
    if (n < 1 || K1 == 0 || K2 == 0 || K1 % 1024 || K2 % 1024 || n_pad % 256 || n_pad < (size_t)n ||
        n_pad > (size_t)INT32_MAX) {
        set_err(err, errlen, "tch_run: bad shape (n=%d n_pad=%zu K1=%zu K2=%zu)", n, n_pad, K1, K2);
        return 1;
    }
    if (!stripe_height_ok(S, err, errlen)) return 1;
    const auto t_start = std::chrono::steady_clock::now();
    const size_t slot_cells = up((size_t)S * (size_t)n, 8);
    const int kw1 = kc_words(K1 / 32), kw2 = kc_words(K2 / 32);
    const int kernel = resolve_kernel(CGH_KERNEL_AUTO), tile = pick_tile(n);
    OwnedPlanes planes;
    CsrRanges csr;
    StripeRing ring;
    size_t free_b = 0, total_b = 0;
    bool on_device = false;

    HIP_TRY(hipSetDevice(device));
    HIP_TRY(upload_planes(O, (int)n_pad, (int)(K1 / 32), kw1, &planes.O));
    HIP_TRY(upload_planes(P, (int)n_pad, (int)(K2 / 32), kw2, &planes.P));
    HIP_TRY(ring.open(slot_cells));
    HIP_TRY(hipMemGetInfo(&free_b, &total_b));
    if (const char* env = getenv("CGH_FREE_BYTES"); env && *env) free_b = std::min(free_b, (size_t)strtoull(env, NULL, 10));
    // rare pairs on the device when their ranges fit beside a 64 MiB margin
    on_device = n_rare == 0 || CsrRanges::device_bytes(n, n_rare) + (64u << 20) <= free_b;
    if (n_rare && on_device) HIP_TRY(csr.upload(n, start, js, n_rare));
    if (rare_on_gpu) *rare_on_gpu = on_device;
    {
        const int rsg = lanes_per_range(n_rare, csr.entries);
        auto produce = [&](int r0, int rows, uint16_t* dst, hipStream_t st) -> hipError_t {
            hipError_t e = launch_tiles(kernel, tile, planes.O, kw1, planes.P, kw2, (int)n_pad, n, r0, S, 1, dst, st);
            if (!e && n_rare && on_device && n - r0 > 1) {
                RareJob jb = rare_job(dst, n, r0, S, csr.rstart, csr.rentry, csr.members, 0, 0, rsg);
                jb.sorted = 0;  // a CSR row lists its partners in any order
                e = launch_rare(jb, rare_layout(jb, rows), st);
            }
            return e;
        };
        auto consume = [&](int r0, int rows, size_t w, uint16_t* cells) {
            if (n_rare && !on_device)  // partner j of row r0 + r is column j - r0 of stripe row r
                host_parallel((size_t)rows, [&](size_t lo, size_t hi) {
                    for (size_t r = lo; r < hi; r++) {
                        uint16_t* row = cells + r * w - r0;
                        const size_t i = (size_t)r0 + r;
                        for (const uint32_t* p = js + start[i]; p != js + start[i + 1]; ++p) row[*p]--;
                    }
                });
            cb(ctx, r0, rows, w, cells);
        };
        HIP_TRY(ring.run(n, S, produce, consume));
    }
    if (gpu_secs) *gpu_secs = seconds_since(t_start);
    return 0;
fail:
    return 1;
}

namespace {
// The device work of one stripe of a dataset: one product launch (with the
// cap of rows that have no rare partners) and one row-local rare launch over
// the dataset's groups, into a rows x (n - r0) buffer.
struct DatasetStripes {
    cgh_dataset* ds;
    int S, kernel, tile, stages, rsg;
    explicit DatasetStripes(cgh_dataset* d, int s)
        : ds(d), S(s), kernel(ds_kernel(d, CGH_KERNEL_AUTO)), tile(ds_tile(d, CGH_KERNEL_AUTO)),
          stages(ds_stages(d, CGH_KERNEL_AUTO)), rsg(lanes_per_range(d->n_rare, d->n_rentries)) {}
    hipError_t operator()(int r0, int rows, uint16_t* dst, hipStream_t st) const {
        const int n = ds->n;
        hipError_t e = launch_tiles(kernel, tile, ds->dO, (int)ds->kw1, ds->dP, (int)ds->kw2, (int)ds->n_pad, n, r0, S,
                                    1, dst, st, ds->drstart, ds->cap, ds_kv1(ds), ds_kv2(ds), stages);
        if (!e && ds->n_rare && n - r0 > 1) {
            RareJob jb = rare_job(dst, n, r0, S, ds->drstart, ds->drentry, ds->dgmem, ds->m16, ds->cap, rsg);
            e = launch_rare(jb, rare_layout(jb, rows), st);
        }
        return e;
    }
};
}  // namespace

// The CLI's path: stripes of a device-resident dataset through the
// StripeRing, each stripe's device work enqueued by the calling thread and
// its callback run there.
extern "C" int cgh_dataset_stream(cgh_dataset* ds, int S, tch_cb cb, void* ctx, double* gpu_secs, char* err,
                                  int errlen) {
    if (!stripe_height_ok(S, err, errlen) || split_unsupported(ds, "cgh_dataset_stream", err, errlen)) return 1;
    const auto t_start = std::chrono::steady_clock::now();
    StripeRing ring;
    HIP_TRY(hipSetDevice(ds->device));
    HIP_TRY(ring.open(up((size_t)S * (size_t)ds->n, 8)));
    {
        const DatasetStripes produce(ds, S);
        auto consume = [&](int r0, int rows, size_t w, const uint16_t* cells) { cb(ctx, r0, rows, w, cells); };
        HIP_TRY(ring.run(ds->n, S, produce, consume));
    }
    if (gpu_secs) *gpu_secs = seconds_since(t_start);
    return 0;
fail:
    return 1;
}

// The same stripes as one captured HIP graph (StripeRing::run_graph).
extern "C" int cgh_dataset_stream_graph(cgh_dataset* ds, int S, tch_cb cb, void* ctx, double* gpu_secs,
                                        double* build_ms, char* err, int errlen) {
    if (!stripe_height_ok(S, err, errlen) || split_unsupported(ds, "cgh_dataset_stream_graph", err, errlen)) return 1;
    const auto t_start = std::chrono::steady_clock::now();
    StripeRing ring;
    HIP_TRY(hipSetDevice(ds->device));
    HIP_TRY(ring.open(up((size_t)S * (size_t)ds->n, 8)));
    (void)rare_grid(1);  // the CU count is looked up before the capture, not inside it
    {
        const DatasetStripes produce(ds, S);
        HIP_TRY(ring.run_graph(ds->n, S, produce, cb, ctx, build_ms));
    }
    if (gpu_secs) *gpu_secs = seconds_since(t_start);
    return 0;
fail:
    return 1;
}
