/*
 * oracle.c -- CPU restatement of the exact cgMLST allelic Hamming distance.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP product
 * path; it is linked or loaded only by tests/, by __graft_entry__.smoke() and
 * by bench.py's cpu_baseline leg.  Nothing in the product library calls it.
 *
 * Functions and the reference lines each one restates:
 *
 *   oracle_hamming_rows    H[i][j] = sum_k [x_ik != 0][x_jk != 0][x_ik != x_jk]
 *                          tensorcore/rs/check.py:39-44 (numpy brute force),
 *                          tensorcore/tc_hamming.cu:244-248 (spot check),
 *                          cuda/hamming.cu:90 (direct-compare op).
 *   oracle_upper_block     the same op over an upper-triangle row block,
 *                          blocked for cache reuse (CPU baseline kernel).
 *   oracle_checksum_upper  u64 checksum h = h*1000003 + H[i][j], i ascending,
 *                          j > i ascending: tensorcore/tc_hamming.cu:241.
 *   oracle_dense_ranks     per-locus dense ranks of u64 ids, 0 stays 0,
 *                          rank = 1 + index of the id in the sorted distinct
 *                          non-zero ids: tensorcore/rs/src/main.rs:135-165.
 *
 * Parity pins (what checks this file against something other than itself):
 *   - tensorcore/README.md:52 checksum KAT 85724a9d7d090f9f of the synthetic
 *     30,000 x 5,000 set (generator python/gen_profiles.py), checked in
 *     tests/test_oracle.py;
 *   - the per-locus edge vector of python/relu_hamming.py:56-57;
 *   - an independent route: H = P P^T - O O^T over a full one-hot (integer
 *     matrix products in numpy) on the check.py datasets;
 *   - hand-counted tiny cases and invariants (symmetry, zero diagonal,
 *     H <= min(called_i, called_j), permutation and id-relabelling invariance).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* One pair: the direct compare of cuda/hamming.cu:90, counted over L loci. */
static inline uint32_t pair_distance(const uint16_t* a, const uint16_t* b, int L) {
    uint32_t d = 0;
    for (int k = 0; k < L; k++) d += (a[k] != 0) & (b[k] != 0) & (a[k] != b[k]);
    return d;
}

/* Full rows i0..i1-1 of H (n columns each), out is (i1-i0) x n u16. */
void oracle_hamming_rows(const uint16_t* X, int n, int L, int i0, int i1, uint16_t* out) {
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = i0; i < i1; i++) {
        const uint16_t* a = X + (size_t)i * L;
        uint16_t* o = out + (size_t)(i - i0) * n;
        for (int j = 0; j < n; j++) o[j] = (uint16_t)(i == j ? 0 : pair_distance(a, X + (size_t)j * L, L));
    }
}

/*
 * Upper-triangle rows i0..i1-1, strict upper part only: row i holds
 * H[i][i+1..n-1] at out + offs[i - i0], offs computed by the caller as the
 * running sum of (n - 1 - i).  Cache-blocked: 16x16 pair tiles over 512-locus
 * chunks so the working set stays in L1/L2.
 */
#define OB 16
#define OKC 512
void oracle_upper_block(const uint16_t* X, int n, int L, int i0, int i1, uint16_t* out) {
    const int nbi = (i1 - i0 + OB - 1) / OB;
    /* offset of row i within out */
#define ROWOFF(i) ((size_t)((i) - i0) * (size_t)(n - 1) - ((size_t)((i) - i0) * (size_t)((i) - i0 - 1)) / 2 - (size_t)((i) - i0) * (size_t)i0)
#pragma omp parallel
    {
        uint32_t acc[OB][OB];
#pragma omp for schedule(dynamic, 1)
        for (int bi = 0; bi < nbi; bi++) {
            const int ia = i0 + bi * OB, ib = ia + OB < i1 ? ia + OB : i1;
            for (int ja = ia; ja < n; ja += OB) {
                const int jb = ja + OB < n ? ja + OB : n;
                memset(acc, 0, sizeof acc);
                for (int k0 = 0; k0 < L; k0 += OKC) {
                    const int kc = k0 + OKC < L ? OKC : L - k0;
                    for (int i = ia; i < ib; i++) {
                        const uint16_t* a = X + (size_t)i * L + k0;
                        for (int j = ja; j < jb; j++) {
                            if (j <= i) continue;
                            acc[i - ia][j - ja] += pair_distance(a, X + (size_t)j * L + k0, kc);
                        }
                    }
                }
                for (int i = ia; i < ib; i++) {
                    const size_t ro = ROWOFF(i);
                    for (int j = ja; j < jb; j++)
                        if (j > i) out[ro + (size_t)(j - i - 1)] = (uint16_t)acc[i - ia][j - ja];
                }
            }
        }
    }
#undef ROWOFF
}

/* Number of u16 cells of the strict upper part of rows i0..i1-1. */
size_t oracle_upper_cells(int n, int i0, int i1) {
    size_t c = 0;
    for (int i = i0; i < i1; i++) c += (size_t)(n - 1 - i);
    return c;
}

/* tc_hamming.cu:241 checksum over the whole strict upper triangle. */
uint64_t oracle_checksum_upper(const uint16_t* X, int n, int L) {
    uint64_t h = 0;
    const int block = 512;
    uint16_t* buf = (uint16_t*)malloc(oracle_upper_cells(n, 0, block < n ? block : n) * 2 + 16);
    if (!buf) return 0;
    for (int i0 = 0; i0 < n; i0 += block) {
        const int i1 = i0 + block < n ? i0 + block : n;
        oracle_upper_block(X, n, L, i0, i1, buf);
        const size_t cells = oracle_upper_cells(n, i0, i1);
        for (size_t c = 0; c < cells; c++) h = h * 1000003u + buf[c];
    }
    free(buf);
    return h;
}

/* Checksum of an already computed strict upper triangle in row order. */
uint64_t oracle_checksum_cells(const uint16_t* cells, size_t count, uint64_t h) {
    for (size_t c = 0; c < count; c++) h = h * 1000003u + cells[c];
    return h;
}

static int cmp_u64(const void* a, const void* b) {
    const uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return x < y ? -1 : x > y;
}

/*
 * main.rs:135-165.  calls is n x L row-major u64 (0 = missing).  Writes X
 * (n x L u16) and returns the largest number of distinct alleles at any locus,
 * or -1 (the reference dies) if a locus has more than 65,535 distinct alleles.
 */
int oracle_dense_ranks(const uint64_t* calls, int n, int L, uint16_t* X) {
    int worst = 0, bad = 0;
#pragma omp parallel
    {
        uint64_t* u = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(n > 0 ? n : 1));
#pragma omp for schedule(dynamic, 4) reduction(max : worst) reduction(| : bad)
        for (int k = 0; k < L; k++) {
            int m = 0;
            for (int i = 0; i < n; i++) {
                uint64_t v = calls[(size_t)i * L + k];
                if (v) u[m++] = v;
            }
            qsort(u, (size_t)m, sizeof(uint64_t), cmp_u64);
            int d = 0;
            for (int t = 0; t < m; t++)
                if (d == 0 || u[d - 1] != u[t]) u[d++] = u[t];
            if (d > 65535) {
                bad = 1;
                continue;
            }
            if (d > worst) worst = d;
            for (int i = 0; i < n; i++) {
                uint64_t v = calls[(size_t)i * L + k];
                uint16_t r = 0;
                if (v) {
                    int lo = 0, hi = d - 1;
                    while (lo < hi) {
                        int mid = (lo + hi) / 2;
                        if (u[mid] < v) lo = mid + 1; else hi = mid;
                    }
                    r = (uint16_t)(lo + 1);
                }
                X[(size_t)i * L + k] = r;
            }
        }
        free(u);
    }
    return bad ? -1 : worst;
}

int oracle_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
