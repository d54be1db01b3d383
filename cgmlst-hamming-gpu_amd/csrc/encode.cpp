// encode.cpp -- host encoder of the exact Hamming identity (C++17 + OpenMP).
//
// Dense ranks, allele counts, the frequent-column map, O/P bit packing, rare
// groups, the reference's rare CSR and the tau cost model.  All of it is
// integer bookkeeping on the host; the device kernels consume its output.
//
// Reference lines restated: tensorcore/rs/src/main.rs:135-165 (dense_ranks),
// 169-198 (choose_tau), 366-385 (column map), 386-403 (bit pack), 404-444
// (rare groups and CSR).

#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <vector>

#include <omp.h>

#include "cgmlst_hamming.h"

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

// Per-locus counts of ranks 0..max (cnt[k][0] = missing), like main.rs:146-153.
void locus_counts(const uint16_t* X, int n, int L, std::vector<std::vector<uint32_t>>& cnt) {
    cnt.assign((size_t)L, {});
#pragma omp parallel for schedule(static)
        for (int k = 0; k < L; k++) {
            uint32_t m = 0;
            for (int i = 0; i < n; i++) m = std::max<uint32_t>(m, X[(size_t)i * L + k]);
            std::vector<uint32_t> c(m + 1, 0);
            for (int i = 0; i < n; i++) c[X[(size_t)i * L + k]]++;
            cnt[k].swap(c);
        }
}

}  // namespace

extern "C" int cgh_dense_ranks(const uint64_t* calls, int n, int L, uint16_t* X, char* err, int errlen) {
    std::atomic<int> bad_locus{-1};
    std::atomic<size_t> bad_count{0};
#pragma omp parallel
    {
        std::vector<uint64_t> u;
        u.reserve((size_t)n);
#pragma omp for schedule(dynamic, 4)
        for (int k = 0; k < L; k++) {
            u.clear();
            for (int i = 0; i < n; i++) {
                const uint64_t v = calls[(size_t)i * L + k];
                if (v) u.push_back(v);
            }
            std::sort(u.begin(), u.end());
            u.erase(std::unique(u.begin(), u.end()), u.end());
            if (u.size() > 65535) {
                int expect = -1;
                if (bad_locus.compare_exchange_strong(expect, k)) bad_count = u.size();
                continue;
            }
            for (int i = 0; i < n; i++) {
                const uint64_t v = calls[(size_t)i * L + k];
                X[(size_t)i * L + k] =
                    v ? (uint16_t)(std::lower_bound(u.begin(), u.end(), v) - u.begin() + 1) : (uint16_t)0;
            }
        }
    }
    if (bad_locus >= 0) {
        set_err(err, errlen, "locus %d has %zu distinct alleles (limit 65535)", bad_locus.load(), bad_count.load());
        return 1;
    }
    return 0;
}

extern "C" size_t cgh_allele_counts(const uint16_t* X, int n, int L, uint32_t* cnt, size_t cap) {
    std::vector<std::vector<uint32_t>> c;
    locus_counts(X, n, L, c);
    size_t need = 0;
    for (auto& v : c) need += v.size() - 1;
    if (cnt && cap >= need) {
        size_t o = 0;
        for (auto& v : c)
            for (size_t a = 1; a < v.size(); a++) cnt[o++] = v[a];
    }
    return need;
}

extern "C" void cgh_encoding_free(cgh_encoding* enc) {
    if (!enc) return;
    free(enc->O);
    free(enc->P);
    free(enc->gstart);
    free(enc->gmem);
    free(enc->gpairs);
    free(enc->rstart);
    free(enc->rentry);
    memset(enc, 0, sizeof *enc);
}

namespace {
int encode_impl(const uint16_t* X, int n, int L, int tau, int k_align, bool bits, cgh_encoding* enc, char* err,
                int errlen);
}

extern "C" int cgh_encode(const uint16_t* X, int n, int L, int tau, int k_align, cgh_encoding* enc, char* err,
                          int errlen) {
    return encode_impl(X, n, L, tau, k_align, true, enc, err, errlen);
}

extern "C" int cgh_encode_meta(const uint16_t* X, int n, int L, int tau, int k_align, cgh_encoding* enc, char* err,
                               int errlen) {
    return encode_impl(X, n, L, tau, k_align, false, enc, err, errlen);
}

extern "C" size_t cgh_column_table(const uint16_t* X, int n, int L, int tau, uint32_t* colstart, uint32_t* colmap,
                                   size_t cap) {
    std::vector<std::vector<uint32_t>> cnt;
    locus_counts(X, n, L, cnt);
    size_t need = 0;
    for (int k = 0; k < L; k++) need += cnt[k].size();
    if (!colstart || !colmap || cap < need) return need;
    size_t o = 0, c = 0;
    for (int k = 0; k < L; k++) {
        colstart[k] = (uint32_t)o;
        for (size_t a = 0; a < cnt[k].size(); a++)
            colmap[o++] = a && cnt[k][a] >= (uint32_t)tau ? (uint32_t)c++ : 0xffffffffu;
    }
    colstart[L] = (uint32_t)o;
    return need;
}

namespace {
int encode_impl(const uint16_t* X, int n, int L, int tau, int k_align, bool bits, cgh_encoding* enc, char* err,
                int errlen) {
    memset(enc, 0, sizeof *enc);
    if (n < 0 || L < 0 || tau < 1 || (k_align != 256 && k_align != 512 && k_align != 1024)) {
        set_err(err, errlen, "cgh_encode: bad arguments (n=%d L=%d tau=%d k_align=%d)", n, L, tau, k_align);
        return 1;
    }
    std::vector<std::vector<uint32_t>> cnt;
    locus_counts(X, n, L, cnt);
    // frequent-column map (main.rs:366-385): col[k][a] local index or -1
    std::vector<std::vector<int32_t>> col((size_t)L);
    std::vector<size_t> base((size_t)L + 1, 0);
    for (int k = 0; k < L; k++) {
        col[k].assign(cnt[k].size(), -1);
        int32_t c = 0;
        for (size_t a = 1; a < cnt[k].size(); a++)
            if (cnt[k][a] >= (uint32_t)tau) col[k][a] = c++;
        base[k + 1] = base[k] + (size_t)c;
    }
    const size_t D = base[L];
    enc->n = n;
    enc->L = L;
    enc->tau = tau;
    enc->D = D;
    enc->n_pad = up((size_t)std::max(n, 1), 256);
    enc->kw1 = up(std::max<size_t>(D, 1), (size_t)k_align) / 32;
    enc->kw2 = up(std::max<size_t>((size_t)L, 1), (size_t)k_align) / 32;
    if (bits) {
        enc->O = (uint32_t*)calloc(enc->n_pad * enc->kw1, 4);
        enc->P = (uint32_t*)calloc(enc->n_pad * enc->kw2, 4);
        if (!enc->O || !enc->P) {
            cgh_encoding_free(enc);
            set_err(err, errlen, "out of host memory for the bit matrices");
            return 1;
        }
    }
    // bit pack (main.rs:386-403), LSB-first
#pragma omp parallel for schedule(static)
    for (int i = 0; i < (bits ? n : 0); i++) {
        uint32_t* orow = enc->O + (size_t)i * enc->kw1;
        uint32_t* prow = enc->P + (size_t)i * enc->kw2;
        for (int k = 0; k < L; k++) {
            const uint16_t a = X[(size_t)i * L + k];
            if (!a) continue;
            prow[k >> 5] |= 1u << (k & 31);
            const int32_t c = col[k][a];
            if (c >= 0) {
                const size_t cc = base[k] + (size_t)c;
                orow[cc >> 5] |= 1u << (cc & 31);
            }
        }
    }
    // rare groups (main.rs:404-417): (locus, allele) with 2 <= count < tau,
    // ordered by locus then rank, members ascending
    std::vector<size_t> lg((size_t)L + 1, 0), lm((size_t)L + 1, 0);
    for (int k = 0; k < L; k++) {
        size_t g = 0, m = 0;
        for (size_t a = 1; a < cnt[k].size(); a++)
            if (col[k][a] < 0 && cnt[k][a] >= 2) {
                g++;
                m += cnt[k][a];
            }
        lg[k + 1] = lg[k] + g;
        lm[k + 1] = lm[k] + m;
    }
    enc->n_groups = lg[L];
    enc->n_members = lm[L];
    if (enc->n_members > 0xffffffffull) {
        cgh_encoding_free(enc);
        set_err(err, errlen, "too many rare-group members: raise --tau");
        return 1;
    }
    enc->gstart = (uint32_t*)malloc((enc->n_groups + 1) * 4);
    enc->gmem = (uint32_t*)malloc(std::max<size_t>(enc->n_members, 1) * 4);
    enc->gpairs = (uint64_t*)malloc((enc->n_groups + 1) * 8);
    if (!enc->gstart || !enc->gmem || !enc->gpairs) {
        cgh_encoding_free(enc);
        set_err(err, errlen, "out of host memory for the rare groups");
        return 1;
    }
#pragma omp parallel
    {
        std::vector<int64_t> slot;   // rank -> write cursor into gmem, -1 if not rare
#pragma omp for schedule(dynamic, 4)
        for (int k = 0; k < L; k++) {
            slot.assign(cnt[k].size(), -1);
            size_t g = lg[k], m = lm[k];
            for (size_t a = 1; a < cnt[k].size(); a++)
                if (col[k][a] < 0 && cnt[k][a] >= 2) {
                    enc->gstart[g] = (uint32_t)m;
                    const uint64_t c = cnt[k][a];
                    enc->gpairs[g] = c * (c - 1) / 2;  // per-group count, prefixed below
                    slot[a] = (int64_t)m;
                    g++;
                    m += c;
                }
            for (int i = 0; i < n; i++) {
                const uint16_t a = X[(size_t)i * L + k];
                if (a && slot[a] >= 0) enc->gmem[slot[a]++] = (uint32_t)i;
            }
        }
    }
    enc->gstart[enc->n_groups] = (uint32_t)enc->n_members;
    uint64_t acc = 0;
    for (size_t g = 0; g < enc->n_groups; g++) {
        const uint64_t c = enc->gpairs[g];
        enc->gpairs[g] = acc;
        acc += c;
    }
    enc->gpairs[enc->n_groups] = acc;
    enc->n_rare = acc;

    // row index of the same pairs: member p of a group of c gets the range
    // gmem[gs+p+1, gs+c) of its later partners, listed under row gmem[gs+p]
    std::vector<std::atomic<uint32_t>> rc((size_t)n + 1);
    for (auto& x : rc) x.store(0, std::memory_order_relaxed);
    const long G = (long)enc->n_groups;
#pragma omp parallel for schedule(dynamic, 256)
    for (long g = 0; g < G; g++) {
        const uint32_t s0 = enc->gstart[g], s1 = enc->gstart[g + 1];
        for (uint32_t p = s0; p + 1 < s1; p++) rc[enc->gmem[p]].fetch_add(1, std::memory_order_relaxed);
    }
    enc->rstart = (uint32_t*)malloc(((size_t)n + 1) * 4);
    size_t tot = 0;
    for (int i = 0; i < n; i++) {
        enc->rstart[i] = (uint32_t)tot;
        tot += rc[i].load(std::memory_order_relaxed);
    }
    enc->rstart[n] = (uint32_t)tot;
    enc->n_rentries = tot;
    enc->rentry = (uint64_t*)malloc(std::max<size_t>(tot, 1) * 8);
    if (!enc->rstart || !enc->rentry) {
        cgh_encoding_free(enc);
        set_err(err, errlen, "out of host memory for the rare row index");
        return 1;
    }
    for (int i = 0; i < n; i++) rc[i].store(enc->rstart[i], std::memory_order_relaxed);
#pragma omp parallel for schedule(dynamic, 256)
    for (long g = 0; g < G; g++) {
        const uint32_t s0 = enc->gstart[g], s1 = enc->gstart[g + 1];
        for (uint32_t p = s0; p + 1 < s1; p++)
            enc->rentry[rc[enc->gmem[p]].fetch_add(1, std::memory_order_relaxed)] =
                ((uint64_t)s1 << 32) | (uint64_t)(p + 1);
    }
    return 0;
}
}  // namespace

extern "C" int cgh_rare_csr(const cgh_encoding* enc, uint32_t* start, uint32_t* js, char* err, int errlen) {
    const int n = enc->n;
    if (enc->n_rare >= (1ull << 32)) {
        set_err(err, errlen, "too many rare pairs: raise --tau");
        return 1;
    }
    std::vector<std::atomic<uint64_t>> counts((size_t)n);
    for (auto& c : counts) c.store(0, std::memory_order_relaxed);
    const long G = (long)enc->n_groups;
#pragma omp parallel for schedule(dynamic, 64)
    for (long g = 0; g < G; g++) {
        const uint32_t s = enc->gstart[g], c = enc->gstart[g + 1] - s;
        for (uint32_t a = 0; a < c; a++)
            counts[enc->gmem[s + a]].fetch_add(c - 1 - a, std::memory_order_relaxed);
    }
    uint64_t tot = 0;
    for (int i = 0; i < n; i++) {
        start[i] = (uint32_t)tot;
        tot += counts[i].load(std::memory_order_relaxed);
    }
    start[n] = (uint32_t)tot;
    std::vector<std::atomic<uint32_t>> fill((size_t)n);
    for (int i = 0; i < n; i++) fill[i].store(start[i], std::memory_order_relaxed);
#pragma omp parallel for schedule(dynamic, 64)
    for (long g = 0; g < G; g++) {
        const uint32_t s = enc->gstart[g], c = enc->gstart[g + 1] - s;
        for (uint32_t a = 0; a < c; a++) {
            const uint32_t i = enc->gmem[s + a];
            for (uint32_t b = a + 1; b < c; b++) js[fill[i].fetch_add(1, std::memory_order_relaxed)] = enc->gmem[s + b];
        }
    }
    return 0;
}

// Device-time model of one upper-triangle pass on MI355X, fitted (least
// squares on relative error) to the product and rare phases timed at
// tau = 16 .. 4096 on three long-tail sets (63,005 x 1,142, 12,000 x 2,000,
// 4,000 x 1,500); it is there to rank the tau candidates of one set, and its
// pick is the measured optimum on the first two sets and within 2% of it on
// the third:
//   MFMA binary product   n^2/2 * (K1 + K2) pair-bits at GEMM_RATE, K padded
//                         to the kernel's 512-bit slabs, plus the u16 output
//                         written once (OUT_BW)
//   rare-pair pass        each output cell staged through LDS once
//                         (read + write, STAGE_BW), R pairs at RARE_RATE and
//                         one member range per (row, rare call) at RANGE_RATE
static const double GEMM_RATE = 2.90e15;   // pair-bits per second (fp4 MFMA kernel)
static const double OUT_BW = 2.39e12;      // output bytes per second (2 B per pair), product epilogue
static const double STAGE_BW = 5.15e12;    // bytes per second (4 B per pair), rare-pass staging
static const double RARE_RATE = 2.54e12;   // rare pairs per second
static const double RANGE_RATE = 2.12e10;  // member ranges per second

extern "C" int cgh_choose_tau(const uint32_t* cnt, size_t ncnt, int n, int L, size_t out_bytes, size_t free_bytes,
                              void (*log)(void* ctx, const char* line), void* log_ctx) {
    const double n_pad = (double)up((size_t)std::max(n, 1), 256);
    const double k2 = (double)up((size_t)std::max(L, 1), 512);
    const double pairs = (double)n * (double)(n - 1) / 2.0;
    int best = 0;
    double best_t = 1e300;
    for (int tau = 2; tau <= 4096; tau *= 2) {
        size_t d = 0;
        double r = 0, members = 0, groups = 0;
        for (size_t t = 0; t < ncnt; t++) {
            const uint32_t x = cnt[t];
            if (!x) continue;
            if (x >= (uint32_t)tau) {
                d++;
            } else if (x >= 2) {
                r += (double)x * ((double)x - 1.0) / 2.0;
                members += x;
                groups += 1;
            }
        }
        const double ranges = members - groups;  // every member but a group's last has partners after it
        const double k1 = (double)up(std::max<size_t>(d, 1), 512);
        // operands + group members (4 B) + row index (8 B per member at most)
        // + row offsets + output buffers + slack
        const double mem = n_pad * (k1 + k2) / 8.0 + 12.0 * members + 4.0 * (n + 1.0) + (double)out_bytes +
                           (double)(64u << 20);
        const double t = pairs * (k1 + k2) / GEMM_RATE + pairs * 2.0 / OUT_BW +
                         (r > 0 ? pairs * 4.0 / STAGE_BW + r / RARE_RATE + ranges / RANGE_RATE : 0.0);
        const bool fits = mem <= (double)free_bytes && members < 4.0e9;
        if (log) {
            char line[256];
            snprintf(line, sizeof line, "  tau=%5d  one-hot=%8zu  rare=%14.0f  est %.4fs  vram %.2f GB%s", tau, d, r, t,
                     mem / 1e9, fits ? "" : "  (does not fit)");
            log(log_ctx, line);
        }
        if (fits && t < best_t) {
            best_t = t;
            best = tau;
        }
    }
    return best;
}

// ---------------------------------------------------------------------------
// Query codec of a resident database
// ---------------------------------------------------------------------------

// Slot of id v in a table of 2^(64 - shift) slots (Fibonacci hashing: the
// top bits of v times 2^64 / golden ratio).
static inline size_t codec_hash(uint64_t v, int shift) { return (size_t)((v * 0x9E3779B97F4A7C15ull) >> shift); }

struct cgh_codec {
    int n = 0, L = 0, tau = 0;
    size_t kw1 = 0, kw2 = 0, D = 0;
    // locus -> open-addressing table of the database's distinct ids:
    // 2^(64 - shift[k]) slots from slot off[k] on, at most half of them used,
    // id 0 (never a call) marking an empty slot.  A slot holds what a query
    // call with that id adds: its one-hot column x (y = ~0), or the member
    // range [x, y) of its rare list.
    struct Slot {
        uint64_t id;
        uint32_t x, y;
    };
    std::vector<Slot> slots;
    std::vector<size_t> off;
    std::vector<int> shift;
    std::vector<size_t> base;                // locus -> first one-hot column
    std::vector<std::vector<int32_t>> col;   // locus, rank -> local frequent column, or -1
    std::vector<std::vector<int64_t>> lst;   // locus, rank -> rare list number, or -1
    std::vector<uint32_t> lstart;            // list g = members[lstart[g], lstart[g+1])
    std::vector<uint32_t> members;           // database rows per rare allele, ascending
};

extern "C" void cgh_codec_free(cgh_codec* c) { delete c; }

extern "C" int cgh_codec_create(const uint64_t* calls, int n, int L, int tau, cgh_encoding* enc, cgh_codec** out,
                                char* err, int errlen) {
    *out = NULL;
    if (n < 1 || L < 1 || L > 65535 || tau < 1) {
        set_err(err, errlen, "cgh_codec_create: bad arguments (n=%d L=%d tau=%d)", n, L, tau);
        return 1;
    }
    std::vector<uint16_t> X((size_t)n * L);
    if (cgh_dense_ranks(calls, n, L, X.data(), err, errlen)) return 1;
    cgh_codec* c = new cgh_codec();
    c->n = n;
    c->L = L;
    c->tau = tau = std::min(tau, n + 1);
    std::vector<std::vector<uint64_t>> ids((size_t)L);  // locus -> sorted distinct ids; rank a <-> ids[k][a-1]
    c->col.resize((size_t)L);
    c->lst.resize((size_t)L);
    std::vector<std::vector<uint32_t>> cnt;
    locus_counts(X.data(), n, L, cnt);
    std::vector<size_t> nf((size_t)L, 0), nl((size_t)L, 0), nm((size_t)L, 0);
#pragma omp parallel for schedule(dynamic, 4)
    for (int k = 0; k < L; k++) {
        std::vector<uint64_t>& u = ids[k];
        u.reserve(cnt[k].size() - 1);
        for (int i = 0; i < n; i++)
            if (calls[(size_t)i * L + k]) u.push_back(calls[(size_t)i * L + k]);
        std::sort(u.begin(), u.end());
        u.erase(std::unique(u.begin(), u.end()), u.end());
        c->col[k].assign(cnt[k].size(), -1);
        c->lst[k].assign(cnt[k].size(), -1);
        // frequent columns as in cgh_encode; every other allele (singletons
        // included: a query may carry a database singleton) gets a list
        for (size_t a = 1; a < cnt[k].size(); a++) {
            if (cnt[k][a] >= (uint32_t)tau) {
                c->col[k][a] = (int32_t)nf[k]++;
            } else {
                c->lst[k][a] = (int64_t)nl[k]++;
                nm[k] += cnt[k][a];
            }
        }
    }
    c->base.assign((size_t)L + 1, 0);
    std::vector<size_t> l0((size_t)L + 1, 0), m0((size_t)L + 1, 0);
    for (int k = 0; k < L; k++) {
        c->base[k + 1] = c->base[k] + nf[k];
        l0[k + 1] = l0[k] + nl[k];
        m0[k + 1] = m0[k] + nm[k];
    }
    if (m0[L] >= 0xffffffffull) {
        delete c;
        set_err(err, errlen, "database rare lists exceed 2^32 members: raise tau");
        return 1;
    }
    if (c->base[L] >= 0xffffffffull) {  // columns are kept as u32 in the id tables
        delete c;
        set_err(err, errlen, "database one-hot exceeds 2^32 columns: raise tau");
        return 1;
    }
    c->D = c->base[L];
    c->kw1 = up(std::max<size_t>(c->D, 1), 256) / 32;
    c->kw2 = up((size_t)L, 256) / 32;
    c->lstart.assign(l0[L] + 1, 0);
    c->members.assign(m0[L], 0);
#pragma omp parallel
    {
        std::vector<uint32_t> cur;
#pragma omp for schedule(dynamic, 4)
        for (int k = 0; k < L; k++) {
            cur.assign(cnt[k].size(), 0);
            size_t m = m0[k];
            for (size_t a = 1; a < cnt[k].size(); a++)
                if (c->lst[k][a] >= 0) {
                    const int64_t g = c->lst[k][a] + (int64_t)l0[k];
                    c->lst[k][a] = g;
                    c->lstart[g] = (uint32_t)m;
                    cur[a] = (uint32_t)m;
                    m += cnt[k][a];
                }
            for (int i = 0; i < n; i++) {
                const uint16_t a = X[(size_t)i * L + k];
                if (a && c->lst[k][a] >= 0) c->members[cur[a]++] = (uint32_t)i;
            }
        }
    }
    c->lstart[l0[L]] = (uint32_t)m0[L];
    c->off.assign((size_t)L + 1, 0);
    c->shift.assign((size_t)L, 63);
    for (int k = 0; k < L; k++) {
        int bits = 1;
        while (((size_t)1 << bits) < 2 * ids[k].size()) bits++;
        c->shift[k] = 64 - bits;
        c->off[k + 1] = c->off[k] + ((size_t)1 << bits);
    }
    c->slots.assign(c->off[L], cgh_codec::Slot{0, 0, 0});
#pragma omp parallel for schedule(dynamic, 4)
    for (int k = 0; k < L; k++) {
        cgh_codec::Slot* t = c->slots.data() + c->off[k];
        const size_t mask = c->off[k + 1] - c->off[k] - 1;
        for (size_t a = 1; a <= ids[k].size(); a++) {
            size_t h = codec_hash(ids[k][a - 1], c->shift[k]);
            while (t[h].id) h = (h + 1) & mask;
            if (c->col[k][a] >= 0) {
                t[h] = cgh_codec::Slot{ids[k][a - 1], (uint32_t)(c->base[k] + (size_t)c->col[k][a]), ~0u};
            } else {
                const size_t g = (size_t)c->lst[k][a];
                t[h] = cgh_codec::Slot{ids[k][a - 1], c->lstart[g], c->lstart[g + 1]};
            }
        }
    }
    if (enc) {
        if (cgh_encode(X.data(), n, L, tau, 256, enc, err, errlen)) {
            delete c;
            return 1;
        }
        if (enc->kw1 != c->kw1 || enc->kw2 != c->kw2 || enc->D != c->D) {  // same column map by construction
            cgh_encoding_free(enc);
            delete c;
            set_err(err, errlen, "cgh_codec_create: column map disagrees with cgh_encode");
            return 1;
        }
    }
    *out = c;
    return 0;
}

extern "C" void cgh_codec_info(const cgh_codec* c, int* n, int* L, int* tau, size_t* kw1, size_t* kw2, size_t* D,
                               size_t* n_members) {
    if (n) *n = c->n;
    if (L) *L = c->L;
    if (tau) *tau = c->tau;
    if (kw1) *kw1 = c->kw1;
    if (kw2) *kw2 = c->kw2;
    if (D) *D = c->D;
    if (n_members) *n_members = c->members.size();
}

extern "C" const uint32_t* cgh_codec_members(const cgh_codec* c) { return c->members.data(); }

namespace {
// Query row r's L calls are row(r).
template <typename Row>
long codec_encode(const cgh_codec* c, Row row, int q, uint32_t* O, uint32_t* P, uint32_t* rstart, uint64_t* rentry,
                  size_t cap, char* err, int errlen) {
    const int L = c->L;
    if (q < 0 || cap < (size_t)q * (size_t)L) {
        set_err(err, errlen, "cgh_codec_encode: %d queries need room for %zu rare entries (have %zu)", q,
                (size_t)q * (size_t)L, cap);
        return -1;
    }
    // Row r writes its entries from slot r * L on; compacted below.  The
    // queries go in blocks of qb, locus by locus within a block, so that a
    // locus's id table is probed qb times while it is in cache; blocks are
    // spread over the threads, qb chosen so that every thread gets one.
    std::atomic<int> bad{0};
    const int qb = std::max(1, std::min(16, q / std::max(1, omp_get_max_threads())));
    const int nblk = (q + qb - 1) / qb;
#pragma omp parallel for schedule(dynamic, 1)
    for (int blk = 0; blk < nblk; blk++) {
        const int r0 = blk * qb, r1 = std::min(q, r0 + qb);
        uint32_t m[16] = {};
        for (int r = r0; r < r1; r++) {
            memset(O + (size_t)r * c->kw1, 0, c->kw1 * 4);
            memset(P + (size_t)r * c->kw2, 0, c->kw2 * 4);
        }
        // each call's home slot is hashed and its line prefetched one locus
        // ahead, so the block's probes of a locus overlap their cache misses
        size_t home[2][16];
        auto prefetch = [&](int k, size_t* hk) {
            const cgh_codec::Slot* t = c->slots.data() + c->off[k];
            for (int r = r0; r < r1; r++) {
                hk[r - r0] = codec_hash(row(r)[k], c->shift[k]);
                __builtin_prefetch(t + hk[r - r0]);
            }
        };
        prefetch(0, home[0]);
        for (int k = 0; k < L; k++) {
            const cgh_codec::Slot* t = c->slots.data() + c->off[k];
            const size_t mask = c->off[k + 1] - c->off[k] - 1;
            const size_t* hk = home[k & 1];
            if (k + 1 < L) prefetch(k + 1, home[(k + 1) & 1]);
            for (int r = r0; r < r1; r++) {
                const uint64_t v = row(r)[k];
                if (!v) continue;
                P[(size_t)r * c->kw2 + (k >> 5)] |= 1u << (k & 31);
                size_t h = hk[r - r0];
                while (t[h].id && t[h].id != v) h = (h + 1) & mask;
                if (!t[h].id) continue;  // an id the database never saw matches no row
                const uint32_t x = t[h].x, y = t[h].y;
                if (y == ~0u) {
                    O[(size_t)r * c->kw1 + (x >> 5)] |= 1u << (x & 31);
                } else {
                    if (x > y || y > c->members.size()) bad = 1;  // consumers read members[x, y)
                    rentry[(size_t)r * L + m[r - r0]++] = ((uint64_t)y << 32) | x;
                }
            }
        }
        for (int r = r0; r < r1; r++) rstart[r + 1] = m[r - r0];  // counts for now
    }
    if (bad) {
        set_err(err, errlen, "cgh_codec_encode: rare list outside the member array");
        return -1;
    }
    size_t tot = 0;
    rstart[0] = 0;
    for (int r = 0; r < q; r++) {
        const uint32_t m = rstart[r + 1];
        if (tot != (size_t)r * L) memmove(rentry + tot, rentry + (size_t)r * L, (size_t)m * 8);
        tot += m;
        rstart[r + 1] = (uint32_t)tot;
    }
    return (long)tot;
}
}  // namespace

extern "C" long cgh_codec_encode(const cgh_codec* c, const uint64_t* calls, int q, uint32_t* O, uint32_t* P,
                                 uint32_t* rstart, uint64_t* rentry, size_t cap, char* err, int errlen) {
    const size_t L = (size_t)c->L;
    return codec_encode(c, [calls, L](int r) { return calls + (size_t)r * L; }, q, O, P, rstart, rentry, cap, err,
                        errlen);
}

extern "C" long cgh_codec_encode_rows(const cgh_codec* c, const uint64_t* const* rows, int q, uint32_t* O,
                                      uint32_t* P, uint32_t* rstart, uint64_t* rentry, size_t cap, char* err,
                                      int errlen) {
    return codec_encode(c, [rows](int r) { return rows[r]; }, q, O, P, rstart, rentry, cap, err, errlen);
}
