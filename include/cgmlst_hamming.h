/*
 * cgmlst_hamming.h -- C ABI of the MI355X (gfx950) exact cgMLST allelic
 * Hamming distance library, libcgmlst_hamming.so.
 *
 *   H[i][j] = #loci where both alleles are called (non-zero) and differ
 *           = P P^T - (O O^T + R)
 *   P: n x L presence bits, O: n x D one-hot bits of alleles seen >= tau
 *   times, R[i][j]: loci where i and j share an allele seen 2..tau-1 times.
 *
 * Conventions of every entry point: plain C types, borrowed pointers (nothing
 * is retained after a call returns unless a handle says so), return 0 on
 * success and 1 on failure with a NUL-terminated message written to err
 * (truncated to errlen bytes; err may be NULL).  Bit matrices are row-major,
 * LSB-first within each byte (equivalently within each little-endian 32-bit
 * word), padding rows and columns zero.
 */
#ifndef CGMLST_HAMMING_H
#define CGMLST_HAMMING_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Reference-compatible boundary: the three symbols the reference host binds. */
/* ------------------------------------------------------------------------ */

/* Number of visible GPUs, 0 on error.
 * Replaces tensorcore/rs/cuda/tch.cu:68-71 (declared main.rs:86). */
int tch_device_count(void);

/* Free device memory in bytes on `device`, 0 on error.  The environment
 * variable CGH_FREE_BYTES, when set, overrides the figure (capacity planning
 * can then be exercised without a GPU).
 * Replaces tensorcore/rs/cuda/tch.cu:73-77 (declared main.rs:87). */
size_t tch_free_vram(int device);

/* Called once per stripe, in order r0 = 0, S, 2S, ...; stripe is rows x w u16
 * row-major (column c <-> j = r0 + c, w = n - r0), holding the final H for
 * c > r (rare pairs already subtracted, no cap, no layout); cells with c <= r
 * hold 0.  The buffer belongs to the library and is valid during the call.
 * Same type as tensorcore/rs/cuda/tch.cu:79 / main.rs:84. */
typedef void (*tch_cb)(void* ctx, int r0, int rows, size_t w, const uint16_t* stripe);

/* Whole upper-triangle pipeline on one GPU.  O: n_pad x K1 bits, P: n_pad x
 * K2 bits (K1, K2 multiples of 1024, n_pad a multiple of 256, n_pad >= n),
 * start[n+1]/js[n_rare]: CSR of rare partners j > i per row i, S: stripe rows
 * (multiple of 8).  *gpu_secs gets the pipeline wall time, *rare_on_gpu 1 if
 * the CSR fit in device memory (else rare pairs are subtracted on the host).
 * Replaces tensorcore/rs/cuda/tch.cu:82-170 (declared main.rs:88-92). */
int tch_run(int device, int n, size_t n_pad,
            const uint8_t* O, size_t K1, const uint8_t* P, size_t K2,
            const uint32_t* start, const uint32_t* js, size_t n_rare,
            int S, tch_cb cb, void* ctx, double* gpu_secs, int* rare_on_gpu,
            char* err, int errlen);

/* ------------------------------------------------------------------------ */
/* Host encoder (CPU, multithreaded).                                        */
/* ------------------------------------------------------------------------ */

/* Dense per-locus ranks of u64 allele ids: 0 stays 0 (missing), id -> 1 +
 * its index among the sorted distinct non-zero ids of its locus.  Fails with
 * "locus K has D distinct alleles (limit 65535)" like the reference.
 * Replaces tensorcore/rs/src/main.rs:135-165 (dense_ranks). */
int cgh_dense_ranks(const uint64_t* calls, int n, int L, uint16_t* X, char* err, int errlen);

/* Encoded operands of one profile set (host memory, owned by the library
 * until cgh_encoding_free).  Words are 32-bit, LSB-first. */
typedef struct cgh_encoding {
    int n, L;            /* profiles, loci */
    int tau;             /* frequency threshold used */
    size_t n_pad;        /* rows allocated in O and P (multiple of 256) */
    size_t D;            /* one-hot columns (frequent alleles) */
    size_t kw1, kw2;     /* 32-bit words per row of O and of P */
    uint32_t* O;         /* n_pad x kw1 words */
    uint32_t* P;         /* n_pad x kw2 words */
    size_t n_groups;     /* rare groups: (locus, allele) with 2 <= count < tau */
    size_t n_members;    /* total members over all groups */
    uint32_t* gstart;    /* n_groups + 1 offsets into gmem */
    uint32_t* gmem;      /* member rows, ascending within each group */
    uint64_t* gpairs;    /* n_groups + 1 prefix sums of c(c-1)/2 */
    uint64_t n_rare;     /* rare equal pairs = gpairs[n_groups] */
    /* The same pairs indexed by their first row: row i's rare calls that
     * have later group members, as ranges of gmem (the members j > i). */
    size_t n_rentries;   /* entries over all rows */
    uint32_t* rstart;    /* n + 1 offsets into rentry */
    uint64_t* rentry;    /* (end << 32) | begin, a range of gmem */
} cgh_encoding;

/* Encode dense ranks X (n x L u16) at threshold tau >= 1: O/P bits and rare
 * groups.  Rows of O and P are padded to a multiple of `k_align` bits (256,
 * 512 or 1024; the reference uses 1024).
 * Restates tensorcore/rs/src/main.rs:366-403 (column map, bit pack) and
 * 404-417 (rare groups; the reference expands them into a CSR). */
int cgh_encode(const uint16_t* X, int n, int L, int tau, int k_align, cgh_encoding* enc,
               char* err, int errlen);

/* cgh_encode without the O/P bit matrices (O = P = NULL; every size field
 * and the rare groups as cgh_encode gives them): the metadata a device-side
 * operand build (cgh_dataset_build) needs. */
int cgh_encode_meta(const uint16_t* X, int n, int L, int tau, int k_align, cgh_encoding* enc,
                    char* err, int errlen);

/* Column table of cgh_encode's one-hot map: colmap[colstart[k] + a] is the
 * O column of rank a at locus k (a = 0..max rank), or 0xFFFFFFFF for missing
 * and for alleles with count < tau.  colstart has L + 1 entries.  Returns the
 * number of colmap entries needed; fills the table only when cap is enough. */
size_t cgh_column_table(const uint16_t* X, int n, int L, int tau, uint32_t* colstart, uint32_t* colmap,
                        size_t cap);

/* Expand the rare groups of `enc` into the reference's CSR of partners j > i
 * per row (start[n+1], js[n_rare]; caller-allocated).  Fails with "too many
 * rare pairs: raise --tau" when n_rare >= 2^32.
 * Restates tensorcore/rs/src/main.rs:418-444. */
int cgh_rare_csr(const cgh_encoding* enc, uint32_t* start, uint32_t* js, char* err, int errlen);

void cgh_encoding_free(cgh_encoding* enc);

/* Per-locus allele counts of dense ranks X: counts for ranks 1..max of each
 * locus, concatenated (zeros included), written to cnt (capacity cap entries);
 * returns the number of entries needed.  Used by the tau cost model. */
size_t cgh_allele_counts(const uint16_t* X, int n, int L, uint32_t* cnt, size_t cap);

/* Cost model: tau in {2, 4, ..., 4096} minimising the modelled device time
 * of the MI355X kernels (MFMA product over D + L bit columns, row-local rare
 * pass over R pairs) subject to operands + rare index + out_bytes of output
 * buffers fitting in free_bytes.  Returns 0 when no tau fits ("no tau fits
 * in X GB of free VRAM").  `log` (may be NULL) receives the per-tau table.
 * Restates the selection rule of tensorcore/rs/src/main.rs:169-198 with
 * rates measured on MI355X instead of the L4's. */
int cgh_choose_tau(const uint32_t* cnt, size_t ncnt, int n, int L, size_t out_bytes, size_t free_bytes,
                   void (*log)(void* ctx, const char* line), void* log_ctx);

/* ------------------------------------------------------------------------ */
/* Device-resident dataset (no counterpart in the reference: its pipeline   */
/* re-uploads operands on every call).                                      */
/* ------------------------------------------------------------------------ */

typedef struct cgh_dataset cgh_dataset;

/* Upload O, P and the rare groups of `enc` to `device`. */
int cgh_dataset_upload(int device, const cgh_encoding* enc, cgh_dataset** out, char* err, int errlen);

/* The same dataset with the one-hot O kept in pinned host memory and split
 * along K: the product streams it through two device buffers of o_bytes / 2
 * (whole 512-bit slabs of every row; at least one slab each), one pass per
 * chunk, the upload of the next chunk overlapping the current pass.  Pass 0
 * stores P P^T - O_0 O_0^T; each later pass subtracts O_c O_c^T from the
 * stored u16 cells in place (every cell only falls towards the final value,
 * so none underflows); the -X cap goes with the last pass, the rare pairs
 * with the rare phase as usual.  The device then holds P, the rare groups,
 * the two chunks and the output, not O.  cgh_dataset_upper and _phase run
 * it with the MFMA kernels (CGH_KERNEL_VALU_POPC is refused); _stream,
 * _stream_graph, _autotune and the batch refuse such a dataset.  Lifts the
 * reference's limit that the one-hot fit in device memory
 * (README.md:126-129); no counterpart in the reference. */
int cgh_dataset_upload_split(int device, const cgh_encoding* enc, size_t o_bytes, cgh_dataset** out, char* err,
                             int errlen);

/* The same dataset from dense ranks X (n x L u16) at threshold tau, with the
 * O/P bits built on the device from X and the column table (no host bit
 * matrices: at 100,000 x 5,000 and tau = 1, O alone is 66 GB).  The rare
 * groups are encoded on the host as cgh_encode does. */
int cgh_dataset_build(int device, const uint16_t* X, int n, int L, int tau, cgh_dataset** out, char* err,
                      int errlen);

/* The encoder on the device: u64 allele ids (n x L, 0 = missing, n * L <
 * 2^31) straight to a dataset.  Per-locus dense ranks (main.rs:134-165, the
 * same error past 65,535 distinct alleles), allele counts, the O/P planes,
 * the rare groups (members ascending) and their row index are computed on
 * the device; only per-allele counts travel to the host, to choose tau
 * (tau = 0: cgh_choose_tau with out_bytes and the device's free memory) and
 * to lay out the column and group tables.  tau_used (may be NULL) receives
 * the threshold; X (may be NULL, n x L u16) receives the dense ranks; log
 * (may be NULL) receives cgh_choose_tau's table. */
int cgh_dataset_encode(int device, const uint64_t* calls, int n, int L, int tau, size_t out_bytes,
                       void (*log)(void* ctx, const char* line), void* log_ctx, cgh_dataset** out, int* tau_used,
                       uint16_t* X, char* err, int errlen);

/* Cap every later output of the dataset (upper, phase = both phases, stream,
 * batch): a value above cap is written as cap, cap = 0 (the default) none;
 * main.rs:251-253 applied on the device (in the product epilogue for rows
 * without rare pairs, in the rare pass for the others).  Returns 1 if
 * cap < 0. */
int cgh_dataset_set_cap(cgh_dataset* ds, int cap);

/* Autotuner: time the product kernel candidates (fp4 and int8 MFMA x the
 * 256-row tile, 256 x 128 with a 2-slab LDS ring, and the 128 x 128 tile with
 * rings of 2, 3 and 4 slabs: 8 configurations) on the
 * dataset's first stripe of S rows and keep the fastest for CGH_KERNEL_AUTO
 * in every later call on this dataset (upper, phase, stream; a batch is
 * tuned as a whole by cgh_batch_autotune).  The stripe height is not a
 * candidate: it fixes the caller's output layout.  Outputs
 * (may be NULL): the kernel (CGH_KERNEL_*), the tile edge and its time in
 * ms.  No counterpart in the reference (its CUTLASS tile is fixed at build
 * time, tensorcore/README.md:31-32). */
int cgh_dataset_autotune(cgh_dataset* ds, int S, int* kernel, int* tile, double* ms, char* err, int errlen);

/* Sizes of a dataset: profiles, loci, tau, one-hot columns, rare equal
 * pairs (any pointer may be NULL). */
void cgh_dataset_info(const cgh_dataset* ds, int* n, int* L, int* tau, size_t* D, uint64_t* n_rare);
void cgh_dataset_free(cgh_dataset* ds);

/* u16 elements of the upper-triangle stripe layout for n profiles and stripe
 * height S: stripe s (rows r0 = s*S .. r0+rows-1) is rows x (n - r0) u16
 * row-major, stripes concatenated; rounded up to a multiple of 8 (whole
 * 16-byte words, which the rare-pair pass stages). */
size_t cgh_upper_elems(int n, int S);

/* Offset (u16 elements) of stripe s in that layout. */
size_t cgh_stripe_offset(int n, int S, int s);

/* Binary-product kernels.  Both are exact integer arithmetic: int8 MFMA with
 * int32 accumulation, or fp4 (e2m1, values 0 and +-1) MFMA with fp32
 * accumulation of integers whose magnitude never exceeds L <= 65,535. */
#define CGH_KERNEL_AUTO 0     /* the faster one on MI355X (fp4) */
#define CGH_KERNEL_MFMA_I8 1  /* v_mfma_i32_32x32x32_i8 */
#define CGH_KERNEL_MFMA_FP4 2 /* v_mfma_scale_f32_32x32x64_f8f6f4, fp4 operands */
#define CGH_KERNEL_VALU_POPC 3 /* VALU AND + v_bcnt_u32_b32 on 32-bit words */

/* Exact H of the whole upper triangle into device memory d_out (at least
 * cgh_upper_elems(n, S) u16), stripe layout, cells with j <= i set to 0.
 * Enqueued on `stream` (a hipStream_t, NULL = default stream); returns after
 * enqueueing.  kernel: one of CGH_KERNEL_*. */
int cgh_dataset_upper(cgh_dataset* ds, int S, uint16_t* d_out, void* stream, int kernel,
                      char* err, int errlen);

/* The two phases of cgh_dataset_upper separately (same stream order rules):
 * CGH_PHASE_PRODUCT writes P P^T - O O^T (every cell of the layout),
 * CGH_PHASE_RARE then subtracts the rare equal pairs in place. */
#define CGH_PHASE_PRODUCT 1
#define CGH_PHASE_RARE 2
int cgh_dataset_phase(cgh_dataset* ds, int S, uint16_t* d_out, void* stream, int kernel, int phases,
                      char* err, int errlen);

/* Exact H streamed to the host in stripes of S rows (multiple of 8): the
 * callback contract of tch_run (stripe order, rows x (n - r0) u16 final H
 * for c > r, 0 for c <= r, library-owned buffer valid during the call), with
 * the rare pairs taken from the dataset's groups instead of a per-pair CSR.
 * *gpu_secs (may be NULL) gets the pipeline wall time.  The CLI's path;
 * no counterpart in the reference beyond tch_run's shape. */
int cgh_dataset_stream(cgh_dataset* ds, int S, tch_cb cb, void* ctx, double* gpu_secs, char* err, int errlen);

/* The stripes of cgh_dataset_stream captured as one HIP graph and launched
 * once: per stripe the product and rare launches, the copy to pinned host
 * memory and a host node that calls cb, with the same three-buffer ring (the
 * compute of stripe s + 3 depends on the host node of stripe s).  Same
 * output and stripe order; the callback runs on a HIP runtime thread, one
 * stripe at a time while the caller waits, and must not call HIP.
 * *build_ms (may be NULL) gets the capture and instantiation time.  No
 * counterpart in the reference (its stripe loop is eager, tch.cu:113-149). */
int cgh_dataset_stream_graph(cgh_dataset* ds, int S, tch_cb cb, void* ctx, double* gpu_secs, double* build_ms,
                             char* err, int errlen);

/* ------------------------------------------------------------------------ */
/* Many independent datasets in one persistent launch (no counterpart in the */
/* reference, which runs one dataset per process).                          */
/* ------------------------------------------------------------------------ */

typedef struct cgh_batch cgh_batch;

/* Batch of `count` uploaded datasets (borrowed: they must outlive the batch)
 * sharing one output arena: dataset d's upper-triangle stripe layout (stripe
 * height S) starts at element cgh_batch_out_offset(b, d).  tile: 0 = chosen
 * per dataset (256 for n >= 4096, else 128), or 128 / 256 for all.  All their
 * tiles go into two device work queues (one per tile size), longest first. */
int cgh_batch_create(int device, cgh_dataset* const* sets, int count, int S, int tile, cgh_batch** out,
                     char* err, int errlen);
size_t cgh_batch_out_elems(const cgh_batch* b);
size_t cgh_batch_out_offset(const cgh_batch* b, int d);
int cgh_batch_tile(const cgh_batch* b, int d);

/* Enqueue the whole batch on `stream`: one persistent product launch per
 * tile size pulling tiles from the queue, then one rare-pair launch per
 * dataset, spread over `stream` and three side streams of the batch (they
 * fork from and join back into `stream` by events).  use_graph != 0 (with a
 * non-NULL stream) captures that sequence into a HIP graph on the first call
 * and replays it while d_out, stream and kernel stay the same.  d_out:
 * cgh_batch_out_elems(b) u16, 16-byte aligned.  Calls on one batch object
 * must not overlap (they share its side streams and events). */
int cgh_batch_run(cgh_batch* b, uint16_t* d_out, void* stream, int kernel, int use_graph, char* err, int errlen);
void cgh_batch_free(cgh_batch* b);

/* Autotuner of a whole batch: times it (eager runs into d_out on `stream`)
 * under each tile rule (256-tiles for every dataset, for datasets from
 * 2,048 / 4,096 / 8,192 / 16,384 rows, or none) and each MFMA product form
 * (fp4, int8), and keeps the fastest for later cgh_batch_run calls with
 * CGH_KERNEL_AUTO (a captured graph is re-captured).  Outputs (may be NULL):
 * the row threshold for 256-tiles (0 = all, INT32_MAX = none), the kernel
 * and its time in ms.  No counterpart in the reference (its tile shape is a
 * compile-time constant, tensorcore/README.md:31-32). */
int cgh_batch_autotune(cgh_batch* b, uint16_t* d_out, void* stream, int* threshold, int* kernel, double* ms,
                       char* err, int errlen);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): cgh_batch_create + cgh_batch_run (eager) in one call
 * that allocates nothing.  The job table, the two tile queues and their
 * counters live in a static device array of CGH_NOALLOC_SCRATCH_BYTES that
 * every GPU the library runs on holds until the process ends; every call on
 * a GPU rewrites its copy from the start, so calls that overlap on one GPU
 * race (serialise them, e.g. on one stream).  At most CGH_NOALLOC_MAX_SETS
 * datasets and CGH_NOALLOC_MAX_TILES tiles (the triangle tiles of every
 * dataset at its batch tile edge, 256 from 4,096 rows, else 128); S a
 * positive multiple of 256.  d_out: 16-byte aligned, dataset d's stripe
 * layout at element sum_{e < d} cgh_upper_elems(n_e, S) (cgh_batch_create's
 * layout); all launches, rare passes included, go to `stream`. */
#define CGH_NOALLOC_MAX_SETS 256
#define CGH_NOALLOC_MAX_TILES (1 << 20)
#define CGH_NOALLOC_SCRATCH_BYTES ((size_t)17 << 20)
int cgh_batch_run_noalloc(cgh_dataset* const* sets, int count, int S, uint16_t* d_out, void* stream, int kernel,
                          char* err, int errlen);

/* ------------------------------------------------------------------------ */
/* Resident database + continuous-batching query server (no counterpart in   */
/* the reference).                                                           */
/* ------------------------------------------------------------------------ */

typedef struct cgh_codec cgh_codec;
typedef struct cgh_db cgh_db;
typedef struct cgh_server cgh_server;

/* Host codec of a database of n profiles x L loci (u64 allele ids, 0 =
 * missing) at threshold tau (clamped to n + 1): per-locus id dictionaries,
 * the frequent-column map of cgh_encode (k_align 256) and, for every other
 * allele (count < tau, singletons included), the ascending list of database
 * rows that carry it, all lists concatenated into one member array.  enc (may
 * be NULL) receives the database's own cgh_encode output. */
int cgh_codec_create(const uint64_t* calls, int n, int L, int tau, cgh_encoding* enc, cgh_codec** out, char* err,
                     int errlen);
void cgh_codec_info(const cgh_codec* c, int* n, int* L, int* tau, size_t* kw1, size_t* kw2, size_t* D,
                    size_t* n_members);
const uint32_t* cgh_codec_members(const cgh_codec* c);

/* Queries (q x L u64 ids) in the database's columns: rows 0..q-1 of O (kw1
 * words each) and P (kw2 words each) are overwritten; rstart[0..q] indexes
 * rentry, whose entries (end << 32 | begin) are member ranges of the rare
 * alleles each query shares with the database (ids the database never saw
 * match no row).  rentry needs room for q * L entries.  Returns the number
 * of entries, or -1 with err set.  Then for query r and database row j:
 * H = |P_r & P_j| - |O_r & O_j| - #{ranges of r holding j}. */
long cgh_codec_encode(const cgh_codec* c, const uint64_t* calls, int q, uint32_t* O, uint32_t* P, uint32_t* rstart,
                      uint64_t* rentry, size_t cap, char* err, int errlen);
/* cgh_codec_encode with query r's L calls at rows[r] (queries from several
 * buffers encoded in one call, as the server does for a batch of requests).
 * The library's own; no counterpart in the reference. */
long cgh_codec_encode_rows(const cgh_codec* c, const uint64_t* const* rows, int q, uint32_t* O, uint32_t* P,
                           uint32_t* rstart, uint64_t* rentry, size_t cap, char* err, int errlen);
void cgh_codec_free(cgh_codec* c);

/* Database of n profiles x L loci (u64 allele ids, 0 = missing) resident on
 * `device`: its P/O bits at threshold tau and the member lists of its rare
 * alleles (count < tau, singletons included) stay in HBM; the per-locus id
 * dictionaries stay on the host to place query ids in the same columns. */
int cgh_db_create(int device, const uint64_t* calls, int n, int L, int tau, cgh_db** out, char* err, int errlen);
int cgh_db_info(const cgh_db* db, int* n, int* L, int* tau, size_t* D);
void cgh_db_free(cgh_db* db);

/* Server over a database: a worker thread keeps up to `slots` query batches
 * in flight (one HIP stream each).  Waiting requests are admitted in arrival
 * order as slots and device memory free up; every admitted batch packs as
 * many whole requests as fit in max_batch queries and in the HBM budget left
 * by the batches in flight (hbm_budget bytes, 0 = what free memory allows),
 * and runs as one query x database product, rare-pair pass and top-10
 * selection.  kernel: CGH_KERNEL_*.  Calls that overlap on one server are
 * safe (submit and wait may come from any thread). */
int cgh_server_create(cgh_db* db, int slots, int max_batch, size_t hbm_budget, int kernel, cgh_server** out,
                      char* err, int errlen);
int cgh_server_slots(const cgh_server* sv);

/* Submit q query profiles (q x L u64 ids, copied); returns a ticket >= 0, or
 * -1 with err set.  flags: CGH_WANT_DIST to get the q x n distances back as
 * well as the nearest rows (without it they stay on the device). */
#define CGH_WANT_DIST 1
long cgh_server_submit(cgh_server* sv, const uint64_t* calls, int q, int flags, char* err, int errlen);

/* Block until the ticket's request is done; copy its q x n u16 distances
 * (dist, may be NULL; non-NULL needs CGH_WANT_DIST at submit), q x 10
 * nearest database rows ascending by (distance, row) (top_d u16 / top_i u32,
 * may be NULL; rows past n read 0xffff / 0xffffffff) and its submit-to-done
 * latency in seconds.  Each ticket is waited on once. */
int cgh_server_wait(cgh_server* sv, long ticket, uint16_t* dist, uint16_t* top_d, uint32_t* top_i,
                    double* latency_s, char* err, int errlen);
void cgh_server_free(cgh_server* sv);

/* Measured peaks of the units the kernels use, from register-only loops on
 * `device`: int8 MFMA (TOP/s), fp4 MFMA (TFLOP/s), VALU AND + popcount
 * (TOP/s, 2 ops per bit pair) and a streaming copy (GB/s, read + write).
 * Any output pointer may be NULL.  Same role as tensorcore/peak.cu:8-54. */
int cgh_measure_peaks(int device, double* i8_tops, double* fp4_tflops, double* popc_tops, double* hbm_gbs,
                      char* err, int errlen);

/* Measured memory rates on `device` (GB/s): a streaming copy of 4 GiB
 * (read + write counted), a read-only stream of 4 GiB, and a read-only
 * stream over 64 MiB repeated 40 times (held in the Infinity Cache, larger
 * than the L2s: the rate at which member lists and other reused data reach
 * the CUs).  Any output pointer may be NULL.  No counterpart in the
 * reference. */
int cgh_measure_memory(int device, double* copy_gbs, double* read_gbs, double* cache_read_gbs, char* err,
                       int errlen);

/* Tile shape (rows = cols) used by cgh_dataset_upper for this dataset. */
int cgh_dataset_tile(const cgh_dataset* ds, int kernel);

/* The product configuration CGH_KERNEL_AUTO runs on this dataset: kernel
 * (CGH_KERNEL_*), tile edge and LDS ring depth in 512-bit slabs (2, 3 or 4;
 * cgh_dataset_autotune chooses among them).  Any pointer may be NULL. */
void cgh_dataset_tuning(const cgh_dataset* ds, int* kernel, int* tile, int* stages);

#ifdef __cplusplus
}
#endif

#endif
