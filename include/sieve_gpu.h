/*
 * sieve_gpu.h — C ABI of the MI355X search path for sieve (IVF + SOAR + RaBitQ 1-bit LUT
 * scan + int8 rerank).
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, returns 0 on success
 * and a negative status on failure (the message is in sieve_last_error()), and never lets
 * a C++ exception cross the boundary. Array arguments may be host memory or device memory
 * (HIP allocations, e.g. a torch tensor on cuda:0); the library detects which with
 * hipPointerGetAttributes and copies host arrays in and out itself. The caller owns every
 * array it passes and the library never keeps a pointer to one after the call returns,
 * except where sieve_index_desc says "borrowed".
 *
 * Threading mirrors the reference's `Index: &self` / per-thread `Scratch` split
 * (src/ivf.rs:498-519, src/bin/search.rs:279-295): a sieve_index is immutable after
 * creation and may be shared by any number of host threads; each host thread owns one
 * sieve_stream (a HIP stream plus its scratch buffers), and calls on one stream are
 * serialised by the caller.
 *
 * Reference citations are into the sieve crate (18Abhinav07/A-Game-Of-Vectors).
 */
#ifndef SIEVE_GPU_H
#define SIEVE_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Compile-time widths (src/lib.rs:21-26, src/lut.rs:16-18, src/ivf.rs:18). */
#define SIEVE_DIM 768
#define SIEVE_PDIM 256
#define SIEVE_CODE_BYTES 32
#define SIEVE_BLOCK 32
#define SIEVE_PAD 0xFFFFFFFFu

/* Status codes. */
#define SIEVE_OK 0
#define SIEVE_EINVAL (-1)   /* bad argument, shape or file contents            */
#define SIEVE_EHIP (-2)     /* a HIP runtime call failed                      */
#define SIEVE_ENOMEM (-3)   /* device allocation failed / over the HBM budget */
#define SIEVE_EIO (-4)      /* index file missing or unreadable               */
#define SIEVE_ELIMIT (-5)   /* a size beyond a documented kernel limit        */

/* Rerank modes: the tags of shard::Header (src/shard.rs:51-52) for Rerank
 * (src/ivf.rs:426-439). */
#define SIEVE_RERANK_NONE 0
#define SIEVE_RERANK_PREFIX 1
#define SIEVE_RERANK_FULL 2
#define SIEVE_RERANK_TIERED 3
#define SIEVE_RERANK_SSD 4

/* Query parameters: ivf::Params (src/ivf.rs:442-460), laid out like shard::Header
 * (src/shard.rs:48-62) plus the two routing options the header leaves out.
 *   prune_m == 0        -> prune = None, else Some((prune_m, prune_w))
 *   margin is NaN       -> margin = None
 *   budget == 0         -> budget = None                                     */
typedef struct sieve_params {
    uint32_t cand;
    uint32_t rerank;
    uint32_t rerank_r;
    uint32_t k;
    uint32_t prune_m;
    float prune_w;
    uint32_t nprobe;
    float margin;
    uint64_t budget;
} sieve_params;

/* Stage counters: ivf::Timing (src/ivf.rs:522-538), same field order. Times are
 * nanoseconds of GPU time measured with HIP events on the call's stream; counts are
 * summed over the queries of a call. */
typedef struct sieve_timing {
    uint64_t route;
    uint64_t scan;
    uint64_t rerank;
    uint64_t tier2;
    uint64_t scanned;
    uint64_t reranked;
    uint64_t probed;
    uint64_t fetch;
    uint64_t wait;
} sieve_timing;

/* An index in the reference's on-disk layout, as arrays. The IVF build output
 * (src/ivf.rs:354-391) and the ingest output (src/bin/ingest.rs:99-166):
 *   centroids  k x 256 f32           centroids.f32
 *   block_off  k + 1 u64             block_off.u64
 *   blocks     nblocks x 1024 B      blocks.bin   (nibble-transposed, src/lut.rs:73-83)
 *   ids        nblocks x 32 u32      ids.u32      (PAD = 0xFFFFFFFF)
 *   factors    nblocks x 32 x 2 f32  factors.f32  ([mul, err] per slot; NULL = global
 *                                                  sign codes, no RaBitQ)
 *   rotation   3 x 256 f32           rotation.f32 (NULL without RaBitQ)
 *   rerank     n x 768 i8            rerank.i8
 *   quant      768 f32               quant.f32    (per-dimension int8 scales)
 *   norm256    n f32                 norm256.f32
 *   tier2_*    optional 2-bit tier   tier2.sigma (256 f32), tier2.bin (n x 64 B),
 *                                    tier2.norm (n f32)
 * All arrays are copied into the library's device layout at creation except `rerank`,
 * which is borrowed (used in place, no copy) when it is device memory and
 * borrow_rerank is nonzero: the caller then keeps it alive until sieve_index_free. */
typedef struct sieve_index_desc {
    uint32_t k;
    uint64_t n;
    uint64_t nblocks;
    const float* centroids;
    const uint64_t* block_off;
    const uint8_t* blocks;
    const uint32_t* ids;
    const float* factors;
    const float* rotation;
    const int8_t* rerank;
    const float* quant;
    const float* norm256;
    const float* tier2_sigma;
    const uint8_t* tier2_codes;
    const float* tier2_norm;
    uint32_t borrow_rerank;
} sieve_index_desc;

/* What a query sees beyond the immutable base: ivf::View (src/ivf.rs:462-496).
 * deleted: bitset over row ids, a set bit hides the row (tombstones, src/live.rs:157-158).
 * allow:   bitset over row ids; when present only rows with a set bit are returned
 *          (filters, src/filter.rs:18-39). allow_stride == 0 means one bitset for every
 *          query of the call; otherwise query i uses allow + i * allow_stride words and
 *          allow_flags[i] (nonzero = filtered) says whether query i carries a filter.
 * Word counts bound the bitsets: a row past the end reads as clear
 * (src/ivf.rs:474-477). */
typedef struct sieve_view {
    const uint64_t* deleted;
    uint64_t deleted_words;
    const uint64_t* allow;
    uint64_t allow_words;
    uint64_t allow_stride;
    const uint8_t* allow_flags;
} sieve_view;

/* Counters of the last call on a stream (no counterpart in the reference): what the kernels
 * read, for roofline accounting. Times are GPU nanoseconds between HIP events recorded on the
 * stream around each stage's kernels. */
typedef struct sieve_stats {
    uint64_t queries;
    uint64_t probed;     /* clusters scanned, summed over queries                        */
    uint64_t scanned;    /* stored slots scanned (32 code bytes + 4 mul bytes each)       */
    uint64_t id_reads;   /* slots past the admission threshold (their id is read)        */
    uint64_t reranked;   /* candidates reranked after pruning                            */
    double route_ns;
    double scan_ns;
    double rerank_ns;
    uint64_t rescanned;  /* cluster-major scan: queries whose appended list overflowed       */
    uint64_t graph;      /* sieve_search_batch on a stream with graphs: 0 eager, 1 captured, 2 replayed */
    uint64_t cm_pairs;   /* cluster-major scan: (query, slot) pairs its MFMA pass scored          */
    double cm_ns;        /* GPU time of the MFMA pass's launches (HIP events around each)         */
    uint64_t cm_launches; /* MFMA pass launches timed (one per work-list phase)                   */
    uint64_t cm_staged;  /* pairs the MFMA pass's integer bound let through to the exact estimate  */
} sieve_stats;

typedef struct sieve_index sieve_index;
typedef struct sieve_stream sieve_stream;

typedef struct sieve_index_info {
    uint32_t k;
    uint64_t n;
    uint64_t nblocks;
    uint32_t rabitq;
    uint32_t tier2;
    uint32_t max_cluster_blocks;
    uint64_t device_bytes;
} sieve_index_info;

/* ---------------------------------------------------------------- runtime --- */

/* Message of the last failing call on this host thread ("" if none). */
const char* sieve_last_error(void);

/* A HIP stream plus per-stream scratch on the current device (one per host thread). */
int sieve_stream_create(sieve_stream** out);
void sieve_stream_destroy(sieve_stream* s);
int sieve_stream_sync(sieve_stream* s);
/* Runs sieve_search_batch on this stream as HIP graphs (enable != 0): a call repeated with the
 * same index settings, batch size, parameters, view and device addresses replays the captured
 * launches (the first such call runs eagerly, the second captures). Results are those of the
 * eager path. Meant for serving loops with fixed staging buffers; no counterpart in the
 * reference (a launch-overhead knob of this library). 0 disables and frees the graphs. */
int sieve_stream_set_graphs(sieve_stream* s, int enable);
/* The underlying hipStream_t, for callers that order their own work with it. */
void* sieve_stream_handle(sieve_stream* s);
/* Counters of the last route / candidates / rerank / search call on this stream. */
int sieve_stream_last_stats(const sieve_stream* s, sieve_stats* out);

/* ---------------------------------------------------------- kernel seams --- */

/* C = A * B^T, row-major f32; a is m x k, b is n x k, c is m x n.
 * Replaces blas::gemm_abt (src/blas.rs:5-10), the crate's only FFI on the path. Computed
 * with f32 MFMA (v_mfma_f32_16x16x4_f32) and f32 accumulation. */
int sieve_gemm_abt(const float* a, const float* b, size_t m, size_t n, size_t k, float* c);

/* Scores every 32-slot block of nibble-transposed codes against G 16-entry u8 tables:
 * out[32*b + j] = sum_g tables[g][nibble_g(slot j of block b)].
 * Replaces lut::scan (src/lut.rs:96-103); bit-exact with score_code (src/lut.rs:58-61).
 * groups is 64 (256-bit codes) or 32 (128-bit codes, src/bigann.rs:53-56). */
int sieve_lut_scan(const uint8_t* tables, uint32_t groups, const uint8_t* blocks,
                   size_t nblocks, uint16_t* out);

/* ------------------------------------------------------------ the index --- */

/* Index::open (src/ivf.rs:553-578) + Corpus::open (src/ivf.rs:75-90): reads the build
 * directory and the ingest directory and uploads them to the current device. The directory of
 * a compacted index (sieve_live_compact) holds a row segment (live.*) whose rows its lists
 * reference: such an index is searched through the live state sieve_live_open_dir opens. */
int sieve_index_open(const char* index_dir, const char* ingest_dir, sieve_index** out);
/* The same from arrays already in memory (host or device). */
int sieve_index_create(const sieve_index_desc* desc, sieve_index** out);
void sieve_index_free(sieve_index* ix);
int sieve_index_get_info(const sieve_index* ix, sieve_index_info* out);

/* Scan-kernel autotuner (no counterpart in the reference): times each variant of the scan
 * kernel (slots per thread per round, prefetch of the next round's codes; variants 10 and 11
 * are the cluster-major scan) on this batch of queries with these parameters, and keeps the
 * fastest for later calls on this index. Call it before sharing the index across threads.
 * chosen (may be NULL) gets the variant, and ms_per_variant (NULL or SIEVE_SCAN_VARIANTS
 * floats) the time of each. */
#define SIEVE_SCAN_VARIANTS 12
int sieve_index_autotune(sieve_index* ix, sieve_stream* s, const float* q768, const float* q256, size_t nq,
                         const sieve_params* p, uint32_t* chosen, float* ms_per_variant);
/* Fixes the scan variant (0 .. SIEVE_SCAN_VARIANTS - 1) instead of tuning it. */
int sieve_index_set_scan_variant(sieve_index* ix, uint32_t variant);
/* Work split of the scan: a query's probed blocks are cut into chunks of about chunk_blocks
 * 32-slot blocks, each scanned by its own workgroup and merged on the device; 0 (the default)
 * sizes chunks per batch (whole queries in large batches, many chunks for a lone query).
 * Results do not depend on the split. No counterpart in the reference (one thread scans a
 * whole query there). */
int sieve_index_set_scan_chunk(sieve_index* ix, uint32_t chunk_blocks);
/* The cluster-major scan (scan variants 10 and 11; no counterpart in the reference, where one
 * thread scans one query): a batch's (cluster, query) pairs are grouped by cluster, and each
 * block chunk of a cluster is read once for up to 32 of its queries and scored with int8
 * MFMA. A seed pass first scans each query's nearest probes (whole clusters, at least
 * seed_slots stored slots) for an admission threshold; later slots above it are appended to a
 * per-query list of ext_cap entries. The work list runs in `phases` launches; after each, a
 * query's best cand entries so far raise its threshold. A query whose list overflows is
 * rescanned alone (sieve_stats.rescanned). Used for batches of at least min_queries queries
 * without an allow filter; smaller unfiltered batches of at most the small-batch limit run the
 * small-batch scan (sieve_index_set_small_batch), the rest scan variant 2. Results equal the
 * per-query scan's. 0 keeps a setting's default (min_queries 64, seed_slots 24 * cand, ext_cap
 * 64 * cand, chunk_blocks 128 for variant 10 and 64 for variant 11, phases 2). */
int sieve_index_set_cluster_scan(sieve_index* ix, uint32_t min_queries, uint32_t seed_slots, uint32_t ext_cap,
                                 uint32_t chunk_blocks, uint32_t phases);
/* The small-batch scan of the cluster-major variants (no counterpart in the reference): for an
 * unfiltered batch of at most max_queries queries, every 32-block chunk of a query's probes
 * keeps its best 16 admissible slots, the (cand + 1)-th best of those is the query's
 * admission threshold, every slot above it is appended to the query's list (ext_cap entries,
 * as above) and the list is merged into the query's candidates; an overflowing list is
 * rescanned. Results equal the per-query scan's. max_queries 0 disables it (default 12). */
int sieve_index_set_small_batch(sieve_index* ix, uint32_t max_queries);
/* Routing (no counterpart in the reference; probe lists are the same in every mode): whether
 * the routing GEMM writes every (query, centroid) score or only the maxima of 16 consecutive
 * scores, the select then recomputing from the BF16 operands the scores of the tiles it
 * reads. mode < 0 (default): fused when the select runs over tile maxima and nprobe * 3072 <=
 * k (the recomputed rows cost less than writing the scores); 0: never; > 0: whenever the select
 * runs over tile maxima. */
int sieve_index_set_route_fuse(sieve_index* ix, int mode);

/* ------------------------------------------------------------- the path --- */

/* Index::route (src/ivf.rs:587-613) for nq queries. q256 is nq x 256 (unit rows).
 * Writes probe lists best first: scores[i*nprobe + j], clusters[i*nprobe + j] for
 * j < counts[i] (counts[i] <= min(nprobe, k)). Routing runs as a BF16 MFMA GEMM with f32
 * accumulation; every centroid within the BF16 error bound of the nprobe-th score is
 * rescored in f32 exactly as the reference's gemm_abt fallback sums (src/blas.rs:47-53),
 * so scores and the probe set are the reference's. */
int sieve_route(const sieve_index* ix, sieve_stream* s, const float* q256, size_t nq,
                const sieve_params* p, float* scores, uint32_t* clusters, uint32_t* counts,
                sieve_timing* t);

/* Index::candidates (src/ivf.rs:643-729) for nq queries over given probe lists (as
 * sieve_route writes them, stride nprobe_stride). Writes, per query, up to p->cand
 * (estimate, row, error half-width) triples deduplicated by row, in row order:
 * est/rows/errs[i*p->cand + j] for j < counts[i]. */
int sieve_candidates(const sieve_index* ix, sieve_stream* s, const float* q256, size_t nq,
                     const float* probe_scores, const uint32_t* probe_clusters,
                     const uint32_t* probe_counts, size_t nprobe_stride,
                     const sieve_params* p, const sieve_view* v, float* est, uint32_t* rows,
                     float* errs, uint32_t* counts, sieve_timing* t);

/* Index::rerank (src/ivf.rs:731-815) for nq queries over candidate lists (as
 * sieve_candidates writes them, stride cand_stride). Writes the top p->k (score, row)
 * best first, ties by row ascending: scores/rows[i*p->k + j] for j < counts[i]. */
int sieve_rerank(const sieve_index* ix, sieve_stream* s, const float* q768, const float* q256,
                 size_t nq, const float* est, const uint32_t* rows, const float* errs,
                 const uint32_t* cand_counts, size_t cand_stride, const sieve_params* p,
                 const sieve_view* v, float* scores, uint32_t* out_rows, uint32_t* counts,
                 sieve_timing* t);

/* Index::search / search_view (src/ivf.rs:627-636) for a batch: route, scan and rerank
 * of nq queries in one call, all stages on the device. Output as sieve_rerank. */
int sieve_search_batch(const sieve_index* ix, sieve_stream* s, const float* q768,
                       const float* q256, size_t nq, const sieve_params* p,
                       const sieve_view* v, float* scores, uint32_t* rows, uint32_t* counts,
                       sieve_timing* t);

/* sieve_search_batch without any device allocation. No counterpart in the reference: it is
 * for callers that cannot call hipMalloc (code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation). Every scratch buffer of the search is carved from one
 * fixed device array of 64 MiB that each GPU the library runs on holds until the process ends;
 * a batch whose buffers do not fit fails with SIEVE_ELIMIT (at 16384 clusters and 1000
 * candidates the limit is around a hundred queries per call; fewer with larger probe lists).
 * Queries, view bitsets and outputs must be device memory. The search is enqueued on s's stream
 * and the call returns without waiting: results are ready once that stream is synchronised
 * (sieve_stream_sync), and no counters are recorded. Results equal sieve_search_batch's. All
 * calls on one GPU share the array, so calls in flight at the same time on one device race:
 * serialise them (one stream, or events). */
int sieve_search_batch_noalloc(const sieve_index* ix, sieve_stream* s, const float* q768,
                               const float* q256, size_t nq, const sieve_params* p,
                               const sieve_view* v, float* scores, uint32_t* rows, uint32_t* counts);

/* ------------------------------------------------------------- serving loop --- */

/* No counterpart in the reference, whose server is a thread per connection over TCP
 * (src/shard.rs:156-201). Continuous batching: requests submitted one at a time queue in
 * host memory; a batcher thread launches everything pending (up to max_batch) as one device
 * batch whenever a batch slot is free, on that slot's own stream, so batches overlap on the
 * device; a completer thread retires them. The number of slots is bounded by hbm_budget
 * bytes of device scratch (0: a quarter of free HBM) and by max_inflight (0: no extra
 * bound), so device memory gates admission. One parameter set per server. */
typedef struct sieve_server_opts {
    uint32_t max_batch;      /* queries per device batch (0: 1024) */
    uint32_t max_inflight;   /* device batches in flight at most (0: as many as fit) */
    uint64_t hbm_budget;     /* device bytes the slots may hold (0: a quarter of free HBM) */
    struct sieve_live* live; /* the index's live state (NULL: none): searches see its delta lists and
                              * tombstones and report external ids, and the server accepts writes */
} sieve_server_opts;

typedef struct sieve_request_info {
    double queue_ms;         /* submit -> batch launch */
    double total_ms;         /* submit -> results on the host */
    uint32_t batch_size;
    uint64_t batch_id;
} sieve_request_info;

typedef struct sieve_server_stats {
    uint64_t batches;
    uint64_t queries;
    uint64_t pending;
    uint64_t inflight;
    uint32_t slots;
    uint64_t slot_bytes;
    uint64_t writes;         /* insert / delete requests applied */
} sieve_server_stats;

typedef struct sieve_server sieve_server;

int sieve_server_create(const sieve_index* ix, const sieve_params* p, const sieve_server_opts* o,
                        sieve_server** out);
/* Copies one query (768 + 256 f32, host memory) into the queue; returns a ticket. */
int sieve_server_submit(sieve_server* s, const float* q768, const float* q256, uint64_t* ticket);
/* A query carrying a metadata filter (filter::search, src/filter.rs:72-110) under `plan`
 * (SIEVE_PLAN_*): it is batched with queued queries that carry the same filter and plan. */
int sieve_server_submit_filtered(sieve_server* s, const float* q768, const float* q256,
                                 const struct sieve_filter* filter, uint32_t plan, uint64_t* ticket);
/* Live writes through the server (needs opts.live): Live::insert / Live::delete
 * (src/live.rs:275-315) of n rows (ext ids, n x 768 unit vectors in host memory). A write is a
 * barrier in submission order, kept on the device without draining it: the rows are encoded and
 * logged while earlier searches run, the tombstones and delta lists are published on the live
 * state's stream after those searches, and every later search waits for the publish on the
 * device. sieve_server_wait on its ticket returns once it is logged and its publish enqueued
 * (count 0); the sieve_live_* entry points wait for a pending publish before they read the
 * live state. */
int sieve_server_insert(sieve_server* s, const uint32_t* ext, const float* vecs, size_t n, uint64_t* ticket);
int sieve_server_delete(sieve_server* s, const uint32_t* ext, size_t n, uint64_t* ticket);
/* Blocks until the ticket's results are ready; scores/rows hold p->k entries (count used). */
int sieve_server_wait(sieve_server* s, uint64_t ticket, float* scores, uint32_t* rows, uint32_t* count,
                      sieve_request_info* info);
int sieve_server_get_stats(sieve_server* s, sieve_server_stats* out);
/* Drains queued and in-flight work, then stops the threads and frees the slots. */
void sieve_server_destroy(sieve_server* s);

/* --------------------------------------------- L2 over 100-d int8 vectors (1B scale) --- */

/* bigann::Params (src/bigann.rs:610-622). refine: on an index with refinement bits > 1, the
 * 1-bit shortlist of max(refine, cand) slots whose members are re-estimated from all planes
 * (src/bigann.rs:683-729); 0, or an index without refinement, ranks by the 1-bit estimate alone.
 * margin <= 0 is off. */
typedef struct sieve_l2_params {
    uint32_t nprobe;
    uint32_t cand;
    uint32_t k;
    uint32_t refine;
    float margin;
} sieve_l2_params;

/* The L2 index as arrays (src/bigann.rs:386-460 postings, :237-266 model):
 *   leaves     k x 100 f32          centroids.f32 (the leaves, cell-major)
 *   block_off  k + 1 u64            block_off.u64
 *   blocks     nblocks x 512 B      32-group nibble-transposed codes of the rotated,
 *                                   128-padded residuals (the postings' code section)
 *   ids        nblocks x 32 u32     (PAD = 0xFFFFFFFF)
 *   factors    nblocks x 32 x 2 f32 [mul, |x|^2] per slot
 *   rotation   3 x 128 f32          rotation.f32
 *   base       n x 100 i8           the base vectors (exact rerank rows)
 * `base` is borrowed when it is device memory and borrow_base is nonzero. */
typedef struct sieve_l2_desc {
    uint32_t k;
    uint64_t n;
    uint64_t nblocks;
    const float* leaves;
    const uint64_t* block_off;
    const uint8_t* blocks;
    const uint32_t* ids;
    const float* factors;
    const float* rotation;
    const int8_t* base;
    uint32_t borrow_base;
    /* Multi-bit refinement (src/bigann.rs:38-113; bits.u64): refine_bits 0 or 1 = none, 2 or 3
     * = the record's per-slot extra section in `refine`, nblocks x 32 x SIEVE_REFINE_BYTES(bits)
     * bytes in slot order: the low bits - 1 code planes (16 B each, bit i = dimension i), then
     * the scale f as a little-endian f32 (src/postings.rs:70-76, src/bigann.rs:430-436). */
    uint32_t refine_bits;
    const uint8_t* refine;
} sieve_l2_desc;

/* refine_bytes (src/bigann.rs:110-113): bytes per slot of the refinement section. */
#define SIEVE_REFINE_BYTES(bits) ((bits) > 1 ? ((bits) - 1) * 16 + 4 : 0)
#define SIEVE_REFINE_MAX_BITS 3

typedef struct sieve_l2_index sieve_l2_index;

int sieve_l2_index_create(const sieve_l2_desc* desc, sieve_l2_index** out);
void sieve_l2_index_free(sieve_l2_index* ix);
/* bigann::Index::open (src/bigann.rs:634-641) over the reference's files: Model::open
 * (src/bigann.rs:257-266: centroids.f32, coarse.f32, k2.u64, rotation.f32, optional bits.u64),
 * Postings::open_ext (src/postings.rs:209-229: postings.off, block_off.u64, postings.bin with
 * 4 KiB-aligned records of code blocks, [mul, |x|^2], ids and the refinement section) in
 * index_dir, and Base::open (src/bigann.rs:477-489: base.00, base.01, ... with .done markers,
 * 100,000,000 rows each) in base_dir; everything is uploaded to the current device. */
int sieve_l2_index_open(const char* index_dir, const char* base_dir, sieve_l2_index** out);
/* hot_leaves (src/bigann.rs:537-563): of the leaves the nq queries route to (nprobe, cut by
 * margin), those worth keeping resident under `budget` bytes of postings records, hottest per
 * byte first (probes over the record's bytes, 4 KiB aligned); leaves (room for k) gets them and
 * *count their number. The index itself is wholly resident in HBM; the policy is for callers
 * that tier the records (host RAM, a second device). */
int sieve_l2_hot_leaves(const sieve_l2_index* ix, sieve_stream* s, const int8_t* q, size_t nq, uint32_t nprobe,
                        float margin, uint64_t budget, uint32_t* leaves, uint32_t* count);
/* As sieve_index_set_scan_chunk, for the L2 index. */
int sieve_l2_index_set_scan_chunk(sieve_l2_index* ix, uint32_t chunk_blocks);
/* refine_encode (src/bigann.rs:70-108) for n rotated, 128-dim residuals r (n x 128 f32, the
 * rotated padded row minus its rotated leaf, host or device): the low bits - 1 planes of each
 * residual's bits-bit code into planes (n x 32 bytes: plane 0, then plane 1 or zeros; bit i of
 * a plane = dimension i) and the scale into f (n f32). Bit-exact with the reference's f32 steps;
 * bits is 2 or 3. */
int sieve_l2_refine_encode(const float* r, size_t n, uint32_t bits, uint8_t* planes, float* f);
/* As sieve_index_set_scan_variant, for the L2 index: one of the per-query scan variants
 * (0 .. 9; the cluster-major scan serves the inner-product index only). */
int sieve_l2_index_set_scan_variant(sieve_l2_index* ix, uint32_t variant);
/* As sieve_index_set_small_batch, for the L2 index (no counterpart in the reference): batches
 * of at most max_queries queries run the small-batch scan, larger ones the selected per-query
 * variant. Results equal the per-query scan's. 0 disables it (default 8). */
int sieve_l2_index_set_small_batch(sieve_l2_index* ix, uint32_t max_queries);
/* As sieve_index_set_route_fuse, for the L2 index. */
int sieve_l2_index_set_route_fuse(sieve_l2_index* ix, int mode);

/* bigann::Index::search (src/bigann.rs:652-657) for nq queries (nq x 100 i8): the top k
 * (score = |q|^2 - |q - x|^2 - |q|^2 = 2 <q, x> - |x|^2, row) best first. Optional outputs
 * (NULL to skip): the probe lists (nq x nprobe leaf ids, best first, and counts), and the
 * candidates (nq x cand ids and estimates, best first, and counts). */
int sieve_l2_search_batch(const sieve_l2_index* ix, sieve_stream* s, const int8_t* q, size_t nq,
                          const sieve_l2_params* p, float* scores, uint32_t* rows, uint32_t* counts,
                          uint32_t* probes, uint32_t* probe_counts, uint32_t* cand_rows,
                          float* cand_est, uint32_t* cand_counts, sieve_timing* t);

/* ----------------------------------------------------------- metadata filters --- */
/* filter::Filter (src/filter.rs:18-39): a bitset of allowed rows over the index's corpus, with
 * the allowed slots of every cluster and the allowed rows, both built on the device. */
typedef struct sieve_filter sieve_filter;

/* filter::Plan (src/filter.rs:40-52). */
#define SIEVE_PLAN_POST 0   /* the unfiltered probes, disallowed rows skipped at admission */
#define SIEVE_PLAN_EXPAND 1 /* clusters in centroid order until expand_target allowed slots */
#define SIEVE_PLAN_BRUTE 2  /* exact int8 scores of every allowed row, no index */
#define SIEVE_PLAN_AUTO 3   /* Plan::resolve by the filter's size */

/* Filter::new (src/filter.rs:27-38). allow: allow_words u64 words (host or device), bit r =
 * row r allowed. */
int sieve_filter_create(const sieve_index* ix, const uint64_t* allow, uint64_t allow_words, sieve_filter** out);
void sieve_filter_free(sieve_filter* f);
/* The number of allowed rows and (per_cluster, NULL to skip: k u32) the allowed slots per
 * cluster. */
int sieve_filter_get_info(const sieve_filter* f, uint64_t* rows, uint32_t* per_cluster);
/* Expand scans over a filter of fewer than 30% of the rows (no counterpart in the reference):
 * the filter keeps each cluster's allowed slots, and the scan of an Expand plan reads only those
 * (the delta lists of a live state as before) instead of every slot of the probed clusters;
 * candidates and results are the same, and sieve_stats.scanned then counts the listed slots
 * (padded to blocks of 32). enable 0 scans whole clusters (default 1). */
int sieve_filter_set_gather(sieve_filter* f, int enable);
/* Plan::resolve (src/filter.rs:60-69): AUTO is BRUTE at <= 250,000 allowed rows, POST at >= 30%
 * of the n corpus rows, EXPAND (200,000 slots) between; other plans are kept. */
uint32_t sieve_plan_resolve(uint32_t plan, uint64_t allowed_rows, uint64_t n);
/* filter::search (src/filter.rs:72-110) for nq (<= 65535) queries sharing one filter: the top
 * p->k allowed rows as (score, row) best first. `expand_target` is EXPAND's slot target (AUTO,
 * and 0, use 200,000); the plan used is written to *resolved (may be NULL). EXPAND sorts every
 * centroid by its exact f32 score (at most 16384 centroids) and fails with SIEVE_ELIMIT if its
 * cut needs more than 8192 clusters (or the scan's LDS at that many exceeds the device's). */
int sieve_filter_search_batch(const sieve_index* ix, const sieve_filter* f, uint32_t plan, uint64_t expand_target,
                              sieve_stream* s, const float* q768, const float* q256, size_t nq, const sieve_params* p,
                              float* scores, uint32_t* rows, uint32_t* counts, uint32_t* resolved, sieve_timing* t);

/* ------------------------------------------------------------ live updates --- */
/* Live (src/live.rs:151-480): inserts, deletes and upserts over a built index without a
 * rebuild. Inserted rows are encoded on the device exactly as the build encodes a row and
 * appended to per-cluster delta lists that a live search scans right after each probed
 * cluster's base list; deletes are tombstones. Row ids: corpus rows keep theirs, delta rows
 * count up from the corpus size; live searches report external ids. Durability as the
 * reference's: with a write-ahead log (sieve_live_open_dir) every insert and delete is logged
 * and made durable before it is applied, a restart replays the log, and compaction writes a new
 * index directory with the delta lists merged into the base lists. Like the reference's
 * &mut Live, inserts and deletes must not overlap searches of the same index, and an index has
 * at most one live state. */
typedef struct sieve_live sieve_live;

typedef struct sieve_live_opts {
    uint64_t max_rows;   /* delta rows the device reserves room for */
    float soar_lambda;   /* SOAR spill weight of inserted rows (the build's soar_lambda);
                          * NaN: one slot per row, as an index built without SOAR */
} sieve_live_opts;

typedef struct sieve_live_info {
    uint64_t rows;          /* delta rows inserted */
    uint64_t slots;         /* delta slots (rows x 2 with SOAR) */
    uint64_t arena_blocks;  /* 32-slot blocks the delta lists occupy */
    uint64_t deleted;       /* tombstones set */
    uint64_t max_rows;
    uint64_t device_bytes;  /* held by the live state (the index's slot arrays grow as well) */
} sieve_live_info;

typedef struct sieve_replay_stats {  /* Live::replay's ReplayStats (src/live.rs:137-143) */
    uint64_t inserts;
    uint64_t deletes;
    uint64_t torn;     /* bytes of a torn final record, cut off the log */
    double seconds;
} sieve_replay_stats;

typedef struct sieve_compact_stats {  /* Live::compact's CompactStats (src/live.rs:145-149) */
    uint64_t slots;    /* slots of the new lists (padding aside) */
    uint64_t dropped;  /* tombstoned slots left out */
    uint64_t blocks;   /* 32-slot blocks of the new lists */
    double seconds;
} sieve_compact_stats;

/* Live::open (src/live.rs:166-207) without a directory or a log: the rotated centroids, room for
 * max_rows delta rows and their slots (the index's slot arrays are re-allocated with room for
 * them). Refused for a compacted index (its lists reference its row segment). */
int sieve_live_open(sieve_index* ix, const sieve_live_opts* o, sieve_live** out);
/* Live::open (src/live.rs:166-207) of an index opened from index_dir: soar_lambda is read from
 * index_dir/meta.json (o->soar_lambda is not used; null or absent: no SOAR spill); a compacted
 * index's row segment (live.rows.i8, live.norm256.f32, live.t2.bin, live.t2.norm, live.ext.u32,
 * live.deleted.u64) is loaded, and its rows count against o->max_rows. With wal_path, the log is
 * replayed (Live::replay, src/live.rs:209-249: consecutive inserts as one batch, a torn final
 * record cut off the file, an unknown op an SIEVE_EIO error naming its byte offset), *replay
 * filled (if not NULL), and every later sieve_live_insert / sieve_live_delete appends its records
 * (op u8 = 1 insert / 2 delete, external id u32 LE, then for an insert the prefix norm f32 LE
 * and the 768 int8 row bytes) and waits for fdatasync before it is applied. index_dir NULL: an
 * index created from arrays, with o->soar_lambda as given. */
int sieve_live_open_dir(sieve_index* ix, const sieve_live_opts* o, const char* index_dir, const char* wal_path,
                        sieve_replay_stats* replay, sieve_live** out);
/* Live::compact (src/live.rs:395-468): writes a new index directory out_dir (created if needed;
 * it must not be the directory the index was opened from): per cluster the base list's then the
 * delta list's slots, padding and tombstoned slots left out, re-packed into blocks (blocks.bin,
 * factors.f32, ids.u32, block_off.u64), the same records as postings.bin / postings.off
 * (postings::pack, src/postings.rs:89-114), meta.json (the source's plus compacted_from,
 * live_rows, deleted), the row segment (live.*), and centroids.f32, rotation.f32 and tier2.*
 * symlinked to the source directory's (written from the device for an index from arrays).
 * sieve_index_open + sieve_live_open_dir on out_dir give the same searches as this live state. */
int sieve_live_compact(const sieve_live* l, const char* out_dir, sieve_compact_stats* stats);
/* Live::truncate_wal (src/live.rs:470-479): empties the log (after a compaction made its
 * records redundant) and syncs it. No-op without a log. */
int sieve_live_truncate_wal(sieve_live* l);
void sieve_live_close(sieve_live* l);
/* Live::insert (src/live.rs:275-294): inserts or replaces ext[i] with the unit 768-d vector
 * vecs[i] (n x 768, host or device memory). A live external id's old row is tombstoned. */
int sieve_live_insert(sieve_live* l, const uint32_t* ext, const float* vecs, size_t n);
/* Live::delete (src/live.rs:296-315): tombstones the rows of external ids; unknown ids are
 * ignored. */
int sieve_live_delete(sieve_live* l, const uint32_t* ext, size_t n);
/* Live::current (src/live.rs:260-266): the row id an external id resolves to, or 0xFFFFFFFF. */
uint32_t sieve_live_current(const sieve_live* l, uint32_t ext);
int sieve_live_get_info(const sieve_live* l, sieve_live_info* out);
/* Copies the delta segment to host arrays (NULL to skip): the delta rows (rows x 768 int8),
 * prefix norms, tier-2 codes (rows x 64) and norms, external ids; the delta lists' block
 * offsets per cluster (k + 1), and their slots (arena_blocks x 32): raw 32-byte codes (bit i =
 * dimension i), mul, err and row ids (0xFFFFFFFF for padding). No counterpart in the
 * reference beyond Live::delta (src/live.rs:387-389): it is the inspection hook for tests and
 * for a caller's own compaction. */
int sieve_live_export(const sieve_live* l, int8_t* rows, float* norm256, uint8_t* t2, float* t2n, uint32_t* ext,
                      uint32_t* dblock_off, uint8_t* codes, float* mul, float* err, uint32_t* ids);
/* Live::search (src/live.rs:381-385) for nq queries: sieve_search_batch over the base index,
 * the delta lists and the tombstones, with results as external ids. `view` may carry
 * allowlists (filters); its tombstones must be NULL (the live state's are used). */
int sieve_live_search_batch(const sieve_live* l, sieve_stream* s, const float* q768, const float* q256, size_t nq,
                            const sieve_params* p, const sieve_view* view, float* scores, uint32_t* ext_ids,
                            uint32_t* counts, sieve_timing* t);

/* ------------------------------------------------------------ index build --- */
/* The GEMM-shaped steps of ivf::build (src/ivf.rs:264-395), on the current device, bit-exact
 * with the reference: every row-centroid score is the reference's sequential f32 dot (the x86
 * gemm_abt, src/blas.rs:43-54), reached through the routing GEMM and the exact rescoring of
 * its error band. Arrays may be host or device memory; each call waits for the device before
 * it starts and returns with its outputs written. */

/* kmeans (src/ivf.rs:188-225) without the balance split: spherical k-means of n unit rows
 * (sample, n x 256 f32) into k centroids (cents, k x 256): initial centroids
 * sample[rng.below(n)] drawn k times from Rng(seed), then iters Lloyd steps of assign (the first
 * best centroid per row), per-cluster sums of the members in sample order and
 * quant::normalize; an empty cluster takes sample[rng.below(n)] (drawn in cluster order). */
int sieve_build_kmeans(const float* sample, size_t n, uint32_t k, uint32_t iters, uint64_t seed, float* cents);
/* The clusters of n unit rows (rows, n x 256) against k centroids (cents, k x 256), as the build
 * assigns them (src/ivf.rs:313-321): pairs[2i] = the best centroid; pairs[2i + 1] =
 * soar_secondary (src/ivf.rs:164-186) over the 16 best of assign_top, or 0xFFFFFFFF when
 * soar_lambda is NaN (no SOAR spill). Ties at equal f32 scores go to the lower centroid id. */
int sieve_build_assign(const float* rows, size_t n, const float* cents, uint32_t k, float soar_lambda,
                       uint32_t* pairs);
/* Tier2::fit (src/tier2.rs:20-27): sigma (256 f32) of n rows (n x 256): per dimension the f64
 * sum of x * x in row order, over n, its square root, at least 1e-6. */
int sieve_build_tier2_fit(const float* rows, size_t n, float* sigma);
/* bigann::train (src/bigann.rs:175-227) without the files, over n 100-d int8 rows (base, n x 100,
 * host or device): the sample of sample_n rows drawn with Rng(42) (indices sorted), k1 coarse
 * cells by kmeans_l2 (src/bigann.rs:139-161, seed 7) over min(sample_n, 512 k1) rows at an even
 * stride through the sample, every sample row to its nearest cell, and k2 leaves per cell by
 * kmeans_l2 over the cell's rows (seed = the cell; a cell without rows gets its centroid k2
 * times). coarse: k1 x 100, leaves: k1 k2 x 100 (cell-major). Every score is top_l2's
 * 2 <x, c> - |c|^2 in the reference's f32 order; the k-means sums are f64 in sample order. */
int sieve_build_l2_train(const int8_t* base, size_t n, uint32_t k1, uint32_t k2, size_t sample_n, uint32_t iters,
                         float* coarse, float* leaves);
/* Model::assign (src/bigann.rs:277-302): each of n rows' leaf (leaf, n u32): the best child of
 * its 4 nearest coarse cells (top_l2), the cells taken in ascending order and a later one
 * winning only when strictly better. k1 >= 4. */
int sieve_build_l2_assign(const int8_t* base, size_t n, const float* coarse, uint32_t k1, const float* leaves,
                          uint32_t k2, uint32_t* leaf);

#ifdef __cplusplus
}
#endif

#endif /* SIEVE_GPU_H */
