// Kernel launchers of the sieve MI355X library. Every launcher enqueues on `st` and
// returns immediately; shapes are checked by the host code in api.hip before a launch.
#pragma once

#include "common.hpp"

namespace sieve {

// Per-query state written by the query-prep kernel (rabitq::Query::new, src/rabitq.rs:119-125,
// and the int8 rerank queries, src/ivf.rs:753-755 / :786-788).
// Rotation::apply (src/rabitq.rs:59-84) on D floats in LDS by one 64-thread workgroup: three
// rounds of (sign flip, unnormalised FWHT), then * d^-1.5 (`rot_norm`).
template <int D>
__device__ inline void rotate_lds(float* v, const float* signs, float rot_norm, int lane) {
    for (int r = 0; r < 3; r++) {
        for (int i = lane; i < D; i += 64) v[i] = v[i] * signs[r * D + i];
        __syncthreads();
        for (int h = 1; h < D; h <<= 1) {
            for (int j = lane; j < D / 2; j += 64) {
                const int blk = j / h, i = j - blk * h;
                const int a = blk * 2 * h + i, b = a + h;
                const float x = v[a], y = v[b];
                v[a] = x + y;
                v[b] = x - y;
            }
            __syncthreads();
        }
    }
    for (int i = lane; i < D; i += 64) v[i] = v[i] * rot_norm;
    __syncthreads();
}

struct QueryPrep {
    uint8_t* tables;    // nq x 1024: 64 groups x 16 u8 entries
    float4* qab;        // nq: (a, b, scale, bias)
    int8_t* wq768;      // nq x 768
    int8_t* wq256;      // nq x 256
    float2* inv;        // nq: (1/s of the 768-d query, 1/s of the 256-d query)
    uint16_t* qbf;      // nq x 256 BF16 copy of q256 (routing GEMM operand)
    float* qnorm;       // nq: |q256|
};

void launch_query_prep(const float* q256, const float* q768, size_t nq, const float* signs, float rot_norm,
                       const float* quant, const QueryPrep& out, hipStream_t st);

// Index re-layout at creation.
void launch_transpose_blocks(const uint8_t* blocks, size_t nblocks, int groups, uint4* codes, hipStream_t st);
void launch_split_factors(const float* factors, size_t nslots, float* mul, float* err, hipStream_t st);
void launch_to_bf16(const float* x, size_t n, uint16_t* y, hipStream_t st);
void launch_cluster_slots(const uint32_t* block_off, uint32_t k, uint32_t* slots, hipStream_t st);
// Counts slot ids that are neither PAD nor a row below n (a corrupt ids.u32).
void launch_count_bad_ids(const uint32_t* ids, size_t nslots, uint64_t n, unsigned long long* bad, hipStream_t st);

// C = A * B^T in f32 (m x k, n x k -> m x n, row-major), f32 MFMA.
void launch_gemm_f32_abt(const float* a, const float* b, float* c, int m, int n, int k, hipStream_t st);
// S = Qbf * Cbf^T (m x 256, n x 256 BF16 -> m x n f32, row stride lds), BF16 MFMA.
// S (row stride lds) = qbf (m x 256 BF16) * cbf (n x 256 BF16)^T; tmax (m x ceil(n / 16)): the
// largest approximate route score of each 16 consecutive centroids (S, or 2 S - cnorm when
// cnorm is given).
void launch_gemm_bf16_route(const uint16_t* qbf, const uint16_t* cbf, const float* cnorm, float* s, float* tmax,
                            int m, int n, size_t lds, hipStream_t st);

struct RouteArgs {
    float* approx;             // nq x k approximate scores (row stride lds); a query whose band
                               // overflows the rescoring buffer gets its exact scores written here
    size_t lds;
    const float* tmax;         // nq x ceil(k / 16) tile maxima of the approximate scores, or null
    int k;
    const float* cents;        // k x 256 f32
    const float* q256;         // nq x 256
    const float* qnorm;        // nq
    const uint32_t* cl_slots;  // k: stored slots per cluster
    float cent_norm_max;
    int nprobe;
    float margin;              // NaN = none
    uint64_t budget;           // 0 = none
    float* out_s;              // route score (inner product), or <q, c> in L2 mode
    uint32_t* out_c;
    uint32_t* out_n;
    int out_stride;
    // metric: inner product over `dims` (0) or L2 route score 2 <q, c> - |c|^2 (1)
    int l2;
    int dims;                  // rescoring dims (256, or 100 for L2)
    int cstride;               // f32 centroid row stride
    const float* cnorm;        // L2: |c|^2 per centroid
    const float* qf;           // L2: nq x 100 f32 queries (rescoring operand)
    const float* qq;           // L2: sum q_i^2 per query
    float err_coef;            // |approx - exact| <= err_coef * qnorm
    // fused: the GEMM wrote only the tile maxima; the select recomputes the approximate scores
    // it reads from the GEMM's BF16 operands (qbf: nq x 256 of this launch, cbf: k x 256).
    int fused;
    const uint16_t* qbf;
    const uint16_t* cbf;
};
void launch_route_select(const RouteArgs& a, size_t nq, hipStream_t st);
// Whether the select over k centroids at nprobe runs over tile maxima.
bool route_select_over_tiles(int k, int nprobe);
// Whether the GEMM should skip writing the scores (RouteArgs.fused). mode < 0: when the select
// runs over tile maxima and recomputing its tiles reads less than the score row it replaces;
// 0: never; > 0: whenever the select runs over tile maxima.
bool route_fuse(int mode, int k, int nprobe);
// (primary, secondary) clusters of n unit rows from their best-first candidate lists (pc, pn,
// row stride pstride): the first candidate and soar_secondary over the first 16 (src/ivf.rs:164-186),
// or PAD as the secondary without SOAR.
void launch_soar_pairs(const float* xs, size_t n, const float* cents, const uint32_t* pc, const uint32_t* pn,
                       int pstride, int soar, float lambda, uint32_t* pairs, hipStream_t st);
// out[0..3] = cnt[0..3], out[4] = sum of pn[0, nq), out[5] = *ovf, out[6] = cnt[4] (one workgroup).
void launch_pack_stats(const unsigned long long* cnt, const uint32_t* pn, size_t nq, const uint32_t* ovf,
                       unsigned long long* out, hipStream_t st);
// p[0, n) = 0 by a kernel (a memset node is not replayed reliably in a captured graph).
void launch_zero_words(uint32_t* p, uint32_t n, hipStream_t st);

struct ViewArgs {
    const uint64_t* deleted;
    uint64_t deleted_words;
    const uint64_t* allow;
    uint64_t allow_words;
    uint64_t allow_stride;
    const uint8_t* allow_flags;
};

struct ScanArgs {
    const uint4* codes;          // slot-major 32-byte codes
    const float* mul;
    const float* err;
    const uint32_t* ids;
    const uint32_t* block_off;   // k + 1
    const uint8_t* tables;       // nq x 1024
    const float4* qab;           // nq
    const float* probe_s;        // nq x probe_stride
    const uint32_t* probe_c;
    const uint32_t* probe_n;     // nq
    int probe_stride;
    int cand;
    ViewArgs view;
    float* out_est;              // nq x out_stride
    uint32_t* out_row;
    float* out_err;
    uint32_t* out_n;
    int out_stride;
    unsigned long long* scanned; // [0] += stored slots scanned, [2] += slots past the threshold,
                                 // [4] += (query, slot) pairs the cluster-major MFMA pass scores,
                                 // [5] += pairs its integer bound stages for the exact estimate
    // Work split: query q's probed blocks (T_q) are cut into clamp(ceil(T_q / chunk_blocks), 1,
    // chunk_max) chunks; workgroup q * chunk_max + c scans chunk c.
    uint32_t chunk_blocks;
    uint32_t chunk_max;
    uint2* part;                 // (nq * chunk_max) x cand partial lists (split queries only)
    uint32_t* part_n;            // nq * chunk_max
    unsigned long long* gthr;    // nq, zero between launches: best shared (key, ~slot) threshold
    uint32_t* qdone;             // nq, zero between launches: chunks finished
    // Live delta lists (null: none): cluster c's delta blocks are [dbase + dblock_off[c],
    // dbase + dblock_off[c + 1]) of the same slot arrays, scanned right after its base blocks.
    const uint32_t* dblock_off;
    uint32_t dbase;
    // Gathered base segments (null: none), for a filter's Expand plan: cluster c's allowed slots,
    // ascending and padded to whole blocks of 32 with PAD, are gather_slots[gather_off[c],
    // gather_off[c + 1]); the scan reads those slots instead of the cluster's blocks.
    const uint32_t* gather_slots;
    const uint32_t* gather_off;
    int seg_stride;              // LDS segment table entries: scan_segments(), set by the launcher
    int trigger;                 // buffered entries that trigger a cut to cand (set by the launcher)
    // Cluster-major scan (launch_scan_cm): the seed pass scans query q's first probe_lim[q]
    // probes and leaves its kept entries in seed (nq x cand, seed_n) and their cand-th best
    // (key, ~slot) in seed_thr (0: fewer than cand); the cluster pass appends every later slot
    // above seed_thr to ext (nq x ext_cap, ext_n); the finish pass merges both lists, or rescans
    // a query whose ext list overflowed.
    const uint32_t* probe_lim;
    uint2* seed;
    uint32_t* seed_n;
    unsigned long long* seed_thr;
    uint2* ext;
    uint32_t* ext_n;
    uint32_t ext_cap;
};
// Block segments per probe: the base list, and with live delta lists the delta list too.
__host__ __device__ inline int scan_segments(const ScanArgs& a) { return a.dblock_off ? 2 * a.probe_stride : a.probe_stride; }
// Chunking of a batch's scan: fills the work-split fields of `a` (a.cand set) for nq queries of
// up to nprobe probes over clusters of at most max_cluster_blocks blocks (mean_blocks on average);
// forced_chunk_blocks != 0 fixes the chunk size; min_chunk_blocks raises its floor.
struct ScanPlan {
    uint32_t chunk_blocks, chunk_max;
};
ScanPlan plan_scan(size_t nq, uint32_t nprobe, double mean_blocks, uint32_t max_cluster_blocks, int cand,
                   uint32_t forced_chunk_blocks, uint32_t min_chunk_blocks = 0);
// Upper bound of the scan's LDS over all variants.
size_t scan_lds_bytes(int cand, int probe_stride);
// Scan kernel variants (slots per thread, prefetch), picked per index by the autotuner; the
// last one is the cluster-major scan (launch_scan_cm).
constexpr int SCAN_VARIANTS = 12;
constexpr int SCAN_VARIANT_CM = 10;  // 10: k_scan_cm2, 11: k_scan_cm (one accumulator chain per column half)
const char* scan_variant_name(int variant);
void launch_scan(const ScanArgs& a, size_t nq, int variant, hipStream_t st);

// Cluster-major scan of a batch: every (segment, 32-query tile, block chunk) work item reads
// the segment's codes once for up to 32 queries and scores them with int8 MFMA (one-hot
// nibbles x (table - 128)). Scratch per batch, all device memory:
struct CmScratch {
    uint32_t k;                  // clusters
    uint32_t nseg;               // k, or 2k with live delta lists (segment k + c: c's delta list)
    size_t nq;                   // set by the launcher
    uint32_t* seg_count;         // nseg
    uint32_t* seg_qoff;          // nseg + 1: (query, segment) pairs per segment, prefix
    uint32_t* seg_ioff;          // nseg + 1: work items per segment, prefix
    uint32_t* seg_q;             // nq * probe_stride (segments x their queries)
    float* seg_cent;
    uint32_t* probe_lim;         // nq
    uint4* items;                // work items: (first slot, slots, first seg_q entry, queries)
    uint32_t items_cap;
    uint32_t* work;              // phases: work-item counters (reset by the launch)
    uint32_t seed_slots;         // slots the seed pass scans per query (whole probes)
    uint32_t chunk_blocks;       // blocks per work item
    uint32_t phases;             // the work list runs in this many launches, thresholds raised between
    hipEvent_t* ev;              // null, or events recorded around phase ph's MFMA launch: ev[2 ph],
    uint32_t nev;                // ev[2 ph + 1], for 2 ph + 1 < nev
};
// Enqueues the seed pass, the work-list build, the cluster pass and the finish pass. `a` carries
// the index, batch and output fields and seed / ext buffers; its chunking is ignored.
// The scan of a small unfiltered batch (64-group inner-product codes, or with l2 the 32-group
// L2 codes): per-chunk bounds, a threshold per query, the appended lists above it, the finish
// merge. part / part_n hold bound_chunks x 16 entries per query; seed, seed_n, seed_thr, ext,
// ext_n, ext_cap as for the cluster pass.
void launch_scan_small(const ScanArgs& a, size_t nq, bool l2, uint32_t bound_chunks, uint32_t append_chunk_blocks,
                       uint32_t append_chunks, hipStream_t st);
void launch_scan_cm(const ScanArgs& a, const CmScratch& cm, size_t nq, int variant, hipStream_t st);

// The standalone lut::scan seam over the reference's block layout.
void launch_lut_scan(const uint8_t* tables, int groups, const uint8_t* blocks, size_t nblocks, uint16_t* out,
                     hipStream_t st);

struct RerankArgs {
    const float* est;            // nq x cand_stride
    const uint32_t* rows;
    const float* errs;
    const uint32_t* cand_n;      // nq
    int cand_stride;
    const int8_t* rerank;        // n x 768
    const float* norm256;
    const uint8_t* t2_codes;     // n x 64 or null
    const float* t2_norm;
    const float* t2_sigma;
    const float* q256;           // nq x 256 (tier-2 tables)
    const int8_t* wq768;
    const int8_t* wq256;
    const float2* inv;
    int mode;                    // SIEVE_RERANK_*
    int r;
    int k;
    int prune_m;
    float prune_w;
    float* out_s;                // nq x k
    uint32_t* out_r;
    uint32_t* out_n;
    unsigned long long* reranked;
    // Live delta rows (drows null: none): row ids >= n_base are delta row (id - n_base), read
    // from drows / dnorm256 / dt2 / dt2n, and reported as dext[id - n_base].
    uint32_t n_base;
    const int8_t* drows;
    const float* dnorm256;
    const uint8_t* dt2;
    const float* dt2n;
    const uint32_t* dext;
};
size_t rerank_lds_bytes(int cand_stride, bool tier2_tables);
void launch_rerank(const RerankArgs& a, size_t nq, hipStream_t st);

// ------------------------------------------------------------------ L2 (100-d int8) ----
struct L2Prep {
    uint8_t* tables;   // nq x 512: 32 groups x 16 u8
    float4* qab;       // nq: (a, b, scale, bias)
    uint16_t* qbf;     // nq x 256 BF16 [q (100), 0 (28), q (100), 0 (28)]
    float* qf;         // nq x 100 f32
    float* qq;         // nq: sum q_i^2 (left to right)
    float* qnorm;      // nq: |q|
    float* qrot;       // nq x 128: the rotated, zero-padded query (refinement tables), or null
};
void launch_query_prep_l2(const int8_t* q, size_t nq, const float* signs, float rot_norm, const L2Prep& out,
                          hipStream_t st);
void launch_split_leaves(const float* leaves, uint32_t k, uint16_t* out, hipStream_t st);
void launch_scan_l2(const ScanArgs& a, size_t nq, int variant, hipStream_t st);

struct L2RerankArgs {
    const int8_t* q;           // nq x 100
    const int8_t* base;        // n x 100
    const uint32_t* cand_ids;  // nq x cand_stride
    const float* cand_nx;
    const uint32_t* cand_n;
    int cand_stride;
    int k;
    float* out_s;
    uint32_t* out_r;
    uint32_t* out_n;
};
size_t rerank_l2_lds_bytes(int cand_stride);
void launch_rerank_l2(const L2RerankArgs& a, size_t nq, hipStream_t st);

// Multi-bit refinement (refine.hip). planes: 2 x 16 bytes per row (plane 0, plane 1 or zeros).
void launch_refine_encode(const float* r, size_t n, int bits, uint4* planes, float* f, hipStream_t st);
// The record's extra section (per slot: bits - 1 planes of 16 B, then f) -> planes, f.
void launch_unpack_refine(const uint8_t* extra, size_t nslots, int bits, uint4* planes, float* f, hipStream_t st);
struct L2RefineArgs {
    const uint4* codes;          // raw 16-byte codes per slot
    const float* mul;
    const float* nx;
    const uint32_t* ids;
    const uint32_t* block_off;   // k + 1
    const uint4* planes;         // 2 per slot: the low planes (plane p = bit p of u)
    const float* rf;             // f per slot
    int bits;
    const uint8_t* tables;       // nq x 512
    const float4* qab;           // nq: (a, b, ...)
    const float* qrot;           // nq x 128 rotated queries
    const float* probe_s;        // <q, c> per probe
    const uint32_t* probe_c;
    const uint32_t* probe_n;
    int probe_stride;
    int cand;
    int K;                       // shortlist length, max(refine, cand); 0: never cut (no storage)
    float* out_est;              // nq x out_stride: refined estimates, best first
    uint32_t* out_row;
    float* out_nx;
    uint32_t* out_n;
    int out_stride;
    unsigned long long* scanned;
};
size_t refine_lds_bytes(int K, int cand, int nprobe);
void launch_scan_l2_refine(const L2RefineArgs& a, size_t nq, hipStream_t st);

}  // namespace sieve
