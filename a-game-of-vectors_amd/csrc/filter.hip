// Metadata filters (src/filter.rs): a filter is a bitset of allowed rows; the scan skips
// disallowed slots at admission (sieve_view allowlists). The planner picks how to search with
// one (Plan::resolve, src/filter.rs:60-69):
//   * Post: the unfiltered probes, disallowed rows skipped at admission;
//   * Expand(target): probe clusters in centroid order, skipping clusters with no allowed
//     slot, until `target` allowed slots are covered (src/filter.rs:80-95);
//   * Brute: exact int8 scores of every allowed row, no index (src/filter.rs:97-107).
// Filter::new's per-cluster allowed counts and the allowed row list are built on the device.
#include <hipcub/hipcub.hpp>

#include "host.hpp"
#include "select.hpp"

using namespace sieve;

namespace {

constexpr uint64_t BRUTE_MAX = 250000;       // src/filter.rs:55
constexpr double POST_MIN = 0.3;             // src/filter.rs:57
constexpr uint64_t EXPAND_TARGET = 200000;   // src/filter.rs:59
constexpr int EXPAND_MAX_K = 16384;          // centroids the expand cut sorts in LDS
constexpr int EXPAND_MAX_PROBE = 8192;       // clusters an expand cut may probe (the scan's LDS segment table)
constexpr int FT = 256;                      // threads of the filter kernels

// Allowed slots of cluster c (Filter::new, src/filter.rs:30-36).
__global__ __launch_bounds__(FT) void k_filter_counts(const uint32_t* __restrict__ block_off,
                                                      const uint32_t* __restrict__ ids,
                                                      const uint64_t* __restrict__ allow, uint64_t words,
                                                      uint32_t* __restrict__ per_cluster) {
    __shared__ uint32_t sum;
    const uint32_t c = blockIdx.x;
    if (threadIdx.x == 0) sum = 0;
    __syncthreads();
    uint32_t cnt = 0;
    for (size_t s = (size_t)block_off[c] * BLOCK + threadIdx.x; s < (size_t)block_off[c + 1] * BLOCK; s += FT) {
        const uint32_t id = ids[s];
        if (id != PAD && bit_set(allow, words, id)) cnt++;
    }
    cnt = wave_sum(cnt);
    if ((threadIdx.x & 63) == 0) atomicAdd(&sum, cnt);
    __syncthreads();
    if (threadIdx.x == 0) per_cluster[c] = sum;
}

// Cluster c's allowed slots rounded up to whole blocks of 32 (the gathered list's length).
__global__ void k_padded_counts(const uint32_t* __restrict__ per_cluster, uint32_t k, uint32_t* __restrict__ padded) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c <= k) padded[c] = c < k ? (per_cluster[c] + BLOCK - 1) / BLOCK * BLOCK : 0u;
}

// Cluster c's allowed slots in slot order, then PAD to the end of its list (one workgroup per
// cluster; an ordered compaction of each 256-slot stretch).
__global__ __launch_bounds__(FT) void k_filter_slots(const uint32_t* __restrict__ block_off,
                                                     const uint32_t* __restrict__ ids,
                                                     const uint64_t* __restrict__ allow, uint64_t words,
                                                     const uint32_t* __restrict__ goff, uint32_t* __restrict__ gslots) {
    __shared__ uint32_t scratch[8];
    const uint32_t c = blockIdx.x;
    const size_t s0 = (size_t)block_off[c] * BLOCK, s1 = (size_t)block_off[c + 1] * BLOCK;
    uint32_t o = goff[c];
    for (size_t s = s0; s < s1; s += FT) {
        const size_t si = s + threadIdx.x;
        const uint32_t id = si < s1 ? ids[si] : PAD;
        const uint32_t f = id != PAD && bit_set(allow, words, id) ? 1u : 0u;
        uint32_t tot;
        const uint32_t ex = block_excl_scan(f, scratch, &tot);
        if (f) gslots[o + ex] = (uint32_t)si;
        o += tot;
    }
    for (uint32_t i = o + threadIdx.x; i < goff[c + 1]; i += FT) gslots[i] = PAD;
}

__global__ void k_word_popc(const uint64_t* __restrict__ allow, uint64_t words, uint32_t* __restrict__ cnt) {
    const size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w < words) cnt[w] = (uint32_t)__popcll(allow[w]);
}

// The allowed rows, ascending.
__global__ void k_word_rows(const uint64_t* __restrict__ allow, uint64_t words, const uint32_t* __restrict__ off,
                            uint32_t* __restrict__ rows) {
    const size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= words) return;
    uint64_t b = allow[w];
    uint32_t o = off[w];
    while (b) {
        const int i = __ffsll((long long)b) - 1;
        rows[o++] = (uint32_t)(w * 64 + i);
        b &= b - 1;
    }
}

// Brute: score[q][i] = dot(wq768[q], rerank[rows[i]]) as f32 * inv (src/filter.rs:99-104), 16
// lanes per row on v_dot4_i32_i8, 1024 rows per workgroup.
__global__ __launch_bounds__(FT) void k_brute_scores(const int8_t* __restrict__ rerank,
                                                     const uint32_t* __restrict__ rows, uint64_t R,
                                                     const int8_t* __restrict__ wq768, const float2* __restrict__ inv,
                                                     float* __restrict__ scores) {
    const size_t q = blockIdx.y;
    const int t = threadIdx.x, lane = t & 63, sub = lane & 15;
    int wv[3][4];
    const int8_t* wq = wq768 + q * DIM;
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const uint4 x = *reinterpret_cast<const uint4*>(wq + sub * 16 + 256 * j);
        wv[j][0] = (int)x.x;
        wv[j][1] = (int)x.y;
        wv[j][2] = (int)x.z;
        wv[j][3] = (int)x.w;
    }
    const float iv = inv[q].x;
    const size_t r0 = (size_t)blockIdx.x * 1024;
    const size_t r1 = min(r0 + 1024, (size_t)R);
    // The 16 lanes of a row share `base`, so each width-16 shuffle below runs on whole groups.
    for (size_t base = r0 + (t >> 4); base < r1; base += FT / 16) {
        const int8_t* p = rerank + (size_t)rows[base] * DIM;
        int acc = 0;
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const uint4 x = *reinterpret_cast<const uint4*>(p + sub * 16 + 256 * j);
            acc = __builtin_amdgcn_sdot4((int)x.x, wv[j][0], acc, false);
            acc = __builtin_amdgcn_sdot4((int)x.y, wv[j][1], acc, false);
            acc = __builtin_amdgcn_sdot4((int)x.z, wv[j][2], acc, false);
            acc = __builtin_amdgcn_sdot4((int)x.w, wv[j][3], acc, false);
        }
#pragma unroll
        for (int d = 8; d >= 1; d >>= 1) acc += __shfl_xor(acc, d, 16);
        if (sub == 0) scores[q * R + base] = (float)acc * iv;
    }
}

// The best k of n entries (key, ~row) kept in buf, threshold raised to the k-th.
__device__ void topk_compact(uint2* buf, uint2* tmp, int n, int k, uint32_t* hist, uint32_t* misc) {
    const int t = threadIdx.x;
    const uint2 kth = block_kth_largest(buf, n, k, hist, misc + 8);
    if (t == 0) misc[3] = 0;
    __syncthreads();
    for (int i = t; i < n; i += FT) {
        const uint2 e = buf[i];
        if (ge2(e, kth)) {
            const uint32_t pos = atomicAdd(&misc[3], 1u);
            if (pos < (uint32_t)k) tmp[pos] = e;
        }
    }
    __syncthreads();
    for (int i = t; i < k; i += FT) buf[i] = tmp[i];
    if (t == 0) {
        misc[0] = (uint32_t)k;
        misc[1] = 1;
        misc[2] = kth.x;
        misc[4 + 8] = kth.y;
    }
    __syncthreads();
}

// TopK over a query's n scores (score desc, row asc; src/topk.rs:4-62), streamed through LDS.
__global__ __launch_bounds__(FT) void k_topk_scores(const float* __restrict__ scores,
                                                    const uint32_t* __restrict__ rows, uint64_t R, int k,
                                                    const uint64_t* __restrict__ deleted, uint64_t deleted_words,
                                                    float* __restrict__ out_s, uint32_t* __restrict__ out_r,
                                                    uint32_t* __restrict__ out_n, int out_stride) {
    extern __shared__ __align__(16) uint8_t lds[];
    const int cap = 2 * k + 4 * FT;
    uint2* buf = reinterpret_cast<uint2*>(lds);
    uint2* tmp = buf + cap;
    uint32_t* hist = reinterpret_cast<uint32_t*>(tmp + k);
    uint32_t* misc = hist + 256;  // 0 count, 1 has thr, 2 thr key, 3 scratch, 8..10 select, 12 thr lo, 13 stamp
    const size_t q = blockIdx.x;
    const int t = threadIdx.x, lane = t & 63;
    if (t < 16) misc[t] = 0;
    if (t == 0) misc[13] = 0xFFFFFFFFu;
    __syncthreads();
    const float* sc = scores + q * R;
    uint32_t round = 0;
    for (size_t i0 = 0; i0 < R; i0 += 4 * FT, round++) {
        const bool has = misc[1] != 0;
        const uint32_t thr = misc[2], thr_lo = misc[12];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const size_t i = i0 + u * FT + t;
            uint2 e = make_uint2(0u, 0u);
            bool pass = false;
            if (i < R) {
                e = make_uint2(f2key(sc[i] + 0.0f), ~rows[i]);
                pass = !has || e.x > thr || (e.x == thr && e.y > thr_lo);
                if (pass && deleted && bit_set(deleted, deleted_words, rows[i])) pass = false;
            }
            const uint64_t bal = __ballot(pass);
            if (bal) {
                const uint32_t nw = (uint32_t)__popcll(bal);
                const uint32_t before = (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
                uint32_t base = 0;
                if (lane == 0) {
                    base = atomicAdd(&misc[0], nw);
                    if (base + nw >= 2u * (uint32_t)k) misc[13] = round;
                }
                base = __shfl(base, 0, 64);
                if (pass) buf[base + before] = e;
            }
        }
        __syncthreads();
        if (misc[13] == round) topk_compact(buf, tmp, (int)misc[0], k, hist, misc);
    }
    if ((int)misc[0] > k) topk_compact(buf, tmp, (int)misc[0], k, hist, misc);
    const int n = (int)misc[0];
    const int n2 = pow2_at_least(max(n, 1));
    for (int i = n + t; i < n2; i += FT) buf[i] = make_uint2(0u, 0u);
    __syncthreads();
    block_sort_pairs(buf, nullptr, n2, false);
    for (int i = t; i < n; i += FT) {
        out_s[q * out_stride + i] = key2f(buf[i].x);
        out_r[q * out_stride + i] = ~buf[i].y;
    }
    if (t == 0) out_n[q] = (uint32_t)n;
}

// Exact route scores of every centroid (src/blas.rs:47-53: products left to right from -0.0).
__global__ __launch_bounds__(FT) void k_exact_route(const float* __restrict__ q256, const float* __restrict__ cents,
                                                    int K, float* __restrict__ out) {
    __shared__ float qs[PDIM];
    const size_t q = blockIdx.y;
    for (int i = threadIdx.x; i < PDIM; i += FT) qs[i] = q256[q * PDIM + i];
    __syncthreads();
    const int c = blockIdx.x * FT + threadIdx.x;
    if (c >= K) return;
    const float4* row = reinterpret_cast<const float4*>(cents + (size_t)c * PDIM);
    float acc = -0.0f;
    for (int j = 0; j < PDIM / 4; j++) {
        const float4 x = row[j];
        acc = acc + qs[4 * j] * x.x;
        acc = acc + qs[4 * j + 1] * x.y;
        acc = acc + qs[4 * j + 2] * x.z;
        acc = acc + qs[4 * j + 3] * x.w;
    }
    out[q * K + c] = acc;
}

// Expand's probe list (src/filter.rs:80-93): all centroids by (score desc, id asc), clusters
// with no allowed slot skipped, taken while the allowed slots covered so far are below target.
// The taken clusters are the weighted ones up to and including the first entry, in that order,
// whose inclusive covered count reaches the target. A weighted radix select over the 64-bit keys
// (score key, ~id) finds that entry: each 8-bit pass sums the weights per digit among the entries
// on the prefix found so far and keeps the digit where the running sum reaches what is left of
// the target. An ordered compaction keeps the weighted entries at or above it, and only those
// are sorted. Thread t holds entries i = t + 256 j (j < 64): the key in LDS, the weight in a
// register.
__global__ __launch_bounds__(FT) void k_expand_cut(const float* __restrict__ scores, int K,
                                                   const uint32_t* __restrict__ per_cluster, uint64_t target,
                                                   int max_probe, float* __restrict__ ps, uint32_t* __restrict__ pc,
                                                   uint32_t* __restrict__ pn, uint32_t* __restrict__ overflow) {
    extern __shared__ __align__(16) uint8_t lds[];  // expand_cut_lds(K) bytes
    int kp = 1;
    while (kp < K) kp <<= 1;
    uint2* e = reinterpret_cast<uint2*>(lds);                    // kp entries
    uint32_t* hist = reinterpret_cast<uint32_t*>(lds) + 2 * kp;  // 256 words
    uint32_t* scratch = hist + 256;                              // 16 words
    __shared__ uint32_t sel[2];                                  // digit, what is left
    static_assert(EXPAND_MAX_K <= 64 * FT, "a thread's entries fit its 64 weight registers");
    const size_t q = blockIdx.x;
    const int t = threadIdx.x;
    uint32_t wv[64];
    uint32_t wsum = 0;
#pragma unroll
    for (int j = 0; j < 64; j++) {
        const int i = t + FT * j;
        wv[j] = i < K ? per_cluster[i] : 0u;
        wsum += wv[j];  // an index holds fewer than 2^32 - 1 slots: the sums fit 32 bits
        // a cluster with no allowed slot gets key (0, 0): below every real key, whose low word
        // ~id is never 0
        if (i < K) e[i] = wv[j] ? make_uint2(f2key(scores[q * K + i] + 0.0f), ~(uint32_t)i) : make_uint2(0u, 0u);
    }
    uint32_t wtot;
    (void)block_excl_scan(wsum, scratch, &wtot);
    // every weighted entry when the total stays below the target; none for a zero target
    const bool all = (uint64_t)wtot < target;
    uint32_t p_hi = 0, m_hi = 0, p_lo = 0, m_lo = 0;
    if (!all && target > 0) {
        uint32_t rem = (uint32_t)target;  // <= wtot
        for (int pass = 0; pass < 8; pass++) {
            const bool hi = pass < 4;
            const int shift = 24 - 8 * (pass & 3);
            hist[t] = 0;
            __syncthreads();
#pragma unroll
            for (int j = 0; j < 64; j++) {
                const int i = t + FT * j;
                if (i >= K || wv[j] == 0) continue;
                const uint2 v = e[i];
                if ((v.x & m_hi) == p_hi && (v.y & m_lo) == p_lo)
                    atomicAdd(&hist[((hi ? v.x : v.y) >> shift) & 255u], wv[j]);
            }
            __syncthreads();
            if (t < 64) {
                // lane l holds digits 255 - 4 l .. 252 - 4 l (top first)
                const uint32_t c0 = hist[255 - 4 * t], c1 = hist[254 - 4 * t], c2 = hist[253 - 4 * t],
                               c3 = hist[252 - 4 * t];
                const uint32_t sum = c0 + c1 + c2 + c3;
                const uint32_t incl = wave_incl_scan(sum);
                const uint64_t hit = __ballot(incl >= rem);
                const int f = __ffsll((long long)hit) - 1;
                if (t == f) {
                    uint32_t cum = incl - sum;
                    const uint32_t cs[4] = {c0, c1, c2, c3};
                    for (int k = 0; k < 4; k++) {
                        if (cum + cs[k] >= rem) {
                            sel[0] = 255 - 4 * t - k;
                            sel[1] = rem - cum;
                            break;
                        }
                        cum += cs[k];
                    }
                }
            }
            __syncthreads();
            if (hi) {
                p_hi |= sel[0] << shift;
                m_hi |= 0xFFu << shift;
            } else {
                p_lo |= sel[0] << shift;
                m_lo |= 0xFFu << shift;
            }
            rem = sel[1];
            __syncthreads();
        }
    }
    // Ordered compaction, chunk j of every thread at once: each kept entry moves to the front,
    // never past an entry not yet read (its new place is at most its old one).
    uint32_t m = 0;
    for (int j = 0; j < 64 && FT * j < K; j++) {
        const int i = t + FT * j;
        const uint2 v = i < K ? e[i] : make_uint2(0u, 0u);
        const bool keep = v.y != 0 && target > 0 && (all || v.x > p_hi || (v.x == p_hi && v.y >= p_lo));
        uint32_t tot;
        const uint32_t ex = block_excl_scan(keep ? 1u : 0u, scratch, &tot);
        if (keep) e[m + ex] = v;
        m += tot;
    }
    __syncthreads();
    if ((int)m > max_probe) {
        if (t == 0) {
            atomicOr(overflow, 1u);
            pn[q] = 0;
        }
        return;
    }
    int n2 = 1;
    while (n2 < (int)m) n2 <<= 1;  // <= kp
    for (int i = (int)m + t; i < n2; i += FT) e[i] = make_uint2(0u, 0u);  // below every real key
    __syncthreads();
    block_sort_pairs(e, nullptr, n2, false);
    for (int i = t; i < (int)m; i += FT) {
        ps[q * max_probe + i] = key2f(e[i].x);
        pc[q * max_probe + i] = ~e[i].y;
    }
    if (t == 0) pn[q] = m;
}

inline unsigned grid(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

// k_expand_cut's dynamic LDS: the K keys padded to a power of two, the digit histogram, the
// block scan's words.
inline size_t expand_cut_lds(int K) {
    size_t kp = 1;
    while (kp < (size_t)K) kp <<= 1;
    return kp * 8 + (256 + 16) * 4;
}

}  // namespace

struct sieve_filter {
    const sieve_index* ix = nullptr;
    uint64_t* allow = nullptr;
    uint64_t words = 0;
    uint32_t* rows = nullptr;
    uint64_t nrows = 0;
    uint32_t* per_cluster = nullptr;
    // Filters below the Post plan's share of the rows keep each cluster's allowed slots (padded
    // to blocks of 32): their Expand scans read only those.
    uint32_t* gslots = nullptr;
    uint32_t* goff = nullptr;      // k + 1
    uint64_t gblocks = 0;
    bool gather = true;            // sieve_filter_set_gather
    size_t device_bytes = 0;
    ~sieve_filter() {
        void* ps[] = {allow, rows, per_cluster, gslots, goff};
        for (void* p : ps)
            if (p) (void)hipFree(p);
    }
};

namespace sieve {

// filter::search for nq queries sharing filter f, enqueued on s with no host synchronisation:
// outputs os / orr / on (device), counters in B_CNT, the plan's overflow flag (an Expand cut
// beyond its probe limit) in B_OVF, events ev[0..3] around its stages. With a live state
// (lv), Post and Expand read the delta lists and tombstones; Brute, as the reference's, scores
// the allowed corpus rows (skipping tombstoned ones).
// Smallest scan chunk for a query filtered to f->nrows of the index's rows: a chunk should hold
// about 2 * cand allowed slots, so each chunk reaches cand admitted entries and publishes a
// threshold, and the query's last chunk merges lists the shared threshold mostly rejects.
static uint32_t filtered_chunk_blocks(const sieve_index* ix, const sieve_filter* f, const sieve_params* p) {
    const double frac = std::max(1e-6, (double)f->nrows / (double)std::max<uint64_t>(ix->n, 1));
    return (uint32_t)std::min(1.0e6, std::ceil(2.0 * p->cand / (frac * BLOCK)));
}

FilterRun enqueue_filter_search(const sieve_index* ix, const sieve_filter* f, uint32_t plan, uint64_t expand_target,
                                sieve_stream* s, const float* q2, const float* q7, size_t nq, const sieve_params* p,
                                float* os, uint32_t* orr, uint32_t* on, const LiveArgs* lv) {
    SIEVE_REQUIRE(f->ix == ix, SIEVE_EINVAL, "the filter belongs to another index");
    SIEVE_REQUIRE(ix->seg_rows == 0 || lv, SIEVE_EINVAL,
                  "the index's lists hold rows of its live segment (live.*): search it through its live state");
    SIEVE_REQUIRE(plan <= SIEVE_PLAN_AUTO, SIEVE_EINVAL, "unknown plan");
    SIEVE_REQUIRE(nq <= 65535, SIEVE_ELIMIT, "a filtered batch holds at most 65535 queries");
    FilterRun fr;
    fr.plan = sieve_plan_resolve(plan, f->nrows, ix->n);
    // AUTO's Expand, and an Expand given no target (the serving loop's), cover 200,000 slots.
    if (fr.plan == SIEVE_PLAN_EXPAND && (plan == SIEVE_PLAN_AUTO || expand_target == 0)) expand_target = EXPAND_TARGET;
    uint32_t* ovf = sgrow<uint32_t>(s, B_OVF, 1);
    ViewArgs view{};
    view.allow = f->allow;
    view.allow_words = f->words;
    if (fr.plan == SIEVE_PLAN_POST) {
        enqueue_search(ix, s, q2, q7, nq, p, view, os, orr, on, lv);
        return fr;
    }
    if (lv && lv->deleted) {
        view.deleted = lv->deleted;
        view.deleted_words = lv->deleted_words;
    }
    unsigned long long* cnt = sgrow<unsigned long long>(s, B_CNT, 8);
    if (!s->sizing) {
        SIEVE_HIP(hipMemsetAsync(ovf, 0, 4, s->st));
        SIEVE_HIP(hipMemsetAsync(cnt, 0, 64, s->st));
    }
    record(s, 0);
    Batch bt;
    prep_queries(ix, s, q2, q7, nq, bt);
    uint32_t* pn = sgrow<uint32_t>(s, B_PN, nq);
    if (fr.plan == SIEVE_PLAN_BRUTE) {
        // All allowed rows scored for a sub-batch of queries at a time (<= 256 MB of scores).
        const uint64_t R = f->nrows;
        fr.brute_rows = R;
        if (!s->sizing) SIEVE_HIP(hipMemsetAsync(pn, 0, nq * 4, s->st));
        record(s, 1);
        record(s, 2);
        if (s->sizing) {
            if (R) sgrow<float>(s, B_TMP1, std::max<size_t>(1, std::min<size_t>(nq, (size_t(64) << 20) / R)) * R);
            return fr;
        }
        if (R == 0) {
            SIEVE_HIP(hipMemsetAsync(on, 0, nq * 4, s->st));
        } else {
            const size_t sub = std::max<size_t>(1, std::min<size_t>(nq, (size_t(64) << 20) / R));
            float* sc = sgrow<float>(s, B_TMP1, sub * R);
            const size_t lds = (size_t)(2 * p->k + 4 * FT) * 8 + (size_t)p->k * 8 + 256 * 4 + 16 * 4;
            SIEVE_REQUIRE(lds <= 160 * 1024, SIEVE_ELIMIT, "k too large for the brute-force top-k");
            if (lds > 64 * 1024)
                (void)hipFuncSetAttribute((const void*)k_topk_scores, hipFuncAttributeMaxDynamicSharedMemorySize,
                                          (int)lds);
            for (size_t q0 = 0; q0 < nq; q0 += sub) {
                const size_t m = std::min(sub, nq - q0);
                hipLaunchKernelGGL(k_brute_scores, dim3(grid(R, 1024), (unsigned)m), dim3(FT), 0, s->st, ix->rerank,
                                   f->rows, R, bt.prep.wq768 + q0 * DIM, bt.prep.inv + q0, sc);
                hipLaunchKernelGGL(k_topk_scores, dim3((unsigned)m), dim3(FT), lds, s->st, sc, f->rows, R, (int)p->k,
                                   view.deleted, view.deleted_words, os + q0 * p->k, orr + q0 * p->k, on + q0,
                                   (int)p->k);
            }
            SIEVE_HIP(hipGetLastError());
        }
        SIEVE_HIP(hipEventRecord(s->ev[3], s->st));
        return fr;
    }
    // Expand: exact scores of every centroid, the weighted cut, then the filtered scan.
    SIEVE_REQUIRE((int)ix->k <= EXPAND_MAX_K, SIEVE_ELIMIT,
                  "expand plan sorts at most " + std::to_string(EXPAND_MAX_K) + " centroids");
    // As many probes as the scan's LDS segment table holds (two segments per probe with live
    // delta lists), at most EXPAND_MAX_PROBE.
    int maxp = (int)std::min<uint32_t>(ix->k, EXPAND_MAX_PROBE);
    const int seg_per_probe = lv && lv->dblock_off ? 2 : 1;
    while (maxp > 1 && scan_lds_bytes((int)p->cand, maxp * seg_per_probe) > 160 * 1024) maxp -= 64;
    maxp = std::max(maxp, 1);
    fr.max_probe = (uint32_t)maxp;
    float* all = sgrow<float>(s, B_APPROX, nq * ix->k);
    float* ps = sgrow<float>(s, B_PS, nq * maxp);
    uint32_t* pc = sgrow<uint32_t>(s, B_PC, nq * maxp);
    float* ce = sgrow<float>(s, B_CE, nq * p->cand);
    uint32_t* cr = sgrow<uint32_t>(s, B_CR, nq * p->cand);
    float* cx = sgrow<float>(s, B_CX, nq * p->cand);
    uint32_t* cn = sgrow<uint32_t>(s, B_CN, nq);
    if (s->sizing) {
        scan_batch(ix, s, bt, nq, ps, pc, pn, maxp, p, view, ce, cr, cx, cn, (int)p->cand, cnt, lv,
                   filtered_chunk_blocks(ix, f, p));
        rerank_batch(ix, s, bt, nq, ce, cr, cx, cn, (int)p->cand, p, os, orr, on, cnt + 1, lv);
        return fr;
    }
    hipLaunchKernelGGL(k_exact_route, dim3(grid(ix->k, FT), (unsigned)nq), dim3(FT), 0, s->st, q2, ix->cents,
                       (int)ix->k, all);
    const size_t lds = expand_cut_lds((int)ix->k);
    if (lds > 64 * 1024)
        (void)hipFuncSetAttribute((const void*)k_expand_cut, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(k_expand_cut, dim3((unsigned)nq), dim3(FT), lds, s->st, all, (int)ix->k, f->per_cluster,
                       expand_target, maxp, ps, pc, pn, ovf);
    SIEVE_HIP(hipGetLastError());
    SIEVE_HIP(hipEventRecord(s->ev[1], s->st));
    GatherArgs g;
    if (f->gather && f->goff) {
        g.slots = f->gslots;
        g.off = f->goff;
        g.blocks = f->gblocks;
    }
    scan_batch(ix, s, bt, nq, ps, pc, pn, maxp, p, view, ce, cr, cx, cn, (int)p->cand, cnt, lv,
               filtered_chunk_blocks(ix, f, p), g);
    SIEVE_HIP(hipEventRecord(s->ev[2], s->st));
    rerank_batch(ix, s, bt, nq, ce, cr, cx, cn, (int)p->cand, p, os, orr, on, cnt + 1, lv);
    SIEVE_HIP(hipEventRecord(s->ev[3], s->st));
    return fr;
}

// The filtered paths a batch of nq queries can take, on a stream in sizing mode (their scratch is
// recorded, nothing is enqueued): Post, Expand over a filter of every row (its scan has the
// smallest chunks, so the most partial lists) and Brute at Plan::resolve's largest filter.
void size_filter_search(const sieve_index* ix, sieve_stream* s, size_t nq, const sieve_params* p,
                        const LiveArgs* lv) {
    SIEVE_REQUIRE(s->sizing, SIEVE_EINVAL, "size_filter_search needs a stream in sizing mode");
    sieve_filter f;
    f.ix = ix;
    f.allow = reinterpret_cast<uint64_t*>(uintptr_t(256));  // never dereferenced in sizing mode
    f.words = (ix->n + 63) / 64;
    const float* q = reinterpret_cast<const float*>(uintptr_t(256));
    float* os = reinterpret_cast<float*>(uintptr_t(256));
    uint32_t* ou = reinterpret_cast<uint32_t*>(uintptr_t(256));
    const std::pair<uint32_t, uint64_t> runs[] = {{SIEVE_PLAN_POST, ix->n},
                                                  {SIEVE_PLAN_EXPAND, ix->n},
                                                  {SIEVE_PLAN_BRUTE, std::min<uint64_t>(ix->n, BRUTE_MAX)}};
    try {
        for (const auto& r : runs) {
            f.nrows = r.second;
            enqueue_filter_search(ix, &f, r.first, EXPAND_TARGET, s, q, q, nq, p, os, ou, ou, lv);
        }
    } catch (...) {
        f.allow = nullptr;
        throw;
    }
    f.allow = nullptr;
}

}  // namespace sieve

extern "C" {

int sieve_filter_create(const sieve_index* ix, const uint64_t* allow, uint64_t allow_words, sieve_filter** out) {
    return guarded([&] {
        SIEVE_REQUIRE(ix && allow && out, SIEVE_EINVAL, "null argument");
        SIEVE_REQUIRE(allow_words >= 1, SIEVE_EINVAL, "empty allowlist");
        SIEVE_HIP(hipSetDevice(ix->device));
        std::unique_ptr<sieve_filter> f(new sieve_filter());
        f->ix = ix;
        f->words = allow_words;
        size_t* acct = &f->device_bytes;
        f->allow = dmalloc<uint64_t>(allow_words, acct);
        SIEVE_HIP(hipMemcpy(f->allow, allow, allow_words * 8,
                            on_device(allow) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice));
        f->per_cluster = dmalloc<uint32_t>(ix->k, acct);
        hipStream_t st = nullptr;
        hipLaunchKernelGGL(k_filter_counts, dim3(ix->k), dim3(FT), 0, st, ix->block_off, ix->ids, f->allow,
                           allow_words, f->per_cluster);
        uint32_t* cnt = dmalloc<uint32_t>(allow_words + 1, nullptr);
        uint32_t* off = dmalloc<uint32_t>(allow_words + 1, nullptr);
        SIEVE_HIP(hipMemset(cnt, 0, (allow_words + 1) * 4));
        hipLaunchKernelGGL(k_word_popc, dim3(grid(allow_words, 256)), dim3(256), 0, st, f->allow, allow_words, cnt);
        size_t tmp_bytes = 0;
        SIEVE_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, cnt, off, (int)(allow_words + 1), st));
        void* tmp = dmalloc<uint8_t>(tmp_bytes, nullptr);
        SIEVE_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, cnt, off, (int)(allow_words + 1), st));
        uint32_t total = 0;
        SIEVE_HIP(hipMemcpy(&total, off + allow_words, 4, hipMemcpyDeviceToHost));
        f->nrows = total;
        f->rows = dmalloc<uint32_t>(std::max<uint64_t>(total, 1), acct);
        hipLaunchKernelGGL(k_word_rows, dim3(grid(allow_words, 256)), dim3(256), 0, st, f->allow, allow_words, off,
                           f->rows);
        SIEVE_HIP(hipGetLastError());
        SIEVE_HIP(hipDeviceSynchronize());
        (void)hipFree(cnt);
        (void)hipFree(off);
        (void)hipFree(tmp);
        if ((double)total < POST_MIN * (double)ix->n) {
            // The allowed-slot lists: padded counts, their offsets, then the slots.
            const uint32_t k = ix->k;
            uint32_t* padded = dmalloc<uint32_t>(k + 1, nullptr);
            f->goff = dmalloc<uint32_t>(k + 1, acct);
            hipLaunchKernelGGL(k_padded_counts, dim3(grid(k + 1, 256)), dim3(256), 0, st, f->per_cluster, k, padded);
            size_t tb = 0;
            SIEVE_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, padded, f->goff, (int)(k + 1), st));
            void* t2 = dmalloc<uint8_t>(tb, nullptr);
            SIEVE_HIP(hipcub::DeviceScan::ExclusiveSum(t2, tb, padded, f->goff, (int)(k + 1), st));
            uint32_t slots = 0;
            SIEVE_HIP(hipMemcpy(&slots, f->goff + k, 4, hipMemcpyDeviceToHost));
            f->gblocks = slots / BLOCK;
            f->gslots = dmalloc<uint32_t>(std::max<uint32_t>(slots, 1), acct);
            hipLaunchKernelGGL(k_filter_slots, dim3(k), dim3(FT), 0, st, ix->block_off, ix->ids, f->allow,
                               allow_words, f->goff, f->gslots);
            SIEVE_HIP(hipGetLastError());
            SIEVE_HIP(hipDeviceSynchronize());
            (void)hipFree(padded);
            (void)hipFree(t2);
        }
        *out = f.release();
    });
}

void sieve_filter_free(sieve_filter* f) { delete f; }

int sieve_filter_set_gather(sieve_filter* f, int enable) {
    return guarded([&] {
        SIEVE_REQUIRE(f, SIEVE_EINVAL, "null filter");
        f->gather = enable != 0;
    });
}

int sieve_filter_get_info(const sieve_filter* f, uint64_t* rows, uint32_t* per_cluster) {
    return guarded([&] {
        SIEVE_REQUIRE(f, SIEVE_EINVAL, "null filter");
        if (rows) *rows = f->nrows;
        if (per_cluster) SIEVE_HIP(hipMemcpy(per_cluster, f->per_cluster, f->ix->k * 4, hipMemcpyDeviceToHost));
    });
}

uint32_t sieve_plan_resolve(uint32_t plan, uint64_t allowed_rows, uint64_t n) {
    if (plan != SIEVE_PLAN_AUTO) return plan;
    if (allowed_rows <= BRUTE_MAX) return SIEVE_PLAN_BRUTE;
    if ((double)allowed_rows >= POST_MIN * (double)n) return SIEVE_PLAN_POST;
    return SIEVE_PLAN_EXPAND;
}

int sieve_filter_search_batch(const sieve_index* ix, const sieve_filter* f, uint32_t plan, uint64_t expand_target,
                              sieve_stream* s, const float* q768, const float* q256, size_t nq, const sieve_params* p,
                              float* scores, uint32_t* rows, uint32_t* counts, uint32_t* resolved, sieve_timing* t) {
    return guarded([&] {
        SIEVE_REQUIRE(ix && f && s && q768 && q256 && p && scores && rows && counts, SIEVE_EINVAL, "null argument");
        check_params(p, true);
        SIEVE_HIP(hipSetDevice(ix->device));
        if (nq == 0) {
            if (resolved) *resolved = sieve_plan_resolve(plan, f->nrows, ix->n);
            return;
        }
        ensure_events(s);
        const float* q2 = dev_in(s, B_Q256, q256, nq * PDIM);
        const float* q7 = dev_in(s, B_Q768, q768, nq * DIM);
        auto os = dev_out(s, B_OS, scores, nq * p->k);
        auto orr = dev_out(s, B_OR, rows, nq * p->k);
        auto on = dev_out(s, B_ON, counts, nq);
        const FilterRun fr = enqueue_filter_search(ix, f, plan, expand_target, s, q2, q7, nq, p, os.dev, orr.dev,
                                                   on.dev, nullptr);
        if (resolved) *resolved = fr.plan;
        out_done(s, os);
        out_done(s, orr);
        out_done(s, on);
        uint32_t h = 0;
        SIEVE_HIP(hipMemcpyAsync(&h, s->b[B_OVF].p, 4, hipMemcpyDeviceToHost, s->st));
        SIEVE_HIP(hipStreamSynchronize(s->st));
        SIEVE_REQUIRE(h == 0, SIEVE_ELIMIT,
                      "expand plan needs more than " + std::to_string(fr.max_probe) + " probed clusters");
        finish_search(s, nq, t);
        if (fr.plan == SIEVE_PLAN_BRUTE) {
            s->last.reranked = fr.brute_rows * nq;  // every allowed row, per query
            if (t) t->reranked += s->last.reranked;
        }
    });
}

}  // extern "C"
