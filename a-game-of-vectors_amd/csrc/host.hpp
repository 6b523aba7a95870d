// Host-side internals shared by the C ABI (api.hip) and the serving loop (server.hip):
// device buffers, the index and stream objects, and the batched search pipeline stages.
#pragma once

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "kernels.hpp"

namespace sieve {


inline bool on_device(const void* p) {
    if (!p) return false;
    hipPointerAttribute_t at;
    hipError_t e = hipPointerGetAttributes(&at, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return at.type == hipMemoryTypeDevice || at.type == hipMemoryTypeManaged;
}

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

// Reallocations of stream scratch so far: a captured graph holds scratch addresses, so one
// captured before the latest reallocation is stale.
inline std::atomic<uint64_t> g_scratch_gen{0};
inline uint64_t next_index_uid() {
    static std::atomic<uint64_t> c{0};
    return ++c;
}

template <class T>
T* grow(DevBuf& b, size_t n) {
    size_t bytes = std::max<size_t>(n * sizeof(T), 16);
    if (bytes > b.cap) {
        g_scratch_gen.fetch_add(1);
        b.release();
        hipError_t e = hipMalloc(&b.p, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            b.p = nullptr;
            throw Error{SIEVE_ENOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed"};
        }
        b.cap = bytes;
    }
    return static_cast<T*>(b.p);
}

// grow() whose new allocations start zeroed (counters that kernels leave at zero). A new
// allocation is told by its capacity: hipMalloc may hand back the address just freed.
template <class T>
T* grow_zeroed(DevBuf& b, size_t n, hipStream_t st) {
    const size_t cap_before = b.cap;
    T* p = grow<T>(b, n);
    if (b.cap != cap_before) {
        hipError_t e = hipMemsetAsync(b.p, 0, b.cap, st);
        if (e != hipSuccess) throw Error{SIEVE_EHIP, std::string("hipMemsetAsync: ") + hipGetErrorString(e)};
    }
    return p;
}

template <class T>
T* dmalloc(size_t n, size_t* acct) {
    void* p = nullptr;
    size_t bytes = std::max<size_t>(n * sizeof(T), 16);
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw Error{SIEVE_ENOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed"};
    }
    if (acct) *acct += bytes;
    return static_cast<T*>(p);
}

}  // namespace sieve

using sieve::DevBuf;


struct sieve_index {
    uint64_t uid = sieve::next_index_uid();  // never reused: keys of captured graphs
    int device = 0;
    int route_fuse = -1;  // sieve_index_set_route_fuse
    uint32_t k = 0;
    uint64_t n = 0, nblocks = 0, nslots = 0;
    uint32_t max_cluster_blocks = 0;
    float* cents = nullptr;
    uint16_t* cents_bf = nullptr;
    float cent_norm_max = 0.0f;
    uint32_t* block_off = nullptr;
    uint32_t* cl_slots = nullptr;
    uint4* codes = nullptr;
    float* mul = nullptr;
    float* err = nullptr;
    uint32_t* ids = nullptr;
    float* rotation = nullptr;
    int8_t* rerank = nullptr;
    bool rerank_owned = true;
    float* quant = nullptr;
    float* norm256 = nullptr;
    float* t2_sigma = nullptr;
    uint8_t* t2_codes = nullptr;
    float* t2_norm = nullptr;
    size_t device_bytes = 0;
    struct sieve_live* live = nullptr;  // attached live state (sieve_live_open)
    // Rows of the live segment (live.*) that the lists of a compacted index reference: row ids
    // n .. n + seg_rows - 1 resolve through the live state, so searches need one.
    uint64_t seg_rows = 0;
    int scan_variant = 0;  // scan kernel variant (sieve_index_autotune)
    uint32_t scan_chunk_blocks = 0;  // 0: chunk size chosen per batch (sieve_index_set_scan_chunk)
    // Cluster-major scan (sieve_index_set_cluster_scan): batches of at least cm_min_queries
    // unfiltered queries; seed pass of cm_seed_slots slots and appended lists of cm_ext_cap
    // entries per query (0: 24 and 64 * cand); cm_chunk_blocks blocks per work item (0: 128
    // for variant 10, 64 for variant 11, the fastest of each on the 113.5M batch); the work list
    // in cm_phases launches.
    uint32_t cm_min_queries = 64, cm_seed_slots = 0, cm_ext_cap = 0, cm_chunk_blocks = 0, cm_phases = 2;
    // Below cm_min_queries, unfiltered batches of at most small_max_queries queries use the
    // small-batch scan (per-chunk bounds, then appended lists); larger ones the per-query scan
    // (on the 113.5M index the small-batch scan is faster up to about 12 queries).
    uint32_t small_max_queries = 12;

    ~sieve_index() {
        void* ps[] = {cents, cents_bf, block_off, cl_slots, codes, mul, err, ids, rotation, quant, norm256, t2_sigma,
                      t2_codes, t2_norm};
        for (void* p : ps)
            if (p) (void)hipFree(p);
        if (rerank && rerank_owned) (void)hipFree(rerank);
    }
};

struct sieve_l2_index {
    int device = 0;
    uint32_t small_max_queries = 8;  // batches of at most this many queries: the small-batch scan
    int route_fuse = -1;             // sieve_l2_index_set_route_fuse
    uint32_t k = 0;
    uint64_t n = 0, nblocks = 0, nslots = 0;
    uint32_t max_cluster_blocks = 0;
    float* leaves = nullptr;       // k x 100 f32
    uint16_t* leaves_bf = nullptr; // k x 256 BF16 [hi | lo]
    float* leaf_norm = nullptr;    // k: |c|^2 summed left to right
    float cent_norm_max = 0.0f;
    uint32_t* block_off = nullptr;
    uint4* codes = nullptr;        // 16 bytes per slot
    float* mul = nullptr;
    float* nx = nullptr;
    uint32_t* ids = nullptr;
    float* rotation = nullptr;     // 3 x 128
    int8_t* base = nullptr;
    bool base_owned = true;
    size_t device_bytes = 0;
    int scan_variant = 0;
    uint32_t scan_chunk_blocks = 0;
    std::vector<uint32_t> hblock_off;  // host copy of block_off (record sizes for the hot-leaf policy)
    int bits = 1;                  // refinement code width (1: none)
    uint4* rplanes = nullptr;      // 2 x 16 bytes per slot
    float* rf = nullptr;           // per slot
    ~sieve_l2_index() {
        void* ps[] = {leaves, leaves_bf, leaf_norm, block_off, codes, mul, nx, ids, rotation, rplanes, rf};
        for (void* p : ps)
            if (p) (void)hipFree(p);
        if (base && base_owned) (void)hipFree(base);
    }
};

enum Slot {
    B_Q256, B_Q768, B_TABLES, B_QAB, B_WQ768, B_WQ256, B_INV, B_QBF, B_QNORM, B_APPROX, B_PS, B_PC, B_PN, B_CE, B_CR,
    B_CX, B_CN, B_OS, B_OR, B_ON, B_DEL, B_ALLOW, B_AFLAGS, B_CNT, B_OVF, B_TMP1, B_TMP2, B_TMP3, B_QQ, B_QF,
    B_PART, B_PARTN, B_GTHR, B_QDONE, B_CM_LIM, B_CM_SEGC, B_CM_QOFF, B_CM_IOFF, B_CM_Q, B_CM_CENT, B_CM_SEED,
    B_CM_SEEDN, B_CM_THR, B_CM_EXT, B_CM_EXTN, B_CM_WORK, B_CM_ITEMS, B_TMAX, B_STATS, B_QROT,
    B_NSLOTS
};

namespace sieve {
// A search captured as a HIP graph: the call's arguments (key), the executable graph and the
// scratch generation its addresses belong to.
struct GraphEntry {
    std::string key;
    hipGraphExec_t exec = nullptr;
    uint64_t gen = 0;
    uint32_t cm_timed = 0;  // the stream's cm_timed for this call (a replay runs no host code)
    bool failed = false;  // capture failed: this call runs eagerly
};
}  // namespace sieve

struct sieve_stream {
    hipStream_t st = nullptr;
    int device = 0;
    DevBuf b[B_NSLOTS];
    // Sizing mode (sizing != nullptr, B_NSLOTS entries): the pipeline functions record the bytes
    // each scratch slot would need and enqueue nothing, so a caller can reserve a worst case.
    size_t* sizing = nullptr;
    uint8_t* sizing_zeroed = nullptr;  // sizing mode: slots whose buffers must start zeroed
    // ev[0..3]: around route, scan and rerank; ev[4 + 2 ph], ev[5 + 2 ph]: around phase ph's
    // cluster-major MFMA launch (cm_timed of them in the last search).
    hipEvent_t ev[4 + 2 * sieve::CM_TIMED_PHASES] = {};
    uint32_t cm_timed = 0;
    sieve_stats last{};
    unsigned long long* hstats = nullptr;  // pinned: the packed counters of the last search
    bool graphs = false;               // sieve_stream_set_graphs
    std::vector<sieve::GraphEntry> gcache;  // most recent last
    ~sieve_stream() {
        for (auto& g : gcache)
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
        for (auto& x : b) x.release();
        if (hstats) (void)hipHostFree(hstats);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (st) (void)hipStreamDestroy(st);
    }
};

namespace sieve {

// Stream scratch slot `slot` holding at least n T's (grow), or in sizing mode the record of that
// need and a placeholder address that no host code dereferences.
template <class T>
T* sgrow(sieve_stream* s, int slot, size_t n) {
    if (s->sizing) {
        s->sizing[slot] = std::max(s->sizing[slot], std::max<size_t>(n * sizeof(T), 16));
        return reinterpret_cast<T*>(uintptr_t(256));
    }
    return grow<T>(s->b[slot], n);
}
template <class T>
T* sgrow_zeroed(sieve_stream* s, int slot, size_t n) {
    if (s->sizing) {
        if (s->sizing_zeroed) s->sizing_zeroed[slot] = 1;
        return sgrow<T>(s, slot, n);
    }
    return grow_zeroed<T>(s->b[slot], n, s->st);
}
// Event records and launches are skipped in sizing mode.
inline void record(sieve_stream* s, int i) {
    if (!s->sizing) SIEVE_HIP(hipEventRecord(s->ev[i], s->st));
}

template <class F>
int guarded(F&& f) {
    try {
        f();
        set_error("");
        return SIEVE_OK;
    } catch (const Error& e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        set_error("host allocation failed");
        return SIEVE_ENOMEM;
    } catch (const std::exception& e) {
        set_error(e.what());
        return SIEVE_EINVAL;
    } catch (...) {
        set_error("unknown error");
        return SIEVE_EINVAL;
    }
}

// A device view of a caller array: the array itself if it is device memory, else a copy in
// stream scratch.
template <class T>
const T* dev_in(sieve_stream* s, int slot, const T* p, size_t n) {
    if (!p || n == 0) return p;
    if (on_device(p)) return p;
    T* d = grow<T>(s->b[slot], n);
    SIEVE_HIP(hipMemcpyAsync(d, p, n * sizeof(T), hipMemcpyHostToDevice, s->st));
    return d;
}

// Where a kernel writes an output: the caller's array if it is device memory, else scratch
// that is copied back by out_done().
template <class T>
struct DevOut {
    T* user;
    T* dev;
    size_t n;
    bool host;
};

template <class T>
DevOut<T> dev_out(sieve_stream* s, int slot, T* p, size_t n) {
    if (on_device(p)) return {p, p, n, false};
    return {p, grow<T>(s->b[slot], n), n, true};
}

template <class T>
inline void out_done(sieve_stream* s, const DevOut<T>& o) {
    if (o.host && o.n) SIEVE_HIP(hipMemcpyAsync(o.user, o.dev, o.n * sizeof(T), hipMemcpyDeviceToHost, s->st));
}

inline void check_params(const sieve_params* p, bool need_probe) {
    SIEVE_REQUIRE(p, SIEVE_EINVAL, "params is null");
    SIEVE_REQUIRE(p->cand >= 1 && p->cand <= (uint32_t)MAX_CAND, SIEVE_ELIMIT,
                  "cand must be in [1, " + std::to_string(MAX_CAND) + "]");
    SIEVE_REQUIRE(p->k >= 1 && p->k <= (uint32_t)MAX_CAND, SIEVE_ELIMIT,
                  "k must be in [1, " + std::to_string(MAX_CAND) + "]");
    SIEVE_REQUIRE(p->rerank <= SIEVE_RERANK_SSD, SIEVE_EINVAL, "unknown rerank tag " + std::to_string(p->rerank));
    if (need_probe)
        SIEVE_REQUIRE(p->nprobe >= 1 && p->nprobe <= (uint32_t)MAX_PROBE, SIEVE_ELIMIT,
                      "nprobe must be in [1, " + std::to_string(MAX_PROBE) + "]");
    SIEVE_REQUIRE(p->prune_m <= (uint32_t)MAX_CAND, SIEVE_ELIMIT, "prune rank beyond cand limit");
}

inline ViewArgs make_view(sieve_stream* s, const sieve_view* v, size_t nq) {
    ViewArgs a{};
    if (!v) return a;
    a.deleted = dev_in(s, B_DEL, v->deleted, v->deleted ? v->deleted_words : 0);
    a.deleted_words = v->deleted ? v->deleted_words : 0;
    if (v->allow) {
        size_t words = v->allow_stride ? v->allow_stride * nq : v->allow_words;
        a.allow = dev_in(s, B_ALLOW, v->allow, words);
        a.allow_words = v->allow_words;
        a.allow_stride = v->allow_stride;
        SIEVE_REQUIRE(!v->allow_stride || v->allow_stride >= v->allow_words, SIEVE_EINVAL,
                      "allow_stride smaller than allow_words");
        a.allow_flags = v->allow_flags ? dev_in(s, B_AFLAGS, v->allow_flags, nq) : nullptr;
    }
    return a;
}

inline float elapsed_ns(hipEvent_t a, hipEvent_t b) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) {
        (void)hipGetLastError();  // leave no error behind for the next launch check
        return 0.0f;
    }
    return ms * 1e6f;
}

// Per-batch device state of the pipeline.
struct Batch {
    QueryPrep prep;
    const float* q256 = nullptr;
    const float* q768 = nullptr;
};

inline void prep_queries(const sieve_index* ix, sieve_stream* s, const float* q256d, const float* q768d, size_t nq,
                  Batch& bt) {
    bt.q256 = q256d;
    bt.q768 = q768d;
    bt.prep.tables = sgrow<uint8_t>(s, B_TABLES, nq * 1024);
    bt.prep.qab = sgrow<float4>(s, B_QAB, nq);
    bt.prep.wq768 = sgrow<int8_t>(s, B_WQ768, nq * DIM);
    bt.prep.wq256 = sgrow<int8_t>(s, B_WQ256, nq * PDIM);
    bt.prep.inv = sgrow<float2>(s, B_INV, nq);
    bt.prep.qbf = sgrow<uint16_t>(s, B_QBF, nq * PDIM);
    bt.prep.qnorm = sgrow<float>(s, B_QNORM, nq);
    // 256^-1.5 = 2^-12 exactly.
    if (!s->sizing) launch_query_prep(q256d, q768d, nq, ix->rotation, 1.0f / 4096.0f, ix->quant, bt.prep, s->st);
}

// Route nq prepared queries into (ps, pc, pn) with row stride `stride`.
inline void route_batch(const sieve_index* ix, sieve_stream* s, const Batch& bt, size_t nq, const sieve_params* p, float* ps,
                 uint32_t* pc, uint32_t* pn, int stride) {
    uint32_t* ovf = sgrow<uint32_t>(s, B_OVF, 1);
    if (!s->sizing) launch_zero_words(ovf, 1, s->st);  // the filter plans' overflow flag; routing never sets it
    const size_t per = std::max<size_t>(1, std::min<size_t>(nq, (size_t(1) << 30) / (4 * (size_t)ix->k)));
    float* approx = sgrow<float>(s, B_APPROX, per * ix->k);
    float* tmax = sgrow<float>(s, B_TMAX, per * ((ix->k + 15) / 16));
    if (s->sizing) return;
    // Fused: the select recomputes the scores it reads and the GEMM writes none.
    const bool fused = route_fuse(ix->route_fuse, (int)ix->k, (int)p->nprobe);
    for (size_t q0 = 0; q0 < nq; q0 += per) {
        size_t m = std::min(per, nq - q0);
        launch_gemm_bf16_route(bt.prep.qbf + q0 * PDIM, ix->cents_bf, nullptr, fused ? nullptr : approx, tmax, (int)m,
                               (int)ix->k, ix->k, s->st);
        RouteArgs ra{};
        ra.fused = fused;
        ra.qbf = bt.prep.qbf + q0 * PDIM;
        ra.cbf = ix->cents_bf;
        ra.approx = approx;
        ra.lds = ix->k;
        ra.tmax = tmax;
        ra.k = (int)ix->k;
        ra.cents = ix->cents;
        ra.q256 = bt.q256 + q0 * PDIM;
        ra.qnorm = bt.prep.qnorm + q0;
        ra.cl_slots = ix->cl_slots;
        ra.cent_norm_max = ix->cent_norm_max;
        ra.nprobe = (int)p->nprobe;
        ra.margin = p->margin;
        ra.budget = p->budget;
        ra.out_s = ps + q0 * stride;
        ra.out_c = pc + q0 * stride;
        ra.out_n = pn + q0;
        ra.out_stride = stride;
        ra.l2 = 0;
        ra.dims = PDIM;
        ra.cstride = PDIM;
        const float u = 1.0f / 256.0f;  // BF16 unit roundoff
        ra.err_coef = (2.0f * u + u * u + 256.0f / 16777216.0f) * ix->cent_norm_max;
        launch_route_select(ra, m, s->st);
    }
    SIEVE_HIP(hipGetLastError());
}

// The live delta segment a search reads (defaults: none).
struct LiveArgs {
    const uint32_t* dblock_off = nullptr;  // k + 1 delta blocks per cluster (prefix)
    uint32_t dbase = 0;                    // first delta block in the index's slot arrays
    uint32_t dblocks = 0;                  // delta blocks in all
    uint32_t n_base = 0xFFFFFFFFu;         // corpus rows; delta row j is row id n_base + j
    const int8_t* drows = nullptr;
    const float* dnorm256 = nullptr;
    const uint8_t* dt2 = nullptr;
    const float* dt2n = nullptr;
    const uint32_t* dext = nullptr;
    const uint64_t* deleted = nullptr;     // tombstones over corpus and delta row ids
    uint64_t deleted_words = 0;
};

// The work split of a batch's scan and its partial-list scratch (per stream).
inline void set_scan_plan(ScanArgs& a, sieve_stream* s, size_t nq, uint32_t nprobe, double mean_blocks,
                          uint32_t max_cluster_blocks, uint32_t forced_chunk_blocks, uint32_t min_chunk_blocks = 0) {
    const ScanPlan pl =
        plan_scan(nq, nprobe, mean_blocks, max_cluster_blocks, a.cand, forced_chunk_blocks, min_chunk_blocks);
    a.chunk_blocks = pl.chunk_blocks;
    a.chunk_max = pl.chunk_max;
    SIEVE_REQUIRE(nq * pl.chunk_max < (1ull << 31), SIEVE_ELIMIT, "scan grid too large");
    if (pl.chunk_max > 1) {
        a.part = sgrow<uint2>(s, B_PART, nq * pl.chunk_max * (size_t)a.cand);
        a.part_n = sgrow<uint32_t>(s, B_PARTN, nq * pl.chunk_max);
        a.gthr = sgrow_zeroed<unsigned long long>(s, B_GTHR, nq);
        a.qdone = sgrow_zeroed<uint32_t>(s, B_QDONE, nq);
    }
}

// The allowed-slot lists of a filter's Expand scan (ScanArgs::gather_slots / gather_off) and
// their blocks in all (the scan's work bound).
struct GatherArgs {
    const uint32_t* slots = nullptr;
    const uint32_t* off = nullptr;
    uint64_t blocks = 0;
};

inline void scan_batch(const sieve_index* ix, sieve_stream* s, const Batch& bt, size_t nq, const float* ps,
                const uint32_t* pc, const uint32_t* pn, int pstride, const sieve_params* p, const ViewArgs& view,
                float* ce, uint32_t* cr, float* cx, uint32_t* cn, int cstride, unsigned long long* scanned,
                const LiveArgs* lv = nullptr, uint32_t min_chunk_blocks = 0, const GatherArgs& gather = GatherArgs{}) {
    if (!s->sizing) s->cm_timed = 0;
    ScanArgs a{};
    if (lv) {
        a.dblock_off = lv->dblock_off;
        a.dbase = lv->dbase;
    }
    a.codes = ix->codes;
    a.mul = ix->mul;
    a.err = ix->err;
    a.ids = ix->ids;
    a.block_off = ix->block_off;
    a.tables = bt.prep.tables;
    a.qab = bt.prep.qab;
    a.probe_s = ps;
    a.probe_c = pc;
    a.probe_n = pn;
    a.probe_stride = pstride;
    a.cand = (int)p->cand;
    a.view = view;
    a.out_est = ce;
    a.out_row = cr;
    a.out_err = cx;
    a.out_n = cn;
    a.out_stride = cstride;
    a.scanned = scanned;
    const size_t lds = scan_lds_bytes(a.cand, scan_segments(a));
    SIEVE_REQUIRE(lds <= 160 * 1024, SIEVE_ELIMIT, "scan workspace exceeds LDS");
    int variant = ix->scan_variant;
    if (variant >= SCAN_VARIANT_CM) {
        // Small or filtered batches share too little of their probed clusters: per-query scan.
        if (nq >= ix->cm_min_queries && !view.allow) {
            const uint32_t cand = (uint32_t)a.cand;
            CmScratch cm{};
            cm.k = ix->k;
            cm.nseg = a.dblock_off ? 2 * ix->k : ix->k;
            cm.seg_count = sgrow<uint32_t>(s, B_CM_SEGC, cm.nseg);
            cm.seg_qoff = sgrow<uint32_t>(s, B_CM_QOFF, cm.nseg + 1);
            cm.seg_ioff = sgrow<uint32_t>(s, B_CM_IOFF, cm.nseg + 1);
            cm.seg_q = sgrow<uint32_t>(s, B_CM_Q, nq * (size_t)pstride * (a.dblock_off ? 2 : 1));
            cm.seg_cent = sgrow<float>(s, B_CM_CENT, nq * (size_t)pstride * (a.dblock_off ? 2 : 1));
            cm.probe_lim = sgrow<uint32_t>(s, B_CM_LIM, nq);
            // Work items: sum over segments of ceil(queries / 32) * ceil(blocks / chunk) <= the
            // bound below (pairs * max blocks / (32 chunk) + pairs / 32 + blocks / chunk + segments).
            const double pairs = (double)nq * pstride * (a.dblock_off ? 2 : 1);
            const uint32_t cm_chunk =
                ix->cm_chunk_blocks ? ix->cm_chunk_blocks : (variant == SCAN_VARIANT_CM ? 128u : 64u);
            const double chunk = (double)cm_chunk;
            const double maxnb = std::max<double>(ix->max_cluster_blocks, lv ? lv->dblocks : 0);
            const double allnb = (double)ix->nblocks + (lv ? lv->dblocks : 0);
            const double bound = pairs * maxnb / (32.0 * chunk) + pairs / 32.0 + allnb / chunk + cm.nseg + 2.0;
            SIEVE_REQUIRE(bound < 4.0e9, SIEVE_ELIMIT, "cluster scan: too many work items");
            cm.items_cap = (uint32_t)bound;
            cm.items = sgrow<uint4>(s, B_CM_ITEMS, cm.items_cap);
            cm.phases = std::max<uint32_t>(1, ix->cm_phases);
            cm.work = sgrow<uint32_t>(s, B_CM_WORK, cm.phases);
            cm.seed_slots = ix->cm_seed_slots ? ix->cm_seed_slots : 24 * cand;
            cm.chunk_blocks = cm_chunk;
            a.ext_cap = ix->cm_ext_cap ? ix->cm_ext_cap : 64 * cand;
            a.seed = sgrow<uint2>(s, B_CM_SEED, nq * (size_t)cand);
            a.seed_n = sgrow<uint32_t>(s, B_CM_SEEDN, nq);
            a.seed_thr = sgrow<unsigned long long>(s, B_CM_THR, nq);
            a.ext = sgrow<uint2>(s, B_CM_EXT, nq * (size_t)a.ext_cap);
            a.ext_n = sgrow_zeroed<uint32_t>(s, B_CM_EXTN, nq);  // the finish pass leaves it zero
            if (s->sizing) return;
            // HIP events around each phase's MFMA launch (the first CM_TIMED_PHASES phases).
            s->cm_timed = std::min<uint32_t>(cm.phases, CM_TIMED_PHASES);
            cm.ev = s->ev[4] ? s->ev + 4 : nullptr;
            cm.nev = 2 * s->cm_timed;
            launch_scan_cm(a, cm, nq, variant, s->st);
            SIEVE_HIP(hipGetLastError());
            return;
        }
        if (nq <= ix->small_max_queries && !view.allow) {
            // Per-chunk bounds, a threshold per query, appended lists, the finish merge.
            const uint32_t cand = (uint32_t)a.cand;
            const double worst = (double)pstride * std::max<uint32_t>(ix->max_cluster_blocks, 1) +
                                 (lv ? (double)lv->dblocks : 0.0);
            // Grids of at most 512 bound chunks (all the threshold select reads) and 256 append
            // chunks per query: a larger query's chunks take more rounds, and the workgroups past
            // a query's own chunk count, which only exit, stay few.
            const uint32_t bound_chunks = (uint32_t)std::min(512.0, std::max(1.0, std::ceil(worst / 32.0)));
            const uint32_t append_blocks = 64;
            const uint32_t append_chunks =
                (uint32_t)std::min(256.0, std::max(1.0, std::ceil(worst / append_blocks)));
            SIEVE_REQUIRE(nq * (double)bound_chunks < 2.0e9 && nq * (double)append_chunks < 2.0e9, SIEVE_ELIMIT,
                          "scan grid too large");
            a.part = sgrow<uint2>(s, B_PART, nq * (size_t)bound_chunks * 16);
            a.part_n = sgrow<uint32_t>(s, B_PARTN, nq * (size_t)bound_chunks);
            a.ext_cap = ix->cm_ext_cap ? ix->cm_ext_cap : 64 * cand;
            a.seed = sgrow<uint2>(s, B_CM_SEED, nq * (size_t)cand);
            a.seed_n = sgrow<uint32_t>(s, B_CM_SEEDN, nq);
            a.seed_thr = sgrow<unsigned long long>(s, B_CM_THR, nq);
            a.ext = sgrow<uint2>(s, B_CM_EXT, nq * (size_t)a.ext_cap);
            a.ext_n = sgrow_zeroed<uint32_t>(s, B_CM_EXTN, nq);  // the finish pass leaves it zero
            if (s->sizing) return;
            launch_scan_small(a, nq, false, bound_chunks, append_blocks, append_chunks, s->st);
            SIEVE_HIP(hipGetLastError());
            return;
        }
        variant = 2;
    }
    if (gather.off) {
        // Gathered lists: a query reads at most the filter's listed blocks and its delta lists.
        a.gather_slots = gather.slots;
        a.gather_off = gather.off;
        const double work = (double)gather.blocks + (lv ? (double)lv->dblocks : 0.0);
        set_scan_plan(a, s, nq, 1, std::max(1.0, work), (uint32_t)std::min(std::max(1.0, work), 4.0e9),
                      ix->scan_chunk_blocks);
    } else {
        set_scan_plan(a, s, nq, (uint32_t)pstride, (double)ix->nblocks / ix->k, ix->max_cluster_blocks,
                      ix->scan_chunk_blocks, min_chunk_blocks);
    }
    if (s->sizing) return;
    launch_scan(a, nq, variant, s->st);
    SIEVE_HIP(hipGetLastError());
}

inline void rerank_batch(const sieve_index* ix, sieve_stream* s, const Batch& bt, size_t nq, const float* ce,
                  const uint32_t* cr, const float* cx, const uint32_t* cn, int cstride, const sieve_params* p,
                  float* os, uint32_t* orr, uint32_t* on, unsigned long long* reranked, const LiveArgs* lv = nullptr) {
    RerankArgs a{};
    a.n_base = 0xFFFFFFFFu;
    if (lv && lv->drows) {
        a.n_base = lv->n_base;
        a.drows = lv->drows;
        a.dnorm256 = lv->dnorm256;
        a.dt2 = lv->dt2;
        a.dt2n = lv->dt2n;
        a.dext = lv->dext;
    }
    a.est = ce;
    a.rows = cr;
    a.errs = cx;
    a.cand_n = cn;
    a.cand_stride = cstride;
    a.rerank = ix->rerank;
    a.norm256 = ix->norm256;
    a.t2_codes = ix->t2_codes;
    a.t2_norm = ix->t2_norm;
    a.t2_sigma = ix->t2_sigma;
    a.q256 = bt.q256;
    a.wq768 = bt.prep.wq768;
    a.wq256 = bt.prep.wq256;
    a.inv = bt.prep.inv;
    a.mode = (int)p->rerank;
    a.r = (int)p->rerank_r;
    a.k = (int)p->k;
    a.prune_m = (int)p->prune_m;
    a.prune_w = p->prune_w;
    a.out_s = os;
    a.out_r = orr;
    a.out_n = on;
    a.reranked = reranked;
    const bool t2 = a.mode == SIEVE_RERANK_TIERED && a.t2_codes;
    SIEVE_REQUIRE(rerank_lds_bytes(cstride, t2) <= 160 * 1024, SIEVE_ELIMIT, "rerank workspace exceeds LDS");
    if (s->sizing) return;
    launch_rerank(a, nq, s->st);
    SIEVE_HIP(hipGetLastError());
}

// The whole batched search (prep, route, scan, rerank) of nq queries already in device
// memory, enqueued on s's stream with no host synchronisation. Counters land in B_CNT
// (scanned, reranked, past-threshold), probe counts in B_PN; events ev[0..3] bracket route,
// scan and rerank.
inline void enqueue_search(const sieve_index* ix, sieve_stream* s, const float* q2, const float* q7, size_t nq,
                           const sieve_params* p, const ViewArgs& view_in, float* os, uint32_t* orr, uint32_t* on,
                           const LiveArgs* lv = nullptr) {
    SIEVE_REQUIRE(ix->seg_rows == 0 || lv, SIEVE_EINVAL,
                  "the index's lists hold rows of its live segment (live.*): search it through its live state");
    ViewArgs view = view_in;
    if (lv && lv->deleted) {
        view.deleted = lv->deleted;
        view.deleted_words = lv->deleted_words;
    }
    const int pstride = (int)std::min<uint32_t>(p->nprobe, ix->k);
    unsigned long long* cnt = sgrow<unsigned long long>(s, B_CNT, 8);
    sgrow<unsigned long long>(s, B_STATS, 8);  // finish_search's packed counters, sized with the rest
    float* ps = sgrow<float>(s, B_PS, nq * pstride);
    uint32_t* pc = sgrow<uint32_t>(s, B_PC, nq * pstride);
    uint32_t* pn = sgrow<uint32_t>(s, B_PN, nq);
    float* ce = sgrow<float>(s, B_CE, nq * p->cand);
    uint32_t* cr = sgrow<uint32_t>(s, B_CR, nq * p->cand);
    float* cx = sgrow<float>(s, B_CX, nq * p->cand);
    uint32_t* cn = sgrow<uint32_t>(s, B_CN, nq);
    if (!s->sizing) launch_zero_words(reinterpret_cast<uint32_t*>(cnt), 16, s->st);  // a kernel: graph-capturable
    record(s, 0);
    Batch bt;
    prep_queries(ix, s, q2, q7, nq, bt);
    route_batch(ix, s, bt, nq, p, ps, pc, pn, pstride);
    record(s, 1);
    scan_batch(ix, s, bt, nq, ps, pc, pn, pstride, p, view, ce, cr, cx, cn, (int)p->cand, cnt, lv);
    record(s, 2);
    rerank_batch(ix, s, bt, nq, ce, cr, cx, cn, (int)p->cand, p, os, orr, on, cnt + 1, lv);
    record(s, 3);
}

// Waits for an enqueued search and records its counters and stage times in s->last (and adds
// them to *t).
inline void finish_search(sieve_stream* s, size_t nq, sieve_timing* t) {
    // One small copy into pinned memory: the counters, the probe total and the overflow flag.
    const unsigned long long* cnt = static_cast<const unsigned long long*>(s->b[B_CNT].p);
    const uint32_t* pn = static_cast<const uint32_t*>(s->b[B_PN].p);
    unsigned long long* packed = grow<unsigned long long>(s->b[B_STATS], 8);
    launch_pack_stats(cnt, pn, nq, static_cast<const uint32_t*>(s->b[B_OVF].p), packed, s->st);
    if (!s->hstats) SIEVE_HIP(hipHostMalloc((void**)&s->hstats, 8 * sizeof(unsigned long long), hipHostMallocDefault));
    SIEVE_HIP(hipMemcpyAsync(s->hstats, packed, 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s->st));
    SIEVE_HIP(hipStreamSynchronize(s->st));
    const unsigned long long* hcnt = s->hstats;
    s->last = sieve_stats{};
    s->last.queries = nq;
    s->last.scanned = hcnt[0];
    s->last.id_reads = hcnt[2];
    s->last.reranked = hcnt[1];
    s->last.rescanned = hcnt[3];
    s->last.probed = hcnt[4];
    s->last.route_ns = elapsed_ns(s->ev[0], s->ev[1]);
    s->last.scan_ns = elapsed_ns(s->ev[1], s->ev[2]);
    s->last.rerank_ns = elapsed_ns(s->ev[2], s->ev[3]);
    s->last.cm_pairs = hcnt[6];
    s->last.cm_staged = hcnt[7];
    s->last.cm_launches = s->cm_timed;
    for (uint32_t ph = 0; ph < s->cm_timed; ph++) s->last.cm_ns += elapsed_ns(s->ev[4 + 2 * ph], s->ev[5 + 2 * ph]);
    if (t) {
        t->route += (uint64_t)s->last.route_ns;
        t->scan += (uint64_t)s->last.scan_ns;
        t->rerank += (uint64_t)s->last.rerank_ns;
        t->scanned += hcnt[0];
        t->reranked += hcnt[1];
        t->probed += s->last.probed;
    }
}

// The delta segment a live state hands to searches (csrc/live.hip), and the same with the delta
// lists at the largest size the live state reserves (for sizing scratch).
LiveArgs live_args(const sieve_live* l);
LiveArgs live_args_max(const sieve_live* l);
// Live::insert / Live::delete (src/live.rs:276-306): the rows encoded into the slot log, the
// write-ahead log appended, the host state updated. With publish the tombstones and the arena
// are published and waited for; without, they wait for live_publish_async.
void live_insert(sieve_live* l, const uint32_t* ext, const float* vecs, size_t n, bool publish);
void live_delete(sieve_live* l, const uint32_t* ext, size_t n, bool publish);
// Publishes the staged writes on the live state's own stream once the events `waits` (the
// batches in flight, which read the current arena and tombstones) have fired, without waiting
// on the host. Returns the event later searches wait on (hipStreamWaitEvent). Until
// live_settle the arena's block count is its upper bound ceil(slots / 32) + k.
hipEvent_t live_publish_async(sieve_live* l, const hipEvent_t* waits, size_t nwaits);
// Waits for a pending live_publish_async and takes the arena's exact block count.
void live_settle(const sieve_live* l);

// filter::search enqueued without host synchronisation (csrc/filter.hip).
struct FilterRun {
    uint32_t plan = 0;        // the plan used (SIEVE_PLAN_*)
    uint64_t brute_rows = 0;  // rows each query scored under the Brute plan
    uint32_t max_probe = 0;   // Expand: clusters a cut may probe
};
FilterRun enqueue_filter_search(const sieve_index* ix, const sieve_filter* f, uint32_t plan, uint64_t expand_target,
                                sieve_stream* s, const float* q2, const float* q7, size_t nq, const sieve_params* p,
                                float* os, uint32_t* orr, uint32_t* on, const LiveArgs* lv);

void size_filter_search(const sieve_index* ix, sieve_stream* s, size_t nq, const sieve_params* p,
                        const LiveArgs* lv);

// Resolves the no-allocation entry point's fixed scratch array on device dev (csrc/noalloc.hip).
void noalloc_setup(int dev);

// The L2 index from arrays (csrc/api.hip), used by sieve_l2_index_create and the reader of the
// on-disk form (csrc/l2io.hip).
sieve_l2_index* create_l2_index(const sieve_l2_desc* d);

// Query prep and routing of nq L2 queries in device memory (csrc/api.hip): probe lists in stream
// scratch (row stride min(nprobe, k)).
void l2_route(const sieve_l2_index* ix, sieve_stream* s, const int8_t* dq, size_t nq, uint32_t nprobe, float margin,
              bool with_qrot, L2Prep& prep, float** ps_out, uint32_t** pc_out, uint32_t** pn_out);

// A read-only memory map of a whole file.
struct Mapped {
    void* p = nullptr;
    size_t len = 0;
    ~Mapped() {
        if (p && len) munmap(p, len);
    }
};

inline bool map_file(const std::string& path, Mapped& m, bool required) {
    int fd = ::open(path.c_str(), O_RDONLY);
    if (fd < 0) {
        SIEVE_REQUIRE(!required, SIEVE_EIO, "cannot open " + path);
        return false;
    }
    struct stat stt;
    if (fstat(fd, &stt) != 0) {
        ::close(fd);
        throw Error{SIEVE_EIO, "cannot stat " + path};
    }
    m.len = (size_t)stt.st_size;
    if (m.len) {
        m.p = mmap(nullptr, m.len, PROT_READ, MAP_PRIVATE, fd, 0);
        if (m.p == MAP_FAILED) {
            m.p = nullptr;
            ::close(fd);
            throw Error{SIEVE_EIO, "cannot map " + path};
        }
    }
    ::close(fd);
    return true;
}



inline void ensure_events(sieve_stream* s) {
    for (auto& e : s->ev)
        if (!e) SIEVE_HIP(hipEventCreate(&e));
}

// ------------------------------------------------------------------ index building ---

}  // namespace sieve
