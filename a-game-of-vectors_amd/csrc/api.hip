// Host side of the C ABI (include/sieve_gpu.h): index creation and the batched search
// pipeline. The index is immutable after creation; each sieve_stream owns a HIP stream and
// grow-only scratch buffers, so host threads share an index and each drive their own stream
// (the reference's shared `Index` + per-thread `Scratch`, src/ivf.rs:498-519).
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "kernels.hpp"

#include "host.hpp"

namespace sieve {
static thread_local std::string g_error;
void set_error(const std::string& msg) { g_error = msg; }
}  // namespace sieve

using namespace sieve;

namespace {

struct HostCopy {
    std::vector<uint8_t> data;
};

template <class T>
const T* host_view(const T* p, size_t n, HostCopy& hc) {
    if (!p || !on_device(p)) return p;
    hc.data.resize(n * sizeof(T));
    SIEVE_HIP(hipMemcpy(hc.data.data(), p, n * sizeof(T), hipMemcpyDeviceToHost));
    return reinterpret_cast<const T*>(hc.data.data());
}

template <class T>
T* upload(const T* src, size_t n, size_t* acct) {
    T* d = dmalloc<T>(n, acct);
    if (n) SIEVE_HIP(hipMemcpy(d, src, n * sizeof(T), on_device(src) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice));
    return d;
}

// seg_rows: rows of a compacted index's live segment, which its lists may reference as row ids
// n .. n + seg_rows - 1.
sieve_index* create_index(const sieve_index_desc* d, uint64_t seg_rows = 0) {
    SIEVE_REQUIRE(d, SIEVE_EINVAL, "desc is null");
    SIEVE_REQUIRE(d->k >= 1, SIEVE_EINVAL, "k must be >= 1");
    SIEVE_REQUIRE(d->centroids && d->block_off && d->blocks && d->ids && d->rerank && d->quant && d->norm256,
                  SIEVE_EINVAL, "desc is missing a required array");
    SIEVE_REQUIRE(d->factors && d->rotation, SIEVE_EINVAL,
                  "the GPU path needs a RaBitQ index (factors.f32 and rotation.f32)");
    SIEVE_REQUIRE(d->nblocks * BLOCK < (1ull << 32) - 1, SIEVE_ELIMIT, "more than 2^32 - 1 slots");
    std::unique_ptr<sieve_index> ix(new sieve_index());
    SIEVE_HIP(hipGetDevice(&ix->device));
    ix->k = d->k;
    ix->n = d->n;
    ix->nblocks = d->nblocks;
    ix->nslots = d->nblocks * BLOCK;
    size_t* acct = &ix->device_bytes;

    HostCopy hb, hc;
    const uint64_t* bo = host_view(d->block_off, d->k + 1, hb);
    SIEVE_REQUIRE(bo[0] == 0 && bo[d->k] == d->nblocks, SIEVE_EINVAL, "block_off does not match nblocks");
    std::vector<uint32_t> bo32(d->k + 1);
    for (uint32_t c = 0; c <= d->k; c++) {
        if (c) SIEVE_REQUIRE(bo[c] >= bo[c - 1], SIEVE_EINVAL, "block_off is not monotone");
        bo32[c] = (uint32_t)bo[c];
        if (c) ix->max_cluster_blocks = std::max<uint32_t>(ix->max_cluster_blocks, (uint32_t)(bo[c] - bo[c - 1]));
    }
    const float* cents = host_view(d->centroids, (size_t)d->k * PDIM, hc);
    double mx = 0.0;
    for (uint32_t c = 0; c < d->k; c++) {
        double s = 0.0;
        for (int i = 0; i < PDIM; i++) s += (double)cents[(size_t)c * PDIM + i] * cents[(size_t)c * PDIM + i];
        mx = std::max(mx, std::sqrt(s));
    }
    ix->cent_norm_max = (float)(mx * (1.0 + 1e-6));

    hipStream_t st = nullptr;
    ix->cents = upload(cents, (size_t)d->k * PDIM, acct);
    ix->cents_bf = dmalloc<uint16_t>((size_t)d->k * PDIM, acct);
    launch_to_bf16(ix->cents, (size_t)d->k * PDIM, ix->cents_bf, st);
    ix->block_off = upload(bo32.data(), d->k + 1, acct);
    ix->cl_slots = dmalloc<uint32_t>(d->k, acct);
    launch_cluster_slots(ix->block_off, d->k, ix->cl_slots, st);

    // codes: blocks.bin -> slot-major, in chunks through a staging buffer
    ix->codes = dmalloc<uint4>(ix->nslots * 2, acct);
    const bool blocks_dev = on_device(d->blocks);
    const size_t chunk = 1 << 18;  // blocks per staging pass (256 MiB)
    uint8_t* stage = nullptr;
    if (!blocks_dev && d->nblocks) stage = dmalloc<uint8_t>(std::min<size_t>(chunk, d->nblocks) * BLOCK_BYTES, nullptr);
    for (size_t b0 = 0; b0 < d->nblocks; b0 += chunk) {
        size_t nb = std::min(chunk, (size_t)d->nblocks - b0);
        const uint8_t* src = d->blocks + b0 * BLOCK_BYTES;
        if (!blocks_dev) {
            SIEVE_HIP(hipMemcpy(stage, src, nb * BLOCK_BYTES, hipMemcpyHostToDevice));
            src = stage;
        }
        launch_transpose_blocks(src, nb, 64, ix->codes + b0 * BLOCK * 2, st);
        SIEVE_HIP(hipDeviceSynchronize());
    }
    if (stage) (void)hipFree(stage);

    // factors -> mul[], err[]
    ix->mul = dmalloc<float>(ix->nslots, acct);
    ix->err = dmalloc<float>(ix->nslots, acct);
    {
        const float* f = d->factors;
        float* tmp = nullptr;
        if (!on_device(f)) {
            tmp = dmalloc<float>(ix->nslots * 2, nullptr);
            SIEVE_HIP(hipMemcpy(tmp, f, ix->nslots * 8, hipMemcpyHostToDevice));
            f = tmp;
        }
        launch_split_factors(f, ix->nslots, ix->mul, ix->err, st);
        SIEVE_HIP(hipDeviceSynchronize());
        if (tmp) (void)hipFree(tmp);
    }
    ix->ids = upload(d->ids, ix->nslots, acct);
    {
        unsigned long long* bad = dmalloc<unsigned long long>(1, nullptr);
        SIEVE_HIP(hipMemset(bad, 0, 8));
        launch_count_bad_ids(ix->ids, ix->nslots, d->n + seg_rows, bad, st);
        unsigned long long hbad = 0;
        SIEVE_HIP(hipMemcpy(&hbad, bad, 8, hipMemcpyDeviceToHost));
        (void)hipFree(bad);
        SIEVE_REQUIRE(hbad == 0, SIEVE_EINVAL,
                      std::to_string(hbad) + " slot ids are beyond the corpus and its live segment (ids.u32)");
    }
    ix->seg_rows = seg_rows;
    ix->rotation = upload(d->rotation, 3 * PDIM, acct);
    if (d->borrow_rerank && on_device(d->rerank)) {
        ix->rerank = const_cast<int8_t*>(d->rerank);
        ix->rerank_owned = false;
    } else {
        ix->rerank = upload(d->rerank, (size_t)d->n * DIM, acct);
    }
    ix->quant = upload(d->quant, DIM, acct);
    ix->norm256 = upload(d->norm256, d->n, acct);
    if (d->tier2_sigma && d->tier2_codes && d->tier2_norm) {
        ix->t2_sigma = upload(d->tier2_sigma, PDIM, acct);
        ix->t2_codes = upload(d->tier2_codes, (size_t)d->n * 64, acct);
        ix->t2_norm = upload(d->tier2_norm, d->n, acct);
    }
    SIEVE_HIP(hipDeviceSynchronize());
    return ix.release();
}

}  // namespace

sieve_l2_index* sieve::create_l2_index(const sieve_l2_desc* d) {
    using namespace sieve;
    constexpr int D = 100, PD = 128, BB = 512;
    SIEVE_REQUIRE(d, SIEVE_EINVAL, "desc is null");
    SIEVE_REQUIRE(d->k >= 1, SIEVE_EINVAL, "k must be >= 1");
    SIEVE_REQUIRE(d->leaves && d->block_off && d->blocks && d->ids && d->factors && d->rotation && d->base,
                  SIEVE_EINVAL, "desc is missing a required array");
    SIEVE_REQUIRE(d->nblocks * BLOCK < (1ull << 32) - 1, SIEVE_ELIMIT, "more than 2^32 - 1 slots");
    std::unique_ptr<sieve_l2_index> ix(new sieve_l2_index());
    SIEVE_HIP(hipGetDevice(&ix->device));
    ix->k = d->k;
    ix->n = d->n;
    ix->nblocks = d->nblocks;
    ix->nslots = d->nblocks * BLOCK;
    size_t* acct = &ix->device_bytes;
    HostCopy hb, hc;
    const uint64_t* bo = host_view(d->block_off, d->k + 1, hb);
    SIEVE_REQUIRE(bo[0] == 0 && bo[d->k] == d->nblocks, SIEVE_EINVAL, "block_off does not match nblocks");
    std::vector<uint32_t> bo32(d->k + 1);
    for (uint32_t c = 0; c <= d->k; c++) {
        if (c) SIEVE_REQUIRE(bo[c] >= bo[c - 1], SIEVE_EINVAL, "block_off is not monotone");
        if (c) ix->max_cluster_blocks = std::max<uint32_t>(ix->max_cluster_blocks, (uint32_t)(bo[c] - bo[c - 1]));
        bo32[c] = (uint32_t)bo[c];
    }
    const float* lv = host_view(d->leaves, (size_t)d->k * D, hc);
    // norms (src/bigann.rs:135-137): sum x*x left to right from -0.0
    std::vector<float> ln(d->k);
    double mx = 0.0;
    for (uint32_t c = 0; c < d->k; c++) {
        float s = -0.0f;
        double s64 = 0.0;
        for (int i = 0; i < D; i++) {
            float x = lv[(size_t)c * D + i];
            s = s + x * x;
            s64 += (double)x * x;
        }
        ln[c] = s;
        mx = std::max(mx, std::sqrt(s64));
    }
    ix->cent_norm_max = (float)(mx * (1.0 + 1e-6));
    hipStream_t st = nullptr;
    ix->leaves = upload(lv, (size_t)d->k * D, acct);
    ix->leaves_bf = dmalloc<uint16_t>((size_t)d->k * PDIM, acct);
    launch_split_leaves(ix->leaves, d->k, ix->leaves_bf, st);
    ix->leaf_norm = upload(ln.data(), d->k, acct);
    ix->block_off = upload(bo32.data(), d->k + 1, acct);
    ix->hblock_off = bo32;
    ix->codes = dmalloc<uint4>(ix->nslots, acct);
    {
        const uint8_t* src = d->blocks;
        uint8_t* tmp = nullptr;
        if (!on_device(src)) {
            tmp = dmalloc<uint8_t>(d->nblocks * BB, nullptr);
            SIEVE_HIP(hipMemcpy(tmp, src, d->nblocks * BB, hipMemcpyHostToDevice));
            src = tmp;
        }
        launch_transpose_blocks(src, d->nblocks, 32, ix->codes, st);
        SIEVE_HIP(hipDeviceSynchronize());
        if (tmp) (void)hipFree(tmp);
    }
    ix->mul = dmalloc<float>(ix->nslots, acct);
    ix->nx = dmalloc<float>(ix->nslots, acct);
    {
        const float* f = d->factors;
        float* tmp = nullptr;
        if (!on_device(f)) {
            tmp = dmalloc<float>(ix->nslots * 2, nullptr);
            SIEVE_HIP(hipMemcpy(tmp, f, ix->nslots * 8, hipMemcpyHostToDevice));
            f = tmp;
        }
        launch_split_factors(f, ix->nslots, ix->mul, ix->nx, st);
        SIEVE_HIP(hipDeviceSynchronize());
        if (tmp) (void)hipFree(tmp);
    }
    ix->ids = upload(d->ids, ix->nslots, acct);
    {
        unsigned long long* bad = dmalloc<unsigned long long>(1, nullptr);
        SIEVE_HIP(hipMemset(bad, 0, 8));
        launch_count_bad_ids(ix->ids, ix->nslots, d->n, bad, st);
        unsigned long long hbad = 0;
        SIEVE_HIP(hipMemcpy(&hbad, bad, 8, hipMemcpyDeviceToHost));
        (void)hipFree(bad);
        SIEVE_REQUIRE(hbad == 0, SIEVE_EINVAL, std::to_string(hbad) + " slot ids are beyond the base");
    }
    ix->rotation = upload(d->rotation, 3 * PD, acct);
    if (d->refine_bits > 1) {
        // The per-slot extra section into two plane words and f per slot, through a staging
        // buffer when it is host memory.
        SIEVE_REQUIRE(d->refine_bits <= SIEVE_REFINE_MAX_BITS && d->refine, SIEVE_EINVAL,
                      "refinement: refine_bits must be 2 or 3, with its extra section");
        ix->bits = (int)d->refine_bits;
        const size_t rb = SIEVE_REFINE_BYTES(d->refine_bits);
        ix->rplanes = dmalloc<uint4>(ix->nslots * 2, acct);
        ix->rf = dmalloc<float>(ix->nslots, acct);
        const bool dev = on_device(d->refine);
        const size_t chunk = size_t(1) << 22;  // slots per pass
        uint8_t* stage = dev ? nullptr : dmalloc<uint8_t>(std::min<size_t>(chunk, ix->nslots) * rb, nullptr);
        for (size_t s0 = 0; s0 < ix->nslots; s0 += chunk) {
            const size_t m = std::min<size_t>(chunk, ix->nslots - s0);
            const uint8_t* src = d->refine + s0 * rb;
            if (!dev) {
                SIEVE_HIP(hipMemcpy(stage, src, m * rb, hipMemcpyHostToDevice));
                src = stage;
            }
            launch_unpack_refine(src, m, ix->bits, ix->rplanes + 2 * s0, ix->rf + s0, st);
            SIEVE_HIP(hipDeviceSynchronize());
        }
        if (stage) (void)hipFree(stage);
    }
    if (d->borrow_base && on_device(d->base)) {
        ix->base = const_cast<int8_t*>(d->base);
        ix->base_owned = false;
    } else {
        ix->base = upload(d->base, (size_t)d->n * D, acct);
    }
    SIEVE_HIP(hipDeviceSynchronize());
    return ix.release();
}

namespace {

}  // namespace

// ====================================================================== C ABI ==========
extern "C" {

const char* sieve_last_error(void) { return g_error.c_str(); }

int sieve_stream_create(sieve_stream** out) {
    return guarded([&] {
        SIEVE_REQUIRE(out, SIEVE_EINVAL, "out is null");
        std::unique_ptr<sieve_stream> s(new sieve_stream());
        SIEVE_HIP(hipGetDevice(&s->device));
        SIEVE_HIP(hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking));
        ensure_events(s.get());
        noalloc_setup(s->device);
        *out = s.release();
    });
}

void sieve_stream_destroy(sieve_stream* s) { delete s; }

int sieve_stream_set_graphs(sieve_stream* s, int enable) {
    return guarded([&] {
        SIEVE_REQUIRE(s, SIEVE_EINVAL, "stream is null");
        s->graphs = enable != 0;
        if (!s->graphs) {
            for (auto& g : s->gcache)
                if (g.exec) (void)hipGraphExecDestroy(g.exec);
            s->gcache.clear();
        }
    });
}

int sieve_stream_sync(sieve_stream* s) {
    return guarded([&] {
        SIEVE_REQUIRE(s, SIEVE_EINVAL, "stream is null");
        SIEVE_HIP(hipStreamSynchronize(s->st));
    });
}

void* sieve_stream_handle(sieve_stream* s) { return s ? (void*)s->st : nullptr; }

int sieve_stream_last_stats(const sieve_stream* s, sieve_stats* out) {
    return guarded([&] {
        SIEVE_REQUIRE(s && out, SIEVE_EINVAL, "null argument");
        *out = s->last;
    });
}

int sieve_gemm_abt(const float* a, const float* b, size_t m, size_t n, size_t k, float* c) {
    return guarded([&] {
        SIEVE_REQUIRE(a && b && c, SIEVE_EINVAL, "null matrix");
        SIEVE_REQUIRE(m < (1u << 31) && n < (1u << 31) && k < (1u << 31), SIEVE_ELIMIT, "matrix too large");
        if (m == 0 || n == 0) return;
        sieve_stream s;
        SIEVE_HIP(hipStreamCreateWithFlags(&s.st, hipStreamNonBlocking));
        const float* da = dev_in(&s, B_TMP1, a, m * k);
        const float* db = dev_in(&s, B_TMP2, b, n * k);
        DevOut<float> dc = dev_out(&s, B_TMP3, c, m * n);
        launch_gemm_f32_abt(da, db, dc.dev, (int)m, (int)n, (int)k, s.st);
        SIEVE_HIP(hipGetLastError());
        out_done(&s, dc);
        SIEVE_HIP(hipStreamSynchronize(s.st));
    });
}

int sieve_lut_scan(const uint8_t* tables, uint32_t groups, const uint8_t* blocks, size_t nblocks, uint16_t* out) {
    return guarded([&] {
        SIEVE_REQUIRE(tables && blocks && out, SIEVE_EINVAL, "null argument");
        SIEVE_REQUIRE(groups == 64 || groups == 32, SIEVE_EINVAL, "groups must be 64 or 32");
        if (nblocks == 0) return;
        sieve_stream s;
        SIEVE_HIP(hipStreamCreateWithFlags(&s.st, hipStreamNonBlocking));
        const uint8_t* dt = dev_in(&s, B_TMP1, tables, (size_t)groups * 16);
        const uint8_t* dbk = dev_in(&s, B_TMP2, blocks, nblocks * groups * 16);
        DevOut<uint16_t> o = dev_out(&s, B_TMP3, out, nblocks * BLOCK);
        launch_lut_scan(dt, (int)groups, dbk, nblocks, o.dev, s.st);
        SIEVE_HIP(hipGetLastError());
        out_done(&s, o);
        SIEVE_HIP(hipStreamSynchronize(s.st));
    });
}

int sieve_index_create(const sieve_index_desc* desc, sieve_index** out) {
    return guarded([&] {
        SIEVE_REQUIRE(out, SIEVE_EINVAL, "out is null");
        *out = create_index(desc);
    });
}

int sieve_index_open(const char* index_dir, const char* ingest_dir, sieve_index** out) {
    return guarded([&] {
        SIEVE_REQUIRE(index_dir && ingest_dir && out, SIEVE_EINVAL, "null argument");
        std::string I(index_dir), C(ingest_dir);
        Mapped cents, bo, blocks, ids, factors, rot, rr, quant, norm, t2s, t2c, t2n, seg;
        map_file(I + "/centroids.f32", cents, true);
        map_file(I + "/block_off.u64", bo, true);
        map_file(I + "/blocks.bin", blocks, true);
        map_file(I + "/ids.u32", ids, true);
        map_file(I + "/factors.f32", factors, true);
        map_file(I + "/rotation.f32", rot, true);
        map_file(C + "/rerank.i8", rr, true);
        map_file(C + "/quant.f32", quant, true);
        map_file(C + "/norm256.f32", norm, true);
        bool t2 = map_file(I + "/tier2.bin", t2c, false);
        if (t2) {
            map_file(I + "/tier2.sigma", t2s, true);
            map_file(I + "/tier2.norm", t2n, true);
        }
        sieve_index_desc d{};
        d.k = (uint32_t)(cents.len / (4 * PDIM));
        d.n = norm.len / 4;
        d.nblocks = blocks.len / BLOCK_BYTES;
        SIEVE_REQUIRE(cents.len == (size_t)d.k * PDIM * 4, SIEVE_EIO, "centroids.f32 size");
        SIEVE_REQUIRE(bo.len == (size_t)(d.k + 1) * 8, SIEVE_EIO, "block_off.u64 does not match centroids");
        SIEVE_REQUIRE(blocks.len == d.nblocks * BLOCK_BYTES, SIEVE_EIO, "blocks.bin size");
        SIEVE_REQUIRE(ids.len == d.nblocks * BLOCK * 4, SIEVE_EIO, "ids.u32 size");
        SIEVE_REQUIRE(factors.len == d.nblocks * BLOCK * 8, SIEVE_EIO, "factors.f32 size");
        SIEVE_REQUIRE(rot.len == 3 * PDIM * 4, SIEVE_EIO, "rotation.f32 size");
        SIEVE_REQUIRE(rr.len == d.n * DIM, SIEVE_EIO, "rerank.i8 size");
        SIEVE_REQUIRE(quant.len == DIM * 4, SIEVE_EIO, "quant.f32 size");
        d.centroids = (const float*)cents.p;
        d.block_off = (const uint64_t*)bo.p;
        d.blocks = (const uint8_t*)blocks.p;
        d.ids = (const uint32_t*)ids.p;
        d.factors = (const float*)factors.p;
        d.rotation = (const float*)rot.p;
        d.rerank = (const int8_t*)rr.p;
        d.quant = (const float*)quant.p;
        d.norm256 = (const float*)norm.p;
        if (t2) {
            SIEVE_REQUIRE(t2c.len == d.n * 64 && t2n.len == d.n * 4 && t2s.len == PDIM * 4, SIEVE_EIO, "tier2 size");
            d.tier2_codes = (const uint8_t*)t2c.p;
            d.tier2_norm = (const float*)t2n.p;
            d.tier2_sigma = (const float*)t2s.p;
        }
        // A compacted index (src/live.rs:395-468) carries its row segment: its lists reference
        // the segment's rows.
        const bool compacted = map_file(I + "/live.ext.u32", seg, false);
        SIEVE_REQUIRE(!compacted || seg.len % 4 == 0, SIEVE_EIO, "live.ext.u32 size");
        *out = create_index(&d, compacted ? seg.len / 4 : 0);
    });
}

void sieve_index_free(sieve_index* ix) { delete ix; }

int sieve_index_get_info(const sieve_index* ix, sieve_index_info* out) {
    return guarded([&] {
        SIEVE_REQUIRE(ix && out, SIEVE_EINVAL, "null argument");
        out->k = ix->k;
        out->n = ix->n;
        out->nblocks = ix->nblocks;
        out->rabitq = 1;
        out->tier2 = ix->t2_codes ? 1 : 0;
        out->max_cluster_blocks = ix->max_cluster_blocks;
        out->device_bytes = ix->device_bytes;
    });
}

int sieve_route(const sieve_index* ix, sieve_stream* s, const float* q256, size_t nq, const sieve_params* p,
                float* scores, uint32_t* clusters, uint32_t* counts, sieve_timing* t) {
    return guarded([&] {
        SIEVE_REQUIRE(ix && s && q256 && scores && clusters && counts, SIEVE_EINVAL, "null argument");
        check_params(p, true);
        if (nq == 0) return;
        const int stride = (int)p->nprobe;
        const float* q2 = dev_in(s, B_Q256, q256, nq * PDIM);
        // query prep needs a q768; routing ignores the int8 parts, so q256 stands in for it
        // only where prep reads 768 values: give it a zero buffer of the right size.
        float* q7 = grow<float>(s->b[B_Q768], nq * DIM);
        SIEVE_HIP(hipMemsetAsync(q7, 0, nq * DIM * 4, s->st));
        SIEVE_HIP(hipEventRecord(s->ev[0], s->st));
        Batch bt;
        prep_queries(ix, s, q2, q7, nq, bt);
        DevOut<float> os = dev_out(s, B_PS, scores, nq * stride);
        DevOut<uint32_t> oc = dev_out(s, B_PC, clusters, nq * stride);
        DevOut<uint32_t> on = dev_out(s, B_PN, counts, nq);
        route_batch(ix, s, bt, nq, p, os.dev, oc.dev, on.dev, stride);
        SIEVE_HIP(hipEventRecord(s->ev[1], s->st));
        out_done(s, os);
        out_done(s, oc);
        out_done(s, on);
        SIEVE_HIP(hipStreamSynchronize(s->st));
        s->last = sieve_stats{};
        s->last.queries = nq;
        s->last.route_ns = elapsed_ns(s->ev[0], s->ev[1]);
        if (t) t->route += (uint64_t)s->last.route_ns;
    });
}

int sieve_candidates(const sieve_index* ix, sieve_stream* s, const float* q256, size_t nq, const float* probe_scores,
                     const uint32_t* probe_clusters, const uint32_t* probe_counts, size_t nprobe_stride,
                     const sieve_params* p, const sieve_view* v, float* est, uint32_t* rows, float* errs,
                     uint32_t* counts, sieve_timing* t) {
    return guarded([&] {
        SIEVE_REQUIRE(ix && s && q256 && probe_scores && probe_clusters && probe_counts && est && rows && errs && counts,
                      SIEVE_EINVAL, "null argument");
        check_params(p, false);
        SIEVE_REQUIRE(ix->seg_rows == 0, SIEVE_EINVAL,
                      "the index's lists hold rows of its live segment (live.*): search it through its live state");
        SIEVE_REQUIRE(nprobe_stride >= 1 && nprobe_stride <= (size_t)MAX_PROBE, SIEVE_ELIMIT, "probe stride limit");
        if (nq == 0) return;
        // Probe lists come from the caller: check them on the host before any kernel reads them.
        std::vector<uint32_t> hn(nq), hc(nq * nprobe_stride);
        if (on_device(probe_counts)) SIEVE_HIP(hipMemcpy(hn.data(), probe_counts, nq * 4, hipMemcpyDeviceToHost));
        else std::memcpy(hn.data(), probe_counts, nq * 4);
        if (on_device(probe_clusters)) SIEVE_HIP(hipMemcpy(hc.data(), probe_clusters, hc.size() * 4, hipMemcpyDeviceToHost));
        else std::memcpy(hc.data(), probe_clusters, hc.size() * 4);
        for (size_t i = 0; i < nq; i++) {
            SIEVE_REQUIRE(hn[i] <= nprobe_stride, SIEVE_EINVAL, "probe count exceeds stride");
            for (uint32_t j = 0; j < hn[i]; j++)
                SIEVE_REQUIRE(hc[i * nprobe_stride + j] < ix->k, SIEVE_EINVAL, "probe cluster out of range");
        }
        const float* q2 = dev_in(s, B_Q256, q256, nq * PDIM);
        float* q7 = grow<float>(s->b[B_Q768], nq * DIM);
        SIEVE_HIP(hipMemsetAsync(q7, 0, nq * DIM * 4, s->st));
        const float* ps = dev_in(s, B_PS, probe_scores, nq * nprobe_stride);
        const uint32_t* pc = dev_in(s, B_PC, probe_clusters, nq * nprobe_stride);
        const uint32_t* pn = dev_in(s, B_PN, probe_counts, nq);
        ViewArgs view = make_view(s, v, nq);
        unsigned long long* cnt = grow<unsigned long long>(s->b[B_CNT], 8);
        SIEVE_HIP(hipMemsetAsync(cnt, 0, 64, s->st));
        Batch bt;
        prep_queries(ix, s, q2, q7, nq, bt);
        DevOut<float> oe = dev_out(s, B_CE, est, nq * p->cand);
        DevOut<uint32_t> orow = dev_out(s, B_CR, rows, nq * p->cand);
        DevOut<float> oerr = dev_out(s, B_CX, errs, nq * p->cand);
        DevOut<uint32_t> on = dev_out(s, B_CN, counts, nq);
        SIEVE_HIP(hipEventRecord(s->ev[0], s->st));
        scan_batch(ix, s, bt, nq, ps, pc, pn, (int)nprobe_stride, p, view, oe.dev, orow.dev, oerr.dev, on.dev,
                   (int)p->cand, cnt);
        SIEVE_HIP(hipEventRecord(s->ev[1], s->st));
        out_done(s, oe);
        out_done(s, orow);
        out_done(s, oerr);
        out_done(s, on);
        unsigned long long hcnt[4] = {0, 0, 0, 0};
        SIEVE_HIP(hipMemcpyAsync(hcnt, cnt, 32, hipMemcpyDeviceToHost, s->st));
        SIEVE_HIP(hipStreamSynchronize(s->st));
        s->last = sieve_stats{};
        s->last.queries = nq;
        s->last.scanned = hcnt[0];
        s->last.id_reads = hcnt[2];
        s->last.rescanned = hcnt[3];
        for (size_t i = 0; i < nq; i++) s->last.probed += hn[i];
        s->last.scan_ns = elapsed_ns(s->ev[0], s->ev[1]);
        if (t) {
            t->scan += (uint64_t)s->last.scan_ns;
            t->scanned += hcnt[0];
            t->probed += s->last.probed;
        }
    });
}

int sieve_rerank(const sieve_index* ix, sieve_stream* s, const float* q768, const float* q256, size_t nq,
                 const float* est, const uint32_t* rows, const float* errs, const uint32_t* cand_counts,
                 size_t cand_stride, const sieve_params* p, const sieve_view* v, float* scores, uint32_t* out_rows,
                 uint32_t* counts, sieve_timing* t) {
    (void)v;
    return guarded([&] {
        SIEVE_REQUIRE(ix && s && q768 && q256 && est && rows && errs && cand_counts && scores && out_rows && counts,
                      SIEVE_EINVAL, "null argument");
        check_params(p, false);
        SIEVE_REQUIRE(ix->seg_rows == 0, SIEVE_EINVAL,
                      "the index's lists hold rows of its live segment (live.*): search it through its live state");
        SIEVE_REQUIRE(cand_stride >= 1 && cand_stride <= (size_t)MAX_CAND, SIEVE_ELIMIT, "candidate stride limit");
        if (nq == 0) return;
        std::vector<uint32_t> hn(nq), hr(nq * cand_stride);
        if (on_device(cand_counts)) SIEVE_HIP(hipMemcpy(hn.data(), cand_counts, nq * 4, hipMemcpyDeviceToHost));
        else std::memcpy(hn.data(), cand_counts, nq * 4);
        if (on_device(rows)) SIEVE_HIP(hipMemcpy(hr.data(), rows, hr.size() * 4, hipMemcpyDeviceToHost));
        else std::memcpy(hr.data(), rows, hr.size() * 4);
        for (size_t i = 0; i < nq; i++) {
            SIEVE_REQUIRE(hn[i] <= cand_stride, SIEVE_EINVAL, "candidate count exceeds stride");
            for (uint32_t j = 0; j < hn[i]; j++)
                SIEVE_REQUIRE(hr[i * cand_stride + j] < ix->n, SIEVE_EINVAL, "candidate row out of range");
        }
        const float* q2 = dev_in(s, B_Q256, q256, nq * PDIM);
        const float* q7 = dev_in(s, B_Q768, q768, nq * DIM);
        const float* ce = dev_in(s, B_CE, est, nq * cand_stride);
        const uint32_t* cr = dev_in(s, B_CR, rows, nq * cand_stride);
        const float* cx = dev_in(s, B_CX, errs, nq * cand_stride);
        const uint32_t* cn = dev_in(s, B_CN, cand_counts, nq);
        unsigned long long* cnt = grow<unsigned long long>(s->b[B_CNT], 8);
        SIEVE_HIP(hipMemsetAsync(cnt, 0, 64, s->st));
        Batch bt;
        prep_queries(ix, s, q2, q7, nq, bt);
        DevOut<float> os = dev_out(s, B_OS, scores, nq * p->k);
        DevOut<uint32_t> orr = dev_out(s, B_OR, out_rows, nq * p->k);
        DevOut<uint32_t> on = dev_out(s, B_ON, counts, nq);
        SIEVE_HIP(hipEventRecord(s->ev[0], s->st));
        rerank_batch(ix, s, bt, nq, ce, cr, cx, cn, (int)cand_stride, p, os.dev, orr.dev, on.dev, cnt + 1);
        SIEVE_HIP(hipEventRecord(s->ev[1], s->st));
        out_done(s, os);
        out_done(s, orr);
        out_done(s, on);
        unsigned long long hcnt[4] = {0, 0, 0, 0};
        SIEVE_HIP(hipMemcpyAsync(hcnt, cnt, 32, hipMemcpyDeviceToHost, s->st));
        SIEVE_HIP(hipStreamSynchronize(s->st));
        s->last = sieve_stats{};
        s->last.queries = nq;
        s->last.reranked = hcnt[1];
        s->last.rerank_ns = elapsed_ns(s->ev[0], s->ev[1]);
        if (t) {
            t->rerank += (uint64_t)s->last.rerank_ns;
            t->reranked += hcnt[1];
        }
    });
}

}  // extern "C"

namespace sieve {

// enqueue_search as a HIP graph. The first call with a given key (index, its scan settings,
// batch size, parameters, view and device addresses) runs eagerly and sizes the stream's
// scratch; the next one captures the launches and every later one replays them, as long as no
// scratch was reallocated since (the graph holds scratch addresses). Returns 0 (eager),
// 1 (captured and launched) or 2 (replayed).
static uint64_t search_graphed(const sieve_index* ix, sieve_stream* s, const float* q2, const float* q7, size_t nq,
                               const sieve_params* p, const ViewArgs& view, float* os, uint32_t* orr, uint32_t* on) {
    std::string key;
    auto put = [&](const auto& v) { key.append(reinterpret_cast<const char*>(&v), sizeof v); };
    put(ix->uid);
    put(ix->scan_variant);
    put(ix->scan_chunk_blocks);
    put(ix->cm_min_queries);
    put(ix->cm_seed_slots);
    put(ix->cm_ext_cap);
    put(ix->cm_chunk_blocks);
    put(ix->cm_phases);
    put(ix->small_max_queries);
    put(ix->route_fuse);
    put(nq);
    put(*p);
    put(view.deleted);
    put(view.deleted_words);
    put(view.allow);
    put(view.allow_words);
    put(view.allow_stride);
    put(view.allow_flags);
    put(q2);
    put(q7);
    put(os);
    put(orr);
    put(on);
    auto it = std::find_if(s->gcache.begin(), s->gcache.end(), [&](const GraphEntry& g) { return g.key == key; });
    auto eager = [&](GraphEntry* g) {
        enqueue_search(ix, s, q2, q7, nq, p, view, os, orr, on);
        if (g) {
            g->gen = g_scratch_gen.load();
            g->cm_timed = s->cm_timed;
        }
        return (uint64_t)0;
    };
    if (it == s->gcache.end()) {
        if (s->gcache.size() >= 32) {
            if (s->gcache.front().exec) (void)hipGraphExecDestroy(s->gcache.front().exec);
            s->gcache.erase(s->gcache.begin());
        }
        s->gcache.push_back(GraphEntry{key});
        return eager(&s->gcache.back());
    }
    GraphEntry& g = *it;
    if (g.failed) return eager(nullptr);
    if (g.gen != g_scratch_gen.load()) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        g.exec = nullptr;
        return eager(&g);
    }
    if (g.exec) {
        SIEVE_HIP(hipGraphLaunch(g.exec, s->st));
        s->cm_timed = g.cm_timed;
        return 2;
    }
    // Capture. Relaxed mode: the stage functions may query function attributes; nothing in
    // them synchronises or allocates at this size (checked by the scratch generation).
    SIEVE_HIP(hipStreamBeginCapture(s->st, hipStreamCaptureModeRelaxed));
    hipGraph_t graph = nullptr;
    try {
        enqueue_search(ix, s, q2, q7, nq, p, view, os, orr, on);
    } catch (...) {
        (void)hipStreamEndCapture(s->st, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        (void)hipGetLastError();
        g.failed = true;
        return eager(nullptr);
    }
    hipError_t e = hipStreamEndCapture(s->st, &graph);
    hipGraphExec_t exec = nullptr;
    if (e == hipSuccess && graph && g.gen == g_scratch_gen.load())
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (graph) (void)hipGraphDestroy(graph);
    if (e != hipSuccess || !exec) {
        if (exec) (void)hipGraphExecDestroy(exec);
        (void)hipGetLastError();
        g.failed = true;
        return eager(nullptr);
    }
    g.exec = exec;
    g.cm_timed = s->cm_timed;
    SIEVE_HIP(hipGraphLaunch(g.exec, s->st));
    return 1;
}

}  // namespace sieve


namespace sieve {
// Query prep and routing of nq L2 queries already in device memory (bigann::Index::route,
// src/bigann.rs:660-678): the probe lists (<q, c> per probe, leaf ids, counts; row stride
// min(nprobe, k)) in stream scratch, enqueued on s's stream. with_qrot keeps the rotated queries
// (the refinement tables' operand) in prep.qrot.
void l2_route(const sieve_l2_index* ix, sieve_stream* s, const int8_t* dq, size_t nq, uint32_t nprobe, float margin,
              bool with_qrot, L2Prep& prep, float** ps_out, uint32_t** pc_out, uint32_t** pn_out) {
    constexpr int D = 100;
    const int pstride = (int)std::min<uint32_t>(nprobe, ix->k);
    prep.qrot = with_qrot ? grow<float>(s->b[B_QROT], nq * 128) : nullptr;
    prep.tables = grow<uint8_t>(s->b[B_TABLES], nq * 512);
    prep.qab = grow<float4>(s->b[B_QAB], nq);
    prep.qbf = grow<uint16_t>(s->b[B_QBF], nq * PDIM);
    prep.qf = grow<float>(s->b[B_QF], nq * D);
    prep.qq = grow<float>(s->b[B_QQ], nq);
    prep.qnorm = grow<float>(s->b[B_QNORM], nq);
    float* ps = grow<float>(s->b[B_PS], nq * pstride);
    uint32_t* pc = grow<uint32_t>(s->b[B_PC], nq * pstride);
    uint32_t* pn = grow<uint32_t>(s->b[B_PN], nq);
    // 128^-1.5 as (d as f32).powf(-1.5) (src/rabitq.rs:67)
    launch_query_prep_l2(dq, nq, ix->rotation, std::pow(128.0f, -1.5f), prep, s->st);
    const size_t per = std::max<size_t>(1, std::min<size_t>(nq, (size_t(1) << 30) / (4 * (size_t)ix->k)));
    float* approx = grow<float>(s->b[B_APPROX], per * ix->k);
    float* tmax = grow<float>(s->b[B_TMAX], per * ((ix->k + 15) / 16));
    // Fused: the select recomputes the scores it reads and the GEMM writes none.
    const bool fused = route_fuse(ix->route_fuse, (int)ix->k, (int)nprobe);
    for (size_t q0 = 0; q0 < nq; q0 += per) {
        size_t m = std::min(per, nq - q0);
        launch_gemm_bf16_route(prep.qbf + q0 * PDIM, ix->leaves_bf, ix->leaf_norm, fused ? nullptr : approx, tmax,
                               (int)m, (int)ix->k, ix->k, s->st);
        RouteArgs ra{};
        ra.fused = fused;
        ra.qbf = prep.qbf + q0 * PDIM;
        ra.cbf = ix->leaves_bf;
        ra.approx = approx;
        ra.lds = ix->k;
        ra.tmax = tmax;
        ra.k = (int)ix->k;
        ra.cents = ix->leaves;
        ra.q256 = nullptr;
        ra.qnorm = prep.qnorm + q0;
        ra.cl_slots = nullptr;
        ra.cent_norm_max = ix->cent_norm_max;
        ra.nprobe = (int)nprobe;
        ra.margin = margin;
        ra.budget = 0;
        ra.out_s = ps + q0 * pstride;
        ra.out_c = pc + q0 * pstride;
        ra.out_n = pn + q0;
        ra.out_stride = pstride;
        ra.l2 = 1;
        ra.dims = D;
        ra.cstride = D;
        ra.cnorm = ix->leaf_norm;
        ra.qf = prep.qf + q0 * D;
        ra.qq = prep.qq + q0;
        // 2 <q, c> from BF16 hi + lo halves: |err| <= 2 (2^-16 + 256 * 2^-24 * 1.01) |q| |c|
        ra.err_coef = 2.0f * (1.0f / 65536.0f + 1.01f * 256.0f / 16777216.0f) * ix->cent_norm_max;
        launch_route_select(ra, m, s->st);
    }
    *ps_out = ps;
    *pc_out = pc;
    *pn_out = pn;
}
}  // namespace sieve

extern "C" {

int sieve_search_batch(const sieve_index* ix, sieve_stream* s, const float* q768, const float* q256, size_t nq,
                       const sieve_params* p, const sieve_view* v, float* scores, uint32_t* rows, uint32_t* counts,
                       sieve_timing* t) {
    return guarded([&] {
        SIEVE_REQUIRE(ix && s && q768 && q256 && scores && rows && counts, SIEVE_EINVAL, "null argument");
        check_params(p, true);
        SIEVE_REQUIRE(ix->seg_rows == 0, SIEVE_EINVAL,
                      "the index's lists hold rows of its live segment (live.*): search it through its live state");
        if (nq == 0) return;
        const float* q2 = dev_in(s, B_Q256, q256, nq * PDIM);
        const float* q7 = dev_in(s, B_Q768, q768, nq * DIM);
        ViewArgs view = make_view(s, v, nq);
        DevOut<float> os = dev_out(s, B_OS, scores, nq * p->k);
        DevOut<uint32_t> orr = dev_out(s, B_OR, rows, nq * p->k);
        DevOut<uint32_t> on = dev_out(s, B_ON, counts, nq);
        uint64_t graph = 0;
        if (s->graphs)
            graph = search_graphed(ix, s, q2, q7, nq, p, view, os.dev, orr.dev, on.dev);
        else
            enqueue_search(ix, s, q2, q7, nq, p, view, os.dev, orr.dev, on.dev);
        out_done(s, os);
        out_done(s, orr);
        out_done(s, on);
        finish_search(s, nq, t);
        s->last.graph = graph;
    });
}

int sieve_l2_index_create(const sieve_l2_desc* desc, sieve_l2_index** out) {
    return guarded([&] {
        SIEVE_REQUIRE(out, SIEVE_EINVAL, "out is null");
        *out = create_l2_index(desc);
    });
}

void sieve_l2_index_free(sieve_l2_index* ix) { delete ix; }

int sieve_l2_search_batch(const sieve_l2_index* ix, sieve_stream* s, const int8_t* q, size_t nq,
                          const sieve_l2_params* p, float* scores, uint32_t* rows, uint32_t* counts,
                          uint32_t* probes, uint32_t* probe_counts, uint32_t* cand_rows, float* cand_est,
                          uint32_t* cand_counts, sieve_timing* t) {
    return guarded([&] {
        constexpr int D = 100;
        SIEVE_REQUIRE(ix && s && q && p && scores && rows && counts, SIEVE_EINVAL, "null argument");
        SIEVE_REQUIRE(p->nprobe >= 1 && p->nprobe <= (uint32_t)MAX_PROBE, SIEVE_ELIMIT,
                      "nprobe must be in [1, " + std::to_string(MAX_PROBE) + "]");
        SIEVE_REQUIRE(p->cand >= 1 && p->cand <= (uint32_t)MAX_CAND, SIEVE_ELIMIT,
                      "cand must be in [1, " + std::to_string(MAX_CAND) + "]");
        SIEVE_REQUIRE(p->k >= 1 && p->k <= (uint32_t)MAX_CAND, SIEVE_ELIMIT, "k out of range");
        if (nq == 0) return;
        const int pstride = (int)std::min<uint32_t>(p->nprobe, ix->k);
        // src/bigann.rs:688: refinement only on an index with bits > 1 and refine > 0.
        const bool refining = ix->bits > 1 && p->refine > 0;
        // A shortlist whose buffer (2 K) outgrows every slot a query can probe is never cut, so
        // it needs no storage: K = 0 tells the kernel so.
        uint64_t K = std::max<uint32_t>(p->refine, p->cand);
        if (2 * K > (uint64_t)pstride * std::max<uint32_t>(ix->max_cluster_blocks, 1) * BLOCK) K = 0;
        if (refining)
            SIEVE_REQUIRE(K <= 2 * (uint64_t)MAX_CAND && refine_lds_bytes((int)K, (int)p->cand, pstride) <= 160 * 1024,
                          SIEVE_ELIMIT,
                          "refine: a shortlist of " + std::to_string(K) + " slots exceeds the scan's LDS (at most " +
                              std::to_string(2 * MAX_CAND) + ")");
        const int8_t* dq = dev_in(s, B_Q768, q, nq * D);
        unsigned long long* cnt = grow<unsigned long long>(s->b[B_CNT], 8);
        uint32_t* ovf = grow<uint32_t>(s->b[B_OVF], 1);
        float* ce = grow<float>(s->b[B_CE], nq * p->cand);
        uint32_t* cr = grow<uint32_t>(s->b[B_CR], nq * p->cand);
        float* cx = grow<float>(s->b[B_CX], nq * p->cand);
        uint32_t* cn = grow<uint32_t>(s->b[B_CN], nq);
        DevOut<float> os = dev_out(s, B_OS, scores, nq * p->k);
        DevOut<uint32_t> orr = dev_out(s, B_OR, rows, nq * p->k);
        DevOut<uint32_t> on = dev_out(s, B_ON, counts, nq);
        SIEVE_HIP(hipMemsetAsync(cnt, 0, 64, s->st));
        SIEVE_HIP(hipMemsetAsync(ovf, 0, 4, s->st));
        SIEVE_HIP(hipEventRecord(s->ev[0], s->st));
        L2Prep prep;
        float* ps;
        uint32_t *pc, *pn;
        l2_route(ix, s, dq, nq, p->nprobe, p->margin, refining, prep, &ps, &pc, &pn);
        SIEVE_HIP(hipEventRecord(s->ev[1], s->st));
        // In L2 mode the probe scores are <q, c> (the estimator's `cent`, src/bigann.rs:701).
        ScanArgs a{};
        a.codes = ix->codes;
        a.mul = ix->mul;
        a.err = ix->nx;
        a.ids = ix->ids;
        a.block_off = ix->block_off;
        a.tables = prep.tables;
        a.qab = prep.qab;
        a.probe_s = ps;
        a.probe_c = pc;
        a.probe_n = pn;
        a.probe_stride = pstride;
        a.cand = (int)p->cand;
        a.out_est = ce;
        a.out_row = cr;
        a.out_err = cx;
        a.out_n = cn;
        a.out_stride = (int)p->cand;
        a.scanned = cnt;
        const size_t lds = scan_lds_bytes(a.cand, pstride);
        SIEVE_REQUIRE(lds <= 160 * 1024, SIEVE_ELIMIT, "scan workspace exceeds LDS");
        if (refining) {
            // The 1-bit shortlist replayed in probe order, refined estimates for its entrants,
            // the best cand of those (src/bigann.rs:683-729).
            L2RefineArgs ra{};
            ra.codes = ix->codes;
            ra.mul = ix->mul;
            ra.nx = ix->nx;
            ra.ids = ix->ids;
            ra.block_off = ix->block_off;
            ra.planes = ix->rplanes;
            ra.rf = ix->rf;
            ra.bits = ix->bits;
            ra.tables = prep.tables;
            ra.qab = prep.qab;
            ra.qrot = prep.qrot;
            ra.probe_s = ps;
            ra.probe_c = pc;
            ra.probe_n = pn;
            ra.probe_stride = pstride;
            ra.cand = (int)p->cand;
            ra.K = (int)K;
            ra.out_est = ce;
            ra.out_row = cr;
            ra.out_nx = cx;
            ra.out_n = cn;
            ra.out_stride = (int)p->cand;
            ra.scanned = cnt;
            launch_scan_l2_refine(ra, nq, s->st);
        } else if (nq <= ix->small_max_queries) {
            // A few queries: per-chunk bounds, a threshold per query, appended lists, the merge.
            const uint32_t cand = (uint32_t)a.cand;
            const double worst = (double)pstride * std::max<uint32_t>(ix->max_cluster_blocks, 1);
            // At most 512 bound and 256 append chunks per query (larger chunks take more rounds).
            const uint32_t bound_chunks = (uint32_t)std::min(512.0, std::max(1.0, std::ceil(worst / 32.0)));
            const uint32_t append_chunks = (uint32_t)std::min(256.0, std::max(1.0, std::ceil(worst / 64.0)));
            SIEVE_REQUIRE(nq * (double)bound_chunks < 2.0e9, SIEVE_ELIMIT, "scan grid too large");
            a.part = grow<uint2>(s->b[B_PART], nq * (size_t)bound_chunks * 16);
            a.part_n = grow<uint32_t>(s->b[B_PARTN], nq * (size_t)bound_chunks);
            a.ext_cap = 64 * cand;
            a.seed = grow<uint2>(s->b[B_CM_SEED], nq * (size_t)cand);
            a.seed_n = grow<uint32_t>(s->b[B_CM_SEEDN], nq);
            a.seed_thr = grow<unsigned long long>(s->b[B_CM_THR], nq);
            a.ext = grow<uint2>(s->b[B_CM_EXT], nq * (size_t)a.ext_cap);
            a.ext_n = grow_zeroed<uint32_t>(s->b[B_CM_EXTN], nq, s->st);  // the finish pass leaves it zero
            launch_scan_small(a, nq, true, bound_chunks, 64, append_chunks, s->st);
        } else {
            set_scan_plan(a, s, nq, (uint32_t)pstride, (double)ix->nblocks / ix->k, ix->max_cluster_blocks,
                          ix->scan_chunk_blocks);
            launch_scan_l2(a, nq, ix->scan_variant, s->st);
        }
        SIEVE_HIP(hipGetLastError());
        SIEVE_HIP(hipEventRecord(s->ev[2], s->st));
        L2RerankArgs r{};
        r.q = dq;
        r.base = ix->base;
        r.cand_ids = cr;
        r.cand_nx = cx;
        r.cand_n = cn;
        r.cand_stride = (int)p->cand;
        r.k = (int)p->k;
        r.out_s = os.dev;
        r.out_r = orr.dev;
        r.out_n = on.dev;
        launch_rerank_l2(r, nq, s->st);
        SIEVE_HIP(hipGetLastError());
        SIEVE_HIP(hipEventRecord(s->ev[3], s->st));
        out_done(s, os);
        out_done(s, orr);
        out_done(s, on);
        auto copy_out = [&](void* dst, const void* src, size_t bytes) {
            if (!dst) return;
            SIEVE_HIP(hipMemcpyAsync(dst, src, bytes, on_device(dst) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                                     s->st));
        };
        copy_out(probes, pc, nq * pstride * 4);
        copy_out(probe_counts, pn, nq * 4);
        copy_out(cand_rows, cr, nq * p->cand * 4);
        copy_out(cand_est, ce, nq * p->cand * 4);
        copy_out(cand_counts, cn, nq * 4);
        // One small copy into pinned memory: the counters and the probe total.
        unsigned long long* packed = grow<unsigned long long>(s->b[B_STATS], 8);
        launch_pack_stats(cnt, pn, nq, ovf, packed, s->st);
        if (!s->hstats)
            SIEVE_HIP(hipHostMalloc((void**)&s->hstats, 8 * sizeof(unsigned long long), hipHostMallocDefault));
        SIEVE_HIP(hipMemcpyAsync(s->hstats, packed, 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s->st));
        SIEVE_HIP(hipStreamSynchronize(s->st));
        const unsigned long long* hcnt = s->hstats;
        s->last = sieve_stats{};
        s->last.queries = nq;
        s->last.scanned = hcnt[0];
        s->last.id_reads = hcnt[2];
        s->last.rescanned = hcnt[3];
        s->last.probed = hcnt[4];
        s->last.route_ns = elapsed_ns(s->ev[0], s->ev[1]);
        s->last.scan_ns = elapsed_ns(s->ev[1], s->ev[2]);
        s->last.rerank_ns = elapsed_ns(s->ev[2], s->ev[3]);
        if (t) {
            t->route += (uint64_t)s->last.route_ns;
            t->scan += (uint64_t)s->last.scan_ns;
            t->rerank += (uint64_t)s->last.rerank_ns;
            t->scanned += hcnt[0];
            t->probed += s->last.probed;
        }
    });
}

int sieve_index_autotune(sieve_index* ix, sieve_stream* s, const float* q768, const float* q256, size_t nq,
                         const sieve_params* p, uint32_t* chosen, float* ms_per_variant) {
    return guarded([&] {
        SIEVE_REQUIRE(ix && s && q768 && q256 && p, SIEVE_EINVAL, "null argument");
        check_params(p, true);
        SIEVE_REQUIRE(nq >= 1, SIEVE_EINVAL, "autotune needs at least one query");
        const int pstride = (int)std::min<uint32_t>(p->nprobe, ix->k);
        const float* q2 = dev_in(s, B_Q256, q256, nq * PDIM);
        const float* q7 = dev_in(s, B_Q768, q768, nq * DIM);
        unsigned long long* cnt = grow<unsigned long long>(s->b[B_CNT], 8);
        float* ps = grow<float>(s->b[B_PS], nq * pstride);
        uint32_t* pc = grow<uint32_t>(s->b[B_PC], nq * pstride);
        uint32_t* pn = grow<uint32_t>(s->b[B_PN], nq);
        float* ce = grow<float>(s->b[B_CE], nq * p->cand);
        uint32_t* cr = grow<uint32_t>(s->b[B_CR], nq * p->cand);
        float* cx = grow<float>(s->b[B_CX], nq * p->cand);
        uint32_t* cn = grow<uint32_t>(s->b[B_CN], nq);
        Batch bt;
        prep_queries(ix, s, q2, q7, nq, bt);
        route_batch(ix, s, bt, nq, p, ps, pc, pn, pstride);
        int best = 0;
        float best_ms = 0.0f;
        const int saved = ix->scan_variant;
        for (int v = 0; v < SCAN_VARIANTS; v++) {
            ix->scan_variant = v;
            float ms = 0.0f;
            for (int rep = 0; rep < 3; rep++) {  // first run warms up; keep the faster of the rest
                SIEVE_HIP(hipEventRecord(s->ev[0], s->st));
                scan_batch(ix, s, bt, nq, ps, pc, pn, pstride, p, ViewArgs{}, ce, cr, cx, cn, (int)p->cand, cnt);
                SIEVE_HIP(hipEventRecord(s->ev[1], s->st));
                SIEVE_HIP(hipEventSynchronize(s->ev[1]));
                float t = elapsed_ns(s->ev[0], s->ev[1]) / 1e6f;
                if (rep == 1 || (rep > 1 && t < ms)) ms = t;
            }
            if (ms_per_variant) ms_per_variant[v] = ms;
            if (v == 0 || ms < best_ms) {
                best = v;
                best_ms = ms;
            }
        }
        (void)saved;
        ix->scan_variant = best;
        if (chosen) *chosen = (uint32_t)best;
    });
}

int sieve_index_set_scan_variant(sieve_index* ix, uint32_t variant) {
    return guarded([&] {
        SIEVE_REQUIRE(ix, SIEVE_EINVAL, "null index");
        SIEVE_REQUIRE(variant < (uint32_t)SCAN_VARIANTS, SIEVE_EINVAL, "no such scan variant");
        ix->scan_variant = (int)variant;
    });
}

int sieve_index_set_scan_chunk(sieve_index* ix, uint32_t chunk_blocks) {
    return guarded([&] {
        SIEVE_REQUIRE(ix, SIEVE_EINVAL, "null index");
        ix->scan_chunk_blocks = chunk_blocks;
    });
}

int sieve_index_set_cluster_scan(sieve_index* ix, uint32_t min_queries, uint32_t seed_slots, uint32_t ext_cap,
                                 uint32_t chunk_blocks, uint32_t phases) {
    return guarded([&] {
        SIEVE_REQUIRE(ix, SIEVE_EINVAL, "null index");
        SIEVE_REQUIRE(chunk_blocks <= (1u << 20), SIEVE_EINVAL, "chunk_blocks too large");
        SIEVE_REQUIRE(phases <= 64, SIEVE_EINVAL, "at most 64 phases");
        ix->cm_min_queries = min_queries ? min_queries : 64;
        ix->cm_seed_slots = seed_slots;
        ix->cm_ext_cap = ext_cap;
        ix->cm_chunk_blocks = chunk_blocks;
        ix->cm_phases = phases ? phases : 2;
    });
}

int sieve_index_set_route_fuse(sieve_index* ix, int mode) {
    return guarded([&] {
        SIEVE_REQUIRE(ix, SIEVE_EINVAL, "null index");
        ix->route_fuse = mode < 0 ? -1 : mode > 0 ? 1 : 0;
    });
}

int sieve_l2_index_set_route_fuse(sieve_l2_index* ix, int mode) {
    return guarded([&] {
        SIEVE_REQUIRE(ix, SIEVE_EINVAL, "null index");
        ix->route_fuse = mode < 0 ? -1 : mode > 0 ? 1 : 0;
    });
}

int sieve_index_set_small_batch(sieve_index* ix, uint32_t max_queries) {
    return guarded([&] {
        SIEVE_REQUIRE(ix, SIEVE_EINVAL, "null index");
        ix->small_max_queries = max_queries;
    });
}

int sieve_l2_index_set_small_batch(sieve_l2_index* ix, uint32_t max_queries) {
    return guarded([&] {
        SIEVE_REQUIRE(ix, SIEVE_EINVAL, "null index");
        ix->small_max_queries = max_queries;
    });
}

int sieve_l2_index_set_scan_variant(sieve_l2_index* ix, uint32_t variant) {
    return guarded([&] {
        SIEVE_REQUIRE(ix, SIEVE_EINVAL, "null index");
        SIEVE_REQUIRE(variant < (uint32_t)SCAN_VARIANT_CM, SIEVE_EINVAL, "no such scan variant for the L2 index");
        ix->scan_variant = (int)variant;
    });
}

int sieve_l2_refine_encode(const float* r, size_t n, uint32_t bits, uint8_t* planes, float* f) {
    return guarded([&] {
        SIEVE_REQUIRE(r && planes && f, SIEVE_EINVAL, "null argument");
        SIEVE_REQUIRE(bits >= 2 && bits <= SIEVE_REFINE_MAX_BITS, SIEVE_EINVAL, "bits must be 2 or 3");
        if (n == 0) return;
        hipStream_t st = nullptr;
        std::vector<void*> tmp;
        auto dev_copy_in = [&](const float* src) -> const float* {
            if (on_device(src)) return src;
            float* d = dmalloc<float>(n * 128, nullptr);
            tmp.push_back(d);
            SIEVE_HIP(hipMemcpy(d, src, n * 128 * 4, hipMemcpyHostToDevice));
            return d;
        };
        try {
            const float* dr = dev_copy_in(r);
            uint4* dp = on_device(planes) ? reinterpret_cast<uint4*>(planes) : nullptr;
            float* df = on_device(f) ? f : nullptr;
            if (!dp) tmp.push_back(dp = dmalloc<uint4>(2 * n, nullptr));
            if (!df) tmp.push_back(df = dmalloc<float>(n, nullptr));
            launch_refine_encode(dr, n, (int)bits, dp, df, st);
            SIEVE_HIP(hipGetLastError());
            SIEVE_HIP(hipDeviceSynchronize());
            if ((void*)dp != (void*)planes) SIEVE_HIP(hipMemcpy(planes, dp, n * 32, hipMemcpyDeviceToHost));
            if (df != f) SIEVE_HIP(hipMemcpy(f, df, n * 4, hipMemcpyDeviceToHost));
        } catch (...) {
            for (void* x : tmp) (void)hipFree(x);
            throw;
        }
        for (void* x : tmp) (void)hipFree(x);
    });
}

int sieve_l2_index_set_scan_chunk(sieve_l2_index* ix, uint32_t chunk_blocks) {
    return guarded([&] {
        SIEVE_REQUIRE(ix, SIEVE_EINVAL, "null index");
        ix->scan_chunk_blocks = chunk_blocks;
    });
}

}  // extern "C"
