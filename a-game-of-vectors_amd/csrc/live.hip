// Live updates on the device: Live (src/live.rs). The built index stays as it is; inserted
// rows are encoded exactly as the build encodes a row (int8 row, prefix from the int8 row,
// nearest centroids plus a SOAR spill, RaBitQ residual codes against each centroid, a 2-bit
// tier code) and their slots go to per-cluster delta lists that the scan reads right after
// each probed cluster's base list (src/ivf.rs:682-684); deletes set tombstones the scan checks
// at admission. Row ids: corpus rows keep theirs, delta rows count up from the corpus size;
// searches report external ids (src/live.rs:11-14).
//
// Device layout. The index's slot arrays (codes, mul, err, ids) get room for an arena after
// the base slots; the arena holds the delta lists cluster by cluster, each padded to whole
// 32-slot blocks, so the scan reads a delta list as one more block range. Every insert batch
// appends its slots to a log (insertion order) and rebuilds the arena from the log with a
// stable sort by cluster, which keeps each list in insertion order as DeltaList::push does.
//
// Concurrency: like the reference's &mut Live (src/live.rs:16-17), inserts and deletes must
// not overlap searches of the same index.
#include <float.h>
#include <cstring>

#include <hipcub/hipcub.hpp>

#include "live.hpp"

using namespace sieve;

namespace {

constexpr int SOAR_CAND = 16;  // src/live.rs:35
constexpr float EPS0 = 1.9f;   // src/rabitq.rs:28
constexpr float T2_T = 0.9816f;
__constant__ float T2_LEVELS[4] = {-1.510f, -0.4528f, 0.4528f, 1.510f};  // src/tier2.rs:10-11

__device__ __forceinline__ float sqrt_rn(float x) { return (float)sqrt((double)x); }

// quant::Int8Quantizer::encode (src/quant.rs:52-56) and the prefix norm of Live::insert
// (src/live.rs:280-283). One thread per row: every sum is sequential, as the reference's.
__global__ void k_live_quant(const float* __restrict__ vecs, size_t n, const float* __restrict__ quant,
                             int8_t* __restrict__ rows, float* __restrict__ norm256) {
    const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const float* v = vecs + j * DIM;
    int8_t* r = rows + j * DIM;
    for (int i = 0; i < DIM; i++) {
        float q = roundf(v[i] * quant[i]);
        q = fminf(fmaxf(q, -127.0f), 127.0f);
        r[i] = (int8_t)q;
    }
    float s = -0.0f;
    for (int i = 0; i < PDIM; i++) s = s + v[i] * v[i];
    norm256[j] = sqrt_rn(s);
}

// The dequantised, normalised prefix of apply_insert (src/live.rs:325-329), from the int8 row
// alone, so a replayed log gives the same codes.
__global__ void k_live_prefix(const int8_t* __restrict__ rows, size_t n, const float* __restrict__ quant,
                              float* __restrict__ xs) {
    const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int8_t* r = rows + j * DIM;
    float* x = xs + j * PDIM;
    float ss = -0.0f;
    for (int i = 0; i < PDIM; i++) {
        const float t = (float)r[i] / quant[i];
        x[i] = t;
        ss = ss + t * t;
    }
    const float nn = sqrt_rn(ss);  // quant::normalize, src/quant.rs:5-13
    if (nn > 0.0f) {
        const float inv = 1.0f / nn;
        for (int i = 0; i < PDIM; i++) x[i] = x[i] * inv;
    }
}

// The (primary, secondary) clusters of each row: assign_top's best centroid and
// soar_secondary over its SOAR_CAND best (src/ivf.rs:164-186); PAD without SOAR. With fewer
// than SOAR_CAND clusters, assign_top leaves the candidate list's tail at cluster 0
// (src/ivf.rs:141, :154), and so does this.
// One wave per row. The row needs 31 dot products: x . c1, then x . c_i and c1 . c_i for the
// other SOAR_CAND - 1 candidates, each summed as quant::dot_f32 (src/quant.rs:86-100) sums: 16
// lane accumulators, accumulator l adding products l, l + 16, ... in order, then the 16 added
// in order from -0.0. Lane group g (16 lanes) computes dots g, g + 4, ...: lane l holds
// accumulator l, and the group's first lane adds the 16, so each dot is bit for bit the
// sequential one, with the centroid rows read 64 bytes per group and step. The wave's first
// lane then takes the loss minimum in candidate order.
__global__ __launch_bounds__(64) void k_live_pairs(const float* __restrict__ xs, size_t n,
                                                   const float* __restrict__ cents, const uint32_t* __restrict__ pc,
                                                   const uint32_t* __restrict__ pn, int pstride, int soar,
                                                   float lambda, uint32_t* __restrict__ pairs) {
    constexpr int ND = 2 * SOAR_CAND - 1;
    __shared__ float part[4][16];
    __shared__ float dots[ND];
    const size_t j = blockIdx.x;  // grid: one workgroup per row
    if (j >= n) return;           // workgroup-uniform
    const int lane = threadIdx.x, g = lane >> 4, l = lane & 15;
    const uint32_t* cand = pc + j * pstride;
    const int m = min((int)pn[j], pstride);
    auto cand_at = [&](int i) -> uint32_t { return i < m ? cand[i] : 0u; };
    if (!soar) {
        if (lane == 0) {
            pairs[2 * j] = cand_at(0);
            pairs[2 * j + 1] = PAD;
        }
        return;
    }
    const float* x = xs + j * PDIM;
    const float* c1 = cents + (size_t)cand_at(0) * PDIM;
    for (int k0 = 0; k0 < ND; k0 += 4) {
        const int k = k0 + g;
        float acc = 0.0f;
        if (k < ND) {
            // k = 0: x . c1; k < SOAR_CAND: x . c_k; else c1 . c_(k - SOAR_CAND + 1)
            const float* u = k < SOAR_CAND ? x : c1;
            const float* v = k == 0 ? c1 : cents + (size_t)cand_at(k < SOAR_CAND ? k : k - SOAR_CAND + 1) * PDIM;
            for (int i = 0; i < PDIM; i += 16) acc = acc + u[i + l] * v[i + l];
        }
        part[g][l] = acc;
        __syncthreads();
        if (l == 0 && k < ND) {
            float sum = -0.0f;
#pragma unroll
            for (int q = 0; q < 16; q++) sum = sum + part[g][q];
            dots[k] = sum + -0.0f;
        }
        __syncthreads();
    }
    if (lane != 0) return;
    const float x_c1 = dots[0];
    const float r2 = fmaxf(2.0f - 2.0f * x_c1, 1e-12f);
    const float r_x = 1.0f - x_c1;
    float best = FLT_MAX;
    uint32_t bc = cand_at(1);
    for (int i = 1; i < SOAR_CAND; i++) {
        const float x_c = dots[i];
        const float r_c = x_c - dots[SOAR_CAND - 1 + i];
        const float along = r_x - r_c;
        const float loss = (2.0f - 2.0f * x_c) + lambda * along * along / r2;
        if (loss < best) {
            best = loss;
            bc = cand_at(i);
        }
    }
    pairs[2 * j] = cand_at(0);
    pairs[2 * j + 1] = bc;
}

// Per row (one 64-thread workgroup): Rotation::apply of the prefix, then for each of its
// clusters rabitq::encode against the rotated centroid (src/rabitq.rs:86-108) into the slot
// log, and tier2::encode of the prefix (src/tier2.rs:36-47).
__global__ __launch_bounds__(64) void k_live_encode(const float* __restrict__ xs, size_t n,
                                                    const float* __restrict__ signs, float rot_norm,
                                                    const float* __restrict__ cents_rot,
                                                    const uint32_t* __restrict__ pairs, int spr, uint32_t row0,
                                                    uint4* __restrict__ lcode, float* __restrict__ lmul,
                                                    float* __restrict__ lerr, uint32_t* __restrict__ lrow,
                                                    uint32_t* __restrict__ lcl, const float* __restrict__ t2_sigma,
                                                    uint8_t* __restrict__ t2, float* __restrict__ t2n) {
    __shared__ float v[PDIM];
    __shared__ float r[PDIM];
    __shared__ uint32_t code[8], chunked[8];
    const size_t j = blockIdx.x;
    const int lane = threadIdx.x;
    const float* x = xs + j * PDIM;
    for (int i = lane; i < PDIM; i += 64) v[i] = x[i];
    __syncthreads();
    if (t2_sigma && lane == 0) {
        uint8_t out[PDIM / 4];
        for (int b = 0; b < PDIM / 4; b++) out[b] = 0;
        float nn = 0.0f;
        for (int i = 0; i < PDIM; i++) {
            const float s = t2_sigma[i], xi = v[i], t = T2_T * s;
            const int l = xi < -t ? 0 : xi < 0.0f ? 1 : xi < t ? 2 : 3;
            out[i >> 2] |= (uint8_t)(l << ((i & 3) * 2));
            const float q = T2_LEVELS[l] * s;
            nn = nn + q * q;
        }
        for (int b = 0; b < PDIM / 4; b++) t2[j * 64 + b] = out[b];
        t2n[j] = sqrt_rn(nn);
    }
    __syncthreads();
    rotate_lds<PDIM>(v, signs, rot_norm, lane);
    for (int h = 0; h < spr; h++) {
        const uint32_t c = pairs[2 * j + h];
        const size_t L = j * spr + h;
        const float* cr = cents_rot + (size_t)c * PDIM;
        for (int i = lane; i < PDIM; i += 64) r[i] = v[i] - cr[i];
        __syncthreads();
        if (lane < 8) {
            uint32_t w = 0;
            for (int b = 0; b < 32; b++)
                if (r[lane * 32 + b] > 0.0f) w |= 1u << b;
            code[lane] = w;
        }
        if (lane == 0) {
            float nr2 = 0.0f, l1 = 0.0f;
            for (int i = 0; i < PDIM; i++) {
                nr2 = nr2 + r[i] * r[i];
                l1 = l1 + fabsf(r[i]);
            }
            float mul = 0.0f, err = 0.0f;
            if (!(l1 < 1e-12f)) {
                const float df = (float)PDIM;
                const float nr = sqrt_rn(nr2);
                const float od = l1 / (nr * sqrt_rn(df));
                float inner = (1.0f - od * od) / (od * od * (df - 1.0f));
                inner = fmaxf(inner, 0.0f);
                err = nr * EPS0 * sqrt_rn(inner);
                mul = nr2 / l1;
            }
            lmul[L] = mul;
            lerr[L] = err;
            lrow[L] = row0 + (uint32_t)j;
            lcl[L] = c;
        }
        __syncthreads();
        if (lane == 0) code_to_chunked(code, chunked);
        __syncthreads();
        if (lane < 2) lcode[2 * L + lane] = make_uint4(chunked[4 * lane], chunked[4 * lane + 1], chunked[4 * lane + 2],
                                                       chunked[4 * lane + 3]);
        __syncthreads();
    }
}

// Rotation::apply of each centroid (src/live.rs:172-180).
__global__ __launch_bounds__(64) void k_rotate_rows(const float* __restrict__ in, const float* __restrict__ signs,
                                                    float rot_norm, float* __restrict__ out) {
    __shared__ float v[PDIM];
    const size_t c = blockIdx.x;
    const int lane = threadIdx.x;
    for (int i = lane; i < PDIM; i += 64) v[i] = in[c * PDIM + i];
    __syncthreads();
    rotate_lds<PDIM>(v, signs, rot_norm, lane);
    for (int i = lane; i < PDIM; i += 64) out[c * PDIM + i] = v[i];
}

__global__ void k_iota(uint32_t* v, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = (uint32_t)i;
}

// counts[c] = slots of cluster c in the log; nb[c] = its delta blocks (counts padded to 32).
__global__ void k_cluster_counts(const uint32_t* __restrict__ sorted_cl, size_t m, uint32_t k,
                                 uint32_t* __restrict__ counts) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) atomicAdd(&counts[sorted_cl[i]], 1u);
}
__global__ void k_blocks_of(const uint32_t* __restrict__ counts, uint32_t k, uint32_t* __restrict__ nb) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c <= k) nb[c] = c < k ? (counts[c] + BLOCK - 1) / BLOCK : 0u;
}

// Arena slots [0, nslots_arena): empty (PAD id, zero code and factors).
__global__ void k_arena_clear(uint4* __restrict__ codes, float* __restrict__ mul, float* __restrict__ err,
                              uint32_t* __restrict__ ids, size_t slots) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= slots) return;
    codes[2 * i] = make_uint4(0u, 0u, 0u, 0u);
    codes[2 * i + 1] = make_uint4(0u, 0u, 0u, 0u);
    mul[i] = 0.0f;
    err[i] = 0.0f;
    ids[i] = PAD;
}

// Sorted log entry i (cluster c, log index L) goes to arena slot dblock_off[c] * 32 + its rank
// within the cluster (i - cstart[c]).
__global__ void k_arena_scatter(const uint32_t* __restrict__ sorted_cl, const uint32_t* __restrict__ sorted_idx,
                                size_t m, const uint32_t* __restrict__ cstart, const uint32_t* __restrict__ dblock_off,
                                const uint4* __restrict__ lcode, const float* __restrict__ lmul,
                                const float* __restrict__ lerr, const uint32_t* __restrict__ lrow,
                                uint4* __restrict__ codes, float* __restrict__ mul, float* __restrict__ err,
                                uint32_t* __restrict__ ids) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint32_t c = sorted_cl[i], L = sorted_idx[i];
    const size_t dst = (size_t)dblock_off[c] * BLOCK + (i - cstart[c]);
    codes[2 * dst] = lcode[2 * (size_t)L];
    codes[2 * dst + 1] = lcode[2 * (size_t)L + 1];
    mul[dst] = lmul[L];
    err[dst] = lerr[L];
    ids[dst] = lrow[L];
}

__global__ void k_set_words(uint64_t* __restrict__ set, const uint64_t* __restrict__ idx,
                            const uint64_t* __restrict__ val, size_t m) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) set[idx[i]] = val[i];
}

inline unsigned grid(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

template <class T>
void dev_realloc(T*& p, size_t old_n, size_t new_n, hipStream_t st, size_t* acct) {
    T* q = dmalloc<T>(new_n, nullptr);
    if (acct) *acct += (new_n - old_n) * sizeof(T);
    if (p && old_n) SIEVE_HIP(hipMemcpyAsync(q, p, old_n * sizeof(T), hipMemcpyDeviceToDevice, st));
    SIEVE_HIP(hipStreamSynchronize(st));
    if (p) (void)hipFree(p);
    p = q;
}

// Rows per encoder pass: bounds the encoder's and the router's scratch. Rows are applied in
// order and each row's codes depend on the row alone, so passes give the state one pass over
// all rows would.
constexpr size_t APPLY_ROWS = 8192;

// Sets the tombstone of `id` on the host; records its word.
void tombstone(sieve_live* l, uint32_t id, std::vector<uint64_t>& changed) {
    const size_t w = id >> 6;
    l->deleted_host[w] |= 1ull << (id & 63);
    changed.push_back(w);
    l->ndeleted++;
}

// Rebuilds the arena from the slot log: a stable sort of the log by cluster, delta blocks per
// cluster, then each slot to its place.
// With exact, waits for the stream and reads the arena's block count back; without, leaves the
// count to a read into blocks_pinned and takes its upper bound meanwhile.
void rebuild_arena(sieve_live* l, hipStream_t st, bool exact = true) {
    sieve_index* ix = l->ix;
    const size_t m = l->nlog;
    const uint32_t k = ix->k;
    uint32_t* counts = grow<uint32_t>(l->counts, k + 1);
    uint32_t* nb = grow<uint32_t>(l->nb, k + 1);
    uint32_t* cstart = grow<uint32_t>(l->cstart, k + 1);
    SIEVE_HIP(hipMemsetAsync(counts, 0, (k + 1) * 4, st));
    uint32_t* keys_out = grow<uint32_t>(l->sort_keys, std::max<size_t>(m, 1));
    uint32_t* vals = grow<uint32_t>(l->sort_vals, 2 * std::max<size_t>(m, 1));
    uint32_t* vals_in = vals + std::max<size_t>(m, 1);
    if (m) {
        hipLaunchKernelGGL(k_iota, dim3(grid(m, 256)), dim3(256), 0, st, vals_in, m);
        int bits = 1;
        while (bits < 32 && (1u << bits) < k) bits++;
        size_t tmp_bytes = 0;
        SIEVE_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, l->lcl, keys_out, vals_in, vals, (int)m, 0,
                                                     bits, st));
        void* tmp = grow<uint8_t>(l->sort_tmp, tmp_bytes);
        SIEVE_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, l->lcl, keys_out, vals_in, vals, (int)m, 0, bits,
                                                     st));
        hipLaunchKernelGGL(k_cluster_counts, dim3(grid(m, 256)), dim3(256), 0, st, keys_out, m, k, counts);
    }
    hipLaunchKernelGGL(k_blocks_of, dim3(grid(k + 1, 256)), dim3(256), 0, st, counts, k, nb);
    size_t tmp_bytes = 0;
    SIEVE_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, counts, cstart, (int)(k + 1), st));
    size_t tmp2 = 0;
    SIEVE_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp2, nb, l->dblock_off, (int)(k + 1), st));
    void* tmp = grow<uint8_t>(l->sort_tmp, std::max(tmp_bytes, tmp2));
    SIEVE_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, counts, cstart, (int)(k + 1), st));
    SIEVE_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tmp2, nb, l->dblock_off, (int)(k + 1), st));
    // The arena's blocks, sum over clusters of ceil(count / 32), are at most m / 32 + k, and it
    // holds 2 max_rows + 32 k slots with m <= 2 max_rows: it cannot overflow, and the clear
    // covers the old arena and the bound of the new one without reading the count back first.
    const uint64_t bound = std::min<uint64_t>(l->cap_slots / BLOCK, (m + BLOCK - 1) / BLOCK + k);
    const size_t base = ix->nslots;
    const size_t clear = std::max<uint64_t>(bound, l->arena_blocks) * BLOCK;
    if (clear)
        hipLaunchKernelGGL(k_arena_clear, dim3(grid(clear, 256)), dim3(256), 0, st, ix->codes + 2 * base,
                           ix->mul + base, ix->err + base, ix->ids + base, clear);
    if (m)
        hipLaunchKernelGGL(k_arena_scatter, dim3(grid(m, 256)), dim3(256), 0, st, keys_out, vals, m, cstart,
                           l->dblock_off, l->lcode, l->lmul, l->lerr, l->lrow, ix->codes + 2 * base, ix->mul + base,
                           ix->err + base, ix->ids + base);
    SIEVE_HIP(hipGetLastError());
    if (!exact) {
        SIEVE_HIP(hipMemcpyAsync(l->blocks_pinned, l->dblock_off + k, 4, hipMemcpyDeviceToHost, st));
        l->arena_blocks = bound;
        return;
    }
    uint32_t blocks = 0;
    SIEVE_HIP(hipMemcpyAsync(&blocks, l->dblock_off + k, 4, hipMemcpyDeviceToHost, st));
    SIEVE_HIP(hipStreamSynchronize(st));
    l->arena_blocks = blocks;
}

// One encoder pass over delta rows r0 .. r0 + n (already in l->rows and l->norm256): prefix,
// clusters, slot codes into the log at l->nlog, tier-2 codes, external ids; then the upserts.
void apply_pass(sieve_live* l, const uint32_t* eh, size_t n, size_t r0, std::vector<uint64_t>& changed) {
    sieve_index* ix = l->ix;
    sieve_stream* s = &l->work;
    const hipStream_t st = s->st;
    float* xs = grow<float>(l->xs, n * PDIM);
    hipLaunchKernelGGL(k_live_prefix, dim3(grid(n, 64)), dim3(64), 0, st, l->rows + r0 * DIM, n, ix->quant, xs);
    SIEVE_HIP(hipGetLastError());
    // assign_top (src/ivf.rs:137-158) through the exact route: BF16 GEMM plus f32 rescoring of
    // every centroid near the cut gives the reference's f32 top SOAR_CAND, ties by id. Routing
    // reads the prefix only; the 768-d query the prep also takes is a zero stand-in.
    const int m = (int)std::min<uint32_t>(l->soar ? SOAR_CAND : 1, ix->k);
    sieve_params rp{};
    rp.cand = 1;
    rp.k = 1;
    rp.nprobe = (uint32_t)m;
    rp.margin = std::numeric_limits<float>::quiet_NaN();
    rp.budget = 0;
    const size_t zq_cap = l->zq.cap;
    float* zq = grow<float>(l->zq, n * DIM);
    if (l->zq.cap != zq_cap) SIEVE_HIP(hipMemsetAsync(zq, 0, l->zq.cap, st));
    Batch bt;
    prep_queries(ix, s, xs, zq, n, bt);
    float* ps = grow<float>(s->b[B_PS], n * m);
    uint32_t* pc = grow<uint32_t>(s->b[B_PC], n * m);
    uint32_t* pn = grow<uint32_t>(s->b[B_PN], n);
    route_batch(ix, s, bt, n, &rp, ps, pc, pn, m);
    uint32_t* pairs = grow<uint32_t>(l->pairs, 2 * n);
    hipLaunchKernelGGL(k_live_pairs, dim3((unsigned)n), dim3(64), 0, st, xs, n, ix->cents, pc, pn, m, l->soar ? 1 : 0,
                       l->lambda, pairs);
    const int spr = l->soar ? 2 : 1;
    const size_t L0 = l->nlog;
    hipLaunchKernelGGL(k_live_encode, dim3((unsigned)n), dim3(64), 0, st, xs, n, ix->rotation, 1.0f / 4096.0f,
                       l->cents_rot, pairs, spr, (uint32_t)(l->n + r0), l->lcode + 2 * L0, l->lmul + L0,
                       l->lerr + L0, l->lrow + L0, l->lcl + L0, ix->t2_sigma, l->t2 + r0 * 64, l->t2n + r0);
    SIEVE_HIP(hipGetLastError());
    SIEVE_HIP(hipMemcpyAsync(l->ext + r0, eh, n * 4, hipMemcpyHostToDevice, st));
    // The encoder's launches finish before anything is committed (the log and row counts advance
    // below).
    SIEVE_HIP(hipStreamSynchronize(st));
    // Upserts, in order (src/live.rs:361-373): a live external id's old row is tombstoned.
    for (size_t j = 0; j < n; j++) {
        uint32_t old;
        if (l->current(eh[j], &old)) tombstone(l, old, changed);
        l->by_ext[eh[j]] = (uint32_t)(l->n + r0 + j);
    }
    l->nrows += n;
    l->nlog += spr * n;
}

// Live::insert (src/live.rs:275-294): quantise on the device, log the records (op, external
// id, prefix norm, int8 row) and wait for them to be durable, then apply.
}  // namespace

namespace sieve {

void live_insert(sieve_live* l, const uint32_t* ext, const float* vecs, size_t n, bool publish) {
    sieve_index* ix = l->ix;
    SIEVE_REQUIRE(l->nrows + n <= l->cap_rows, SIEVE_ELIMIT,
                  "live capacity: " + std::to_string(l->cap_rows) + " delta rows");
    if (n == 0) return;
    sieve_stream* s = &l->work;
    const hipStream_t st = s->st;
    const float* dv = dev_in(s, B_Q768, vecs, n * DIM);
    const size_t r0 = l->nrows;
    hipLaunchKernelGGL(k_live_quant, dim3(grid(n, 64)), dim3(64), 0, st, dv, n, ix->quant, l->rows + r0 * DIM,
                       l->norm256 + r0);
    SIEVE_HIP(hipGetLastError());
    std::vector<uint32_t> he;
    const uint32_t* eh = ext;
    if (on_device(ext)) {
        he.resize(n);
        SIEVE_HIP(hipMemcpyAsync(he.data(), ext, n * 4, hipMemcpyDeviceToHost, st));
        eh = he.data();
    }
    if (l->wal_fd >= 0) {
        std::vector<int8_t> rows(n * DIM);
        std::vector<float> norms(n);
        SIEVE_HIP(hipMemcpyAsync(rows.data(), l->rows + r0 * DIM, n * DIM, hipMemcpyDeviceToHost, st));
        SIEVE_HIP(hipMemcpyAsync(norms.data(), l->norm256 + r0, n * 4, hipMemcpyDeviceToHost, st));
        SIEVE_HIP(hipStreamSynchronize(st));
        std::vector<uint8_t> rec(n * WAL_INSERT_BYTES);
        for (size_t j = 0; j < n; j++) wal_insert_record(eh[j], norms[j], rows.data() + j * DIM, &rec[j * WAL_INSERT_BYTES]);
        live_log(l, rec);
    } else {
        SIEVE_HIP(hipStreamSynchronize(st));
    }
    live_apply_rows(l, eh, n, publish);
}

void live_delete(sieve_live* l, const uint32_t* ext, size_t n, bool publish) {
    std::vector<uint32_t> he;
    const uint32_t* eh = ext;
    if (n && on_device(ext)) {
        he.resize(n);
        SIEVE_HIP(hipMemcpy(he.data(), ext, n * 4, hipMemcpyDeviceToHost));
        eh = he.data();
    }
    if (l->wal_fd >= 0 && n) {
        std::vector<uint8_t> rec(n * WAL_DELETE_BYTES);
        for (size_t j = 0; j < n; j++) wal_delete_record(eh[j], &rec[j * WAL_DELETE_BYTES]);
        live_log(l, rec);
    }
    live_apply_delete(l, eh, n, publish);
}


void launch_soar_pairs(const float* xs, size_t n, const float* cents, const uint32_t* pc, const uint32_t* pn,
                       int pstride, int soar, float lambda, uint32_t* pairs, hipStream_t st) {
    // 64 work-items per row, and a launch counts its work-items in 32 bits
    constexpr size_t ROWS = size_t(1) << 24;
    for (size_t r0 = 0; r0 < n; r0 += ROWS) {
        const size_t c = std::min(ROWS, n - r0);
        hipLaunchKernelGGL(k_live_pairs, dim3((unsigned)c), dim3(64), 0, st, xs + r0 * PDIM, c, cents,
                           pc + r0 * pstride, pn + r0, pstride, soar, lambda, pairs + 2 * r0);
    }
}

LiveArgs live_args(const sieve_live* l) { return l->args(); }
LiveArgs live_args_max(const sieve_live* l) {
    LiveArgs a = l->args();
    a.dblocks = (uint32_t)(l->cap_slots / BLOCK);  // every delta list padded to whole blocks fits
    return a;
}

void live_push_words(sieve_live* l, hipStream_t st, const std::vector<uint64_t>& changed) {
    if (changed.empty()) return;
    std::vector<uint64_t> vals(changed.size());
    for (size_t i = 0; i < changed.size(); i++) vals[i] = l->deleted_host[changed[i]];
    uint64_t* di = grow<uint64_t>(l->words_i, changed.size());
    uint64_t* dv = grow<uint64_t>(l->words_v, changed.size());
    SIEVE_HIP(hipMemcpyAsync(di, changed.data(), changed.size() * 8, hipMemcpyHostToDevice, st));
    SIEVE_HIP(hipMemcpyAsync(dv, vals.data(), vals.size() * 8, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_set_words, dim3(grid(changed.size(), 256)), dim3(256), 0, st, l->deleted, di, dv,
                       changed.size());
    SIEVE_HIP(hipGetLastError());
    SIEVE_HIP(hipStreamSynchronize(st));
}

void live_apply_rows(sieve_live* l, const uint32_t* eh, size_t n, bool publish) {
    SIEVE_REQUIRE(l->nrows + n <= l->cap_rows, SIEVE_ELIMIT,
                  "live capacity: " + std::to_string(l->cap_rows) + " delta rows");
    if (n == 0) return;
    const int spr = l->soar ? 2 : 1;
    SIEVE_REQUIRE(l->nlog + spr * n <= 2 * l->cap_rows, SIEVE_ELIMIT, "live slot log full");
    std::vector<uint64_t> changed;
    for (size_t j0 = 0; j0 < n; j0 += APPLY_ROWS) {
        const size_t m = std::min(APPLY_ROWS, n - j0);
        apply_pass(l, eh + j0, m, l->nrows, changed);
    }
    l->pend.insert(l->pend.end(), changed.begin(), changed.end());
    l->arena_dirty = true;
    if (publish) live_publish(l);
}

void live_settle(const sieve_live* l) {
    std::lock_guard<std::mutex> lk(l->pub_mu);
    if (!l->pub_pending) return;
    SIEVE_HIP(hipEventSynchronize(l->pub_ev));
    l->pub_pending = false;
    if (l->blocks_from_pub) {
        l->arena_blocks = *l->blocks_pinned;
        l->blocks_from_pub = false;
    }
}

void live_publish(sieve_live* l) {
    live_settle(l);
    live_push_words(l, l->work.st, l->pend);
    l->pend.clear();
    if (l->arena_dirty) {
        rebuild_arena(l, l->work.st);
        l->arena_dirty = false;
    }
}

hipEvent_t live_publish_async(sieve_live* l, const hipEvent_t* waits, size_t nwaits) {
    live_settle(l);  // one publish pending at most: its staging and scratch are free again
    std::lock_guard<std::mutex> lk(l->pub_mu);
    if (!l->pub) {
        SIEVE_HIP(hipStreamCreate(&l->pub));
        SIEVE_HIP(hipEventCreateWithFlags(&l->pub_ev, hipEventDisableTiming));
        SIEVE_HIP(hipHostMalloc(reinterpret_cast<void**>(&l->blocks_pinned), 16, hipHostMallocDefault));
    }
    const hipStream_t st = l->pub;
    const size_t nc = l->pend.size();
    uint64_t* di = nullptr;
    uint64_t* dv = nullptr;
    if (nc) {
        // Staged before the waits: the copies need not wait for the searches.
        std::vector<uint64_t> vals(nc);
        for (size_t i = 0; i < nc; i++) vals[i] = l->deleted_host[l->pend[i]];
        di = grow<uint64_t>(l->words_i, nc);
        dv = grow<uint64_t>(l->words_v, nc);
        SIEVE_HIP(hipMemcpyAsync(di, l->pend.data(), nc * 8, hipMemcpyHostToDevice, st));
        SIEVE_HIP(hipMemcpyAsync(dv, vals.data(), nc * 8, hipMemcpyHostToDevice, st));
        SIEVE_HIP(hipStreamSynchronize(st));
    }
    for (size_t i = 0; i < nwaits; i++) SIEVE_HIP(hipStreamWaitEvent(st, waits[i], 0));
    if (nc) {
        hipLaunchKernelGGL(k_set_words, dim3(grid(nc, 256)), dim3(256), 0, st, l->deleted, di, dv, nc);
        SIEVE_HIP(hipGetLastError());
    }
    l->pend.clear();
    if (l->arena_dirty) {
        rebuild_arena(l, st, false);
        l->arena_dirty = false;
        l->blocks_from_pub = true;
    }
    SIEVE_HIP(hipEventRecord(l->pub_ev, st));
    l->pub_pending = true;
    return l->pub_ev;
}

void live_apply_delete(sieve_live* l, const uint32_t* eh, size_t n, bool publish) {
    std::vector<uint64_t> changed;
    for (size_t j = 0; j < n; j++) {  // Live::apply_delete, src/live.rs:308-315
        uint32_t id;
        if (l->current(eh[j], &id)) {
            if (!l->bit_host(id)) tombstone(l, id, changed);
            l->by_ext.erase(eh[j]);
        }
    }
    l->pend.insert(l->pend.end(), changed.begin(), changed.end());
    if (publish) live_publish(l);
}

sieve_live* live_create(sieve_index* ix, const sieve_live_opts& opts) {
    SIEVE_REQUIRE(ix->rotation && ix->mul, SIEVE_EINVAL, "live updates need a RaBitQ index");
    SIEVE_REQUIRE(!ix->live, SIEVE_EINVAL, "the index already has a live state");
    SIEVE_REQUIRE(opts.max_rows >= 1, SIEVE_EINVAL, "max_rows must be >= 1");
    SIEVE_REQUIRE(ix->n + opts.max_rows < 0xFFFFFFFFull, SIEVE_ELIMIT, "row ids exceed 32 bits");
    std::unique_ptr<sieve_live> l(new sieve_live());
    SIEVE_HIP(hipSetDevice(ix->device));
    l->ix = ix;
    l->soar = !(opts.soar_lambda != opts.soar_lambda);
    l->lambda = l->soar ? opts.soar_lambda : 0.0f;
    l->n = ix->n;
    l->cap_rows = opts.max_rows;
    l->cap_slots = 2 * opts.max_rows + (uint64_t)BLOCK * ix->k;
    SIEVE_REQUIRE(ix->nslots + l->cap_slots < 0xFFFFFFFFull, SIEVE_ELIMIT, "slot ids exceed 32 bits");
    SIEVE_HIP(hipStreamCreate(&l->work.st));
    l->work.device = ix->device;
    const hipStream_t st = l->work.st;
    size_t* acct = &l->device_bytes;
    // Room for the arena after the base slots.
    const size_t ns = ix->nslots, total = ns + l->cap_slots;
    dev_realloc(ix->codes, 2 * ns, 2 * total, st, &ix->device_bytes);
    dev_realloc(ix->mul, ns, total, st, &ix->device_bytes);
    dev_realloc(ix->err, ns, total, st, &ix->device_bytes);
    dev_realloc(ix->ids, ns, total, st, &ix->device_bytes);
    l->cents_rot = dmalloc<float>((size_t)ix->k * PDIM, acct);
    hipLaunchKernelGGL(k_rotate_rows, dim3(ix->k), dim3(64), 0, st, ix->cents, ix->rotation, 1.0f / 4096.0f,
                       l->cents_rot);
    const size_t R = l->cap_rows;
    l->rows = dmalloc<int8_t>(R * DIM, acct);
    l->norm256 = dmalloc<float>(R, acct);
    l->t2 = dmalloc<uint8_t>(R * 64, acct);
    l->t2n = dmalloc<float>(R, acct);
    l->ext = dmalloc<uint32_t>(R, acct);
    l->deleted_words = (l->n + R + 63) / 64;
    l->deleted = dmalloc<uint64_t>(l->deleted_words, acct);
    SIEVE_HIP(hipMemsetAsync(l->deleted, 0, l->deleted_words * 8, st));
    l->deleted_host.assign(l->deleted_words, 0);
    l->lcode = dmalloc<uint4>(2 * 2 * R, acct);
    l->lmul = dmalloc<float>(2 * R, acct);
    l->lerr = dmalloc<float>(2 * R, acct);
    l->lrow = dmalloc<uint32_t>(2 * R, acct);
    l->lcl = dmalloc<uint32_t>(2 * R, acct);
    l->dblock_off = dmalloc<uint32_t>(ix->k + 1, acct);
    SIEVE_HIP(hipGetLastError());
    rebuild_arena(l.get(), st);
    ix->live = l.get();
    return l.release();
}

}  // namespace sieve

extern "C" {

int sieve_live_open(sieve_index* ix, const sieve_live_opts* o, sieve_live** out) {
    return guarded([&] {
        SIEVE_REQUIRE(ix && out, SIEVE_EINVAL, "null argument");
        SIEVE_REQUIRE(ix->seg_rows == 0, SIEVE_EINVAL,
                      "the index's lists hold rows of its live segment (live.*): open it with sieve_live_open_dir");
        *out = live_create(ix, o ? *o : sieve_live_opts{});
    });
}

void sieve_live_close(sieve_live* l) {
    if (!l) return;
    if (l->ix) l->ix->live = nullptr;
    delete l;
}

int sieve_live_insert(sieve_live* l, const uint32_t* ext, const float* vecs, size_t n) {
    return guarded([&] {
        SIEVE_REQUIRE(l && (n == 0 || (ext && vecs)), SIEVE_EINVAL, "null argument");
        SIEVE_HIP(hipSetDevice(l->ix->device));
        live_insert(l, ext, vecs, n, true);
    });
}

// Live::delete (src/live.rs:296-306): logs the records, then applies them.
int sieve_live_delete(sieve_live* l, const uint32_t* ext, size_t n) {
    return guarded([&] {
        SIEVE_REQUIRE(l && (n == 0 || ext), SIEVE_EINVAL, "null argument");
        SIEVE_HIP(hipSetDevice(l->ix->device));
        live_delete(l, ext, n, true);
    });
}

int sieve_live_get_info(const sieve_live* l, sieve_live_info* out) {
    return guarded([&] {
        SIEVE_REQUIRE(l && out, SIEVE_EINVAL, "null argument");
        SIEVE_HIP(hipSetDevice(l->ix->device));
        live_settle(l);
        out->rows = l->nrows;
        out->slots = l->nlog;
        out->arena_blocks = l->arena_blocks;
        out->deleted = l->ndeleted;
        out->max_rows = l->cap_rows;
        out->device_bytes = l->device_bytes;
    });
}

uint32_t sieve_live_current(const sieve_live* l, uint32_t ext) {
    uint32_t id;
    return l && l->current(ext, &id) ? id : 0xFFFFFFFFu;
}

int sieve_live_export(const sieve_live* l, int8_t* rows, float* norm256, uint8_t* t2, float* t2n, uint32_t* ext,
                      uint32_t* dblock_off, uint8_t* codes, float* mul, float* err, uint32_t* ids) {
    return guarded([&] {
        SIEVE_REQUIRE(l, SIEVE_EINVAL, "null argument");
        SIEVE_HIP(hipSetDevice(l->ix->device));
        live_settle(l);
        const size_t r = l->nrows, base = l->ix->nslots, slots = l->arena_blocks * BLOCK;
        auto get = [&](void* dst, const void* src, size_t bytes) {
            if (dst && bytes) SIEVE_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
        };
        get(rows, l->rows, r * DIM);
        get(norm256, l->norm256, r * 4);
        get(t2, l->t2, r * 64);
        get(t2n, l->t2n, r * 4);
        get(ext, l->ext, r * 4);
        get(dblock_off, l->dblock_off, (l->ix->k + 1) * 4);
        get(codes, l->ix->codes + 2 * base, slots * 32);
        if (codes)  // the device keeps them lane-chunked; the export gives raw codes
            for (size_t i = 0; i < slots; i++) {
                uint32_t ch[8], raw[8];
                std::memcpy(ch, codes + 32 * i, 32);
                code_from_chunked(ch, raw);
                std::memcpy(codes + 32 * i, raw, 32);
            }
        get(mul, l->ix->mul + base, slots * 4);
        get(err, l->ix->err + base, slots * 4);
        get(ids, l->ix->ids + base, slots * 4);
    });
}

int sieve_live_search_batch(const sieve_live* l, sieve_stream* s, const float* q768, const float* q256, size_t nq,
                            const sieve_params* p, const sieve_view* v, float* scores, uint32_t* ext_ids,
                            uint32_t* counts, sieve_timing* t) {
    return guarded([&] {
        SIEVE_REQUIRE(l && s && q768 && q256 && p && scores && ext_ids && counts, SIEVE_EINVAL, "null argument");
        check_params(p, true);
        if (nq == 0) return;
        const sieve_index* ix = l->ix;
        SIEVE_HIP(hipSetDevice(ix->device));
        live_settle(l);
        ensure_events(s);
        const float* q2 = dev_in(s, B_Q256, q256, nq * PDIM);
        const float* q7 = dev_in(s, B_Q768, q768, nq * DIM);
        ViewArgs view = make_view(s, v, nq);
        SIEVE_REQUIRE(!view.deleted, SIEVE_EINVAL, "a live search uses the live tombstones; pass allowlists only");
        auto os = dev_out(s, B_OS, scores, nq * p->k);
        auto orr = dev_out(s, B_OR, ext_ids, nq * p->k);
        auto on = dev_out(s, B_ON, counts, nq);
        const LiveArgs la = l->args();
        enqueue_search(ix, s, q2, q7, nq, p, view, os.dev, orr.dev, on.dev, &la);
        out_done(s, os);
        out_done(s, orr);
        out_done(s, on);
        finish_search(s, nq, t);
    });
}

}  // extern "C"
