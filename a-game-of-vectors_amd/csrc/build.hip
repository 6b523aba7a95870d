// The index build's GEMM-shaped steps, bit-exact with the reference: kmeans (src/ivf.rs:188-225,
// balance off), assign_top + soar_secondary (src/ivf.rs:137-186) and Tier2::fit
// (src/tier2.rs:20-27). The element-wise steps of the build (prefixes, rotation, RaBitQ codes
// and factors, tier-2 codes, block packing) stay with the caller.
//
// Scores. The reference computes every row-centroid score as one sequential f32 dot (the x86
// gemm_abt, src/blas.rs:43-54). Here they run on the routing GEMM (BF16 MFMA, f32 accumulation),
// and the route's select rescores every centroid within the error band of the m-th best
// exactly as gemm_abt sums, so the best m are the reference's, ties by id: the first maximum
// for m = 1 (assign's fold keeps the first, src/ivf.rs:146-148) and TopK::into_sorted's order
// for m = 16.
//
// k-means step. The sample is sorted by cluster with a stable radix sort, so each cluster's
// members keep sample order; one workgroup per cluster sums them dimension by dimension in
// that order from +0.0 (src/ivf.rs:201-205) and normalises the sum (quant::normalize: the
// sequential sum of squares, its correctly rounded square root, 1 / n, one multiply per
// element). An empty cluster takes a sample row the host draws with the reference's Rng, in
// cluster order, as do the initial centroids.
//
// Traffic: per k-means iteration the routing GEMM over n x k x 256 (BF16) writes n x k f32
// scores; the sums read the sample once. Per assigned row: its 1 KiB prefix, 16 candidates'
// exact rescoring (1 KiB each from the cache-resident centroids).
#include <float.h>

#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "host.hpp"

using namespace sieve;

namespace {

constexpr int SOAR_CAND = 16;  // src/ivf.rs:296

// splitmix64 (src/ivf.rs:120-135).
struct Rng {
    uint64_t s;
    uint64_t next() {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    uint64_t below(uint64_t n) { return next() % n; }
};

inline unsigned blocks_of(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

// A device buffer owned by one call (not stream scratch: captured graphs stay valid).
struct Tmp {
    void* p = nullptr;
    size_t cap = 0;
    ~Tmp() {
        if (p) (void)hipFree(p);
    }
    template <class T>
    T* get(size_t n) {
        const size_t bytes = std::max<size_t>(n * sizeof(T), 16);
        if (bytes > cap) {
            if (p) (void)hipFree(p);
            p = nullptr;
            cap = 0;
            p = dmalloc<uint8_t>(bytes, nullptr);
            cap = bytes;
        }
        return static_cast<T*>(p);
    }
};

// |x| of each 256-d row (the route's error bound scales with it). One wave per row.
__global__ __launch_bounds__(64) void k_row_norms(const float* __restrict__ x, size_t n, float* __restrict__ out) {
    const size_t r = blockIdx.x;
    const int lane = threadIdx.x;
    float nn = 0.0f;
    for (int i = lane; i < PDIM; i += 64) {
        const float v = x[r * PDIM + i];
        nn += v * v;
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) nn += __shfl_xor(nn, d, 64);
    if (lane == 0) out[r] = sqrtf(nn);
}

// The largest |c| over the centroids, as the bits of a positive float (atomicMax orders them).
__global__ __launch_bounds__(64) void k_norm_max(const float* __restrict__ c, uint32_t k, uint32_t* __restrict__ out) {
    const size_t r = blockIdx.x;
    const int lane = threadIdx.x;
    float nn = 0.0f;
    for (int i = lane; i < PDIM; i += 64) {
        const float v = c[r * PDIM + i];
        nn += v * v;
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) nn += __shfl_xor(nn, d, 64);
    if (lane == 0) atomicMax(out, __float_as_uint(sqrtf(nn)));
}

__global__ void k_gather_rows(const float* __restrict__ src, const uint32_t* __restrict__ idx, size_t n,
                              float* __restrict__ dst) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * PDIM) return;
    dst[i] = src[(size_t)idx[i / PDIM] * PDIM + (i % PDIM)];
}

__global__ void k_iota_u32(uint32_t* __restrict__ v, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = (uint32_t)i;
}

__global__ void k_zero_u32(uint32_t* __restrict__ v, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = 0u;
}

// The best candidate of each row's list (m = 1: the route's first entry).
__global__ void k_first(const uint32_t* __restrict__ pc, size_t n, int stride, uint32_t* __restrict__ a) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = pc[i * stride];
}

__global__ void k_count(const uint32_t* __restrict__ a, size_t n, uint32_t* __restrict__ counts) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicAdd(&counts[a[i]], 1u);
}

__device__ __forceinline__ float sqrt_rn(float x) { return (float)sqrt((double)x); }

// One workgroup per cluster, one thread per dimension: the members' sum in sample order, then
// quant::normalize (src/quant.rs:5-13). Empty clusters are left to the host.
__global__ __launch_bounds__(PDIM) void k_cluster_mean(const float* __restrict__ sample,
                                                       const uint32_t* __restrict__ order,
                                                       const uint32_t* __restrict__ off,
                                                       const uint32_t* __restrict__ counts,
                                                       float* __restrict__ cents) {
    __shared__ float v[PDIM];
    __shared__ float inv_s;
    __shared__ int scale_s;
    const uint32_t c = blockIdx.x;
    const int d = threadIdx.x;
    const uint32_t cnt = counts[c];
    if (cnt == 0) return;
    const uint32_t m0 = off[c];
    float s = 0.0f;
    for (uint32_t i = 0; i < cnt; i++) s = s + sample[(size_t)order[m0 + i] * PDIM + d];
    v[d] = s;
    __syncthreads();
    if (d == 0) {
        float ss = -0.0f;
        for (int i = 0; i < PDIM; i++) ss = ss + v[i] * v[i];
        const float nn = sqrt_rn(ss);
        scale_s = nn > 0.0f;
        inv_s = nn > 0.0f ? 1.0f / nn : 0.0f;
    }
    __syncthreads();
    cents[(size_t)c * PDIM + d] = scale_s ? s * inv_s : s;
}

// Tier2::fit: sequential f64 sums of x * x per dimension, in row order. One workgroup.
__global__ __launch_bounds__(PDIM) void k_tier2_fit(const float* __restrict__ rows, size_t n,
                                                    float* __restrict__ sigma) {
    const int d = threadIdx.x;
    double ss = 0.0;
    for (size_t r = 0; r < n; r++) {
        const float x = rows[r * PDIM + d];
        ss += (double)(x * x);
    }
    const float nf = (float)(n > 0 ? n : 1);
    sigma[d] = fmaxf((float)sqrt(ss / (double)nf), 1e-6f);
}

// Centroids on the device with what the route needs.
struct Cents {
    const float* c = nullptr;
    uint16_t* bf = nullptr;
    uint32_t k = 0;
    float norm_max = 0.0f;
};

void cents_prepare(Cents& C, Tmp& bf, Tmp& nm, hipStream_t st) {
    C.bf = bf.get<uint16_t>((size_t)C.k * PDIM);
    launch_to_bf16(C.c, (size_t)C.k * PDIM, C.bf, st);
    uint32_t* d = nm.get<uint32_t>(1);
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, d, (size_t)1);
    hipLaunchKernelGGL(k_norm_max, dim3(C.k), dim3(64), 0, st, C.c, C.k, d);
    uint32_t bits = 0;
    SIEVE_HIP(hipMemcpyAsync(&bits, d, 4, hipMemcpyDeviceToHost, st));
    SIEVE_HIP(hipStreamSynchronize(st));
    float mx;
    std::memcpy(&mx, &bits, 4);
    C.norm_max = mx * (1.0f + 1e-6f);
}

// Scratch of exact_top.
struct TopScratch {
    Tmp approx, tmax, xbf, xn, ps;
};

// The best m (<= 16) centroids of each of n device rows by the reference's f32 inner product,
// best first, ties by id, into pc (n x m) and their number into pn (min(m, k)).
void exact_top(const Cents& C, const float* rows, size_t n, int m, uint32_t* pc, uint32_t* pn, TopScratch& ts,
               hipStream_t st) {
    const size_t per = std::max<size_t>(1, std::min<size_t>(n, (size_t(1) << 30) / (4 * (size_t)C.k)));
    float* approx = ts.approx.get<float>(per * C.k);
    float* tmax = ts.tmax.get<float>(per * ((C.k + 15) / 16));
    uint16_t* xbf = ts.xbf.get<uint16_t>(per * PDIM);
    float* xn = ts.xn.get<float>(per);
    float* ps = ts.ps.get<float>(per * m);
    for (size_t q0 = 0; q0 < n; q0 += per) {
        const size_t mm = std::min(per, n - q0);
        const float* x = rows + q0 * PDIM;
        launch_to_bf16(x, mm * PDIM, xbf, st);
        hipLaunchKernelGGL(k_row_norms, dim3((unsigned)mm), dim3(64), 0, st, x, mm, xn);
        const bool fused = route_fuse(-1, (int)C.k, m);
        launch_gemm_bf16_route(xbf, C.bf, nullptr, fused ? nullptr : approx, tmax, (int)mm, (int)C.k, C.k, st);
        RouteArgs ra{};
        ra.fused = fused;
        ra.qbf = xbf;
        ra.cbf = C.bf;
        ra.approx = approx;
        ra.lds = C.k;
        ra.tmax = tmax;
        ra.k = (int)C.k;
        ra.cents = C.c;
        ra.q256 = x;
        ra.qnorm = xn;
        ra.cl_slots = nullptr;  // read only under a budget
        ra.cent_norm_max = C.norm_max;
        ra.nprobe = m;
        ra.margin = std::numeric_limits<float>::quiet_NaN();
        ra.budget = 0;
        ra.out_s = ps;
        ra.out_c = pc + q0 * m;
        ra.out_n = pn + q0;
        ra.out_stride = m;
        ra.l2 = 0;
        ra.dims = PDIM;
        ra.cstride = PDIM;
        const float u = 1.0f / 256.0f;  // BF16 unit roundoff
        ra.err_coef = (2.0f * u + u * u + 256.0f / 16777216.0f) * C.norm_max;
        launch_route_select(ra, mm, st);
    }
    SIEVE_HIP(hipGetLastError());
}

// A device copy of a caller array (or the array itself).
template <class T>
const T* to_device(const T* p, size_t n, Tmp& t, hipStream_t st) {
    if (on_device(p)) return p;
    T* d = t.get<T>(n);
    SIEVE_HIP(hipMemcpyAsync(d, p, n * sizeof(T), hipMemcpyHostToDevice, st));
    return d;
}

struct Stream {
    hipStream_t st = nullptr;
    Stream() { SIEVE_HIP(hipStreamCreate(&st)); }
    ~Stream() {
        if (st) (void)hipStreamDestroy(st);
    }
};

// ------------------------------------------------------------------ the 1B build -------
// bigann's train (src/bigann.rs:175-227) and Model::assign (:277-302) over 100-d int8 rows.
// Every score is top_l2's 2 <x, c> - |c|^2 with <x, c> the x86 gemm_abt's sequential f32 dot
// from -0.0 and |c|^2 norms' sequential sum: each thread computes them for its own row exactly
// so, against centroids staged in LDS. The rows are int8, so their f32 values are exact.

constexpr int L2D = 100;
constexpr int L2_TOP_COARSE = 4;  // src/bigann.rs:60
constexpr int L2_CT = 64;         // centroids per LDS tile

// norms (src/bigann.rs:135-137).
__global__ void k_l2_norms(const float* __restrict__ c, size_t k, float* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= k) return;
    float s = -0.0f;
    for (int d = 0; d < L2D; d++) s = s + c[i * L2D + d] * c[i * L2D + d];
    out[i] = s;
}

// Byte b of a packed int8 word as f32 (exact).
__device__ __forceinline__ float i8f(uint32_t w, int b) { return (float)(int8_t)(w >> (8 * b)); }

// Streaming top-M in the order of TopK::into_sorted: score desc, then id asc. Ids arrive in
// increasing order, so a later equal score never displaces an earlier one.
template <int M>
__device__ __forceinline__ void topm_push(float (&bs)[M], uint32_t (&bi)[M], float s, uint32_t id) {
    if (!(s > bs[M - 1])) return;
    int p = M - 1;
#pragma unroll
    for (int j = M - 1; j > 0; j--) {
        if (s > bs[j - 1]) {
            bs[j] = bs[j - 1];
            bi[j] = bi[j - 1];
            p = j - 1;
        }
    }
    bs[p] = s;
    bi[p] = id;
}

// The work of one workgroup of k_l2_top: rows [r0, r0 + nr) of a row list (or of the base),
// scored against centroids [c0, c0 + kc); ids reported relative to c0 plus id_base.
struct L2TopItem {
    uint32_t r0, nr, c0, kc;
};

// top_l2 (src/bigann.rs:117-133) for the rows of each work item: one thread per row, its 100
// values in registers, the item's centroids through LDS 64 at a time, four dots at once. The
// best M (ids relative to the item's first centroid) and their scores go to out[row slot].
template <int M>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_l2_top(const int8_t* __restrict__ base, const uint32_t* __restrict__ rows,
                                               const L2TopItem* __restrict__ items, const float* __restrict__ cents,
                                               const float* __restrict__ cnorm, uint32_t* __restrict__ out_id,
                                               float* __restrict__ out_s) {
    __shared__ __align__(16) float cl[L2_CT * L2D];
    __shared__ float cn[L2_CT];
    const L2TopItem it = items[blockIdx.x];
    const uint32_t t = threadIdx.x;
    const bool ok = t < it.nr;
    const size_t slot = (size_t)it.r0 + t;
    // The row packed, 4 int8 values a word (rows are 4-byte aligned: 100 = 4 x 25), converted
    // 4 dimensions at a time for all four dots.
    uint32_t xw[L2D / 4];
    if (ok) {
        const size_t r = rows ? rows[slot] : slot;
        const uint32_t* w = reinterpret_cast<const uint32_t*>(base + r * L2D);
#pragma unroll
        for (int q = 0; q < L2D / 4; q++) xw[q] = w[q];
    } else {
#pragma unroll
        for (int q = 0; q < L2D / 4; q++) xw[q] = 0u;
    }
    float bs[M];
    uint32_t bi[M];
#pragma unroll
    for (int j = 0; j < M; j++) {
        bs[j] = -INFINITY;
        bi[j] = 0u;
    }
    const float* cb = cents + (size_t)it.c0 * L2D;
    for (uint32_t t0 = 0; t0 < it.kc; t0 += L2_CT) {
        const uint32_t nt = min((uint32_t)L2_CT, it.kc - t0);
        __syncthreads();
        for (uint32_t e = t; e < (uint32_t)L2_CT * L2D; e += 256) cl[e] = e < nt * L2D ? cb[(size_t)t0 * L2D + e] : 0.0f;
        if (t < (uint32_t)L2_CT) cn[t] = t < nt ? cnorm[it.c0 + t0 + t] : 0.0f;
        __syncthreads();
        if (!ok) continue;
#pragma unroll 1
        for (uint32_t j = 0; j < nt; j += 4) {
            const float* c0 = cl + j * L2D;
            float a0 = -0.0f, a1 = -0.0f, a2 = -0.0f, a3 = -0.0f;
#pragma unroll
            for (int q = 0; q < L2D / 4; q++) {
                const float4 v0 = *reinterpret_cast<const float4*>(c0 + 4 * q);
                const float4 v1 = *reinterpret_cast<const float4*>(c0 + L2D + 4 * q);
                const float4 v2 = *reinterpret_cast<const float4*>(c0 + 2 * L2D + 4 * q);
                const float4 v3 = *reinterpret_cast<const float4*>(c0 + 3 * L2D + 4 * q);
                // Converted here, not hoisted out of the centroid loop: 100 live floats would spill.
                uint32_t w = xw[q];
                asm volatile("" : "+v"(w));
                const float x0 = i8f(w, 0), x1 = i8f(w, 1), x2 = i8f(w, 2), x3 = i8f(w, 3);
                a0 = a0 + x0 * v0.x; a0 = a0 + x1 * v0.y; a0 = a0 + x2 * v0.z; a0 = a0 + x3 * v0.w;
                a1 = a1 + x0 * v1.x; a1 = a1 + x1 * v1.y; a1 = a1 + x2 * v1.z; a1 = a1 + x3 * v1.w;
                a2 = a2 + x0 * v2.x; a2 = a2 + x1 * v2.y; a2 = a2 + x2 * v2.z; a2 = a2 + x3 * v2.w;
                a3 = a3 + x0 * v3.x; a3 = a3 + x1 * v3.y; a3 = a3 + x2 * v3.z; a3 = a3 + x3 * v3.w;
                // The group's products are added here, and its LDS reads are not scheduled ahead of
                // the previous group's: otherwise all 100 reads of a step would be live at once.
                asm volatile("" : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3));
                __builtin_amdgcn_sched_barrier(0);
            }
            const float a[4] = {a0, a1, a2, a3};
#pragma unroll
            for (int u = 0; u < 4; u++)
                if (j + u < nt) topm_push<M>(bs, bi, 2.0f * a[u] - cn[j + u], t0 + j + u);
        }
    }
    if (ok) {
#pragma unroll
        for (int j = 0; j < M; j++) {
            out_id[slot * M + j] = bi[j];
            if (out_s) out_s[slot * M + j] = bs[j];
        }
    }
}

// One workgroup per cluster, one thread per dimension: kmeans_l2's f64 sum of the members in
// sample order, then (s / m) as f32 (src/bigann.rs:148-158). Empty clusters are left to the host.
__global__ __launch_bounds__(128) void k_l2_cluster_mean(const int8_t* __restrict__ base,
                                                         const uint32_t* __restrict__ rows,
                                                         const uint32_t* __restrict__ order,
                                                         const uint32_t* __restrict__ off,
                                                         const uint32_t* __restrict__ counts,
                                                         float* __restrict__ cents) {
    const uint32_t c = blockIdx.x;
    const int d = threadIdx.x;
    const uint32_t cnt = counts[c];
    if (cnt == 0 || d >= L2D) return;
    const uint32_t m0 = off[c];
    double s = 0.0;
    for (uint32_t i = 0; i < cnt; i++) s += (double)(float)base[(size_t)rows[order[m0 + i]] * L2D + d];
    cents[(size_t)c * L2D + d] = (float)(s / (double)cnt);
}

// cents[c] = the f32 values of row r, for each (c, r) pair.
__global__ void k_l2_set_rows(const int8_t* __restrict__ base, const uint2* __restrict__ cr, size_t n,
                              float* __restrict__ cents) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * L2D) return;
    const uint2 p = cr[i / L2D];
    cents[(size_t)p.x * L2D + (i % L2D)] = (float)base[(size_t)p.y * L2D + (i % L2D)];
}

// Model::assign's pick: of a row's TOP_COARSE cells taken in ascending order, a later cell's
// best leaf wins only if strictly better (src/bigann.rs:294-299).
__global__ void k_l2_pick(const uint32_t* __restrict__ cells, const uint32_t* __restrict__ leaf,
                          const float* __restrict__ score, size_t n, uint32_t k2, uint32_t* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t c[L2_TOP_COARSE], lf[L2_TOP_COARSE];
    float sc[L2_TOP_COARSE];
#pragma unroll
    for (int t = 0; t < L2_TOP_COARSE; t++) {
        c[t] = cells[i * L2_TOP_COARSE + t];
        lf[t] = leaf[i * L2_TOP_COARSE + t];
        sc[t] = score[i * L2_TOP_COARSE + t];
    }
    uint32_t best = 0;
    float bs = -FLT_MAX;  // the reference's (0, f32::MIN) start
    // The (distinct) cells in ascending order.
    int ord[L2_TOP_COARSE] = {0, 1, 2, 3};
    for (int a = 0; a < L2_TOP_COARSE; a++)
        for (int b = a + 1; b < L2_TOP_COARSE; b++)
            if (c[ord[b]] < c[ord[a]]) {
                const int tmp = ord[a];
                ord[a] = ord[b];
                ord[b] = tmp;
            }
    for (int r = 0; r < L2_TOP_COARSE; r++) {
        const int t = ord[r];
        if (sc[t] > bs) {
            bs = sc[t];
            best = c[t] * k2 + lf[t];
        }
    }
    out[i] = best;
}

// The base row of each sorted (row, cell) pair: pair p is local row p / 4 of the chunk at r0.
__global__ void k_pair_rows(const uint32_t* __restrict__ sv, size_t n, uint32_t r0, uint32_t* __restrict__ rows) {
    const size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) rows[s] = r0 + sv[s] / L2_TOP_COARSE;
}

// Each sorted pair's best leaf and score back to the pair's own position.
__global__ void k_scatter_pairs(const uint32_t* __restrict__ sv, const uint32_t* __restrict__ lid,
                                const float* __restrict__ ls, size_t n, uint32_t* __restrict__ leaf,
                                float* __restrict__ score) {
    const size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    leaf[sv[s]] = lid[s];
    score[sv[s]] = ls[s];
}

// Work items: rows [0, n) of a list in tiles of 256 against centroids [c0, c0 + kc).
void l2_items(std::vector<L2TopItem>& v, uint32_t r0, uint32_t n, uint32_t c0, uint32_t kc) {
    for (uint32_t a = 0; a < n; a += 256) v.push_back({r0 + a, std::min<uint32_t>(256, n - a), c0, kc});
}

template <int M>
void l2_top(const int8_t* base, const uint32_t* rows, const std::vector<L2TopItem>& items, const float* cents,
            const float* cnorm, uint32_t* out_id, float* out_s, Tmp& t_items, hipStream_t st) {
    if (items.empty()) return;
    L2TopItem* di = t_items.get<L2TopItem>(items.size());
    SIEVE_HIP(hipMemcpyAsync(di, items.data(), items.size() * sizeof(L2TopItem), hipMemcpyHostToDevice, st));
    for (size_t i0 = 0; i0 < items.size(); i0 += (1u << 30)) {
        const size_t m = std::min<size_t>(items.size() - i0, 1u << 30);
        hipLaunchKernelGGL(k_l2_top<M>, dim3((unsigned)m), dim3(256), 0, st, base, rows, di + i0, cents, cnorm, out_id,
                           out_s);
    }
    SIEVE_HIP(hipGetLastError());
    SIEVE_HIP(hipStreamSynchronize(st));  // the item list is reused by the next call
}

// kmeans_l2 (src/bigann.rs:139-161) over the rows base[rows[i]], i < n (a device list), into
// cents (device, k x 100). rng: the reference's Rng(seed).
void kmeans_l2_dev(const int8_t* base, const uint32_t* rows, uint32_t n, uint32_t k, uint32_t iters, uint64_t seed,
                   float* cents, hipStream_t st) {
    Tmp t_pairs, t_cn, t_a, t_keys, t_order, t_iota, t_counts, t_off, t_cub, t_items;
    Rng rng{seed};
    std::vector<uint32_t> hrows(n);
    SIEVE_HIP(hipMemcpyAsync(hrows.data(), rows, 4 * (size_t)n, hipMemcpyDeviceToHost, st));
    SIEVE_HIP(hipStreamSynchronize(st));
    std::vector<uint2> set(k);
    for (uint32_t c = 0; c < k; c++) set[c] = make_uint2(c, hrows[rng.below(n)]);
    uint2* dset = t_pairs.get<uint2>(k);
    auto set_rows = [&](size_t m) {
        if (m == 0) return;
        SIEVE_HIP(hipMemcpyAsync(dset, set.data(), m * sizeof(uint2), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_l2_set_rows, dim3(blocks_of(m * L2D, 256)), dim3(256), 0, st, base, dset, m, cents);
    };
    set_rows(k);
    float* cn = t_cn.get<float>(k);
    uint32_t* a = t_a.get<uint32_t>(n);
    uint32_t* keys = t_keys.get<uint32_t>(n);
    uint32_t* order = t_order.get<uint32_t>(n);
    uint32_t* iota = t_iota.get<uint32_t>(n);
    uint32_t* counts = t_counts.get<uint32_t>(k);
    uint32_t* off = t_off.get<uint32_t>(k);
    hipLaunchKernelGGL(k_iota_u32, dim3(blocks_of(n, 256)), dim3(256), 0, st, iota, (size_t)n);
    int bits = 1;
    while (bits < 32 && (1ull << bits) < k) bits++;
    size_t cub_sort = 0, cub_scan = 0;
    SIEVE_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, cub_sort, a, keys, iota, order, (int)n, 0, bits, st));
    SIEVE_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, cub_scan, counts, off, (int)k, st));
    void* cub = t_cub.get<uint8_t>(std::max(cub_sort, cub_scan));
    std::vector<L2TopItem> items;
    l2_items(items, 0, n, 0, k);
    std::vector<uint32_t> hc(k);
    for (uint32_t it = 0; it < iters; it++) {
        hipLaunchKernelGGL(k_l2_norms, dim3(blocks_of(k, 256)), dim3(256), 0, st, cents, (size_t)k, cn);
        l2_top<1>(base, rows, items, cents, cn, a, nullptr, t_items, st);
        hipLaunchKernelGGL(k_zero_u32, dim3(blocks_of(k, 256)), dim3(256), 0, st, counts, (size_t)k);
        hipLaunchKernelGGL(k_count, dim3(blocks_of(n, 256)), dim3(256), 0, st, a, (size_t)n, counts);
        size_t b1 = cub_sort, b2 = cub_scan;
        SIEVE_HIP(hipcub::DeviceRadixSort::SortPairs(cub, b1, a, keys, iota, order, (int)n, 0, bits, st));
        SIEVE_HIP(hipcub::DeviceScan::ExclusiveSum(cub, b2, counts, off, (int)k, st));
        hipLaunchKernelGGL(k_l2_cluster_mean, dim3(k), dim3(128), 0, st, base, rows, order, off, counts, cents);
        SIEVE_HIP(hipGetLastError());
        SIEVE_HIP(hipMemcpyAsync(hc.data(), counts, 4 * (size_t)k, hipMemcpyDeviceToHost, st));
        SIEVE_HIP(hipStreamSynchronize(st));
        size_t m = 0;
        for (uint32_t c = 0; c < k; c++)
            if (hc[c] == 0) set[m++] = make_uint2(c, hrows[rng.below(n)]);
        set_rows(m);
    }
    SIEVE_HIP(hipStreamSynchronize(st));
}

}  // namespace

extern "C" {

int sieve_build_kmeans(const float* sample, size_t n, uint32_t k, uint32_t iters, uint64_t seed, float* cents) {
    return guarded([&] {
        SIEVE_REQUIRE(sample && cents, SIEVE_EINVAL, "null argument");
        SIEVE_REQUIRE(n >= 1 && k >= 1, SIEVE_EINVAL, "kmeans needs a sample and k >= 1");
        SIEVE_REQUIRE(n <= 0xFFFFFFFFull, SIEVE_ELIMIT, "kmeans: at most 2^32 - 1 sample rows");
        SIEVE_HIP(hipDeviceSynchronize());  // device inputs written on other streams are complete
        Stream S;
        const hipStream_t st = S.st;
        Tmp t_sample, t_cents, t_bf, t_nm, t_init, t_a, t_pc, t_pn, t_keys, t_order, t_iota, t_counts, t_off, t_cub;
        TopScratch ts;
        const float* X = to_device(sample, n * PDIM, t_sample, st);
        float* Cd = t_cents.get<float>((size_t)k * PDIM);
        // Initial centroids: sample[rng.below(n)] for each of the k (src/ivf.rs:196).
        Rng rng{seed};
        std::vector<uint32_t> init(k);
        for (uint32_t c = 0; c < k; c++) init[c] = (uint32_t)rng.below(n);
        uint32_t* di = t_init.get<uint32_t>(k);
        SIEVE_HIP(hipMemcpyAsync(di, init.data(), 4 * (size_t)k, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_gather_rows, dim3(blocks_of((size_t)k * PDIM, 256)), dim3(256), 0, st, X, di, (size_t)k, Cd);
        uint32_t* a = t_a.get<uint32_t>(n);
        uint32_t* pc = t_pc.get<uint32_t>(n);
        uint32_t* pn = t_pn.get<uint32_t>(n);
        uint32_t* keys = t_keys.get<uint32_t>(n);
        uint32_t* order = t_order.get<uint32_t>(n);
        uint32_t* iota = t_iota.get<uint32_t>(n);
        uint32_t* counts = t_counts.get<uint32_t>(k);
        uint32_t* off = t_off.get<uint32_t>(k);
        hipLaunchKernelGGL(k_iota_u32, dim3(blocks_of(n, 256)), dim3(256), 0, st, iota, n);
        int bits = 1;
        while (bits < 32 && (1ull << bits) < k) bits++;
        size_t cub_sort = 0, cub_scan = 0;
        SIEVE_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, cub_sort, a, keys, iota, order, (int)n, 0, bits, st));
        SIEVE_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, cub_scan, counts, off, (int)k, st));
        void* cub = t_cub.get<uint8_t>(std::max(cub_sort, cub_scan));
        std::vector<uint32_t> hc(k);
        for (uint32_t it = 0; it < iters; it++) {
            Cents C;
            C.c = Cd;
            C.k = k;
            cents_prepare(C, t_bf, t_nm, st);
            // assign (src/ivf.rs:160-162): the first best centroid of every sample row.
            exact_top(C, X, n, 1, pc, pn, ts, st);
            hipLaunchKernelGGL(k_first, dim3(blocks_of(n, 256)), dim3(256), 0, st, pc, n, 1, a);
            hipLaunchKernelGGL(k_zero_u32, dim3(blocks_of(k, 256)), dim3(256), 0, st, counts, (size_t)k);
            hipLaunchKernelGGL(k_count, dim3(blocks_of(n, 256)), dim3(256), 0, st, a, n, counts);
            size_t b1 = cub_sort, b2 = cub_scan;
            SIEVE_HIP(hipcub::DeviceRadixSort::SortPairs(cub, b1, a, keys, iota, order, (int)n, 0, bits, st));
            SIEVE_HIP(hipcub::DeviceScan::ExclusiveSum(cub, b2, counts, off, (int)k, st));
            hipLaunchKernelGGL(k_cluster_mean, dim3(k), dim3(PDIM), 0, st, X, order, off, counts, Cd);
            SIEVE_HIP(hipGetLastError());
            // Empty clusters, in cluster order: a fresh sample row each (src/ivf.rs:208-211).
            SIEVE_HIP(hipMemcpyAsync(hc.data(), counts, 4 * (size_t)k, hipMemcpyDeviceToHost, st));
            SIEVE_HIP(hipStreamSynchronize(st));
            for (uint32_t c = 0; c < k; c++)
                if (hc[c] == 0) {
                    const uint64_t r = rng.below(n);
                    SIEVE_HIP(hipMemcpyAsync(Cd + (size_t)c * PDIM, X + r * PDIM, PDIM * 4, hipMemcpyDeviceToDevice, st));
                }
        }
        SIEVE_HIP(hipMemcpyAsync(cents, Cd, (size_t)k * PDIM * 4, hipMemcpyDefault, st));
        SIEVE_HIP(hipStreamSynchronize(st));
    });
}

int sieve_build_assign(const float* rows, size_t n, const float* cents, uint32_t k, float soar_lambda,
                       uint32_t* pairs) {
    return guarded([&] {
        SIEVE_REQUIRE(rows && cents && pairs, SIEVE_EINVAL, "null argument");
        SIEVE_REQUIRE(k >= 1, SIEVE_EINVAL, "k >= 1");
        if (n == 0) return;
        SIEVE_HIP(hipDeviceSynchronize());
        Stream S;
        const hipStream_t st = S.st;
        Tmp t_rows, t_cents, t_bf, t_nm, t_pc, t_pn, t_pairs;
        TopScratch ts;
        const float* X = to_device(rows, n * PDIM, t_rows, st);
        Cents C;
        C.c = to_device(cents, (size_t)k * PDIM, t_cents, st);
        C.k = k;
        cents_prepare(C, t_bf, t_nm, st);
        const bool soar = soar_lambda == soar_lambda;
        // assign_top(SOAR_CAND) with SOAR, else assign (src/ivf.rs:313-321).
        const int m = (int)std::min<uint32_t>(soar ? SOAR_CAND : 1, k);
        uint32_t* pc = t_pc.get<uint32_t>(n * m);
        uint32_t* pn = t_pn.get<uint32_t>(n);
        exact_top(C, X, n, m, pc, pn, ts, st);
        uint32_t* dp = on_device(pairs) ? pairs : t_pairs.get<uint32_t>(2 * n);
        launch_soar_pairs(X, n, C.c, pc, pn, m, soar ? 1 : 0, soar ? soar_lambda : 0.0f, dp, st);
        SIEVE_HIP(hipGetLastError());
        if (dp != pairs) SIEVE_HIP(hipMemcpyAsync(pairs, dp, 8 * n, hipMemcpyDeviceToHost, st));
        SIEVE_HIP(hipStreamSynchronize(st));
    });
}

int sieve_build_tier2_fit(const float* rows, size_t n, float* sigma) {
    return guarded([&] {
        SIEVE_REQUIRE(rows && sigma, SIEVE_EINVAL, "null argument");
        SIEVE_HIP(hipDeviceSynchronize());
        Stream S;
        const hipStream_t st = S.st;
        Tmp t_rows, t_sigma;
        const float* X = n ? to_device(rows, n * PDIM, t_rows, st) : nullptr;
        float* ds = on_device(sigma) ? sigma : t_sigma.get<float>(PDIM);
        hipLaunchKernelGGL(k_tier2_fit, dim3(1), dim3(PDIM), 0, st, X, n, ds);
        SIEVE_HIP(hipGetLastError());
        if (ds != sigma) SIEVE_HIP(hipMemcpyAsync(sigma, ds, PDIM * 4, hipMemcpyDeviceToHost, st));
        SIEVE_HIP(hipStreamSynchronize(st));
    });
}


int sieve_build_l2_train(const int8_t* base, size_t n, uint32_t k1, uint32_t k2, size_t sample_n, uint32_t iters,
                         float* coarse, float* leaves) {
    return guarded([&] {
        SIEVE_REQUIRE(base && coarse && leaves, SIEVE_EINVAL, "null argument");
        SIEVE_REQUIRE(n >= 1 && sample_n >= 1 && k1 >= 1 && k2 >= 1, SIEVE_EINVAL, "empty base, sample or model");
        SIEVE_REQUIRE(n <= 0xFFFFFFFFull && sample_n <= 0xFFFFFFFFull, SIEVE_ELIMIT, "at most 2^32 - 1 rows");
        SIEVE_HIP(hipDeviceSynchronize());
        Stream S;
        const hipStream_t st = S.st;
        Tmp t_base, t_idx, t_crow, t_coarse, t_cn, t_a, t_keys, t_order, t_iota, t_items, t_leaves, t_cells;
        const int8_t* B = to_device(base, n * L2D, t_base, st);
        // The sample: Rng(42) draws, sorted (src/bigann.rs:181-185).
        Rng rng{42};
        std::vector<uint32_t> idx(sample_n);
        for (auto& i : idx) i = (uint32_t)rng.below(n);
        std::sort(idx.begin(), idx.end());
        uint32_t* didx = t_idx.get<uint32_t>(sample_n);
        SIEVE_HIP(hipMemcpyAsync(didx, idx.data(), 4 * sample_n, hipMemcpyHostToDevice, st));
        // Coarse rows at an even stride through the sample (:186-189), k1 cells (seed 7).
        const size_t coarse_n = std::min<size_t>(sample_n, (size_t)k1 * 512);
        const size_t stride = sample_n / coarse_n;
        std::vector<uint32_t> crow(coarse_n);
        for (size_t r = 0; r < coarse_n; r++) crow[r] = idx[r * stride];
        uint32_t* dcrow = t_crow.get<uint32_t>(coarse_n);
        SIEVE_HIP(hipMemcpyAsync(dcrow, crow.data(), 4 * coarse_n, hipMemcpyHostToDevice, st));
        float* C = t_coarse.get<float>((size_t)k1 * L2D);
        kmeans_l2_dev(B, dcrow, (uint32_t)coarse_n, k1, iters, 7, C, st);
        // Every sample row to its cell (:191-195), members in sample order.
        float* cn = t_cn.get<float>(k1);
        hipLaunchKernelGGL(k_l2_norms, dim3(blocks_of(k1, 256)), dim3(256), 0, st, C, (size_t)k1, cn);
        uint32_t* a = t_a.get<uint32_t>(sample_n);
        std::vector<L2TopItem> items;
        l2_items(items, 0, (uint32_t)sample_n, 0, k1);
        l2_top<1>(B, didx, items, C, cn, a, nullptr, t_items, st);
        std::vector<uint32_t> ha(sample_n);
        SIEVE_HIP(hipMemcpyAsync(ha.data(), a, 4 * sample_n, hipMemcpyDeviceToHost, st));
        SIEVE_HIP(hipStreamSynchronize(st));
        std::vector<std::vector<uint32_t>> members(k1);
        for (size_t i = 0; i < sample_n; i++) members[ha[i]].push_back(idx[i]);
        std::vector<float> hcoarse((size_t)k1 * L2D);
        SIEVE_HIP(hipMemcpyAsync(hcoarse.data(), C, hcoarse.size() * 4, hipMemcpyDeviceToHost, st));
        SIEVE_HIP(hipStreamSynchronize(st));
        // k2 leaves per cell over its rows (:196-208; seed = the cell). A cell without rows has
        // only its centroid: every leaf is that centroid.
        float* L = on_device(leaves) ? leaves : t_leaves.get<float>((size_t)k1 * k2 * L2D);
        for (uint32_t c = 0; c < k1; c++) {
            float* Lc = L + (size_t)c * k2 * L2D;
            if (members[c].empty()) {
                std::vector<float> rep((size_t)k2 * L2D);
                for (uint32_t j = 0; j < k2; j++)
                    std::copy(hcoarse.begin() + (size_t)c * L2D, hcoarse.begin() + (size_t)(c + 1) * L2D,
                              rep.begin() + (size_t)j * L2D);
                SIEVE_HIP(hipMemcpyAsync(Lc, rep.data(), rep.size() * 4, hipMemcpyHostToDevice, st));
                SIEVE_HIP(hipStreamSynchronize(st));
                continue;
            }
            uint32_t* dm = t_cells.get<uint32_t>(members[c].size());
            SIEVE_HIP(hipMemcpyAsync(dm, members[c].data(), 4 * members[c].size(), hipMemcpyHostToDevice, st));
            kmeans_l2_dev(B, dm, (uint32_t)members[c].size(), k2, iters, (uint64_t)c, Lc, st);
        }
        SIEVE_HIP(hipMemcpyAsync(coarse, C, (size_t)k1 * L2D * 4, hipMemcpyDefault, st));
        if (L != leaves) SIEVE_HIP(hipMemcpyAsync(leaves, L, (size_t)k1 * k2 * L2D * 4, hipMemcpyDeviceToHost, st));
        SIEVE_HIP(hipStreamSynchronize(st));
    });
}

int sieve_build_l2_assign(const int8_t* base, size_t n, const float* coarse, uint32_t k1, const float* leaves,
                          uint32_t k2, uint32_t* leaf) {
    return guarded([&] {
        SIEVE_REQUIRE(base && coarse && leaves && leaf, SIEVE_EINVAL, "null argument");
        SIEVE_REQUIRE(k1 >= (uint32_t)L2_TOP_COARSE && k2 >= 1, SIEVE_EINVAL, "at least 4 coarse cells and 1 leaf each");
        SIEVE_REQUIRE(n <= 0xFFFFFFFFull, SIEVE_ELIMIT, "at most 2^32 - 1 rows");
        if (n == 0) return;
        SIEVE_HIP(hipDeviceSynchronize());
        Stream S;
        const hipStream_t st = S.st;
        Tmp t_base, t_coarse, t_leaves, t_cn, t_ln, t_cells, t_cs, t_keys, t_vals, t_sk, t_sv, t_cub, t_items, t_lid, t_ls,
            t_rows, t_out, t_counts, t_off;
        const int8_t* B = to_device(base, n * L2D, t_base, st);
        const float* C = to_device(coarse, (size_t)k1 * L2D, t_coarse, st);
        const float* L = to_device(leaves, (size_t)k1 * k2 * L2D, t_leaves, st);
        float* cn = t_cn.get<float>(k1);
        float* ln = t_ln.get<float>((size_t)k1 * k2);
        hipLaunchKernelGGL(k_l2_norms, dim3(blocks_of(k1, 256)), dim3(256), 0, st, C, (size_t)k1, cn);
        hipLaunchKernelGGL(k_l2_norms, dim3(blocks_of((size_t)k1 * k2, 256)), dim3(256), 0, st, L, (size_t)k1 * k2, ln);
        uint32_t* out = on_device(leaf) ? leaf : t_out.get<uint32_t>(n);
        // Rows in chunks: their 4 nearest cells, the (cell, row) pairs sorted by cell, each cell's
        // best leaf for its rows, then the pick.
        const size_t CH = std::min<size_t>(n, size_t(1) << 25);
        const size_t P = CH * L2_TOP_COARSE;
        uint32_t* cells = t_cells.get<uint32_t>(P);
        float* cs = t_cs.get<float>(P);
        uint32_t* keys = t_keys.get<uint32_t>(P);
        uint32_t* vals = t_vals.get<uint32_t>(P);
        uint32_t* sk = t_sk.get<uint32_t>(P);
        uint32_t* sv = t_sv.get<uint32_t>(P);
        uint32_t* prow = t_rows.get<uint32_t>(P);
        uint32_t* lid = t_lid.get<uint32_t>(P);
        float* ls = t_ls.get<float>(P);
        uint32_t* counts = t_counts.get<uint32_t>(k1);
        int bits = 1;
        while (bits < 32 && (1ull << bits) < k1) bits++;
        size_t cub_sort = 0;
        SIEVE_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, cub_sort, keys, sk, vals, sv, (int)P, 0, bits, st));
        void* cub = t_cub.get<uint8_t>(cub_sort);
        std::vector<uint32_t> hk(k1);
        for (size_t r0 = 0; r0 < n; r0 += CH) {
            const size_t m = std::min(CH, n - r0);
            std::vector<L2TopItem> items;
            l2_items(items, 0, (uint32_t)m, 0, k1);
            l2_top<L2_TOP_COARSE>(B + r0 * L2D, nullptr, items, C, cn, cells, cs, t_items, st);
            // pairs p = (row, t): key its cell, value p; sorted by cell, stable.
            SIEVE_HIP(hipMemcpyAsync(keys, cells, 4 * m * L2_TOP_COARSE, hipMemcpyDeviceToDevice, st));
            hipLaunchKernelGGL(k_iota_u32, dim3(blocks_of(m * L2_TOP_COARSE, 256)), dim3(256), 0, st, vals,
                               m * L2_TOP_COARSE);
            size_t b1 = cub_sort;
            SIEVE_HIP(hipcub::DeviceRadixSort::SortPairs(cub, b1, keys, sk, vals, sv, (int)(m * L2_TOP_COARSE), 0, bits,
                                                         st));
            hipLaunchKernelGGL(k_zero_u32, dim3(blocks_of(k1, 256)), dim3(256), 0, st, counts, (size_t)k1);
            hipLaunchKernelGGL(k_count, dim3(blocks_of(m * L2_TOP_COARSE, 256)), dim3(256), 0, st, cells,
                               m * L2_TOP_COARSE, counts);
            hipLaunchKernelGGL(k_pair_rows, dim3(blocks_of(m * L2_TOP_COARSE, 256)), dim3(256), 0, st, sv,
                               m * L2_TOP_COARSE, (uint32_t)r0, prow);
            SIEVE_HIP(hipMemcpyAsync(hk.data(), counts, 4 * (size_t)k1, hipMemcpyDeviceToHost, st));
            SIEVE_HIP(hipStreamSynchronize(st));
            // Each cell's rows (a run of the sorted pairs) against its k2 leaves.
            items.clear();
            uint32_t o = 0;
            for (uint32_t c = 0; c < k1; c++) {
                l2_items(items, o, hk[c], c * k2, k2);
                o += hk[c];
            }
            l2_top<1>(B, prow, items, L, ln, lid, ls, t_items, st);
            hipLaunchKernelGGL(k_scatter_pairs, dim3(blocks_of(m * L2_TOP_COARSE, 256)), dim3(256), 0, st, sv, lid, ls,
                               m * L2_TOP_COARSE, keys, cs);
            hipLaunchKernelGGL(k_l2_pick, dim3(blocks_of(m, 256)), dim3(256), 0, st, cells, keys, cs, m, k2, out + r0);
            SIEVE_HIP(hipGetLastError());
        }
        if (out != leaf) SIEVE_HIP(hipMemcpyAsync(leaf, out, 4 * n, hipMemcpyDeviceToHost, st));
        SIEVE_HIP(hipStreamSynchronize(st));
    });
}

}  // extern "C"
