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
        launch_gemm_bf16_route(xbf, C.bf, nullptr, approx, tmax, (int)mm, (int)C.k, C.k, st);
        RouteArgs ra{};
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

}  // extern "C"
