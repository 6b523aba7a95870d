// Centroid routing: Index::route (src/ivf.rs:587-613) over a batch of queries.
//
// 1. k_gemm_bf16_route: S = Q * C^T with v_mfma_f32_16x16x32_bf16 (BF16 inputs, f32
//    accumulation) for every (query, centroid).
// 2. k_route_select, one workgroup per query: the nprobe-th largest approximate score by a
//    4-pass radix select; every centroid whose approximate score lies within twice the BF16
//    error bound of it is rescored in f32 exactly as the reference's gemm_abt fallback sums
//    (src/blas.rs:47-53: products left to right from -0.0, no FMA); the rescored scores are
//    sorted (score desc, id asc) as TopK::into_sorted (src/topk.rs:57-61); then the margin
//    and budget cuts (src/ivf.rs:598-611).
//
// Error bound: BF16 round-to-nearest has unit roundoff u = 2^-8, so each product carries a
// relative error below 2u + u^2, and |S - q.c| <= (2u + u^2 + 256 * 2^-24) |q| |c| (the last
// term covers f32 accumulation in any order). Any centroid in the exact top-nprobe then has an
// approximate score within 2e of the approximate nprobe-th score, so the rescored set contains
// the reference's probe set and the reference's f32 scores are reproduced bit for bit.
//
// sieve_gemm_abt (blas::gemm_abt, src/blas.rs:5-10) uses k_gemm_f32_abt: v_mfma_f32_16x16x4_f32.
#include <float.h>

#include "kernels.hpp"
#include "select.hpp"

namespace sieve {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------------ BF16 routing GEMM ---
// Workgroup tile 128 x 128 (4 waves in 2 x 2, each 64 x 64 = 4 x 4 MFMA tiles), K = 256.
// MFMA 16x16x32 operand layout: lane l supplies row (l % 16) and k-octet (l / 16) of each
// operand; the accumulator holds rows 4 * (l / 16) + i, column l % 16.
__global__ __launch_bounds__(256) void k_gemm_bf16_route(const uint16_t* __restrict__ A, const uint16_t* __restrict__ B,
                                                         float* __restrict__ S, int m, int n, size_t lds) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int tm = blockIdx.y * 128 + (wave >> 1) * 64;
    const int tn = blockIdx.x * 128 + (wave & 1) * 64;
    const int r16 = lane & 15, ko = (lane >> 4) * 8;
    f32x4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const uint16_t* arow[4];
    const uint16_t* brow[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        int r = min(tm + i * 16 + r16, m - 1);
        int c = min(tn + i * 16 + r16, n - 1);
        arow[i] = A + (size_t)r * PDIM + ko;
        brow[i] = B + (size_t)c * PDIM + ko;
    }
#pragma unroll 2
    for (int k0 = 0; k0 < PDIM; k0 += 32) {
        bf16x8 a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            a[i] = *reinterpret_cast<const bf16x8*>(arow[i] + k0);
            b[i] = *reinterpret_cast<const bf16x8*>(brow[i] + k0);
        }
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            int col = tn + j * 16 + r16;
            if (col >= n) continue;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                int row = tm + i * 16 + (lane >> 4) * 4 + v;
                if (row < m) S[(size_t)row * lds + col] = acc[i][j][v];
            }
        }
}

void launch_gemm_bf16_route(const uint16_t* qbf, const uint16_t* cbf, float* s, int m, int n, size_t lds,
                            hipStream_t st) {
    if (m == 0 || n == 0) return;
    dim3 grid((n + 127) / 128, (m + 127) / 128);
    hipLaunchKernelGGL(k_gemm_bf16_route, grid, dim3(256), 0, st, qbf, cbf, s, m, n, lds);
}

// ------------------------------------------------------------------ f32 GEMM -----------
// C = A * B^T, A m x k, B n x k, f32 MFMA 16x16x4 (lane l supplies row l % 16 at k-offset
// l / 16). Workgroup tile 64 x 64: 4 waves, each 32 x 32 (2 x 2 MFMA tiles).
typedef float f32x4v __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_gemm_f32_abt(const float* __restrict__ A, const float* __restrict__ B,
                                                      float* __restrict__ C, int m, int n, int k) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int tm = blockIdx.y * 64 + (wave >> 1) * 32;
    const int tn = blockIdx.x * 64 + (wave & 1) * 32;
    const int r16 = lane & 15, kq = lane >> 4;
    f32x4v acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = f32x4v{0.f, 0.f, 0.f, 0.f};
    const float* ar[2];
    const float* br[2];
    bool av[2], bv[2];
#pragma unroll
    for (int i = 0; i < 2; i++) {
        int r = tm + i * 16 + r16, c = tn + i * 16 + r16;
        av[i] = r < m;
        bv[i] = c < n;
        ar[i] = A + (size_t)min(r, m - 1) * k;
        br[i] = B + (size_t)min(c, n - 1) * k;
    }
    for (int k0 = 0; k0 < k; k0 += 4) {
        int kk = k0 + kq;
        float a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; i++) {
            a[i] = (av[i] && kk < k) ? ar[i][kk] : 0.0f;
            b[i] = (bv[i] && kk < k) ? br[i][kk] : 0.0f;
        }
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
            int col = tn + j * 16 + r16;
            if (col >= n) continue;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                int row = tm + i * 16 + (lane >> 4) * 4 + v;
                if (row < m) C[(size_t)row * n + col] = acc[i][j][v];
            }
        }
}

void launch_gemm_f32_abt(const float* a, const float* b, float* c, int m, int n, int k, hipStream_t st) {
    if (m == 0 || n == 0) return;
    dim3 grid((n + 63) / 64, (m + 63) / 64);
    hipLaunchKernelGGL(k_gemm_f32_abt, grid, dim3(256), 0, st, a, b, c, m, n, k);
}

// ------------------------------------------------------------------ route select -------
// Radix select of the k-th largest u32 key among keys(i), i < n, by a 256-thread block.
// Returns the key; all threads get it.
template <class KeyFn>
__device__ uint32_t block_kth_u32(int n, int k, KeyFn key, uint32_t* hist, uint32_t* sh) {
    const int t = threadIdx.x;
    uint32_t prefix = 0, mask = 0;
    int kk = k;
    for (int shift = 24; shift >= 0; shift -= 8) {
        hist[t] = 0;
        __syncthreads();
        for (int i = t; i < n; i += 256) {
            uint32_t x = key(i);
            if ((x & mask) == prefix) atomicAdd(&hist[(x >> shift) & 255], 1u);
        }
        __syncthreads();
        if (t < 64) {
            // lane l holds digits 255-4l .. 252-4l (top digits first)
            uint32_t c0 = hist[255 - 4 * t], c1 = hist[254 - 4 * t], c2 = hist[253 - 4 * t], c3 = hist[252 - 4 * t];
            uint32_t s = c0 + c1 + c2 + c3;
            uint32_t incl = wave_incl_scan(s);
            uint64_t hit = __ballot(incl >= (uint32_t)kk);
            int f = __ffsll((long long)hit) - 1;
            if (t == f) {
                uint32_t cum = incl - s;
                uint32_t cs[4] = {c0, c1, c2, c3};
                for (int j = 0; j < 4; j++) {
                    if (cum + cs[j] >= (uint32_t)kk) {
                        sh[0] = 255 - 4 * t - j;
                        sh[1] = kk - cum;
                        break;
                    }
                    cum += cs[j];
                }
            }
        }
        __syncthreads();
        prefix |= sh[0] << shift;
        mask |= 0xFFu << shift;
        kk = (int)sh[1];
        __syncthreads();
    }
    return prefix;
}

// The same over keys the threads hold in registers: thread t holds keys of i = t + 256 j
// (j < VALS, i < n), each read once from memory.
template <int VALS>
__device__ uint32_t block_kth_regs(const uint32_t (&x)[VALS], int n, int k, uint32_t* hist, uint32_t* sh) {
    const int t = threadIdx.x;
    uint32_t prefix = 0, mask = 0;
    int kk = k;
    for (int shift = 24; shift >= 0; shift -= 8) {
        hist[t] = 0;
        __syncthreads();
#pragma unroll
        for (int j = 0; j < VALS; j++)
            if (t + 256 * j < n && (x[j] & mask) == prefix) atomicAdd(&hist[(x[j] >> shift) & 255], 1u);
        __syncthreads();
        if (t < 64) {
            uint32_t c0 = hist[255 - 4 * t], c1 = hist[254 - 4 * t], c2 = hist[253 - 4 * t], c3 = hist[252 - 4 * t];
            uint32_t s = c0 + c1 + c2 + c3;
            uint32_t incl = wave_incl_scan(s);
            uint64_t hit = __ballot(incl >= (uint32_t)kk);
            int f = __ffsll((long long)hit) - 1;
            if (t == f) {
                uint32_t cum = incl - s;
                uint32_t cs[4] = {c0, c1, c2, c3};
                for (int j = 0; j < 4; j++) {
                    if (cum + cs[j] >= (uint32_t)kk) {
                        sh[0] = 255 - 4 * t - j;
                        sh[1] = kk - cum;
                        break;
                    }
                    cum += cs[j];
                }
            }
        }
        __syncthreads();
        prefix |= sh[0] << shift;
        mask |= 0xFFu << shift;
        kk = (int)sh[1];
        __syncthreads();
    }
    return prefix;
}

// Bitonic sort, descending, of n2 (a power of two) (key, ~id) pairs in LDS.
__device__ void block_sort_desc(uint2* a, int n2) {
    for (int size = 2; size <= n2; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = threadIdx.x; i < n2 / 2; i += blockDim.x) {
                int lo = 2 * i - (i & (stride - 1));
                int hi = lo + stride;
                bool desc = ((lo & size) == 0);
                uint2 x = a[lo], y = a[hi];
                bool x_less = (x.x < y.x) || (x.x == y.x && x.y < y.y);
                if (x_less == desc) {
                    a[lo] = y;
                    a[hi] = x;
                }
            }
            __syncthreads();
        }
    }
}

// REG_VALS > 0: the approximate scores (k <= 256 REG_VALS) are read once into registers and
// the select and the band pass run over them; 0: over memory (any k).
template <int REG_VALS>
__global__ __launch_bounds__(256) void k_route_select(RouteArgs a) {
    __shared__ float qs[PDIM];
    __shared__ uint32_t hist[256];
    __shared__ uint32_t sh[4];
    __shared__ uint32_t ncand;
    __shared__ uint32_t cids[MAX_RESCORE];
    __shared__ uint2 sorted[MAX_RESCORE];
    const size_t qi = blockIdx.x;
    const int t = threadIdx.x;
    const float* S = a.approx + qi * a.lds;
    const int n = min(a.nprobe, a.k);
    if (a.l2) {
        for (int i = t; i < PDIM; i += 256) qs[i] = i < a.dims ? a.qf[qi * a.dims + i] : 0.0f;
    } else {
        for (int i = t; i < PDIM; i += 256) qs[i] = a.q256[qi * PDIM + i];
    }
    if (t == 0) ncand = 0;
    __syncthreads();
    if (n <= 0) {
        if (t == 0) a.out_n[qi] = 0;
        return;
    }
    // approximate route score of centroid i
    auto approx = [&](int i) -> float { return a.l2 ? 2.0f * S[i] - a.cnorm[i] : S[i]; };
    uint32_t kth;
    float thr;
    if constexpr (REG_VALS > 0) {
        uint32_t x[REG_VALS];
#pragma unroll
        for (int j = 0; j < REG_VALS; j++) {
            const int i = t + 256 * j;
            x[j] = i < a.k ? f2key(approx(i) + 0.0f) : 0u;
        }
        kth = block_kth_regs<REG_VALS>(x, a.k, n, hist, sh);
        const float e = a.err_coef * a.qnorm[qi];
        const float kv = key2f(kth);
        thr = kv - 2.2f * e - 1e-6f * (1.0f + fabsf(kv));
#pragma unroll
        for (int j = 0; j < REG_VALS; j++) {
            const int i = t + 256 * j;
            // key2f(key(v)) == v + 0.0f, which compares with thr as v does
            if (i < a.k && key2f(x[j]) >= thr) {
                uint32_t jj = atomicAdd(&ncand, 1u);
                if (jj < MAX_RESCORE) cids[jj] = (uint32_t)i;
            }
        }
    } else {
        // Large K: a 4096-bin histogram of the keys' top 12 bits (in cids) finds the bin of the
        // n-th largest score. Every score the band keeps (>= kth - 2.2 e - 1e-6 (1 + |kth|)) is
        // at least that bin's lower edge minus the same margin, so one more pass collects a
        // superset of the top n and the band into `sorted`; the n-th and the band then come from
        // LDS: two passes over the K scores. A collection that does not fit falls back to the
        // 8-bit radix over memory and a band pass.
        static_assert(MAX_RESCORE == 4096, "the histogram lives in cids");
        auto key = [&](int i) { return f2key(approx(i) + 0.0f); };
        for (int i = t; i < 4096; i += 256) cids[i] = 0;
        __syncthreads();
        for (int i = t; i < a.k; i += 256) atomicAdd(&cids[key(i) >> 20], 1u);
        __syncthreads();
        if (t < 64) {
            // lane l sums bins 4095 - 64 l .. 4032 - 64 l (top first)
            uint32_t sl = 0;
            for (int b = 0; b < 64; b++) sl += cids[4095 - 64 * t - b];
            const uint32_t incl = wave_incl_scan(sl);
            const uint64_t hit = __ballot(incl >= (uint32_t)n);
            const int f = __ffsll((long long)hit) - 1;
            if (t == f) {
                uint32_t cum = incl - sl;
                for (int b = 0; b < 64; b++) {
                    const uint32_t c = cids[4095 - 64 * t - b];
                    if (cum + c >= (uint32_t)n) {
                        sh[0] = 4095 - 64 * t - b;  // the bin
                        sh[1] = cum + c;            // scores in it and above
                        break;
                    }
                    cum += c;
                }
            }
        }
        if (t == 0) ncand = 0;
        __syncthreads();
        const uint32_t bin = sh[0], upto = sh[1];
        const float e = a.err_coef * a.qnorm[qi];
        const float lo = key2f(bin << 20), hi = key2f((bin << 20) | 0xFFFFFu);
        const float lim = lo - 2.2f * e - 1e-6f * (1.0f + fmaxf(fabsf(lo), fabsf(hi)));
        for (int i = t; i < a.k; i += 256) {
            const float v = approx(i);
            if (v >= lim) {
                const uint32_t j = atomicAdd(&ncand, 1u);
                if (j < (uint32_t)MAX_RESCORE) sorted[j] = make_uint2(f2key(v + 0.0f), ~(uint32_t)i);
            }
        }
        __syncthreads();
        const uint32_t got = ncand;
        __syncthreads();
        if (got <= (uint32_t)MAX_RESCORE && got >= upto) {
            kth = block_kth_largest(sorted, (int)got, n, hist, sh).x;
            const float kv = key2f(kth);
            thr = kv - 2.2f * e - 1e-6f * (1.0f + fabsf(kv));
            if (t == 0) ncand = 0;
            __syncthreads();
            for (uint32_t j = t; j < got; j += 256)
                if (key2f(sorted[j].x) >= thr) cids[atomicAdd(&ncand, 1u)] = ~sorted[j].y;
        } else {
            kth = block_kth_u32(a.k, n, key, hist, sh);
            const float kv = key2f(kth);
            thr = kv - 2.2f * e - 1e-6f * (1.0f + fabsf(kv));
            if (t == 0) ncand = 0;
            __syncthreads();
            for (int i = t; i < a.k; i += 256) {
                if (approx(i) >= thr) {
                    uint32_t j = atomicAdd(&ncand, 1u);
                    if (j < MAX_RESCORE) cids[j] = (uint32_t)i;
                }
            }
        }
    }
    __syncthreads();
    const int nc = (int)ncand;
    if (nc > MAX_RESCORE) {
        if (t == 0) {
            atomicOr(a.overflow, 1u);
            a.out_n[qi] = 0;
        }
        return;
    }
    int n2 = 1;
    while (n2 < nc) n2 <<= 1;
    // Exact f32 rescoring: products summed left to right from -0.0 (src/blas.rs:51).
    for (int i = t; i < n2; i += 256) {
        if (i < nc) {
            uint32_t c = cids[i];
            const float4* row = reinterpret_cast<const float4*>(a.cents + (size_t)c * a.cstride);
            float acc = -0.0f;
            for (int j = 0; j < a.dims / 4; j++) {
                float4 x = row[j];
                acc = acc + qs[4 * j] * x.x;
                acc = acc + qs[4 * j + 1] * x.y;
                acc = acc + qs[4 * j + 2] * x.z;
                acc = acc + qs[4 * j + 3] * x.w;
            }
            // L2 route score 2 <q, c> - |c|^2 (src/bigann.rs:667)
            float sc = a.l2 ? 2.0f * acc - a.cnorm[c] : acc;
            sorted[i] = make_uint2(f2key(sc + 0.0f), ~c);
        } else {
            sorted[i] = make_uint2(0u, 0u);
        }
    }
    __syncthreads();
    block_sort_desc(sorted, n2);
    if (t == 0 && nc == 0) {
        a.out_n[qi] = 0;
        sh[2] = 0;
    } else if (t == 0) {
        // The top n, then the margin cut (keep s >= best - margin) and the scan budget
        // (take clusters while the stored slots so far fit; at least one).
        int keep = min(n, nc);  // nc < n only for non-finite scores
        const float best = key2f(sorted[0].x);
        if (a.l2) {
            // keep |q - c|^2 <= (1 + margin) |q - c_1|^2 (src/bigann.rs:669-674)
            if (a.margin > 0.0f) {
                const float qq = a.qq[qi];
                const float lim = (qq - best) * (1.0f + a.margin);
                int kk = 0;
                while (kk < keep && qq - key2f(sorted[kk].x) <= lim) kk++;
                keep = kk;
            }
        } else if (a.margin == a.margin) {
            const float lim = best - a.margin;
            int kk = 0;
            while (kk < keep && key2f(sorted[kk].x) >= lim) kk++;
            keep = kk;
        }
        if (a.budget > 0) {
            uint64_t used = 0;
            int kk = 0;
            while (kk < keep) {
                used += a.cl_slots[~sorted[kk].y];
                if (used > a.budget) break;
                kk++;
            }
            keep = max(kk, 1);
        }
        a.out_n[qi] = (uint32_t)keep;
        sh[2] = (uint32_t)keep;
    }
    __syncthreads();
    const int keep = (int)sh[2];
    for (int i = t; i < keep; i += 256) {
        const uint32_t c = ~sorted[i].y;
        float outv = key2f(sorted[i].x);
        if (a.l2) {
            // the probe list carries <q, c> itself: the estimator's `cent` (src/bigann.rs:701)
            const float4* row = reinterpret_cast<const float4*>(a.cents + (size_t)c * a.cstride);
            float acc = -0.0f;
            for (int j = 0; j < a.dims / 4; j++) {
                float4 x = row[j];
                acc = acc + qs[4 * j] * x.x;
                acc = acc + qs[4 * j + 1] * x.y;
                acc = acc + qs[4 * j + 2] * x.z;
                acc = acc + qs[4 * j + 3] * x.w;
            }
            outv = acc;
        }
        a.out_s[qi * a.out_stride + i] = outv;
        a.out_c[qi * a.out_stride + i] = c;
    }
}

void launch_route_select(const RouteArgs& a, size_t nq, hipStream_t st) {
    if (nq == 0) return;
    if (a.k <= 256 * 16)
        hipLaunchKernelGGL(k_route_select<16>, dim3((unsigned)nq), dim3(256), 0, st, a);
    else if (a.k <= 256 * 64)
        hipLaunchKernelGGL(k_route_select<64>, dim3((unsigned)nq), dim3(256), 0, st, a);
    else
        hipLaunchKernelGGL(k_route_select<0>, dim3((unsigned)nq), dim3(256), 0, st, a);
}

}  // namespace sieve
