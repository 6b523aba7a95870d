// The 1B-scale L2 path over 100-d int8 vectors: bigann::Index::{route, candidates, rerank}
// (src/bigann.rs:652-747); the multi-bit refinement of the candidates is in refine.hip.
//
//   prep:   q as f32 zero-padded to 128 dims, Rotation (128-d, src/rabitq.rs:59-69), the
//           32-group LUT and (a, b) (rabitq::Query::<32>::new); the routing operand
//           [q | q] in BF16 (int8 values are exact in BF16); sum of q_i^2 for the margin.
//   route:  the BF16 MFMA GEMM against [c_hi | c_lo] (each leaf split into two BF16 halves,
//           so c_hi + c_lo carries 16 significant bits) gives <q, c> to ~2^-16 relative;
//           k_route_select rescores the band in f32 as gemm_abt sums and applies
//           2 <q, c> - |c|^2 and the (1 + margin) cut (src/bigann.rs:660-678).
//   scan:   k_scan_t<32, true> (scan.hip).
//   rerank: exact i32 dot over the 100 int8 dims, score 2 dot - |x|^2, best k
//           (src/bigann.rs:736-747).
#include <float.h>

#include "kernels.hpp"
#include "select.hpp"

namespace sieve {

constexpr int L2_D = 100;
constexpr int L2_PD = 128;
constexpr int L2_G = 32;

__device__ __forceinline__ uint16_t bf16_rne(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// One wave per query.
__global__ __launch_bounds__(64) void k_query_prep_l2(const int8_t* __restrict__ q, const float* __restrict__ signs,
                                                      float rot_norm, L2Prep out) {
    __shared__ float v[L2_PD];
    __shared__ float mins[L2_G];
    const size_t qi = blockIdx.x;
    const int lane = threadIdx.x;
    for (int i = lane; i < L2_PD; i += WAVE) v[i] = i < L2_D ? (float)q[qi * L2_D + i] : 0.0f;
    __syncthreads();
    for (int i = lane; i < PDIM; i += WAVE) {
        int d = i & (L2_PD - 1);
        out.qbf[qi * PDIM + i] = d < L2_D ? bf16_rne(v[d]) : (uint16_t)0;
    }
    if (lane == 0) {
        float qq = -0.0f, nn = 0.0f;
        for (int i = 0; i < L2_D; i++) {
            qq = qq + v[i] * v[i];
            nn += v[i] * v[i];
        }
        out.qq[qi] = qq;
        out.qnorm[qi] = sqrtf(nn);
        for (int i = 0; i < L2_D; i++) out.qf[qi * L2_D + i] = v[i];
    }
    __syncthreads();
    for (int r = 0; r < 3; r++) {
        for (int i = lane; i < L2_PD; i += WAVE) v[i] = v[i] * signs[r * L2_PD + i];
        __syncthreads();
        for (int h = 1; h < L2_PD; h <<= 1) {
            for (int j = lane; j < L2_PD / 2; j += WAVE) {
                int blk = j / h, i = j - blk * h;
                int a = blk * 2 * h + i, b = a + h;
                float x = v[a], y = v[b];
                v[a] = x + y;
                v[b] = x - y;
            }
            __syncthreads();
        }
    }
    for (int i = lane; i < L2_PD; i += WAVE) v[i] = v[i] * rot_norm;
    __syncthreads();
    if (out.qrot)
        for (int i = lane; i < L2_PD; i += WAVE) out.qrot[qi * L2_PD + i] = v[i];
    const int g = lane;
    float mn = FLT_MAX, mx = -FLT_MAX;
    float f[16];
    if (g < L2_G) {
#pragma unroll
        for (int c = 0; c < 16; c++) {
            float s = -0.0f;
#pragma unroll
            for (int j = 0; j < 4; j++)
                if ((c >> j) & 1) s = s + v[4 * g + j];
            f[c] = s;
        }
#pragma unroll
        for (int c = 0; c < 16; c++) mn = fminf(mn, f[c]);
#pragma unroll
        for (int c = 0; c < 16; c++) mx = fmaxf(mx, f[c]);
    }
    const float max_range = fmaxf(0.0f, wave_max(g < L2_G ? mx - mn : 0.0f));
    const float scale = max_range > 0.0f ? 255.0f / max_range : 1.0f;
    if (g < L2_G) {
        uint32_t packed[4] = {0, 0, 0, 0};
#pragma unroll
        for (int c = 0; c < 16; c++) {
            float t = roundf((f[c] - mn) * scale);
            t = fminf(fmaxf(t, 0.0f), 255.0f);
            packed[c >> 2] |= ((uint32_t)t) << ((c & 3) * 8);
        }
        reinterpret_cast<uint4*>(out.tables + qi * L2_G * 16)[g] = make_uint4(packed[0], packed[1], packed[2], packed[3]);
        mins[g] = mn;
    }
    __syncthreads();
    if (lane == 0) {
        float bias = -0.0f;
        for (int i = 0; i < L2_G; i++) bias = bias + mins[i];
        float sum = -0.0f;
        for (int i = 0; i < L2_PD; i++) sum = sum + v[i];
        out.qab[qi] = make_float4(2.0f / scale, 2.0f * bias - sum, scale, bias);
    }
}

void launch_query_prep_l2(const int8_t* q, size_t nq, const float* signs, float rot_norm, const L2Prep& out,
                          hipStream_t st) {
    if (nq == 0) return;
    hipLaunchKernelGGL(k_query_prep_l2, dim3((unsigned)nq), dim3(WAVE), 0, st, q, signs, rot_norm, out);
}

// Leaves k x 100 f32 -> k x 256 BF16 [c_hi (100), 0 (28), c_lo (100), 0 (28)] with
// c_lo = bf16(c - f32(c_hi)).
__global__ void k_split_leaves(const float* __restrict__ leaves, uint32_t k, uint16_t* __restrict__ out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)k * PDIM) return;
    size_t c = i / PDIM;
    int col = (int)(i % PDIM), d = col & (L2_PD - 1);
    uint16_t v = 0;
    if (d < L2_D) {
        float x = leaves[c * L2_D + d];
        uint16_t hi = bf16_rne(x);
        v = col < L2_PD ? hi : bf16_rne(x - bf16_to_f32(hi));
    }
    out[i] = v;
}

void launch_split_leaves(const float* leaves, uint32_t k, uint16_t* out, hipStream_t st) {
    if (k == 0) return;
    size_t n = (size_t)k * PDIM;
    hipLaunchKernelGGL(k_split_leaves, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, leaves, k, out);
}

// Exact rerank: one 256-thread workgroup per query; 8 lanes per row, each summing up to four
// of the row's 25 dwords with v_dot4_i32_i8.
__global__ __launch_bounds__(256) void k_rerank_l2(L2RerankArgs a) {
    extern __shared__ __align__(16) uint8_t lds[];
    const int cs = a.cand_stride;
    const int p2 = pow2_at_least(max(cs, 1));
    float* score = reinterpret_cast<float*>(lds);
    uint32_t* ids = reinterpret_cast<uint32_t*>(lds + 4 * (size_t)cs);
    uint2* keys = reinterpret_cast<uint2*>(lds + ((8 * (size_t)cs + 15) & ~(size_t)15));
    const size_t qi = blockIdx.x;
    const int t = threadIdx.x, sub = t & 7;
    const int n = (int)a.cand_n[qi];
    const int32_t* qw = reinterpret_cast<const int32_t*>(a.q + qi * L2_D);
    int wq[4];
#pragma unroll
    for (int j = 0; j < 4; j++) wq[j] = (sub + 8 * j) < 25 ? qw[sub + 8 * j] : 0;
    for (int base = 0; base < n; base += 256 / 8) {
        const int i = base + (t >> 3);
        int acc = 0;
        uint32_t id = 0;
        if (i < n) {
            id = a.cand_ids[qi * cs + i];
            const int32_t* row = reinterpret_cast<const int32_t*>(a.base + (size_t)id * L2_D);
#pragma unroll
            for (int j = 0; j < 4; j++)
                if (sub + 8 * j < 25) acc = __builtin_amdgcn_sdot4(row[sub + 8 * j], wq[j], acc, false);
        }
        acc = group_sum_dpp<8>(acc);  // the 8 lanes of the row (integer: any order)
        if (sub == 0 && i < n) {
            score[i] = 2.0f * (float)acc - a.cand_nx[qi * cs + i];
            ids[i] = id;
        }
    }
    __syncthreads();
    // The top k by (score desc, row asc): k rounds of a block maximum for small k (rows are
    // unique: one slot per row), else a full sort.
    const int kk = min(a.k, n);
    if (a.k <= 32) {
        __shared__ unsigned long long red[8];
        __shared__ uint2 best[32];
        for (int i = t; i < n; i += 256) keys[i] = make_uint2(f2key(score[i] + 0.0f), ~ids[i]);
        __syncthreads();
        block_topk_small(keys, n, a.k, best, red);
        for (int i = t; i < kk; i += 256) keys[i] = best[i];
        __syncthreads();
    } else {
        const int n2 = pow2_at_least(max(n, 1));
        for (int i = t; i < n2; i += 256)
            keys[i] = i < n ? make_uint2(f2key(score[i] + 0.0f), ~ids[i]) : make_uint2(0u, 0u);
        __syncthreads();
        block_sort_pairs(keys, nullptr, n2, false);
    }
    for (int i = t; i < kk; i += 256) {
        a.out_s[qi * a.k + i] = key2f(keys[i].x);
        a.out_r[qi * a.k + i] = ~keys[i].y;
    }
    if (t == 0) a.out_n[qi] = (uint32_t)kk;
    (void)p2;
}

size_t rerank_l2_lds_bytes(int cand_stride) {
    int p2 = 1;
    while (p2 < cand_stride) p2 <<= 1;
    return ((8 * (size_t)cand_stride + 15) & ~(size_t)15) + 8 * (size_t)p2;
}

void launch_rerank_l2(const L2RerankArgs& a, size_t nq, hipStream_t st) {
    if (nq == 0) return;
    size_t lds = rerank_l2_lds_bytes(a.cand_stride);
    if (lds > 64 * 1024)
        (void)hipFuncSetAttribute((const void*)k_rerank_l2, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(k_rerank_l2, dim3((unsigned)nq), dim3(256), lds, st, a);
}

}  // namespace sieve
