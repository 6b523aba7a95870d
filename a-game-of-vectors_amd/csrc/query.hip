// Per-query preparation and index re-layout kernels.
//
// Query prep restates, per query, rabitq::Query::new (src/rabitq.rs:119-125): the rotation
// (src/rabitq.rs:59-84), the u8 LUT (lut::Lut::new, src/lut.rs:33-56) and the affine map
// (a, b); and the int8 rerank queries (quant::query_i8, src/quant.rs:104-111, over
// w_i = q_i / scale_i, src/ivf.rs:753-755 and :786-788). Every f32 operation is the
// reference's, in the reference's order, and the library is built with -ffp-contract=off, so
// tables, a, b and the int8 queries are bit-identical to the reference's.
#include <float.h>

#include "kernels.hpp"

namespace sieve {

__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// One wave per query.
__global__ __launch_bounds__(64) void k_query_prep(const float* __restrict__ q256, const float* __restrict__ q768,
                                                   const float* __restrict__ signs, float rot_norm,
                                                   const float* __restrict__ quant, QueryPrep out) {
    __shared__ float v[PDIM];
    __shared__ float mins[GROUPS];
    const size_t qi = blockIdx.x;
    const int lane = threadIdx.x;
    const float* q = q256 + qi * PDIM;

    float nn = 0.0f;
    for (int i = lane; i < PDIM; i += WAVE) {
        float x = q[i];
        v[i] = x;
        out.qbf[qi * PDIM + i] = f32_to_bf16_rne(x);
        nn += x * x;
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) nn += __shfl_xor(nn, d, 64);
    if (lane == 0) out.qnorm[qi] = sqrtf(nn);
    __syncthreads();

    rotate_lds<PDIM>(v, signs, rot_norm, lane);

    // Lut::new: lane g builds group g's 16 sums (each a left-to-right sum from -0.0).
    const int g = lane;
    float f[16];
#pragma unroll
    for (int c = 0; c < 16; c++) {
        float s = -0.0f;
#pragma unroll
        for (int j = 0; j < 4; j++)
            if ((c >> j) & 1) s = s + v[4 * g + j];
        f[c] = s;
    }
    float mn = FLT_MAX, mx = -FLT_MAX;
#pragma unroll
    for (int c = 0; c < 16; c++) mn = fminf(mn, f[c]);
#pragma unroll
    for (int c = 0; c < 16; c++) mx = fmaxf(mx, f[c]);
    // max over groups is exact and order-free: fold(0, max) == max(0, wave max).
    float max_range = fmaxf(0.0f, wave_max(mx - mn));
    float scale = max_range > 0.0f ? 255.0f / max_range : 1.0f;
    uint32_t packed[4] = {0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < 16; c++) {
        float t = roundf((f[c] - mn) * scale);
        t = fminf(fmaxf(t, 0.0f), 255.0f);
        packed[c >> 2] |= ((uint32_t)t) << ((c & 3) * 8);
    }
    reinterpret_cast<uint4*>(out.tables + qi * 1024)[g] = make_uint4(packed[0], packed[1], packed[2], packed[3]);
    mins[g] = mn;
    __syncthreads();
    if (lane == 0) {
        float bias = -0.0f;
        for (int i = 0; i < GROUPS; i++) bias = bias + mins[i];
        float sum = -0.0f;
        for (int i = 0; i < PDIM; i++) sum = sum + v[i];
        float a = 2.0f / scale;
        float b = 2.0f * bias - sum;
        out.qab[qi] = make_float4(a, b, scale, bias);
    }

    // int8 rerank queries: w = q / scale (division), s = 127 / max(1e-12, max |w|).
    const float* q7 = q768 + qi * DIM;
    float w[DIM / WAVE];
    float m = 1e-12f;
#pragma unroll
    for (int j = 0; j < DIM / WAVE; j++) {
        int i = lane + j * WAVE;
        w[j] = q7[i] / quant[i];
        m = fmaxf(m, fabsf(w[j]));
    }
    m = wave_max(m);
    float s = 127.0f / m;
#pragma unroll
    for (int j = 0; j < DIM / WAVE; j++) out.wq768[qi * DIM + lane + j * WAVE] = sat_i8(roundf(w[j] * s));
    float p[PDIM / WAVE];
    float mp = 1e-12f;
#pragma unroll
    for (int j = 0; j < PDIM / WAVE; j++) {
        int i = lane + j * WAVE;
        p[j] = q[i] / quant[i];
        mp = fmaxf(mp, fabsf(p[j]));
    }
    mp = wave_max(mp);
    float sp = 127.0f / mp;
#pragma unroll
    for (int j = 0; j < PDIM / WAVE; j++) out.wq256[qi * PDIM + lane + j * WAVE] = sat_i8(roundf(p[j] * sp));
    if (lane == 0) out.inv[qi] = make_float2(1.0f / s, 1.0f / sp);
}

void launch_query_prep(const float* q256, const float* q768, size_t nq, const float* signs, float rot_norm,
                       const float* quant, const QueryPrep& out, hipStream_t st) {
    if (nq == 0) return;
    hipLaunchKernelGGL(k_query_prep, dim3((unsigned)nq), dim3(WAVE), 0, st, q256, q768, signs, rot_norm, quant, out);
}

// ------------------------------------------------------------------ re-layout --------
// blocks.bin (byte g*16+k = doc k nibble g | doc k+16 nibble g << 4) -> G/2 code bytes per
// slot, so a lane reads its slot's code with 16-byte loads: G = 32 codes raw (bit i of dim i at
// byte i/8, src/quant.rs:24-34), G = 64 codes lane-chunked (code_word_group). One workgroup of
// 4G threads per block; thread t writes code word t % (G / 8) of slot t / (G / 8).
template <int G>
__global__ __launch_bounds__(4 * G) void k_transpose_blocks(const uint8_t* __restrict__ blocks, size_t nblocks,
                                                            uint8_t* __restrict__ codes) {
    constexpr int BB = G * 16, W = G / 8;  // block bytes, code words per slot
    __shared__ uint8_t blk[BB];
    const size_t b = blockIdx.x;
    const int t = threadIdx.x;
    reinterpret_cast<uint32_t*>(blk)[t] = reinterpret_cast<const uint32_t*>(blocks + b * BB)[t];
    __syncthreads();
    const int k = t / W, d = t % W;
    uint32_t word = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) {  // nibble j of word d
        const int g = G == 64 ? code_word_group(d, j) : 8 * d + j;
        const uint8_t x = blk[g * 16 + (k & 15)];
        word |= (uint32_t)(k < 16 ? x & 0x0F : x >> 4) << (4 * j);
    }
    reinterpret_cast<uint32_t*>(codes + (b * BLOCK + k) * (G / 2))[d] = word;
}

void launch_transpose_blocks(const uint8_t* blocks, size_t nblocks, int groups, uint4* codes, hipStream_t st) {
    if (nblocks == 0) return;
    if (groups == 64)
        hipLaunchKernelGGL(k_transpose_blocks<64>, dim3((unsigned)nblocks), dim3(256), 0, st, blocks, nblocks,
                           reinterpret_cast<uint8_t*>(codes));
    else
        hipLaunchKernelGGL(k_transpose_blocks<32>, dim3((unsigned)nblocks), dim3(128), 0, st, blocks, nblocks,
                           reinterpret_cast<uint8_t*>(codes));
}

__global__ void k_split_factors(const float2* __restrict__ f, size_t n, float* __restrict__ mul,
                                float* __restrict__ err) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        float2 x = f[i];
        mul[i] = x.x;
        err[i] = x.y;
    }
}

void launch_split_factors(const float* factors, size_t nslots, float* mul, float* err, hipStream_t st) {
    if (nslots == 0) return;
    hipLaunchKernelGGL(k_split_factors, dim3((unsigned)((nslots + 255) / 256)), dim3(256), 0, st,
                       reinterpret_cast<const float2*>(factors), nslots, mul, err);
}

__global__ void k_to_bf16(const float* __restrict__ x, size_t n, uint16_t* __restrict__ y) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = f32_to_bf16_rne(x[i]);
}

void launch_to_bf16(const float* x, size_t n, uint16_t* y, hipStream_t st) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, x, n, y);
}

__global__ void k_cluster_slots(const uint32_t* __restrict__ block_off, uint32_t k, uint32_t* __restrict__ slots) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < k) slots[c] = (block_off[c + 1] - block_off[c]) * BLOCK;
}

void launch_cluster_slots(const uint32_t* block_off, uint32_t k, uint32_t* slots, hipStream_t st) {
    if (k == 0) return;
    hipLaunchKernelGGL(k_cluster_slots, dim3((k + 255) / 256), dim3(256), 0, st, block_off, k, slots);
}

__global__ void k_count_bad_ids(const uint32_t* __restrict__ ids, size_t nslots, uint64_t n,
                                unsigned long long* __restrict__ bad) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nslots) {
        uint32_t id = ids[i];
        if (id != PAD && (uint64_t)id >= n) atomicAdd(bad, 1ull);
    }
}

void launch_count_bad_ids(const uint32_t* ids, size_t nslots, uint64_t n, unsigned long long* bad, hipStream_t st) {
    if (nslots == 0) return;
    hipLaunchKernelGGL(k_count_bad_ids, dim3((unsigned)((nslots + 255) / 256)), dim3(256), 0, st, ids, nslots, n, bad);
}

}  // namespace sieve
