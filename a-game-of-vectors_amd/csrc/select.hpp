// Block-wide (256-thread) selection and sorting helpers shared by the scan and rerank
// kernels. All of them must be called by every thread of the block.
#pragma once

#include "common.hpp"

namespace sieve {

constexpr int SEL_THREADS = 256;

// Exclusive prefix sum of one value per thread; returns this thread's offset and writes
// the block total to *total. `scratch` holds 8 u32.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* scratch, uint32_t* total) {
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    uint32_t incl = wave_incl_scan(v);
    if (lane == 63) scratch[wave] = incl;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < SEL_THREADS / 64; w++) {
        uint32_t s = scratch[w];
        if (w < wave) off += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return off + incl - v;
}

// One 8-bit radix pass over uint2 entries (hi, lo): histogram of the digit at `shift` of the
// selected half among entries matching (prefix_hi, prefix_lo) under the masks; picks the digit
// that holds the kk-th largest. Returns (digit, kk within that digit, count of that digit).
struct PassResult {
    uint32_t digit, kk, count;
};

__device__ __forceinline__ PassResult radix_pass(const uint2* buf, int n, bool use_hi, int shift, uint32_t p_hi,
                                                 uint32_t m_hi, uint32_t p_lo, uint32_t m_lo, uint32_t kk,
                                                 uint32_t* hist, uint32_t* sh) {
    const int t = threadIdx.x;
    hist[t] = 0;
    __syncthreads();
    for (int i = t; i < n; i += SEL_THREADS) {
        uint2 e = buf[i];
        if ((e.x & m_hi) == p_hi && (e.y & m_lo) == p_lo) {
            uint32_t x = use_hi ? e.x : e.y;
            atomicAdd(&hist[(x >> shift) & 255u], 1u);
        }
    }
    __syncthreads();
    if (t < 64) {
        uint32_t c0 = hist[255 - 4 * t], c1 = hist[254 - 4 * t], c2 = hist[253 - 4 * t], c3 = hist[252 - 4 * t];
        uint32_t s = c0 + c1 + c2 + c3;
        uint32_t incl = wave_incl_scan(s);
        uint64_t hit = __ballot(incl >= kk);
        int f = __ffsll((long long)hit) - 1;
        if (t == f) {
            uint32_t cum = incl - s;
            uint32_t cs[4] = {c0, c1, c2, c3};
            for (int j = 0; j < 4; j++) {
                if (cum + cs[j] >= kk) {
                    sh[0] = 255 - 4 * t - j;
                    sh[1] = kk - cum;
                    sh[2] = cs[j];
                    break;
                }
                cum += cs[j];
            }
        }
    }
    __syncthreads();
    PassResult r{sh[0], sh[1], sh[2]};
    __syncthreads();
    return r;
}

// The k-th largest entry of buf[0, n) under the order (hi, lo), entries unique, 1 <= k <= n.
// Returns the (hi, lo) bound such that exactly k entries compare >= it. When every entry tied
// with the k-th on `hi` is needed, the lo half is not searched and the bound's lo is 0.
__device__ __forceinline__ uint2 block_kth_largest(const uint2* buf, int n, int k, uint32_t* hist, uint32_t* sh) {
    uint32_t p_hi = 0, m_hi = 0, p_lo = 0, m_lo = 0, kk = (uint32_t)k;
    PassResult r{0, 0, 0};
    for (int shift = 24; shift >= 0; shift -= 8) {
        r = radix_pass(buf, n, true, shift, p_hi, m_hi, p_lo, m_lo, kk, hist, sh);
        p_hi |= r.digit << shift;
        m_hi |= 0xFFu << shift;
        kk = r.kk;
    }
    if (r.count == kk) return make_uint2(p_hi, 0u);
    for (int shift = 24; shift >= 0; shift -= 8) {
        r = radix_pass(buf, n, false, shift, p_hi, m_hi, p_lo, m_lo, kk, hist, sh);
        p_lo |= r.digit << shift;
        m_lo |= 0xFFu << shift;
        kk = r.kk;
    }
    return make_uint2(p_hi, p_lo);
}

__device__ __forceinline__ bool ge2(uint2 a, uint2 b) { return a.x > b.x || (a.x == b.x && a.y >= b.y); }
__device__ __forceinline__ bool lt2(uint2 a, uint2 b) { return a.x < b.x || (a.x == b.x && a.y < b.y); }

// Bitonic sort of n2 (a power of two) entries with a u32 payload; ascending if `asc`.
__device__ __forceinline__ void block_sort_pairs(uint2* key, uint32_t* pay, int n2, bool asc) {
    for (int size = 2; size <= n2; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = threadIdx.x; i < n2 / 2; i += SEL_THREADS) {
                int lo = 2 * i - (i & (stride - 1));
                int hi = lo + stride;
                bool up = ((lo & size) == 0) == asc;
                uint2 x = key[lo], y = key[hi];
                bool swap = up ? lt2(y, x) : lt2(x, y);
                if (swap) {
                    key[lo] = y;
                    key[hi] = x;
                    if (pay) {
                        uint32_t px = pay[lo];
                        pay[lo] = pay[hi];
                        pay[hi] = px;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// The best min(k, n) of n unique entries buf[0, n) (n <= 2048) under the order (hi, lo), best
// first, into out: k rounds of a block-wide maximum over the entries the threads hold in
// registers (i = t + 256 j), one barrier per round. For small k it replaces a radix select or
// a full sort. `red` holds 8 u64 (two rounds' per-wave maxima). An entry (0, 0) cannot occur
// (a key of 0 is the bit pattern of a NaN).
__device__ __forceinline__ int block_topk_small(const uint2* buf, int n, int k, uint2* out, unsigned long long* red) {
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    unsigned long long v[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const int i = t + SEL_THREADS * j;
        v[j] = i < n ? (((unsigned long long)buf[i].x << 32) | buf[i].y) : 0ull;
    }
    const int m = min(k, n);
    for (int r = 0; r < m; r++) {
        unsigned long long b = 0;
#pragma unroll
        for (int j = 0; j < 8; j++) b = v[j] > b ? v[j] : b;
        b = wave_max_u64(b);
        unsigned long long* slot = red + 4 * (r & 1);
        if (lane == 0) slot[wave] = b;
        __syncthreads();  // the other parity's slots were read before the previous barrier
        b = slot[0];
#pragma unroll
        for (int w = 1; w < SEL_THREADS / 64; w++) b = slot[w] > b ? slot[w] : b;
#pragma unroll
        for (int j = 0; j < 8; j++)
            if (v[j] == b) v[j] = 0;
        if (t == 0) out[r] = make_uint2((uint32_t)(b >> 32), (uint32_t)b);
    }
    __syncthreads();
    return m;
}

__device__ __forceinline__ int pow2_at_least(int n) {
    int p = 1;
    while (p < n) p <<= 1;
    return p;
}

}  // namespace sieve
