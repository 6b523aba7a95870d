// Shared definitions of the sieve MI355X library (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/sieve_gpu.h"

namespace sieve {

constexpr int DIM = 768;
constexpr int PDIM = 256;
constexpr int GROUPS = 64;           // 4-dim groups of the 256-dim prefix
constexpr int BLOCK = 32;            // slots per code block
constexpr int BLOCK_BYTES = 1024;    // GROUPS * 16
constexpr uint32_t PAD = 0xFFFFFFFFu;
constexpr int WAVE = 64;

// Largest candidate list a query may keep (p->cand) and largest probe count (p->nprobe):
// both bound LDS arrays of the per-query kernels.
constexpr int MAX_CAND = 2048;
constexpr int MAX_PROBE = 1024;
// Centroids within the BF16 error band of the nprobe-th routing score that one query may
// rescore in f32.
constexpr int MAX_RESCORE = 4096;
// Phases of the cluster-major scan whose MFMA launches are timed with HIP events.
constexpr uint32_t CM_TIMED_PHASES = 8;

// ------------------------------------------------------------------ errors (host) ---
void set_error(const std::string& msg);

struct Error {
    int code;
    std::string msg;
};

#define SIEVE_HIP(expr)                                                                    \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            throw ::sieve::Error{SIEVE_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)}; \
    } while (0)

#define SIEVE_REQUIRE(cond, code, msg)                                                     \
    do {                                                                                   \
        if (!(cond)) throw ::sieve::Error{(code), (msg)};                                  \
    } while (0)

// ------------------------------------------------------------------ device helpers ---
// Order-preserving map of an f32 to u32: a > b  <=>  key(a) > key(b) (no NaNs reach here;
// -0.0 maps below +0.0, which only orders values that compare equal).
__device__ __forceinline__ uint32_t f2key(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key2f(uint32_t k) {
    uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
    return __uint_as_float(u);
}

__device__ __forceinline__ bool bit_set(const uint64_t* set, uint64_t words, uint32_t id) {
    uint64_t w = id >> 6;
    return w < words && ((set[w] >> (id & 63)) & 1ull);
}

// Rust `f32 as i8`: saturating.
__device__ __forceinline__ int8_t sat_i8(float v) {
    if (v != v) return 0;
    if (v >= 127.0f) return 127;
    if (v <= -128.0f) return -128;
    return (int8_t)(int)v;
}

__device__ __forceinline__ int wave_lane() { return __lane_id(); }

// Inclusive prefix sum over the 64 lanes of a wave.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const int lane = __lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    return v;
}

// Sum over aligned groups of G lanes (G = 2, 4, 8 or 16) with DPP row operations, in registers
// (a shuffle goes through LDS): the quad permutes pair lane l with l ^ 1 and l ^ 2, the half-row
// mirror (lane i with 7 - i of its 8) the two quads of a half row, the row mirror (i with
// 15 - i of its 16) the two halves of a row. Every lane of a group gets the group's sum; all
// lanes must be active.
template <int G>
__device__ __forceinline__ int group_sum_dpp(int v) {
    static_assert(G == 2 || G == 4 || G == 8 || G == 16, "groups of 2 to 16 lanes");
    v += __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);  // quad_perm [1, 0, 3, 2]
    if constexpr (G >= 4) v += __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);  // quad_perm [2, 3, 0, 1]
    if constexpr (G >= 8) v += __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, false);  // row_half_mirror
    if constexpr (G >= 16) v += __builtin_amdgcn_update_dpp(0, v, 0x140, 0xF, 0xF, false);  // row_mirror
    return v;
}
template <int CTRL>
__device__ __forceinline__ float dpp_f32(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xF, 0xF, false));
}
template <int G>
__device__ __forceinline__ float group_sum_dpp(float v) {
    static_assert(G == 2 || G == 4 || G == 8 || G == 16, "groups of 2 to 16 lanes");
    v += dpp_f32<0xB1>(v);
    if constexpr (G >= 4) v += dpp_f32<0x4E>(v);
    if constexpr (G >= 8) v += dpp_f32<0x141>(v);
    if constexpr (G >= 16) v += dpp_f32<0x140>(v);
    return v;
}

// The maximum of a u64 over the wave, in registers: DPP quad permutes and the half-row and row
// mirrors give every lane its row's maximum, v_permlane16_swap pairs each row with another (of
// its two results one holds the lane's own value and one its partner's, in both halves alike)
// and v_permlane32_swap the two halves of the wave. All lanes must be active.
template <int CTRL>
__device__ __forceinline__ unsigned long long dpp_u64(unsigned long long x) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)x, CTRL, 0xF, 0xF, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(x >> 32), CTRL, 0xF, 0xF, false);
    return ((unsigned long long)hi << 32) | lo;
}
__device__ __forceinline__ unsigned long long max_u64(unsigned long long a, unsigned long long b) {
    return a > b ? a : b;
}
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long b) {
    b = max_u64(b, dpp_u64<0xB1>(b));   // quad_perm [1, 0, 3, 2]
    b = max_u64(b, dpp_u64<0x4E>(b));   // quad_perm [2, 3, 0, 1]
    b = max_u64(b, dpp_u64<0x141>(b));  // row_half_mirror
    b = max_u64(b, dpp_u64<0x140>(b));  // row_mirror
    {
        const uint32_t lo = (uint32_t)b, hi = (uint32_t)(b >> 32);
        const auto l = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
        const auto h = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
        b = max_u64(((unsigned long long)h[0] << 32) | l[0], ((unsigned long long)h[1] << 32) | l[1]);
    }
    {
        const uint32_t lo = (uint32_t)b, hi = (uint32_t)(b >> 32);
        const auto l = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
        const auto h = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
        b = max_u64(((unsigned long long)h[0] << 32) | l[0], ((unsigned long long)h[1] << 32) | l[1]);
    }
    return b;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = fmaxf(v, __shfl_xor(v, d, 64));
    return v;
}

// Slot codes of 64 groups (256-bit RaBitQ codes) are stored lane-chunked: 8-byte chunk r
// (r = 0..3) holds the nibbles of groups 4 st + r, nibble st at bits 4 st (st = 0..15). A
// lane that scores groups 4 st + r for one slot (the int8 MFMA operand layout) reads 8
// contiguous bytes. code_word_group(d, j): the group of nibble j of 32-bit code word d.
__host__ __device__ constexpr int code_word_group(int d, int j) { return 4 * (8 * (d & 1) + j) + (d >> 1); }

// Raw code (bit i = sign of dimension i, group g = bits 4g .. 4g + 3) <-> lane-chunked code.
__host__ __device__ inline void code_to_chunked(const uint32_t raw[8], uint32_t out[8]) {
    for (int d = 0; d < 8; d++) {
        uint32_t w = 0;
        for (int j = 0; j < 8; j++) {
            const int g = code_word_group(d, j);
            w |= ((raw[g >> 3] >> (4 * (g & 7))) & 15u) << (4 * j);
        }
        out[d] = w;
    }
}
__host__ __device__ inline void code_from_chunked(const uint32_t ch[8], uint32_t raw[8]) {
    for (int d = 0; d < 8; d++) raw[d] = 0;
    for (int d = 0; d < 8; d++)
        for (int j = 0; j < 8; j++) {
            const int g = code_word_group(d, j);
            raw[g >> 3] |= ((ch[d] >> (4 * j)) & 15u) << (4 * (g & 7));
        }
}

__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        const unsigned long long o = __shfl_xor(v, d, 64);
        v = o < v ? o : v;
    }
    return v;
}

}  // namespace sieve
