// Centroid routing: Index::route (src/ivf.rs:587-613) over a batch of queries.
//
// 1. k_gemm_bf16_route: S = Q * C^T with v_mfma_f32_16x16x32_bf16 (BF16 inputs, f32
//    accumulation) for every (query, centroid), and the maximum of every 16 consecutive
//    scores of a query. Fused (route_fuse: few probes over many centroids) it writes only the
//    maxima, and the select recomputes the scores of the tiles it reads from the same BF16
//    operands (BF16 x BF16 products are exact in f32, so any summation order keeps the bound
//    below).
// 2. k_route_select, one workgroup per query: the nprobe-th largest approximate score (a
//    select over the tile maxima, then over the scores of the tiles that can hold the top
//    nprobe); every centroid whose approximate score lies within twice the BF16
//    error bound of it is rescored in f32 exactly as the reference's gemm_abt fallback sums
//    (src/blas.rs:47-53: products left to right from -0.0, no FMA); the rescored scores are
//    sorted (score desc, id asc) as TopK::into_sorted (src/topk.rs:47-66); then the margin
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

#include <algorithm>
#include <type_traits>

#include "kernels.hpp"
#include "select.hpp"

namespace sieve {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------------ BF16 routing GEMM ---
// The maximum of x over lanes l, l ^ 16, l ^ 32 and l ^ 48 (the four 16-lane rows of a wave):
// v_permlane32_swap pairs each lane with lane l ^ 32 (one of its two results holds each), then
// v_permlane16_swap pairs an odd row with an even one. Every lane gets the maximum.
__device__ __forceinline__ float max_over_rows(float x) {
    const auto a = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(x), false, false);
    const float m = fmaxf(__uint_as_float(a[0]), __uint_as_float(a[1]));
    const auto b = __builtin_amdgcn_permlane16_swap(__float_as_uint(m), __float_as_uint(m), false, false);
    return fmaxf(__uint_as_float(b[0]), __uint_as_float(b[1]));
}

// S = Q C^T for every (query, centroid) with v_mfma_f32_16x16x32_bf16, K = 256. Each wave
// owns 64 centroids, whose BF16 rows stay in registers (4 MFMA tiles x 8 k-steps, 128 VGPRs),
// and walks the 16-query tiles of its workgroup's query range (staged in LDS, below): centroid
// rows are read once per workgroup, query tiles once per workgroup and tile. The centroids are the MFMA's first operand, so lane
// l holds centroids 4 (l / 16) + v, v < 4, of tile j for query l % 16 (one float4 store).
// tmax (row stride ceil(n / 16)) receives the largest approximate route score of each 16
// consecutive centroids: S itself, or the L2 route score 2 S - |c|^2 (L2). With S null only the
// maxima are written (the select recomputes the scores it needs): the GEMM is then no longer
// bound by writing 4 bytes per (query, centroid).
template <bool L2>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_gemm_bf16_route(
    const uint16_t* __restrict__ Q, const uint16_t* __restrict__ C, const float* __restrict__ cnorm,
    float* __restrict__ S, float* __restrict__ tmax, int m, int n, size_t lds, int qtiles_per) {
    // The workgroup's 16-query tiles pass through LDS, two buffers taking turns: each tile is
    // read from memory once per workgroup (each thread two 16-byte chunks, two tiles ahead)
    // and each wave reads its operand from LDS. A row's 32 chunks are stored XOR-swizzled by
    // the row (chunk c of row r at r * 32 + (c ^ r)): the 16 rows a fragment read touches at
    // one chunk index fall in 16 distinct bank groups.
    __shared__ uint4 qtile[2][16 * 32];
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const int cb = blockIdx.x * 256 + wave * 64;
    const int qtiles = (m + 15) / 16;
    const int qt0 = blockIdx.y * qtiles_per, qt1 = min(qt0 + qtiles_per, qtiles);
    if (qt0 >= qt1) return;  // workgroup-uniform
    const bool active = cb < n;  // wave-uniform: a wave past the last centroid only stages tiles
    const int r16 = lane & 15, ko = (lane >> 4) * 8, c4 = (lane >> 4) * 4;
    const int ntile = (n + 15) / 16;
    bf16x8 bc[4][8];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const uint16_t* row = C + (size_t)min(cb + j * 16 + r16, n - 1) * PDIM + ko;
#pragma unroll
        for (int s = 0; s < 8; s++) bc[j][s] = *reinterpret_cast<const bf16x8*>(row + 32 * s);
    }
    float cn[4][4];
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int c = cb + j * 16 + c4 + v;
            cn[j][v] = (L2 && c < n) ? cnorm[c] : 0.0f;
        }
    const bool vec = (lds & 3) == 0;
    // thread t stages chunk t & 31 of tile rows t >> 5 and (t >> 5) + 8 (rows past m repeat
    // row m - 1; tiles past the range repeat the last one)
    const int sr = t >> 5, sc = t & 31;
    auto gload = [&](int qt, uint4& x, uint4& y) {
        const uint4* base = reinterpret_cast<const uint4*>(Q);
        const int q = min(qt, qt1 - 1) * 16;
        x = base[(size_t)min(q + sr, m - 1) * 32 + sc];
        y = base[(size_t)min(q + sr + 8, m - 1) * 32 + sc];
    };
    auto lstore = [&](int b, const uint4& x, const uint4& y) {
        qtile[b][sr * 32 + (sc ^ sr)] = x;
        qtile[b][(sr + 8) * 32 + (sc ^ (sr + 8))] = y;
    };
    auto lload = [&](bf16x8(&dst)[8], int b) {
#pragma unroll
        for (int s = 0; s < 8; s++) {
            const uint4 v = qtile[b][r16 * 32 + (((ko >> 3) + 4 * s) ^ r16)];
            dst[s] = __builtin_bit_cast(bf16x8, v);
        }
    };
    // One 16-query tile: 32 MFMAs, then S and the tile maxima. FULL: every query and centroid
    // of the tile exists and rows are float4-aligned, so every store is unconditional (all
    // four lanes of a query store its same tile maximum), which keeps the memory waits exact.
    auto tile = [&](auto full, const bf16x8(&cur)[8], int qt) {
        constexpr bool FULL = decltype(full)::value;
        f32x4 acc[4];
#pragma unroll
        for (int j = 0; j < 4; j++) acc[j] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < 8; s++)
#pragma unroll
            for (int j = 0; j < 4; j++) acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bc[j][s], cur[s], acc[j], 0, 0, 0);
        const int q = qt * 16 + r16;
        float mxs[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int c0 = cb + j * 16 + c4;
            float* dst = S + (size_t)q * lds + c0;
            float mx = -INFINITY;
            if constexpr (FULL) {
                if (S) *reinterpret_cast<f32x4*>(dst) = acc[j];
#pragma unroll
                for (int v = 0; v < 4; v++) mx = fmaxf(mx, L2 ? 2.0f * acc[j][v] - cn[j][v] : acc[j][v]);
            } else {
                if (S && q < m) {
                    if (vec && c0 + 3 < n) {
                        *reinterpret_cast<f32x4*>(dst) = acc[j];
                    } else {
#pragma unroll
                        for (int v = 0; v < 4; v++)
                            if (c0 + v < n) dst[v] = acc[j][v];
                    }
                }
#pragma unroll
                for (int v = 0; v < 4; v++)
                    if (c0 + v < n) mx = fmaxf(mx, L2 ? 2.0f * acc[j][v] - cn[j][v] : acc[j][v]);
            }
            mxs[j] = mx;
        }
        // The four lanes of a query (rows of 16 lanes) hold a quarter of each tile: their
        // maximum in registers (permlane swaps, no LDS round trip), four tiles side by side.
#pragma unroll
        for (int j = 0; j < 4; j++) mxs[j] = max_over_rows(mxs[j]);
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int tl = (cb >> 4) + j;
            if (FULL || (lane < 16 && q < m && tl < ntile)) tmax[(size_t)q * ntile + tl] = mxs[j];
        }
    };
    const int qfull = (vec && cb + 64 <= n) ? min(qt1, m / 16) : qt0;  // tiles before qfull are FULL
    uint4 xa, ya, xb, yb;
    gload(qt0, xa, ya);
    lstore(0, xa, ya);
    gload(qt0 + 1, xa, ya);
    __syncthreads();
    int b = 0;
    for (int qt = qt0; qt < qt1; qt++, b ^= 1) {
        gload(qt + 2, xb, yb);
        if (active) {
            bf16x8 cur[8];
            lload(cur, b);
            if (qt < qfull)
                tile(std::true_type{}, cur, qt);
            else
                tile(std::false_type{}, cur, qt);
        }
        // buffer b ^ 1 was last read in the previous tile, before the barrier that ended it
        lstore(b ^ 1, xa, ya);
        xa = xb;
        ya = yb;
        __syncthreads();
    }
}

void launch_gemm_bf16_route(const uint16_t* qbf, const uint16_t* cbf, const float* cnorm, float* s, float* tmax,
                            int m, int n, size_t lds, hipStream_t st) {
    if (m == 0 || n == 0) return;
    const int gx = (n + 255) / 256;
    const int qtiles = (m + 15) / 16;
    // about 1024 workgroups: two rounds of two per CU
    int gy = std::max(1, std::min(qtiles, (1024 + gx - 1) / gx));
    const int per = (qtiles + gy - 1) / gy;
    gy = (qtiles + per - 1) / per;
    if (cnorm)
        hipLaunchKernelGGL(k_gemm_bf16_route<true>, dim3(gx, gy), dim3(256), 0, st, qbf, cbf, cnorm, s, tmax, m, n,
                           lds, per);
    else
        hipLaunchKernelGGL(k_gemm_bf16_route<false>, dim3(gx, gy), dim3(256), 0, st, qbf, cbf, cnorm, s, tmax, m,
                           n, lds, per);
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
        for (int j = 0; j < VALS; j++) {
            // Scores near the top share their high digits: the lanes that hold the first active
            // lane's digit add once for all of them.
            const bool in = t + 256 * j < n && (x[j] & mask) == prefix;
            const uint32_t d = (x[j] >> shift) & 255;
            const uint64_t act = __ballot(in);
            if (act) {
                const int first = __ffsll((long long)act) - 1;
                const uint32_t d0 = __shfl(d, first, 64);
                const uint64_t same = __ballot(in && d == d0);
                const bool lead = (t & 63) == first;
                if (in && (lead || d != d0)) atomicAdd(&hist[d], lead ? (uint32_t)__popcll(same) : 1u);
            }
        }
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

// Bitonic sort of one key per lane across a wave, descending: lane 0 gets the largest.
__device__ __forceinline__ uint32_t wave_sort_desc(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 2; k <= 64; k <<= 1)
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            const uint32_t o = __shfl_xor(v, j, 64);
            // the lower lane of a pair keeps the larger key in the blocks that end descending
            const bool keep_max = ((lane & j) == 0) == ((lane & k) == 0);
            v = keep_max ? max(v, o) : min(v, o);
        }
    return v;
}

// block_kth_regs for k <= 64 without a radix pass over every key. Each wave sorts its lanes'
// largest keys; a wave's k-th largest lane key L_w has k keys >= it, so the largest L_w, L,
// is at most the k-th largest key. The keys >= L (about k to 4k of them) are gathered into
// buf and their k-th largest is the answer: one more wave sort when they fit a wave, else
// the radix over buf; more than `cap` of them (equal keys), the radix over key(i) (the same
// keys, read again). k <= 64, 1 <= k <= n. ws: 6 words of LDS. Ends with a barrier. The
// callers take it for k <= 16: at larger k the bound is loose (at k = 64 it is the smallest lane
// maximum of a wave), hundreds of keys pass it and the radix over registers is faster.
template <int VALS, class KeyFn>
__device__ uint32_t block_kth_regs_small(const uint32_t (&x)[VALS], int n, int k, KeyFn key, uint2* buf, int cap,
                                         uint32_t* ws, uint32_t* hist, uint32_t* sh) {
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    uint32_t mk = 0;
#pragma unroll
    for (int j = 0; j < VALS; j++)
        if (t + 256 * j < n) mk = max(mk, x[j]);
    mk = wave_sort_desc(mk);
    const uint32_t lw = __shfl(mk, k - 1, 64);
    if (lane == 0) ws[1 + wave] = lw;
    if (t == 0) ws[0] = 0;
    __syncthreads();
    const uint32_t L = max(max(ws[1], ws[2]), max(ws[3], ws[4]));
#pragma unroll
    for (int j = 0; j < VALS; j++) {
        const bool in = t + 256 * j < n && x[j] >= L;
        const uint64_t b = __ballot(in);
        if (b) {  // wave-uniform
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(&ws[0], (uint32_t)__popcll(b));
            base = (uint32_t)__builtin_amdgcn_readlane((int)base, 0) + (uint32_t)__popcll(b & ((1ull << lane) - 1));
            if (in && base < (uint32_t)cap) buf[base] = make_uint2(x[j], ~(uint32_t)(t + 256 * j));
        }
    }
    __syncthreads();
    const uint32_t got = ws[0];  // >= k unless keys are 0 (a NaN's)
    uint32_t res;
    if (got > (uint32_t)cap || got < (uint32_t)k) {
        res = block_kth_u32(n, k, key, hist, sh);
    } else if (got <= 64) {
        if (wave == 0) {
            const uint32_t v = wave_sort_desc(lane < (int)got ? buf[lane].x : 0u);
            if (lane == k - 1) ws[5] = v;
        }
        __syncthreads();
        res = ws[5];
    } else {
        res = block_kth_largest(buf, (int)got, k, hist, sh).x;
    }
    __syncthreads();
    return res;
}

// Route select modes. ROUTE_REGS: the approximate scores (k <= 256 VALS) are read once into
// registers and the select and the band pass run over them. ROUTE_MEM: over memory (any k).
// ROUTE_TILES: over the GEMM's maxima of 16 consecutive scores (ceil(k / 16) <= 256 VALS,
// at least nprobe of them), then only the tiles that can hold the top nprobe or the band.
enum { ROUTE_REGS = 0, ROUTE_MEM = 1, ROUTE_TILES = 2 };

template <int MODE, int VALS>
__global__ __launch_bounds__(256) void k_route_select(RouteArgs a) {
    __shared__ float qs[PDIM];
    __shared__ uint32_t hist[256];
    __shared__ uint32_t sh[4];
    __shared__ float red[4];
    __shared__ uint32_t ncand, ntl;
    __shared__ uint32_t ws[6];
    __shared__ uint32_t cids[MAX_RESCORE];
    __shared__ uint2 sorted[MAX_RESCORE];
    __shared__ __align__(16) float qb[PDIM];
    const size_t qi = blockIdx.x;
    const int t = threadIdx.x;
    const float* S = a.approx + qi * a.lds;
    const int n = min(a.nprobe, a.k);
    if (a.l2) {
        for (int i = t; i < PDIM; i += 256) qs[i] = i < a.dims ? a.qf[qi * a.dims + i] : 0.0f;
    } else {
        for (int i = t; i < PDIM; i += 256) qs[i] = a.q256[qi * PDIM + i];
    }
    if (a.fused)
        for (int i = t; i < PDIM; i += 256) qb[i] = __uint_as_float((uint32_t)a.qbf[qi * PDIM + i] << 16);
    if (t == 0) {
        ncand = 0;
        ntl = 0;
    }
    __syncthreads();
    if (n <= 0) {
        if (t == 0) a.out_n[qi] = 0;
        return;
    }
    // approximate route score of centroid i: the GEMM's, or (fused) recomputed from its BF16
    // operands. Products of BF16 values are exact in f32, so the recomputed sum, in any order,
    // is within the same error bound of the exact score as the GEMM's.
    auto approx = [&](int i) -> float {
        float s;
        if (!a.fused) {
            s = S[i];
        } else {
            const uint4* row = reinterpret_cast<const uint4*>(a.cbf + (size_t)i * PDIM);
            const float4* qb4 = reinterpret_cast<const float4*>(qb);
            s = 0.0f;
            for (int j = 0; j < PDIM / 8; j++) {
                const uint4 w = row[j];
                const float4 x0 = qb4[2 * j], x1 = qb4[2 * j + 1];
                s = s + x0.x * __uint_as_float(w.x << 16);
                s = s + x0.y * __uint_as_float(w.x & 0xFFFF0000u);
                s = s + x0.z * __uint_as_float(w.y << 16);
                s = s + x0.w * __uint_as_float(w.y & 0xFFFF0000u);
                s = s + x1.x * __uint_as_float(w.z << 16);
                s = s + x1.y * __uint_as_float(w.z & 0xFFFF0000u);
                s = s + x1.z * __uint_as_float(w.w << 16);
                s = s + x1.w * __uint_as_float(w.w & 0xFFFF0000u);
            }
        }
        return a.l2 ? 2.0f * s - a.cnorm[i] : s;
    };
    auto key = [&](int i) { return f2key(approx(i) + 0.0f); };
    const float e = a.err_coef * a.qnorm[qi];
    // The band of the n-th largest approximate score kv: every centroid whose exact score can
    // reach the exact top n.
    auto band_of = [&](float kv) { return kv - 2.2f * e - 1e-6f * (1.0f + fabsf(kv)); };
    // 8-bit radix over memory, then a band pass: any k.
    auto radix_over_memory = [&]() {
        const uint32_t kth = block_kth_u32(a.k, n, key, hist, sh);
        const float thr = band_of(key2f(kth));
        if (t == 0) ncand = 0;
        __syncthreads();
        for (int i = t; i < a.k; i += 256) {
            if (approx(i) >= thr) {
                const uint32_t j = atomicAdd(&ncand, 1u);
                if (j < MAX_RESCORE) cids[j] = (uint32_t)i;
            }
        }
    };
    // `got` candidates (key, ~id) in `sorted` hold the top n and every score >= the band's
    // floor: the n-th largest and the band from LDS into cids.
    auto band_from_sorted = [&](uint32_t got) {
        const float thr = band_of(key2f(block_kth_largest(sorted, (int)got, n, hist, sh).x));
        if (t == 0) ncand = 0;
        __syncthreads();
        for (uint32_t j = t; j < got; j += 256)
            if (key2f(sorted[j].x) >= thr) cids[atomicAdd(&ncand, 1u)] = ~sorted[j].y;
    };
    if constexpr (MODE == ROUTE_REGS) {
        uint32_t x[VALS];
#pragma unroll
        for (int j = 0; j < VALS; j++) {
            const int i = t + 256 * j;
            x[j] = i < a.k ? key(i) : 0u;
        }
        const float thr = band_of(key2f(n <= 16 ? block_kth_regs_small<VALS>(x, a.k, n, key, sorted, MAX_RESCORE, ws, hist, sh)
                                                : block_kth_regs<VALS>(x, a.k, n, hist, sh)));
#pragma unroll
        for (int j = 0; j < VALS; j++) {
            const int i = t + 256 * j;
            // key2f(key(v)) == v + 0.0f, which compares with thr as v does
            if (i < a.k && key2f(x[j]) >= thr) {
                uint32_t jj = atomicAdd(&ncand, 1u);
                if (jj < MAX_RESCORE) cids[jj] = (uint32_t)i;
            }
        }
    } else if constexpr (MODE == ROUTE_TILES) {
        // T0, the n-th largest tile maximum, is at most the n-th largest score (n tiles hold a
        // score >= T0), and the band's floor over kth in [T0, max] is >= lim. Every score the
        // band or the top n needs is >= lim, and so is its tile's maximum: the tiles whose
        // maximum is >= lim hold all of them.
        const int nt = (a.k + 15) >> 4;
        const float* T = a.tmax + qi * (size_t)nt;
        uint32_t tk[VALS];
        float mx = -INFINITY;
#pragma unroll
        for (int j = 0; j < VALS; j++) {
            const int i = t + 256 * j;
            const float v = i < nt ? T[i] : -INFINITY;
            tk[j] = i < nt ? f2key(v + 0.0f) : 0u;
            mx = fmaxf(mx, v);
        }
        mx = wave_max(mx);
        if ((t & 63) == 0) red[t >> 6] = mx;
        // ends with a barrier
        auto tkey = [&](int i) { return f2key(T[i] + 0.0f); };
        const float t0 = key2f(n <= 16 ? block_kth_regs_small<VALS>(tk, nt, n, tkey, sorted, MAX_RESCORE, ws, hist, sh)
                                       : block_kth_regs<VALS>(tk, nt, n, hist, sh));
        const float top = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
        const float lim = t0 - 2.2f * e - 1e-6f * (1.0f + fmaxf(fabsf(t0), fabsf(top)));
#pragma unroll
        for (int j = 0; j < VALS; j++) {
            const int i = t + 256 * j;
            if (i < nt && key2f(tk[j]) >= lim) {
                const uint32_t p = atomicAdd(&ntl, 1u);
                if (p < MAX_RESCORE) cids[p] = (uint32_t)i;
            }
        }
        __syncthreads();
        const uint32_t nl = ntl;
        if (nl <= (uint32_t)MAX_RESCORE && a.fused) {
            // The collected tiles' scores from the BF16 operands, 8 lanes per centroid: lane l8
            // takes the 16-byte pieces l8, l8 + 8, l8 + 16, l8 + 24 of the 512-byte row, so each
            // load of a lane group reads one whole 128-byte line.
            const int g = t >> 3, l8 = t & 7;
            const float4* qb4 = reinterpret_cast<const float4*>(qb);
            for (uint32_t f = g; f < nl * 16; f += 32) {
                const int c = (int)cids[f >> 4] * 16 + (int)(f & 15);
                if (c >= a.k) continue;  // the whole lane group
                const uint4* row = reinterpret_cast<const uint4*>(a.cbf + (size_t)c * PDIM);
                float s = 0.0f;
#pragma unroll
                for (int m = 0; m < 4; m++) {
                    const int j = l8 + 8 * m;
                    const uint4 w = row[j];
                    const float4 x0 = qb4[2 * j], x1 = qb4[2 * j + 1];
                    s = s + x0.x * __uint_as_float(w.x << 16);
                    s = s + x0.y * __uint_as_float(w.x & 0xFFFF0000u);
                    s = s + x0.z * __uint_as_float(w.y << 16);
                    s = s + x0.w * __uint_as_float(w.y & 0xFFFF0000u);
                    s = s + x1.x * __uint_as_float(w.z << 16);
                    s = s + x1.y * __uint_as_float(w.z & 0xFFFF0000u);
                    s = s + x1.z * __uint_as_float(w.w << 16);
                    s = s + x1.w * __uint_as_float(w.w & 0xFFFF0000u);
                }
                s = group_sum_dpp<8>(s);
                const float v = a.l2 ? 2.0f * s - a.cnorm[c] : s;
                if (l8 == 0 && v >= lim) {
                    const uint32_t j = atomicAdd(&ncand, 1u);
                    if (j < (uint32_t)MAX_RESCORE) sorted[j] = make_uint2(f2key(v + 0.0f), ~(uint32_t)c);
                }
            }
        } else if (nl <= (uint32_t)MAX_RESCORE) {
            for (uint32_t f = t; f < nl * 16; f += 256) {
                const int c = (int)cids[f >> 4] * 16 + (int)(f & 15);
                if (c < a.k) {
                    const float v = approx(c);
                    if (v >= lim) {
                        const uint32_t j = atomicAdd(&ncand, 1u);
                        if (j < (uint32_t)MAX_RESCORE) sorted[j] = make_uint2(f2key(v + 0.0f), ~(uint32_t)c);
                    }
                }
            }
        }
        __syncthreads();
        const uint32_t got = ncand;
        __syncthreads();
        if (nl <= (uint32_t)MAX_RESCORE && got <= (uint32_t)MAX_RESCORE && got >= (uint32_t)n)
            band_from_sorted(got);
        else
            radix_over_memory();
    } else {
        // Large K: a 4096-bin histogram of the keys' top 12 bits (in cids) finds the bin of the
        // n-th largest score. Every score the band keeps (>= kth - 2.2 e - 1e-6 (1 + |kth|)) is
        // at least that bin's lower edge minus the same margin, so one more pass collects a
        // superset of the top n and the band into `sorted`; the n-th and the band then come from
        // LDS: two passes over the K scores. A collection that does not fit falls back to the
        // 8-bit radix over memory and a band pass.
        static_assert(MAX_RESCORE == 4096, "the histogram lives in cids");
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
        __syncthreads();
        const uint32_t bin = sh[0], upto = sh[1];
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
        if (got <= (uint32_t)MAX_RESCORE && got >= upto)
            band_from_sorted(got);
        else
            radix_over_memory();
    }
    __syncthreads();
    if (ncand > (uint32_t)MAX_RESCORE) {
        // More centroids in the band than the rescoring buffer holds (a degenerate query: a zero
        // q256 ties every score): this query alone takes the exact route. Its row of S gets the
        // exact f32 score of every centroid (summed as below), and the top n by (score desc, id
        // asc) are selected from it directly: the n-th largest score, then among the scores equal
        // to it the smallest ids. The result is the one the rescoring path would give.
        __syncthreads();
        float* Sw = a.approx + qi * a.lds;
        for (int i = t; i < a.k; i += 256) {
            const float4* row = reinterpret_cast<const float4*>(a.cents + (size_t)i * a.cstride);
            float acc = -0.0f;
            for (int j = 0; j < a.dims / 4; j++) {
                float4 x = row[j];
                acc = acc + qs[4 * j] * x.x;
                acc = acc + qs[4 * j + 1] * x.y;
                acc = acc + qs[4 * j + 2] * x.z;
                acc = acc + qs[4 * j + 3] * x.w;
            }
            Sw[i] = acc;
        }
        __syncthreads();
        // keys of the exact route scores in the row
        auto skey = [&](int i) { return f2key((a.l2 ? 2.0f * Sw[i] - a.cnorm[i] : Sw[i]) + 0.0f); };
        const uint32_t kth = block_kth_u32(a.k, n, skey, hist, sh);
        if (t == 0) {
            ncand = 0;
            ntl = 0;
        }
        __syncthreads();
        for (int i = t; i < a.k; i += 256)
            if (skey(i) > kth) atomicAdd(&ntl, 1u);
        __syncthreads();
        const int need = n - (int)ntl;  // >= 1: at least n keys are >= kth
        const uint32_t idk = block_kth_u32(a.k, need, [&](int i) { return skey(i) == kth ? ~(uint32_t)i : 0u; }, hist, sh);
        for (int i = t; i < a.k; i += 256) {
            const uint32_t x = skey(i);
            if (x > kth || (x == kth && ~(uint32_t)i >= idk)) cids[atomicAdd(&ncand, 1u)] = (uint32_t)i;
        }
        __syncthreads();
    }
    const int nc = (int)ncand;
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

bool route_select_over_tiles(int k, int nprobe) {
    const int nt = (k + 15) / 16;
    return nt >= std::min(nprobe, k) && nt <= 256 * 64;
}

// Unfused, a query's scores cost 4 k bytes written (and the collected tiles' 64 bytes read);
// fused, the select reads the BF16 rows of about 1.5 nprobe collected tiles, 16 x 512 bytes
// each, mostly from cache: 12 KiB per probe against 4 k bytes.
bool route_fuse(int mode, int k, int nprobe) {
    if (mode == 0 || !route_select_over_tiles(k, nprobe)) return false;
    return mode > 0 || (int64_t)std::min(nprobe, k) * 3072 <= (int64_t)k;
}

void launch_route_select(const RouteArgs& a, size_t nq, hipStream_t st) {
    if (nq == 0) return;
    const int nt = (a.k + 15) / 16;
    const dim3 g((unsigned)nq), b(256);
    if (a.tmax && route_select_over_tiles(a.k, a.nprobe)) {
        if (nt <= 256 * 4)
            hipLaunchKernelGGL((k_route_select<ROUTE_TILES, 4>), g, b, 0, st, a);
        else if (nt <= 256 * 16)
            hipLaunchKernelGGL((k_route_select<ROUTE_TILES, 16>), g, b, 0, st, a);
        else if (nt <= 256 * 32)
            hipLaunchKernelGGL((k_route_select<ROUTE_TILES, 32>), g, b, 0, st, a);
        else
            hipLaunchKernelGGL((k_route_select<ROUTE_TILES, 64>), g, b, 0, st, a);
    } else if (a.k <= 256 * 16) {
        hipLaunchKernelGGL((k_route_select<ROUTE_REGS, 16>), g, b, 0, st, a);
    } else if (a.k <= 256 * 64) {
        hipLaunchKernelGGL((k_route_select<ROUTE_REGS, 64>), g, b, 0, st, a);
    } else {
        hipLaunchKernelGGL((k_route_select<ROUTE_MEM, 0>), g, b, 0, st, a);
    }
}

}  // namespace sieve
