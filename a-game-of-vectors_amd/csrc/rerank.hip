// The rerank: Index::rerank (src/ivf.rs:731-815) for a batch, one 256-thread workgroup per
// query over its deduplicated candidates (est, row, err).
//
//   prune (src/ivf.rs:736-739, prune_threshold :540-550): tau = m-th largest est - w*err;
//         keep est + w*err >= tau, order kept;
//   none:   score = est;
//   full:   score = (i32 dot of the int8 query and the 768-byte row) * (1/s);
//   prefix: score = (i32 dot over the first 256 bytes) * (1/s) / norm256[row];
//   ssd(r): the best r by estimate, then full;
//   tiered(r): the best r by the 2-bit tier score (Tier2::tables / score, src/tier2.rs:52-75)
//         when the index has a 2-bit tier, else by the prefix score; then full;
//   final: the best k by (score desc, row asc) (TopK::into_sorted, src/topk.rs:47-66).
// Integer dots run on v_dot4_i32_i8 with 16 lanes per row (3 x 16 coalesced bytes per lane),
// and are exact; the f32 steps after them are the reference's, in its order.
#include "kernels.hpp"
#include "select.hpp"

namespace sieve {

constexpr int RR_THREADS = 256;

struct RerankLds {
    size_t off_est, off_row, off_err, off_score, off_keys, off_list, off_t2, off_hist, off_misc, total;
};

__host__ __device__ inline RerankLds rerank_layout(int cs, bool t2) {
    auto al = [](size_t x) { return (x + 15) & ~(size_t)15; };
    int p2 = 1;
    while (p2 < cs) p2 <<= 1;
    RerankLds l;
    size_t o = 0;
    l.off_est = o;
    o = al(o + 4 * (size_t)cs);
    l.off_row = o;
    o = al(o + 4 * (size_t)cs);
    l.off_err = o;
    o = al(o + 4 * (size_t)cs);
    l.off_score = o;
    o = al(o + 4 * (size_t)cs);
    l.off_keys = o;
    o = al(o + 8 * (size_t)p2);
    l.off_list = o;
    o = al(o + 4 * (size_t)cs);
    l.off_t2 = o;
    o = al(o + (t2 ? 4 * 64 * 256 : 0));
    l.off_hist = o;
    o = al(o + 4 * 256);
    l.off_misc = o;
    o = al(o + 4 * 16);
    l.total = o;
    return l;
}

size_t rerank_lds_bytes(int cand_stride, bool tier2_tables) { return rerank_layout(cand_stride, tier2_tables).total; }

// score[i] = dot(wq, rerank[row_of(i)][0, dims)) as f32 * inv (dims = 768 or 256), for the
// n entries rows[0..n).
// The int8 row of `row`: a corpus row, or (live) a delta row.
__device__ __forceinline__ const int8_t* row_ptr(const RerankArgs& a, uint32_t row) {
    return row < a.n_base ? a.rerank + (size_t)row * DIM : a.drows + (size_t)(row - a.n_base) * DIM;
}
__device__ __forceinline__ float row_norm256(const RerankArgs& a, uint32_t row) {
    return row < a.n_base ? a.norm256[row] : a.dnorm256[row - a.n_base];
}

template <int DIMS>
__device__ void int8_scores(const RerankArgs& a, const int8_t* __restrict__ wq, const uint32_t* rows, int n,
                            float inv, float* score) {
    constexpr int CH = DIMS / 256;  // 16-byte chunks per lane
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6, sub = lane & 15;
    int wv[CH][4];
#pragma unroll
    for (int j = 0; j < CH; j++) {
        uint4 x = *reinterpret_cast<const uint4*>(wq + sub * 16 + 256 * j);
        wv[j][0] = (int)x.x;
        wv[j][1] = (int)x.y;
        wv[j][2] = (int)x.z;
        wv[j][3] = (int)x.w;
    }
    // 16 lanes per row, four rows per lane group at a time (their loads in flight together):
    // rows 64 it + 16 r + 4 wave + lane / 16, r < 4. The i32 sums are exact in any order.
    constexpr int R = 4;
    for (int b0 = wave * 4; b0 < n; b0 += R * (RR_THREADS / 16)) {  // wave-uniform
        const int base = b0 + (lane >> 4);
        uint4 x[R][CH];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int i = base + r * (RR_THREADS / 16);
            if (i < n) {
                const int8_t* p = row_ptr(a, rows[i]);
#pragma unroll
                for (int j = 0; j < CH; j++) x[r][j] = *reinterpret_cast<const uint4*>(p + sub * 16 + 256 * j);
            }
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int i = base + r * (RR_THREADS / 16);
            int acc = 0;
            if (i < n) {
#pragma unroll
                for (int j = 0; j < CH; j++) {
                    acc = __builtin_amdgcn_sdot4((int)x[r][j].x, wv[j][0], acc, false);
                    acc = __builtin_amdgcn_sdot4((int)x[r][j].y, wv[j][1], acc, false);
                    acc = __builtin_amdgcn_sdot4((int)x[r][j].z, wv[j][2], acc, false);
                    acc = __builtin_amdgcn_sdot4((int)x[r][j].w, wv[j][3], acc, false);
                }
            }
            acc = group_sum_dpp<16>(acc);  // the 16 lanes of the row (integer: any order)
            if (sub == 0 && i < n) score[i] = (float)acc * inv;
        }
    }
}

// The best `r` of n scored entries by (score desc, row asc): writes their rows to out_rows
// (best first) and returns how many.
__device__ int best_rows(const float* score, const uint32_t* rows, int n, int r, uint2* keys, uint32_t* out_rows) {
    const int t = threadIdx.x;
    const int n2 = pow2_at_least(max(n, 1));
    for (int i = t; i < n2; i += RR_THREADS)
        keys[i] = i < n ? make_uint2(f2key(score[i] + 0.0f), ~rows[i]) : make_uint2(0u, 0u);
    __syncthreads();
    block_sort_pairs(keys, nullptr, n2, false);
    const int m = min(r, n);
    for (int i = t; i < m; i += RR_THREADS) out_rows[i] = ~keys[i].y;
    __syncthreads();
    return m;
}

__global__ __launch_bounds__(RR_THREADS) void k_rerank(RerankArgs a) {
    extern __shared__ __align__(16) uint8_t lds[];
    const bool use_t2 = a.mode == SIEVE_RERANK_TIERED && a.t2_codes != nullptr;
    const RerankLds L = rerank_layout(a.cand_stride, use_t2);
    float* est = reinterpret_cast<float*>(lds + L.off_est);
    uint32_t* row = reinterpret_cast<uint32_t*>(lds + L.off_row);
    float* err = reinterpret_cast<float*>(lds + L.off_err);
    float* score = reinterpret_cast<float*>(lds + L.off_score);
    uint2* keys = reinterpret_cast<uint2*>(lds + L.off_keys);
    uint32_t* list = reinterpret_cast<uint32_t*>(lds + L.off_list);
    float* t2 = reinterpret_cast<float*>(lds + L.off_t2);
    uint32_t* hist = reinterpret_cast<uint32_t*>(lds + L.off_hist);
    uint32_t* misc = reinterpret_cast<uint32_t*>(lds + L.off_misc);
    const size_t qi = blockIdx.x;
    const int t = threadIdx.x;
    int n = (int)a.cand_n[qi];
    const float* ge = a.est + qi * a.cand_stride;
    const uint32_t* gr = a.rows + qi * a.cand_stride;
    const float* gerr = a.errs + qi * a.cand_stride;
    for (int i = t; i < n; i += RR_THREADS) {
        est[i] = ge[i];
        row[i] = gr[i];
        err[i] = gerr[i];
    }
    __syncthreads();

    // Bound pruning (src/ivf.rs:736-739).
    if (a.prune_m > 0 && n >= a.prune_m) {
        const float w = a.prune_w;
        const int n2 = n;
        for (int i = t; i < n2; i += RR_THREADS) {
            float lo = est[i] - w * err[i];
            keys[i] = make_uint2(f2key(lo + 0.0f), (uint32_t)i);
        }
        __syncthreads();
        uint2 kth;
        if (a.prune_m <= 32) {  // the m-th largest by m rounds of a block maximum
            // hist (1 KiB): 8 u64 of maxima, then up to 32 entries
            uint2* best = reinterpret_cast<uint2*>(hist) + 8;
            block_topk_small(keys, n, a.prune_m, best, reinterpret_cast<unsigned long long*>(hist));
            kth = best[a.prune_m - 1];
            __syncthreads();
        } else {
            kth = block_kth_largest(keys, n, a.prune_m, hist, misc + 8);
        }
        const float tau = key2f(kth.x);
        // keep est + w*err >= tau, in order: each thread a contiguous run
        const int per = (n + RR_THREADS - 1) / RR_THREADS;
        const int i0 = t * per;
        float ke[16];
        uint32_t kr[16];
        float kx[16];
        uint32_t mine = 0;
        for (int i = i0; i < i0 + per && i < n; i++) {
            if (est[i] + w * err[i] >= tau) {
                ke[mine] = est[i];
                kr[mine] = row[i];
                kx[mine] = err[i];
                mine++;
            }
        }
        uint32_t tot;
        uint32_t off = block_excl_scan(mine, misc + 4, &tot);
        __syncthreads();
        for (uint32_t j = 0; j < mine; j++) {
            est[off + j] = ke[j];
            row[off + j] = kr[j];
            err[off + j] = kx[j];
        }
        n = (int)tot;
        __syncthreads();
    }
    if (t == 0 && a.reranked) atomicAdd(a.reranked, (unsigned long long)n);

    const int8_t* wq768 = a.wq768 + qi * DIM;
    const int8_t* wq256 = a.wq256 + qi * PDIM;
    const float2 inv = a.inv[qi];
    int m = n;              // entries in (score, list) to rank at the end
    const uint32_t* fin_rows = row;
    switch (a.mode) {
        case SIEVE_RERANK_NONE:
            for (int i = t; i < n; i += RR_THREADS) score[i] = est[i];
            __syncthreads();
            break;
        case SIEVE_RERANK_FULL:
            int8_scores<DIM>(a, wq768, row, n, inv.x, score);
            __syncthreads();
            break;
        case SIEVE_RERANK_PREFIX:
            int8_scores<PDIM>(a, wq256, row, n, inv.y, score);
            __syncthreads();
            for (int i = t; i < n; i += RR_THREADS) score[i] = score[i] / row_norm256(a, row[i]);
            __syncthreads();
            break;
        default: {
            // ssd(r) / tiered(r): a first ranking picks r rows, then the full rerank.
            if (a.mode == SIEVE_RERANK_SSD) {
                for (int i = t; i < n; i += RR_THREADS) score[i] = est[i];
            } else if (use_t2) {
                const float* q = a.q256 + qi * PDIM;
                const float LV[4] = {-1.510f, -0.4528f, 0.4528f, 1.510f};
                for (int idx = t; idx < 64 * 256; idx += RR_THREADS) {
                    const int g = idx >> 8, b = idx & 255;
                    float h[2];
#pragma unroll
                    for (int hh = 0; hh < 2; hh++) {
                        const int bp = hh == 0 ? (b & 15) : (b >> 4);
                        const int i0 = 4 * g + 2 * hh;
                        float s = -0.0f;
                        s = s + q[i0] * LV[bp & 3] * a.t2_sigma[i0];
                        s = s + q[i0 + 1] * LV[(bp >> 2) & 3] * a.t2_sigma[i0 + 1];
                        h[hh] = s;
                    }
                    t2[idx] = h[0] + h[1];
                }
                __syncthreads();
                for (int i = t; i < n; i += RR_THREADS) {
                    const uint32_t ri = row[i];
                    const bool dr = ri >= a.n_base;
                    const uint4* cp = reinterpret_cast<const uint4*>(dr ? a.dt2 + (size_t)(ri - a.n_base) * 64
                                                                        : a.t2_codes + (size_t)ri * 64);
                    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                    for (int c4 = 0; c4 < 4; c4++) {
                        uint4 x = cp[c4];
                        const uint32_t wd[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
                        for (int j = 0; j < 16; j++) {
                            const int g = c4 * 16 + j;
                            const uint32_t byte = (wd[j >> 2] >> ((j & 3) * 8)) & 255u;
                            acc[g & 3] = acc[g & 3] + t2[g * 256 + byte];
                        }
                    }
                    score[i] = (acc[0] + acc[1] + acc[2] + acc[3]) / (dr ? a.dt2n[ri - a.n_base] : a.t2_norm[ri]);
                }
            } else {
                int8_scores<PDIM>(a, wq256, row, n, inv.y, score);
                __syncthreads();
                for (int i = t; i < n; i += RR_THREADS) score[i] = score[i] / row_norm256(a, row[i]);
            }
            __syncthreads();
            m = best_rows(score, row, n, a.r, keys, list);
            int8_scores<DIM>(a, wq768, list, m, inv.x, score);
            __syncthreads();
            fin_rows = list;
            break;
        }
    }
    // Final top k by (score desc, row asc).
    const int kk = min(a.k, m);
    if (a.k <= 32) {
        for (int i = t; i < m; i += RR_THREADS) keys[i] = make_uint2(f2key(score[i] + 0.0f), ~fin_rows[i]);
        __syncthreads();
        uint2* best = reinterpret_cast<uint2*>(hist) + 8;  // after the 8 u64 of maxima
        block_topk_small(keys, m, a.k, best, reinterpret_cast<unsigned long long*>(hist));
        for (int i = t; i < kk; i += RR_THREADS) keys[i] = best[i];
        __syncthreads();
    } else {
        const int n2 = pow2_at_least(max(m, 1));
        for (int i = t; i < n2; i += RR_THREADS)
            keys[i] = i < m ? make_uint2(f2key(score[i] + 0.0f), ~fin_rows[i]) : make_uint2(0u, 0u);
        __syncthreads();
        block_sort_pairs(keys, nullptr, n2, false);
    }
    for (int i = t; i < kk; i += RR_THREADS) {
        a.out_s[qi * a.k + i] = key2f(keys[i].x);
        const uint32_t r = ~keys[i].y;
        a.out_r[qi * a.k + i] = (a.dext && r >= a.n_base) ? a.dext[r - a.n_base] : r;
    }
    if (t == 0) a.out_n[qi] = (uint32_t)kk;
}

void launch_rerank(const RerankArgs& a, size_t nq, hipStream_t st) {
    if (nq == 0) return;
    const bool t2 = a.mode == SIEVE_RERANK_TIERED && a.t2_codes != nullptr;
    const size_t lds = rerank_layout(a.cand_stride, t2).total;
    if (lds > 64 * 1024) (void)hipFuncSetAttribute((const void*)k_rerank, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(k_rerank, dim3((unsigned)nq), dim3(RR_THREADS), lds, st, a);
}

}  // namespace sieve
