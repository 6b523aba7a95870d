// Multi-bit refinement of the 1B L2 path (src/bigann.rs:38-113, 565-602, 683-729).
//
// An index trained with bits > 1 stores, per slot, the low bits - 1 planes of a uniform
// bits-bit code of the rotated residual (its top plane is the sign, the 1-bit code) and a scale
// f. A search with refine > 0 keeps a 1-bit shortlist of max(refine, cand) slots; every slot that
// enters it when it is scanned is re-estimated from all planes, and the best cand by the refined
// estimate go to the exact rerank.
//
// k_refine_encode: refine_encode (src/bigann.rs:70-108) for the builder, one thread per row,
// every f32 step in the reference's order.
//
// k_scan_l2_refine: the refined candidates of one query per 256-thread workgroup. The
// shortlist's admission depends on the order slots are pushed (its threshold is the
// max(refine, cand)-th best so far, raised each time its buffer reaches twice that), so the
// workgroup replays the reference's sequence exactly: the query's probes in route order, each
// leaf's slots in order, 256 slots a step. A step's 1-bit estimates are computed in parallel;
// admission then runs in slot order: a block prefix count places the passing slots, the one
// that fills the buffer triggers the cut (the radix-selected K-th best becomes the threshold and
// the buffer keeps the best K), and the slots after it are tested again against the new
// threshold. Admitted slots get their refined estimate from per-query tables in LDS (for each of
// the 16 code bytes and byte value, the sum of the rotated query over its set bits) and go to a
// second list of the best cand by refined estimate, the candidates. Ties there are broken by
// admission order, as TopK::into_sorted breaks them by the index each slot is pushed with.
#include <float.h>

#include "kernels.hpp"
#include "select.hpp"

namespace sieve {

constexpr int RF_PD = 128;
constexpr int RF_CODE = 16;
constexpr int RF_T = 256;

__global__ __launch_bounds__(RF_T) void k_refine_encode(const float* __restrict__ r, size_t n, int bits,
                                                        uint4* __restrict__ planes, float* __restrict__ fout) {
    const size_t row = (size_t)blockIdx.x * RF_T + threadIdx.x;
    if (row >= n) return;
    float v[RF_PD];
    const float4* x = reinterpret_cast<const float4*>(r + row * RF_PD);
#pragma unroll
    for (int i = 0; i < RF_PD / 4; i++) {
        const float4 a = x[i];
        v[4 * i] = a.x;
        v[4 * i + 1] = a.y;
        v[4 * i + 2] = a.z;
        v[4 * i + 3] = a.w;
    }
    const uint32_t top = 1u << (bits - 1);
    const float h = (float)(2 * top - 1) / 2.0f;
    const float umax = (float)(2 * top - 1);
    float amax = 0.0f;
#pragma unroll
    for (int i = 0; i < RF_PD; i++) amax = fmaxf(amax, fabsf(v[i]));
    uint32_t pw[2][4] = {{0u, 0u, 0u, 0u}, {0u, 0u, 0u, 0u}};
    float f = 0.0f;
    if (amax >= 1e-12f) {
        // The nearest level for step delta, its top bit forced to the sign.
        auto level = [&](float vi, float delta) -> uint32_t {
            float t = floorf(vi / delta + h + 0.5f);
            t = t < 0.0f ? 0.0f : (t > umax ? umax : t);
            const uint32_t u = (uint32_t)t;
            return vi > 0.0f ? (u | top) : (u & (top - 1));
        };
        float best = -FLT_MAX, best_delta = amax / h;
        for (int st = 0; st < 24; st++) {
            const float delta = amax / h * (0.3f + 0.035f * (float)st);
            float dot = 0.0f, n2 = 0.0f;
#pragma unroll
            for (int i = 0; i < RF_PD; i++) {
                const float w = (float)level(v[i], delta) - h;
                dot = dot + v[i] * w;
                n2 = n2 + w * w;
            }
            const float c = dot / sqrtf(n2);
            if (c > best) {
                best = c;
                best_delta = delta;
            }
        }
        float rr = 0.0f, dot = 0.0f;
#pragma unroll
        for (int i = 0; i < RF_PD; i++) {
            const uint32_t u = level(v[i], best_delta);
            if (bits > 1) pw[0][i >> 5] |= (u & 1u) << (i & 31);
            if (bits > 2) pw[1][i >> 5] |= ((u >> 1) & 1u) << (i & 31);
            rr = rr + v[i] * v[i];
            dot = dot + v[i] * ((float)u - h);
        }
        f = dot > 0.0f ? rr / dot : 0.0f;
    }
    planes[2 * row] = make_uint4(pw[0][0], pw[0][1], pw[0][2], pw[0][3]);
    planes[2 * row + 1] = make_uint4(pw[1][0], pw[1][1], pw[1][2], pw[1][3]);
    fout[row] = f;
}

void launch_refine_encode(const float* r, size_t n, int bits, uint4* planes, float* f, hipStream_t st) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_refine_encode, dim3((unsigned)((n + RF_T - 1) / RF_T)), dim3(RF_T), 0, st, r, n, bits, planes,
                       f);
}

// The per-slot extra section of a record (bits - 1 planes of 16 B, then f as LE f32; stride
// refine_bytes(bits)) into the device layout: two 16-byte plane words per slot and f.
__global__ __launch_bounds__(RF_T) void k_unpack_refine(const uint8_t* __restrict__ extra, size_t nslots, int bits,
                                                        uint4* __restrict__ planes, float* __restrict__ f) {
    const size_t s = (size_t)blockIdx.x * RF_T + threadIdx.x;
    if (s >= nslots) return;
    const int rb = (bits - 1) * 16 + 4;
    const uint8_t* e = extra + s * rb;
    uint32_t w[8] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
    for (int b = 0; b < (bits - 1) * 16; b++) w[b >> 2] |= (uint32_t)e[b] << (8 * (b & 3));
    planes[2 * s] = make_uint4(w[0], w[1], w[2], w[3]);
    planes[2 * s + 1] = make_uint4(w[4], w[5], w[6], w[7]);
    const uint8_t* fb = e + rb - 4;
    f[s] = __uint_as_float((uint32_t)fb[0] | ((uint32_t)fb[1] << 8) | ((uint32_t)fb[2] << 16) | ((uint32_t)fb[3] << 24));
}

void launch_unpack_refine(const uint8_t* extra, size_t nslots, int bits, uint4* planes, float* f, hipStream_t st) {
    if (nslots == 0) return;
    hipLaunchKernelGGL(k_unpack_refine, dim3((unsigned)((nslots + RF_T - 1) / RF_T)), dim3(RF_T), 0, st, extra,
                       nslots, bits, planes, f);
}

// ------------------------------------------------------------------ the refined scan ---
struct RefineLayout {
    size_t lut, rtab, pbase, pcent, poff, cbuf, bkey, bpay, misc, total;
};

__host__ __device__ inline RefineLayout refine_layout(int K, int cand, int nprobe) {
    RefineLayout l{};
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o = (o + bytes + 15) & ~size_t(15);
        return at;
    };
    const int bcap = 2 * cand + RF_T;
    l.lut = take(32 * 16);
    l.rtab = take((size_t)RF_CODE * 256 * 4);
    l.pbase = take((size_t)nprobe * 4);
    l.pcent = take((size_t)nprobe * 4);
    l.poff = take(((size_t)nprobe + 1) * 4);
    // the shortlist (2K entries), reused at the end for the final sort of the candidates
    size_t p2 = 1;
    while (p2 < (size_t)cand) p2 <<= 1;
    const size_t sl = (size_t)2 * K * 8, fs = p2 * 12;
    l.cbuf = take(sl > fs ? sl : fs);
    l.bkey = take((size_t)bcap * 8);
    l.bpay = take((size_t)bcap * 8);
    l.misc = take(64 * 4 + 256 * 4);
    l.total = o;
    return l;
}

size_t refine_lds_bytes(int K, int cand, int nprobe) { return refine_layout(K, cand, nprobe).total; }

// In-place compaction of entries [0, n) of key (and, when pay is set, their payload) to those
// >= bound in (hi, lo) order, keeping their relative order: chunk by chunk, a chunk is read
// before any of it is written, and a kept entry only moves down. Returns the count kept.
__device__ int block_compact_ge(uint2* key, uint2* pay, int n, uint2 bound, uint32_t* scratch) {
    int kept = 0;
    for (int c0 = 0; c0 < n; c0 += RF_T) {
        const int i = c0 + (int)threadIdx.x;
        uint2 k = make_uint2(0u, 0u), p = make_uint2(0u, 0u);
        bool keep = false;
        if (i < n) {
            k = key[i];
            if (pay) p = pay[i];
            keep = ge2(k, bound);
        }
        uint32_t tot;
        const uint32_t at = block_excl_scan(keep ? 1u : 0u, scratch, &tot);  // barriers: reads done
        if (keep) {
            key[kept + at] = k;
            if (pay) pay[kept + at] = p;
        }
        kept += (int)tot;
        __syncthreads();
    }
    return kept;
}


__global__ __launch_bounds__(RF_T) void k_scan_l2_refine(L2RefineArgs a) {
    extern __shared__ __align__(16) uint8_t lds[];
    const int np = (int)min(a.probe_n[blockIdx.x], (uint32_t)a.probe_stride);
    const RefineLayout L = refine_layout(a.K, a.cand, a.probe_stride);
    uint8_t* lut = lds + L.lut;
    float* rtab = reinterpret_cast<float*>(lds + L.rtab);
    uint32_t* pbase = reinterpret_cast<uint32_t*>(lds + L.pbase);
    float* pcent = reinterpret_cast<float*>(lds + L.pcent);
    uint32_t* poff = reinterpret_cast<uint32_t*>(lds + L.poff);
    uint2* cbuf = reinterpret_cast<uint2*>(lds + L.cbuf);
    uint2* bkey = reinterpret_cast<uint2*>(lds + L.bkey);
    uint2* bpay = reinterpret_cast<uint2*>(lds + L.bpay);
    uint32_t* misc = reinterpret_cast<uint32_t*>(lds + L.misc);
    uint32_t* scratch = misc;            // 16 words: block scans
    uint32_t* sh = misc + 16;            // 8 words: select results, the cutting thread
    uint32_t* hist = misc + 64;          // 256 words
    float* offp = reinterpret_cast<float*>(misc + 32);
    const size_t qi = blockIdx.x;
    const int t = threadIdx.x;
    const int K = a.K, cand = a.cand, bits = a.bits;

    for (int i = t; i < 32 * 16; i += RF_T) lut[i] = a.tables[qi * 512 + i];
    // Refiner tables: entry (k, v) sums the query over v's set bits, highest bit first from
    // +0.0, which is t[v] = t[v & (v - 1)] + q[8 k + lowest set bit] unrolled.
    const float* qr = a.qrot + qi * RF_PD;
    for (int e = t; e < RF_CODE * 256; e += RF_T) {
        const int k = e >> 8, v = e & 255;
        float s = 0.0f;
        for (int j = 7; j >= 0; j--)
            if ((v >> j) & 1) s = s + qr[8 * k + j];
        rtab[e] = s;
    }
    if (t == 0) {
        sh[5] = 0;  // the candidate list's staging counter
        const float hh = (float)((1u << bits) - 1) / 2.0f;
        float sum = -0.0f;
        for (int i = 0; i < RF_PD; i++) sum = sum + qr[i];
        offp[0] = hh * sum;
    }
    // The probes' slot ranges, concatenated in route order.
    for (int p = t; p < np; p += RF_T) {
        const uint32_t c = a.probe_c[qi * a.probe_stride + p];
        pbase[p] = a.block_off[c] * BLOCK;
        pcent[p] = a.probe_s[qi * a.probe_stride + p];
    }
    __syncthreads();
    if (t == 0) {
        uint32_t acc = 0;
        for (int p = 0; p < np; p++) {
            poff[p] = acc;
            const uint32_t c = a.probe_c[qi * a.probe_stride + p];
            acc += (a.block_off[c + 1] - a.block_off[c]) * BLOCK;
        }
        poff[np] = acc;
    }
    __syncthreads();
    const uint32_t total = poff[np];
    const float qa = a.qab[qi].x, qb = a.qab[qi].y, off = offp[0];
    const float ftop = (float)(1u << (bits - 1));

    // Block-uniform state: the shortlist's count and threshold, the candidate list's count and
    // threshold (keys), the admission counter.
    int cnt = 0;
    bool has_thr = false;
    float thr = 0.0f;
    int nb = 0;
    bool has_thr2 = false;
    uint32_t thr2 = 0;
    uint32_t info = 0;
    auto plane = [&](const uint4 w) -> float {
        const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
        float s = -0.0f;
#pragma unroll
        for (int k = 0; k < RF_CODE; k++) s = s + rtab[k * 256 + ((ws[k >> 2] >> (8 * (k & 3))) & 255u)];
        return s;
    };
    for (uint32_t s0 = 0; s0 < total; s0 += RF_T) {
        const uint32_t pos = s0 + (uint32_t)t;
        bool valid = false;
        float est = 0.0f, cent = 0.0f, xn = 0.0f;
        uint32_t slot = 0, id = PAD;
        uint4 code = make_uint4(0u, 0u, 0u, 0u);
        if (pos < total) {
            int lo = 0, hi = np - 1;  // the probe holding pos: poff[p] <= pos < poff[p + 1]
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (poff[mid] <= pos) lo = mid;
                else hi = mid - 1;
            }
            slot = pbase[lo] + (pos - poff[lo]);
            id = a.ids[slot];
            if (id != PAD) {
                valid = true;
                cent = pcent[lo];
                code = a.codes[slot];
                xn = a.nx[slot];
                const uint32_t w[4] = {code.x, code.y, code.z, code.w};
                uint32_t sc = 0;
#pragma unroll
                for (int g = 0; g < 32; g++) sc += lut[g * 16 + ((w[g >> 3] >> (4 * (g & 7))) & 15u)];
                est = 2.0f * (cent + a.mul[slot] * (qa * (float)sc + qb)) - xn;
            }
        }
        // Admission in slot order, with cuts as the shortlist fills.
        int cursor = 0;
        for (;;) {
            const bool pass = valid && t >= cursor && (!has_thr || est > thr);
            uint32_t P;
            const uint32_t rank = block_excl_scan(pass ? 1u : 0u, scratch, &P);
            const int room = K ? 2 * K - cnt : RF_T + 1;  // K == 0: the shortlist is never cut
            const int take = min((int)P, room);
            const bool adm = pass && (int)rank < take;
            if (adm) {
                if (K) cbuf[cnt + rank] = make_uint2(f2key(est + 0.0f), (uint32_t)(cnt + (int)rank));
                if (rank == (uint32_t)(room - 1)) sh[4] = (uint32_t)t;
                // the refined estimate, and the candidate list's own threshold
                float u = ftop * plane(code);
                for (int p = 0; p < bits - 1; p++) u = u + (float)(1u << p) * plane(a.planes[2 * (size_t)slot + p]);
                const float e2 = 2.0f * (cent + a.rf[slot] * (u - off)) - xn;
                const uint32_t k2 = f2key(e2 + 0.0f);
                if (!has_thr2 || k2 > thr2) {
                    const uint32_t at = atomicAdd(&sh[5], 1u);
                    bkey[nb + at] = make_uint2(k2, ~(info + rank));
                    bpay[nb + at] = make_uint2(id, __float_as_uint(xn));
                }
            }
            __syncthreads();
            nb += (int)sh[5];
            info += (uint32_t)take;
            __syncthreads();
            if (t == 0) sh[5] = 0;
            if (nb >= 2 * cand) {
                // the candidate list keeps its best cand (TopK::compact): its threshold rises
                const uint2 kb = block_kth_largest(bkey, nb, cand, hist, sh);
                nb = block_compact_ge(bkey, bpay, nb, kb, scratch);
                has_thr2 = true;
                thr2 = kb.x;
            }
            if (take < room) {
                if (K) cnt += take;
                break;
            }
            // The shortlist is full at the slot with rank room - 1: its K-th best becomes the
            // threshold and it keeps its best K; later slots of this step are tested again.
            const int cut_t = (int)sh[4];
            const uint2 kb = block_kth_largest(cbuf, 2 * K, K, hist, sh);
            cnt = block_compact_ge(cbuf, nullptr, 2 * K, kb, scratch);
            for (int i = t; i < cnt; i += RF_T) cbuf[i].y = (uint32_t)i;  // keep positions unique
            __syncthreads();
            has_thr = true;
            thr = key2f(kb.x);
            cursor = cut_t + 1;
        }
    }
    if (t == 0 && a.scanned) atomicAdd(a.scanned, (unsigned long long)total);
    // The candidates: the best cand by (refined estimate desc, admission order), sorted.
    if (nb > cand) {
        const uint2 kb = block_kth_largest(bkey, nb, cand, hist, sh);
        nb = block_compact_ge(bkey, bpay, nb, kb, scratch);
    }
    const int n2 = pow2_at_least(max(nb, 1));
    uint2* skey = cbuf;                                          // the shortlist is no longer needed
    uint32_t* spay = reinterpret_cast<uint32_t*>(cbuf + n2);
    for (int i = t; i < n2; i += RF_T) {
        skey[i] = i < nb ? bkey[i] : make_uint2(0u, 0u);
        spay[i] = (uint32_t)i;
    }
    __syncthreads();
    block_sort_pairs(skey, spay, n2, false);
    for (int i = t; i < nb; i += RF_T) {
        const uint2 pv = bpay[spay[i]];
        a.out_est[qi * a.out_stride + i] = key2f(skey[i].x);
        a.out_row[qi * a.out_stride + i] = pv.x;
        a.out_nx[qi * a.out_stride + i] = __uint_as_float(pv.y);
    }
    if (t == 0) a.out_n[qi] = (uint32_t)nb;
}

void launch_scan_l2_refine(const L2RefineArgs& a, size_t nq, hipStream_t st) {
    if (nq == 0) return;
    const size_t lds = refine_lds_bytes(a.K, a.cand, a.probe_stride);
    if (lds > 64 * 1024)
        (void)hipFuncSetAttribute((const void*)k_scan_l2_refine, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(k_scan_l2_refine, dim3((unsigned)nq), dim3(RF_T), lds, st, a);
}

}  // namespace sieve
