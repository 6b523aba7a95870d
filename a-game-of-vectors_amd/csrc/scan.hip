// The scan: Index::candidates (src/ivf.rs:643-729) for a batch, plus the standalone lut::scan
// seam (src/lut.rs:121-152).
//
// Work items. A query's probed clusters, taken in probe order, form one sequence of blocks;
// it is cut into C_q = clamp(ceil(T_q / chunk_blocks), 1, chunk_max) chunks of nearly equal
// size (T_q blocks in all), and one 256-thread workgroup scans each chunk, 256 * S slots per
// round (S slots per thread, a compile-time variant), one lane per slot:
//   * score: the 64 nibbles of the slot's 32-byte code index the query's 64 16-entry u8
//     tables, held in LDS (all lanes of a wave read the same 16-byte table row at a time, so
//     the reads are broadcast and conflict-free); the u16 sum equals score_code
//     (src/lut.rs:83-86) exactly;
//   * estimate: cent + mul * (a * sc + b), f32, no FMA (rabitq::Query::estimate,
//     src/rabitq.rs:127-130);
//   * admission: rejected if (est, slot) is not above the threshold, if the slot is padding,
//     or if the view hides its row (src/ivf.rs:669); otherwise appended to an LDS buffer;
//   * TopK (src/topk.rs:27-44): when the buffer holds 2*cand entries it is cut to the best
//     cand by a radix select in LDS and the threshold rises to the cand-th entry.
// Entries are ordered by (estimate desc, slot asc), a total order, and the threshold only
// drops entries that cand others beat, so the kept set is exactly the best cand admitted slots
// whatever the visiting order: results do not depend on how a query is cut into chunks or
// batched with others. (The reference's select_nth_unstable keeps an arbitrary one of tied
// estimates.) With C_q > 1 the chunks of a query share the best threshold any of them has
// reached (a 64-bit atomic max per query), write their kept entries to a partial list, and the
// last chunk to finish merges the lists through the same buffer.
// The final cand slots are sorted by (row asc, est desc) and deduplicated by row
// (src/ivf.rs:725-726).
//
// Traffic: 32 code bytes + 4 mul bytes per stored slot, read once per (query, probed
// cluster); ids and err only for admitted slots; 8 bytes per partial entry, written and read
// once when a query is split.
#include <algorithm>
#include <cmath>

#include "kernels.hpp"
#include "select.hpp"

namespace sieve {

constexpr int SCAN_THREADS = 256;

struct ScanLds {
    int cap;
    size_t off_pref, off_start, off_cent, off_buf, off_tmp, off_hist, off_misc, total;
};

// slots: slots per thread per round (a round covers 256 * slots slots); trigger: buffered
// entries at which the buffer is cut back to cand (TopK cuts at 2 * cand; a lower trigger cuts
// more often but needs less LDS, so more workgroups fit a CU).
__host__ __device__ inline ScanLds scan_layout(int cand, int probe_stride, int slots, int trigger) {
    ScanLds l;
    l.cap = trigger + SCAN_THREADS * slots;
    size_t o = 0;
    auto al = [](size_t x) { return (x + 15) & ~(size_t)15; };
    l.off_pref = o;
    o = al(o + 4 * (size_t)(probe_stride + 1));
    l.off_start = o;
    o = al(o + 4 * (size_t)probe_stride);
    l.off_cent = o;
    o = al(o + 4 * (size_t)probe_stride);
    l.off_buf = o;
    o = al(o + 8 * (size_t)l.cap);
    l.off_tmp = o;
    o = al(o + 8 * (size_t)cand);
    l.off_hist = o;
    o = al(o + 4 * 256);
    l.off_misc = o;
    o = al(o + 4 * 32);
    l.total = o;
    return l;
}

size_t scan_lds_bytes(int cand, int probe_stride) {
    return scan_layout(cand, probe_stride, 4, 2 * cand).total + 1024;  // largest variant, plus the static tables
}

// misc words: 0 count, 1 has threshold, 2 threshold key, 3 scratch count, 4..7 scan scratch,
// 8..10 select scratch, 12..13 round stamps, 14 threshold ~slot, 15 last-chunk flag,
// 16..17 adopt stamps.
constexpr int M_COUNT = 0, M_HAS = 1, M_THR = 2, M_TMP = 3, M_SCAN = 4, M_SEL = 8, M_STAMP = 12, M_THR_LO = 14,
              M_LAST = 15, M_ADOPT = 16;

__device__ __forceinline__ unsigned long long pack_pair(uint2 e) {
    return ((unsigned long long)e.x << 32) | e.y;
}

// Cuts buf[0, n) to its best k entries (under (est key, ~slot)) and raises the threshold; with
// a per-query shared threshold, publishes it there.
__device__ void scan_compact(uint2* buf, uint2* tmp, int n, int k, uint32_t* hist, uint32_t* misc,
                             unsigned long long* gthr) {
    const int t = threadIdx.x;
    uint2 kth = block_kth_largest(buf, n, k, hist, misc + M_SEL);
    if (t == 0) misc[M_TMP] = 0;
    __syncthreads();
    for (int i = t; i < n; i += SCAN_THREADS) {
        uint2 e = buf[i];
        if (ge2(e, kth)) {
            uint32_t pos = atomicAdd(&misc[M_TMP], 1u);
            if (pos < (uint32_t)k) tmp[pos] = e;  // k exactly when entries are unique
        }
    }
    __syncthreads();
    for (int i = t; i < k; i += SCAN_THREADS) buf[i] = tmp[i];
    if (t == 0) {
        misc[M_COUNT] = (uint32_t)k;
        const bool raise = !misc[M_HAS] || kth.x > misc[M_THR] || (kth.x == misc[M_THR] && kth.y > misc[M_THR_LO]);
        if (raise) {
            misc[M_HAS] = 1;
            misc[M_THR] = kth.x;
            misc[M_THR_LO] = kth.y;
        }
        if (gthr) atomicMax(gthr, pack_pair(kth));
    }
    __syncthreads();
}

// G = 64 (256-bit codes, inner-product estimate) or G = 32 (128-bit codes of the padded
// 100-d int8 vectors, L2 estimate 2 (cent + mul (a sc + b)) - |x|^2, src/bigann.rs:705-711).
// SLOTS_PER_THREAD slots per thread per round; PF: prefetch the next round's codes into
// registers while scoring the current one (more bytes in flight, more VGPRs); WPE: minimum
// waves per SIMD the register allocation must allow (0: the compiler's choice).
template <int G, bool L2, int SLOTS_PER_THREAD, bool PF, int WPE>
__global__ __launch_bounds__(SCAN_THREADS) __attribute__((amdgpu_waves_per_eu(WPE > 0 ? WPE : 1)))
void k_scan_t(ScanArgs a) {
    constexpr int CW = G / 32;  // 16-byte words per slot code
    constexpr int ROUND_BLOCKS = SCAN_THREADS * SLOTS_PER_THREAD / BLOCK;
    // The query's tables sit in static LDS at a compile-time address, so each lookup is one
    // ds_read_u8 with the table offset in its immediate field.
    __shared__ __align__(16) uint8_t tab[G * 16];
    extern __shared__ __align__(16) uint8_t lds[];
    const ScanLds L = scan_layout(a.cand, a.seg_stride, SLOTS_PER_THREAD, a.trigger);
    uint32_t* pref = reinterpret_cast<uint32_t*>(lds + L.off_pref);
    uint32_t* pstart = reinterpret_cast<uint32_t*>(lds + L.off_start);
    float* pcent = reinterpret_cast<float*>(lds + L.off_cent);
    uint2* buf = reinterpret_cast<uint2*>(lds + L.off_buf);
    uint2* tmp = reinterpret_cast<uint2*>(lds + L.off_tmp);
    uint32_t* hist = reinterpret_cast<uint32_t*>(lds + L.off_hist);
    uint32_t* misc = reinterpret_cast<uint32_t*>(lds + L.off_misc);

    const size_t qi = blockIdx.x / a.chunk_max;
    const uint32_t ci = blockIdx.x % a.chunk_max;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    // Block segments in probe order: each probed cluster's base list, then (live) its delta
    // list (src/ivf.rs:682-684).
    const bool live = a.dblock_off != nullptr;
    const int np = (live ? 2 : 1) * (int)a.probe_n[qi];
    for (int j = t; j < np; j += SCAN_THREADS) {
        const int pj = live ? j >> 1 : j;
        const uint32_t c = a.probe_c[qi * a.probe_stride + pj];
        if (!live || (j & 1) == 0) {
            const uint32_t b0 = a.block_off[c];
            pstart[j] = b0;
            pref[j] = a.block_off[c + 1] - b0;
        } else {
            const uint32_t d0 = a.dblock_off[c];
            pstart[j] = a.dbase + d0;
            pref[j] = a.dblock_off[c + 1] - d0;
        }
        pcent[j] = a.probe_s[qi * a.probe_stride + pj];
    }
    if (t < 32) misc[t] = 0;
    __syncthreads();
    if (wave == 0) {
        uint32_t run = 0;
        for (int base = 0; base < np; base += 64) {
            int j = base + lane;
            uint32_t v = j < np ? pref[j] : 0u;
            uint32_t incl = wave_incl_scan(v);
            if (j < np) pref[j] = run + incl - v;
            run += __shfl(incl, 63, 64);
        }
        if (lane == 0) pref[np] = run;
    }
    __syncthreads();
    const uint32_t total = pref[np];
    const uint32_t nch = min(a.chunk_max, max(1u, (uint32_t)((total + (uint64_t)a.chunk_blocks - 1) / a.chunk_blocks)));
    if (ci >= nch) return;  // the whole workgroup: this query has fewer chunks
    const uint32_t vb_lo = (uint32_t)((uint64_t)total * ci / nch);
    const uint32_t vb_hi = (uint32_t)((uint64_t)total * (ci + 1) / nch);
    if (t < G * 4) reinterpret_cast<uint32_t*>(tab)[t] = reinterpret_cast<const uint32_t*>(a.tables + qi * G * 16)[t];
    const float4 qab = a.qab[qi];
    if (t == 0 && a.scanned) atomicAdd(a.scanned, (unsigned long long)(vb_hi - vb_lo) * BLOCK);
    unsigned long long* gthr = nch > 1 ? a.gthr + qi : nullptr;

    const uint64_t* allow = nullptr;
    if (a.view.allow && (a.view.allow_stride == 0 || !a.view.allow_flags || a.view.allow_flags[qi]))
        allow = a.view.allow + qi * a.view.allow_stride;
    const int cand = a.cand;
    uint32_t id_reads = 0;  // slots past the threshold: their id is read

    // Appends the wave's passing entries; the wave whose append takes the buffer to 2 * cand
    // stamps the round so that every wave compacts after the barrier.
    auto admit = [&](bool pass, uint32_t key, uint32_t lo, uint32_t round) {
        const uint64_t bal = __ballot(pass);
        if (bal) {
            const uint32_t nw = (uint32_t)__popcll(bal);
            const uint32_t before = (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
            uint32_t base = 0;
            if (lane == 0) {
                base = atomicAdd(&misc[M_COUNT], nw);
                if (base + nw >= (uint32_t)a.trigger) misc[M_STAMP + (round & 1)] = round;
            }
            base = __shfl(base, 0, 64);
            if (pass) buf[base + before] = make_uint2(key, lo);
        }
    };
    // After the barrier of round `round`: compact if stamped, then adopt a better shared
    // threshold that thread 0 read during the round (`g`, valid in thread 0 only).
    auto settle = [&](uint32_t round, unsigned long long g) {
        if (misc[M_STAMP + (round & 1)] == round)
            scan_compact(buf, tmp, (int)misc[M_COUNT], cand, hist, misc, gthr);
        if (misc[M_ADOPT + (round & 1)] == round) {
            if (t == 0) {
                const uint2 e = make_uint2((uint32_t)(g >> 32), (uint32_t)g);
                if (!misc[M_HAS] || e.x > misc[M_THR] || (e.x == misc[M_THR] && e.y > misc[M_THR_LO])) {
                    misc[M_HAS] = 1;
                    misc[M_THR] = e.x;
                    misc[M_THR_LO] = e.y;
                }
            }
            __syncthreads();
        }
    };
    // Thread 0 reads the shared threshold during the round and flags the round if it beats
    // the local one (called before the round's barrier).
    auto peek = [&](uint32_t round, unsigned long long& g) {
        if (gthr && t == 0) {
            g = __hip_atomic_load(gthr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned long long mine =
                misc[M_HAS] ? (((unsigned long long)misc[M_THR] << 32) | misc[M_THR_LO]) : 0ull;
            if (g > mine) misc[M_ADOPT + (round & 1)] = round;
        }
    };

    // A thread's blocks only move forward (vb0 + (t >> 5) + 8u, rounds in order), so a cursor
    // over the probe prefix sums finds each block's probed cluster in amortised O(1).
    int cur = 0;
    struct Regs {
        uint4 cw[SLOTS_PER_THREAD][CW];
        float mul[SLOTS_PER_THREAD], cent[SLOTS_PER_THREAD];
        uint32_t gs[SLOTS_PER_THREAD];
        bool ok[SLOTS_PER_THREAD];
    };
    auto fetch = [&](uint32_t vb0, Regs& R) {
#pragma unroll
        for (int u = 0; u < SLOTS_PER_THREAD; u++) {
            const uint32_t vb = vb0 + (t >> 5) + 8 * u;
            R.ok[u] = vb < vb_hi;
            R.gs[u] = 0;
            R.cent[u] = 0.0f;
            if (R.ok[u]) {
                while (cur + 1 < np && pref[cur + 1] <= vb) cur++;
                R.gs[u] = (pstart[cur] + (vb - pref[cur])) * BLOCK + (t & 31);
                R.cent[u] = pcent[cur];
                const uint4* cp = a.codes + CW * (size_t)R.gs[u];
#pragma unroll
                for (int w = 0; w < CW; w++) R.cw[u][w] = cp[w];
                R.mul[u] = a.mul[R.gs[u]];
            }
        }
    };
    auto score = [&](const Regs& R, uint32_t round) {
        const bool has_thr = misc[M_HAS] != 0;
        const uint32_t thr = misc[M_THR], thr_lo = misc[M_THR_LO];
#pragma unroll
        for (int u = 0; u < SLOTS_PER_THREAD; u++) {
            bool pass = false;
            uint32_t key = 0;
            if (R.ok[u]) {
                uint32_t w[4 * CW];
#pragma unroll
                for (int x = 0; x < CW; x++) {
                    w[4 * x] = R.cw[u][x].x;
                    w[4 * x + 1] = R.cw[u][x].y;
                    w[4 * x + 2] = R.cw[u][x].z;
                    w[4 * x + 3] = R.cw[u][x].w;
                }
                uint32_t sc = 0;
#pragma unroll
                for (int d = 0; d < G / 8; d++)
#pragma unroll
                    for (int j = 0; j < 8; j++) sc += tab[(8 * d + j) * 16 + ((w[d] >> (4 * j)) & 15u)];
                float est = R.cent[u] + R.mul[u] * (qab.x * (float)sc + qab.y);
                if constexpr (L2) est = 2.0f * est - a.err[R.gs[u]];
                est = est + 0.0f;  // -0.0 -> +0.0 so keys order like the floats compare
                key = f2key(est);
                pass = !has_thr || key > thr || (key == thr && ~R.gs[u] > thr_lo);
                if (pass) {
                    id_reads++;
                    const uint32_t id = a.ids[R.gs[u]];
                    pass = id != PAD;
                    if (pass && a.view.deleted && bit_set(a.view.deleted, a.view.deleted_words, id)) pass = false;
                    if (pass && allow && !bit_set(allow, a.view.allow_words, id)) pass = false;
                }
            }
            admit(pass, key, ~R.gs[u], round);
        }
    };
    // One round: prefetch the next round's codes, score this one, then a single barrier. The
    // stamps of round r live in the words of parity r & 1: a wave already appending in round
    // r + 1 writes the other word, and round r + 2 cannot start before every wave passed
    // barrier r + 1.
    unsigned long long g = 0;
    auto step = [&](Regs& now, Regs& next, uint32_t vb0, uint32_t round) {
        if constexpr (PF) {
            if (vb0 + ROUND_BLOCKS < vb_hi) fetch(vb0 + ROUND_BLOCKS, next);
        } else {
            fetch(vb0, now);
        }
        peek(round, g);
        score(now, round);
        __syncthreads();
        settle(round, g);
    };
    if (t == 0) {
        misc[M_STAMP] = misc[M_STAMP + 1] = 0xFFFFFFFFu;
        misc[M_ADOPT] = misc[M_ADOPT + 1] = 0xFFFFFFFFu;
    }
    __syncthreads();
    Regs ra, rb;
    uint32_t round = 0;
    if (PF && vb_lo < vb_hi) fetch(vb_lo, ra);
    for (uint32_t vb0 = vb_lo; vb0 < vb_hi;) {
        step(ra, rb, vb0, round++);
        vb0 += ROUND_BLOCKS;
        if (vb0 >= vb_hi) break;
        step(rb, ra, vb0, round++);
        vb0 += ROUND_BLOCKS;
    }
    id_reads = wave_sum(id_reads);
    if (lane == 0 && a.scanned) atomicAdd(a.scanned + 2, (unsigned long long)id_reads);
    if ((int)misc[M_COUNT] > cand) scan_compact(buf, tmp, (int)misc[M_COUNT], cand, hist, misc, gthr);

    if (nch > 1) {
        // Publish this chunk's list; the last chunk of the query to finish merges them all.
        const size_t slot0 = qi * a.chunk_max;
        const int n = (int)misc[M_COUNT];
        uint2* mine = a.part + (slot0 + ci) * (size_t)cand;
        for (int i = t; i < n; i += SCAN_THREADS) mine[i] = buf[i];
        if (t == 0) a.part_n[slot0 + ci] = (uint32_t)n;
        __threadfence();
        __syncthreads();
        if (t == 0) misc[M_LAST] = atomicAdd(a.qdone + qi, 1u) == nch - 1 ? 1u : 0u;
        __syncthreads();
        if (!misc[M_LAST]) return;
        __threadfence();
        // Stream the other chunks' entries through the buffer (own entries are in it already).
        for (uint32_t c = 0; c < nch; c++) {
            if (c == ci) continue;
            const uint2* lst = a.part + (slot0 + c) * (size_t)cand;
            const int m = (int)min((uint32_t)cand, __hip_atomic_load(a.part_n + slot0 + c, __ATOMIC_RELAXED,
                                                                     __HIP_MEMORY_SCOPE_AGENT));
            for (int i0 = 0; i0 < m; i0 += SCAN_THREADS * SLOTS_PER_THREAD) {
                const bool has_thr = misc[M_HAS] != 0;
                const uint32_t thr = misc[M_THR], thr_lo = misc[M_THR_LO];
#pragma unroll
                for (int u = 0; u < SLOTS_PER_THREAD; u++) {
                    const int i = i0 + u * SCAN_THREADS + t;
                    uint2 e = make_uint2(0u, 0u);
                    bool pass = false;
                    if (i < m) {
                        e = lst[i];
                        pass = !has_thr || e.x > thr || (e.x == thr && e.y > thr_lo);
                    }
                    admit(pass, e.x, e.y, round);
                }
                __syncthreads();
                if (misc[M_STAMP + (round & 1)] == round)
                    scan_compact(buf, tmp, (int)misc[M_COUNT], cand, hist, misc, nullptr);
                round++;
            }
        }
        if ((int)misc[M_COUNT] > cand) scan_compact(buf, tmp, (int)misc[M_COUNT], cand, hist, misc, nullptr);
        if (t == 0) {  // ready for the next launch
            __hip_atomic_store(a.qdone + qi, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(gthr, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }

    const int n = (int)misc[M_COUNT];
    if constexpr (L2) {
        // Best first by estimate, ties by slot: (id, est, |x|^2) (src/bigann.rs:732-733).
        const int n2 = pow2_at_least(max(n, 1));
        for (int i = n + t; i < n2; i += SCAN_THREADS) buf[i] = make_uint2(0u, 0u);
        __syncthreads();
        block_sort_pairs(buf, nullptr, n2, false);
        for (int i = t; i < n; i += SCAN_THREADS) {
            const uint32_t slot = ~buf[i].y;
            a.out_est[qi * a.out_stride + i] = key2f(buf[i].x);
            a.out_row[qi * a.out_stride + i] = a.ids[slot];
            a.out_err[qi * a.out_stride + i] = a.err[slot];
        }
        if (t == 0) a.out_n[qi] = (uint32_t)n;
        return;
    }
    // Sort the kept slots by (row asc, est desc) and keep the first of each row.
    const int n2 = pow2_at_least(max(n, 1));
    uint32_t* pay = reinterpret_cast<uint32_t*>(tmp);
    for (int i = t; i < n2; i += SCAN_THREADS) {
        if (i < n) {
            uint2 e = buf[i];
            uint32_t slot = ~e.y;
            buf[i] = make_uint2(a.ids[slot], ~e.x);
            pay[i] = slot;
        } else {
            buf[i] = make_uint2(0xFFFFFFFFu, 0xFFFFFFFFu);
            pay[i] = 0;
        }
    }
    __syncthreads();
    block_sort_pairs(buf, pay, n2, true);
    // Each thread takes a contiguous run of the sorted entries.
    const int per = (n2 + SCAN_THREADS - 1) / SCAN_THREADS;
    const int i0 = t * per;
    uint32_t mine = 0;
    for (int i = i0; i < i0 + per && i < n; i++)
        if (i == 0 || buf[i].x != buf[i - 1].x) mine++;
    uint32_t tot;
    uint32_t off = block_excl_scan(mine, misc + M_SCAN, &tot);
    float* oe = a.out_est + qi * a.out_stride;
    uint32_t* orow = a.out_row + qi * a.out_stride;
    float* oerr = a.out_err + qi * a.out_stride;
    for (int i = i0; i < i0 + per && i < n; i++) {
        if (i == 0 || buf[i].x != buf[i - 1].x) {
            oe[off] = key2f(~buf[i].y);
            orow[off] = buf[i].x;
            oerr[off] = a.err[pay[i]];
            off++;
        }
    }
    if (t == 0) a.out_n[qi] = tot;
}

template <int G, bool L2, int S, bool PF, int WPE>
static void launch_variant(const ScanArgs& a_in, size_t nq, hipStream_t st, int trig4) {
    // The segment stride travels as a plain kernel argument: derived in the kernel from a
    // pointer test, the LDS offsets would be recomputed in vector registers.
    ScanArgs a = a_in;
    a.seg_stride = scan_segments(a);
    a.trigger = std::max(a.cand + 1, a.cand * trig4 / 4);
    const size_t lds = scan_layout(a.cand, a.seg_stride, S, a.trigger).total;
    if (lds > 64 * 1024)
        (void)hipFuncSetAttribute((const void*)k_scan_t<G, L2, S, PF, WPE>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL((k_scan_t<G, L2, S, PF, WPE>), dim3((unsigned)(nq * a.chunk_max)), dim3(SCAN_THREADS), lds,
                       st, a);
}

// (slots per thread, prefetch, minimum waves per SIMD; cut trigger in quarters of cand)
template <int G, bool L2>
static void launch_any(const ScanArgs& a, size_t nq, int variant, hipStream_t st) {
    switch (variant) {
        case 0: launch_variant<G, L2, 4, false, 0>(a, nq, st, 8); break;
        case 1: launch_variant<G, L2, 2, true, 0>(a, nq, st, 8); break;
        case 2: launch_variant<G, L2, 2, false, 0>(a, nq, st, 6); break;
        case 3: launch_variant<G, L2, 1, false, 0>(a, nq, st, 6); break;
        case 4: launch_variant<G, L2, 2, true, 5>(a, nq, st, 8); break;
        case 5: launch_variant<G, L2, 2, false, 0>(a, nq, st, 8); break;
        case 6: launch_variant<G, L2, 1, true, 0>(a, nq, st, 8); break;
        case 7: launch_variant<G, L2, 1, false, 0>(a, nq, st, 8); break;
        case 8: launch_variant<G, L2, 1, true, 0>(a, nq, st, 6); break;
        default: launch_variant<G, L2, 1, false, 0>(a, nq, st, 5); break;
    }
}

ScanPlan plan_scan(size_t nq, uint32_t nprobe, double mean_blocks, uint32_t max_cluster_blocks, int cand,
                   uint32_t forced_chunk_blocks) {
    // A chunk holds at least 8 * cand slots (and 64 blocks), so a split query's merge reads at
    // most an eighth of what its chunks scanned.
    const double chmin = std::max(64.0, std::ceil(8.0 * cand / BLOCK));
    // About 2048 workgroups per batch (8 per CU), and in large batches queries of more than
    // twice the expected work split, so the batch's tail is not one heavy query.
    const double expect = std::max(1.0, (double)nprobe * mean_blocks);
    const double ch = forced_chunk_blocks ? (double)forced_chunk_blocks
                                          : std::max(chmin, std::min(expect * (double)nq / 2048.0, 2.0 * expect));
    const double worst = (double)nprobe * std::max<uint32_t>(max_cluster_blocks, 1);
    double cmax = std::max(1.0, std::ceil(worst / ch));
    // Partial lists stay within 1 GiB per stream.
    const double cap = std::floor((double)(1u << 30) / (8.0 * cand * (double)std::max<size_t>(nq, 1)));
    cmax = std::max(1.0, std::min({cmax, cap, 4096.0}));
    ScanPlan p;
    p.chunk_max = (uint32_t)cmax;
    p.chunk_blocks = (uint32_t)std::min(ch, 4.0e9);
    return p;
}

const char* scan_variant_name(int v) {
    static const char* names[SCAN_VARIANTS] = {
        "4 slots/thread", "2 slots/thread + prefetch", "2 slots/thread, cut at 1.5 cand", "1 slot/thread, cut at 1.5 cand",
        "2 slots/thread + prefetch, >= 5 waves/SIMD", "2 slots/thread", "1 slot/thread + prefetch", "1 slot/thread",
        "1 slot/thread + prefetch, cut at 1.5 cand", "1 slot/thread, cut at 1.25 cand"};
    return v >= 0 && v < SCAN_VARIANTS ? names[v] : "?";
}

void launch_scan(const ScanArgs& a, size_t nq, int variant, hipStream_t st) {
    if (nq == 0) return;
    launch_any<64, false>(a, nq, variant, st);
}

void launch_scan_l2(const ScanArgs& a, size_t nq, int variant, hipStream_t st) {
    if (nq == 0) return;
    launch_any<32, true>(a, nq, variant, st);
}

// ------------------------------------------------------------------ lut::scan seam -----
// Over the reference's nibble-transposed blocks: one 256-thread workgroup scores 8 blocks;
// the blocks are staged in LDS with a 16-byte pad per block so the two half-waves of
// different blocks fall in different banks.
__global__ __launch_bounds__(256) void k_lut_scan(const uint8_t* __restrict__ tables, int G,
                                                  const uint8_t* __restrict__ blocks, size_t nblocks,
                                                  uint16_t* __restrict__ out) {
    __shared__ uint8_t tab[64 * 16];
    __shared__ __align__(16) uint8_t blk[8][1024 + 16];
    const int t = threadIdx.x;
    const size_t b0 = (size_t)blockIdx.x * 8;
    const int bb = G * 16;
    for (int i = t; i < bb; i += 256) tab[i] = tables[i];
    const int nb = (int)min((size_t)8, nblocks - b0);
    for (int i = t; i < nb * bb / 16; i += 256) {
        int b = i / (bb / 16), o = i % (bb / 16);
        *reinterpret_cast<uint4*>(&blk[b][o * 16]) = reinterpret_cast<const uint4*>(blocks + (b0 + b) * bb)[o];
    }
    __syncthreads();
    const int b = t >> 5, k = t & 31;
    if (b < nb) {
        uint32_t sc = 0;
        for (int g = 0; g < G; g++) {
            uint8_t x = blk[b][g * 16 + (k & 15)];
            uint8_t nib = k < 16 ? (x & 15) : (x >> 4);
            sc += tab[g * 16 + nib];
        }
        out[(b0 + b) * BLOCK + k] = (uint16_t)sc;
    }
}

void launch_lut_scan(const uint8_t* tables, int groups, const uint8_t* blocks, size_t nblocks, uint16_t* out,
                     hipStream_t st) {
    if (nblocks == 0) return;
    hipLaunchKernelGGL(k_lut_scan, dim3((unsigned)((nblocks + 7) / 8)), dim3(256), 0, st, tables, groups, blocks,
                       nblocks, out);
}

}  // namespace sieve
