// The scan: Index::candidates (src/ivf.rs:643-729) for a batch, plus the standalone lut::scan
// seam (src/lut.rs:96-127).
//
// Work items. A query's probed clusters, taken in probe order, form one sequence of blocks;
// it is cut into C_q = clamp(ceil(T_q / chunk_blocks), 1, chunk_max) chunks of nearly equal
// size (T_q blocks in all), and one 256-thread workgroup scans each chunk, 256 * S slots per
// round (S slots per thread, a compile-time variant), one lane per slot:
//   * score: the 64 nibbles of the slot's 32-byte code index the query's 64 16-entry u8
//     tables, held in LDS (all lanes of a wave read the same 16-byte table row at a time, so
//     the reads are broadcast and conflict-free); the u16 sum equals score_code
//     (src/lut.rs:58-61) exactly;
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
// 16..17 adopt stamps, 18 entries at the buffer's front whose ids were checked, 19 live count.
constexpr int M_COUNT = 0, M_HAS = 1, M_THR = 2, M_TMP = 3, M_SCAN = 4, M_SEL = 8, M_STAMP = 12, M_THR_LO = 14,
              M_LAST = 15, M_ADOPT = 16, M_VALID = 18, M_LIVE = 19;

__device__ __forceinline__ unsigned long long pack_pair(uint2 e) {
    return ((unsigned long long)e.x << 32) | e.y;
}

// Cuts buf[0, n) to its best k entries (under (est key, ~slot)) and, with `raise`, raises the
// threshold to the k-th; with a per-query shared threshold, publishes it there. Dead entries
// (0, 0) rank below every live one; k never exceeds the live entries, so none is kept.
__device__ void scan_compact(uint2* buf, uint2* tmp, int n, int k, uint32_t* hist, uint32_t* misc,
                             unsigned long long* gthr, bool raise_thr = true) {
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
        misc[M_VALID] = (uint32_t)k;
        if (raise_thr) {
            const bool raise =
                !misc[M_HAS] || kth.x > misc[M_THR] || (kth.x == misc[M_THR] && kth.y > misc[M_THR_LO]);
            if (raise) {
                misc[M_HAS] = 1;
                misc[M_THR] = kth.x;
                misc[M_THR_LO] = kth.y;
            }
            if (gthr) atomicMax(gthr, pack_pair(kth));
        }
    }
    __syncthreads();
}

// G = 64 (256-bit codes, inner-product estimate) or G = 32 (128-bit codes of the padded
// 100-d int8 vectors, L2 estimate 2 (cent + mul (a sc + b)) - |x|^2, src/bigann.rs:705-711).
// SLOTS_PER_THREAD slots per thread per round; PF: prefetch the next round's codes into
// registers while scoring the current one (more bytes in flight, more VGPRs); WPE: minimum
// waves per SIMD the register allocation must allow (0: the compiler's choice).
// MODE: SCAN_FULL scans every probe; SCAN_SEED scans the first probe_lim[q] probes and leaves
// the kept list for the cluster pass; SCAN_TIGHTEN (between cluster-pass phases) merges the
// kept list with the entries appended since and keeps the best cand, raising the threshold;
// SCAN_FINISH merges them one last time (or rescans the query when an appended list
// overflowed) and writes the output.
// SCAN_BOUND (small batches): one round per 32-block chunk (more for queries with more chunks
// than the grid holds); writes the chunk's best min(16, cand) admitted entries to part (16 per
// chunk), a subset of the query's admissible entries.
// SCAN_APPEND (small batches): appends every slot above the query's threshold (seed_thr) to its
// appended list unchecked, for the finish pass.
constexpr int SCAN_FULL = 0, SCAN_SEED = 1, SCAN_FINISH = 2, SCAN_TIGHTEN = 3, SCAN_BOUND = 4, SCAN_APPEND = 5;
// A segment's start in the LDS table with this bit set is an offset into gather_slots (its
// blocks are gathered blocks of 32 listed slots), else a block index of the slot arrays.
constexpr uint32_t GATHER_BIT = 0x80000000u;
constexpr int BOUND_PER_CHUNK = 16;
// seed_n flag: an appended list overflowed, so the query is rescanned by the finish pass.
constexpr uint32_t SEED_LOST = 0x80000000u;

template <int G, bool L2, int SLOTS_PER_THREAD, bool PF, int WPE, int MODE = SCAN_FULL>
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
    int nprobed = (int)a.probe_n[qi];
    bool ext_ok = false;  // SCAN_FINISH / SCAN_TIGHTEN: no appended entry was lost
    if constexpr (MODE == SCAN_SEED) nprobed = min(nprobed, (int)a.probe_lim[qi]);
    if constexpr (MODE == SCAN_FINISH || MODE == SCAN_TIGHTEN) {
        ext_ok = a.ext_n[qi] <= a.ext_cap && !(a.seed_n[qi] & SEED_LOST);
        if (ext_ok || MODE == SCAN_TIGHTEN) nprobed = 0;
    }
    const int np = (live ? 2 : 1) * nprobed;
    for (int j = t; j < np; j += SCAN_THREADS) {
        const int pj = live ? j >> 1 : j;
        const uint32_t c = a.probe_c[qi * a.probe_stride + pj];
        if (!L2 && (!live || (j & 1) == 0) && a.gather_off) {  // filters serve the inner-product index
            const uint32_t g0 = a.gather_off[c];
            pstart[j] = GATHER_BIT | g0;
            pref[j] = (a.gather_off[c + 1] - g0) / BLOCK;
        } else if (!live || (j & 1) == 0) {
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
    if (ci >= nch) {  // the whole workgroup: this query has fewer chunks
        if (MODE == SCAN_BOUND && t == 0) a.part_n[qi * a.chunk_max + ci] = 0;
        return;
    }
    const uint32_t vb_lo = (uint32_t)((uint64_t)total * ci / nch);
    const uint32_t vb_hi = (uint32_t)((uint64_t)total * (ci + 1) / nch);
    if (t < G * 4) reinterpret_cast<uint32_t*>(tab)[t] = reinterpret_cast<const uint32_t*>(a.tables + qi * G * 16)[t];
    const float4 qab = a.qab[qi];
    if (MODE != SCAN_BOUND && t == 0 && a.scanned) atomicAdd(a.scanned, (unsigned long long)(vb_hi - vb_lo) * BLOCK);
    unsigned long long* gthr = MODE == SCAN_FULL && nch > 1 ? a.gthr + qi : nullptr;
    // SCAN_APPEND: the query's threshold (0: none)
    const unsigned long long athr = MODE == SCAN_APPEND ? a.seed_thr[qi] : 0ull;

    const uint64_t* allow = nullptr;
    if (a.view.allow && (a.view.allow_stride == 0 || !a.view.allow_flags || a.view.allow_flags[qi]))
        allow = a.view.allow + qi * a.view.allow_stride;
    const int cand = a.cand;
    uint32_t id_reads = 0;  // slots past the threshold: their id is read
    // Without a per-query allowlist the admission's id check (src/ivf.rs:669) is deferred: an
    // entry enters the buffer on its (estimate, slot) alone, and the ids of the entries past the
    // checked prefix are read in one batch when the buffer is cut and at the end; padding slots
    // and rows the view hides become dead entries (0, 0), which rank below every live entry.
    // The kept set is the best cand live entries either way, and no id load sits between one
    // round's code loads and the next's. (With an allowlist most passing slots are usually
    // hidden, so they are checked at once and never fill the buffer.)
    const bool lazy = allow == nullptr;

    // Checks the buffer's unchecked entries; returns the live count.
    auto validate = [&]() -> int {
        const int n = (int)misc[M_COUNT], v0 = (int)misc[M_VALID];
        if (!lazy || v0 >= n) return n;
        uint32_t dead = 0;
        for (int i0 = v0; i0 < n; i0 += 4 * SCAN_THREADS) {
            uint32_t id[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int i = i0 + u * SCAN_THREADS + t;
                id[u] = i < n ? a.ids[~buf[i].y] : PAD;
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int i = i0 + u * SCAN_THREADS + t;
                if (i >= n) continue;
                id_reads++;
                const bool ok = id[u] != PAD && !(a.view.deleted && bit_set(a.view.deleted, a.view.deleted_words, id[u]));
                if (!ok) {
                    buf[i] = make_uint2(0u, 0u);
                    dead++;
                }
            }
        }
        dead = wave_sum(dead);
        if (lane == 0 && dead) atomicAdd(&misc[M_LIVE], dead);
        __syncthreads();
        const int live = n - (int)misc[M_LIVE];
        __syncthreads();
        if (t == 0) {
            misc[M_LIVE] = 0;
            misc[M_VALID] = (uint32_t)n;
        }
        return live;
    };
    // Cuts the buffer to its best k live entries, raising the threshold (and publishing it to
    // `gt`), when more than k are live; otherwise only drops the dead entries.
    auto cut = [&](int k, unsigned long long* gt) {
        const int n = (int)misc[M_COUNT];
        const int live = validate();
        if (live > k) {
            scan_compact(buf, tmp, n, k, hist, misc, gt);
        } else if (live < n) {
            if (live > 0) {
                scan_compact(buf, tmp, n, live, hist, misc, nullptr, false);
            } else {
                if (t == 0) misc[M_COUNT] = misc[M_VALID] = 0;
                __syncthreads();
            }
        }
    };

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
            base = (uint32_t)__builtin_amdgcn_readlane((int)base, 0);  // every lane is active here
            if (pass) buf[base + before] = make_uint2(key, lo);
        }
    };
    // After the barrier of round `round`: compact if stamped, then adopt a better shared
    // threshold that thread 0 read during the round (`g`, valid in thread 0 only).
    auto settle = [&](uint32_t round, unsigned long long g) {
        if (misc[M_STAMP + (round & 1)] == round) cut(cand, gthr);
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
    // the local one (before the round's barrier). The load is issued at the start of the round, before the round's code loads: loads retire
    // in order, so a wait for it issued after them would also wait for the next round's data.
    auto peek_load = [&](unsigned long long& g) {
        if (gthr && t == 0) g = __hip_atomic_load(gthr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    };
    auto peek = [&](uint32_t round, const unsigned long long g) {
        if (gthr && t == 0) {
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
        float x2[L2 ? SLOTS_PER_THREAD : 1];  // L2: the slot's |x|^2, loaded with its code
        uint32_t gs[SLOTS_PER_THREAD];
        bool ok[SLOTS_PER_THREAD];
    };
    auto fetch = [&](uint32_t vb0, Regs& R) {
#pragma unroll
        for (int u = 0; u < SLOTS_PER_THREAD; u++) {
            const uint32_t vb = vb0 + (t >> 5) + 8 * u;
            R.ok[u] = vb < vb_hi;
            // Past the chunk's end a lane re-reads the chunk's last block (in bounds): no load is
            // conditional, so the wait for one round's data leaves the next round's loads in flight.
            const uint32_t vbc = min(vb, vb_hi - 1);
            while (cur + 1 < np && pref[cur + 1] <= vbc) cur++;
            const uint32_t ps = pstart[cur], off = (vbc - pref[cur]) * BLOCK + (t & 31);
            R.gs[u] = ps * BLOCK + off;
            if constexpr (!L2) {
                if (ps & GATHER_BIT) {
                    // a listed slot; a padding entry reads slot 0 and is not scored
                    const uint32_t sl = a.gather_slots[(ps & ~GATHER_BIT) + off];
                    R.ok[u] = R.ok[u] && sl != PAD;
                    R.gs[u] = sl != PAD ? sl : 0u;
                }
            }
            R.cent[u] = pcent[cur];
            const uint4* cp = a.codes + CW * (size_t)R.gs[u];
#pragma unroll
            for (int w = 0; w < CW; w++) R.cw[u][w] = cp[w];
            R.mul[u] = a.mul[R.gs[u]];
            if constexpr (L2) R.x2[u] = a.err[R.gs[u]];
        }
    };
    auto score = [&](const Regs& R, uint32_t round) {
        const bool has_thr = misc[M_HAS] != 0;
        const uint32_t thr = misc[M_THR], thr_lo = misc[M_THR_LO];
        // Every slot's estimate and threshold test, then the id reads of all passing slots at
        // once (and the view words they need), then the admissions.
        uint32_t key[SLOTS_PER_THREAD];
        bool pass[SLOTS_PER_THREAD];
#pragma unroll
        for (int u = 0; u < SLOTS_PER_THREAD; u++) {
            pass[u] = false;
            key[u] = 0;
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
                    for (int j = 0; j < 8; j++) {
                        constexpr bool chunked = G == 64;  // lane-chunked codes
                        const int g = chunked ? code_word_group(d, j) : 8 * d + j;
                        sc += tab[g * 16 + ((w[d] >> (4 * j)) & 15u)];
                    }
                float est = R.cent[u] + R.mul[u] * (qab.x * (float)sc + qab.y);
                if constexpr (L2) est = 2.0f * est - R.x2[u];
                est = est + 0.0f;  // -0.0 -> +0.0 so keys order like the floats compare
                key[u] = f2key(est);
                if constexpr (MODE == SCAN_APPEND) {
                    pass[u] = ((((unsigned long long)key[u]) << 32) | (unsigned long long)(~R.gs[u])) > athr;
                } else {
                    pass[u] = !has_thr || key[u] > thr || (key[u] == thr && ~R.gs[u] > thr_lo);
                }
            }
        }
        if (MODE != SCAN_APPEND && !lazy) {
            uint32_t id[SLOTS_PER_THREAD];
#pragma unroll
            for (int u = 0; u < SLOTS_PER_THREAD; u++) id[u] = pass[u] ? a.ids[R.gs[u]] : PAD;
#pragma unroll
            for (int u = 0; u < SLOTS_PER_THREAD; u++) {
                if (pass[u]) id_reads++;
                pass[u] = pass[u] && id[u] != PAD;
            }
            if (a.view.deleted || allow) {
                uint64_t dw[SLOTS_PER_THREAD], aw[SLOTS_PER_THREAD];
#pragma unroll
                for (int u = 0; u < SLOTS_PER_THREAD; u++) {
                    const uint64_t wi = id[u] >> 6;
                    dw[u] = pass[u] && a.view.deleted && wi < a.view.deleted_words ? a.view.deleted[wi] : 0ull;
                    aw[u] = pass[u] && allow && wi < a.view.allow_words ? allow[wi] : 0ull;
                }
#pragma unroll
                for (int u = 0; u < SLOTS_PER_THREAD; u++) {
                    const uint64_t bit = 1ull << (id[u] & 63);
                    if (dw[u] & bit) pass[u] = false;
                    if (allow && !(aw[u] & bit)) pass[u] = false;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < SLOTS_PER_THREAD; u++) {
            if constexpr (MODE == SCAN_APPEND) {
                const uint64_t bal = __ballot(pass[u]);
                if (bal) {
                    const uint32_t nw = (uint32_t)__popcll(bal);
                    const uint32_t before = (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
                    uint32_t base = 0;
                    if (lane == 0) base = atomicAdd(a.ext_n + qi, nw);
                    base = (uint32_t)__builtin_amdgcn_readlane((int)base, 0);
                    if (pass[u] && base + before < a.ext_cap)
                        a.ext[qi * (size_t)a.ext_cap + base + before] = make_uint2(key[u], ~R.gs[u]);
                }
            } else {
                admit(pass[u], key[u], ~R.gs[u], round);
            }
        }
    };
    // One round: prefetch the next round's codes, score this one, then a single barrier. The
    // stamps of round r live in the words of parity r & 1: a wave already appending in round
    // r + 1 writes the other word, and round r + 2 cannot start before every wave passed
    // barrier r + 1.
    unsigned long long g = 0;
    auto step = [&](Regs& now, Regs& next, uint32_t vb0, uint32_t round) {
        peek_load(g);
        // The last round's prefetch re-reads the chunk's last block (unused): no condition
        // around the loads.
        if constexpr (PF) fetch(vb0 + ROUND_BLOCKS, next);
        else fetch(vb0, now);
        score(now, round);
        peek(round, g);
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
    // Streams m listed entries through the buffer under the running threshold; with check_ids
    // (entries the cluster pass appended unchecked) an entry past the threshold is admitted only
    // if its slot holds a row the view shows (src/ivf.rs:669).
    // `coherent`: the list was written by another workgroup of this launch, with agent-scope
    // atomic stores; its entries are read with agent-scope atomic loads.
    auto merge_list = [&](const uint2* lst, int m, bool check_ids, bool coherent = false) {
        for (int i0 = 0; i0 < m; i0 += SCAN_THREADS * SLOTS_PER_THREAD) {
            const bool has_thr = misc[M_HAS] != 0;
            const uint32_t thr = misc[M_THR], thr_lo = misc[M_THR_LO];
            // The round's entries, then all their id reads at once, then the admissions.
            uint2 e[SLOTS_PER_THREAD];
            bool pass[SLOTS_PER_THREAD];
#pragma unroll
            for (int u = 0; u < SLOTS_PER_THREAD; u++) {
                const int i = i0 + u * SCAN_THREADS + t;
                if (i < m && coherent) {
                    const unsigned long long v =
                        __hip_atomic_load(reinterpret_cast<unsigned long long*>(const_cast<uint2*>(lst)) + i,
                                          __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    e[u] = make_uint2((uint32_t)v, (uint32_t)(v >> 32));
                } else {
                    e[u] = i < m ? lst[i] : make_uint2(0u, 0u);
                }
                pass[u] = i < m && (!has_thr || e[u].x > thr || (e[u].x == thr && e[u].y > thr_lo));
            }
            if (check_ids) {
                uint32_t id[SLOTS_PER_THREAD];
#pragma unroll
                for (int u = 0; u < SLOTS_PER_THREAD; u++) id[u] = pass[u] ? a.ids[~e[u].y] : PAD;
#pragma unroll
                for (int u = 0; u < SLOTS_PER_THREAD; u++) {
                    if (!pass[u]) continue;
                    id_reads++;
                    pass[u] = id[u] != PAD &&
                              !(a.view.deleted && bit_set(a.view.deleted, a.view.deleted_words, id[u]));
                }
            }
#pragma unroll
            for (int u = 0; u < SLOTS_PER_THREAD; u++) admit(pass[u], e[u].x, e[u].y, round);
            __syncthreads();
            if (misc[M_STAMP + (round & 1)] == round)
                scan_compact(buf, tmp, (int)misc[M_COUNT], cand, hist, misc, nullptr);
            round++;
        }
        // Every listed entry was checked (a published list, or the ids read above).
        if (t == 0) misc[M_VALID] = misc[M_COUNT];
        __syncthreads();
    };
    if (PF && vb_lo < vb_hi) fetch(vb_lo, ra);
    for (uint32_t vb0 = vb_lo; vb0 < vb_hi;) {
        step(ra, rb, vb0, round++);
        vb0 += ROUND_BLOCKS;
        if (vb0 >= vb_hi) break;
        step(rb, ra, vb0, round++);
        vb0 += ROUND_BLOCKS;
    }
    if constexpr (MODE == SCAN_APPEND) return;
    if constexpr (MODE == SCAN_BOUND) {
        // The chunk's best entries: every one is an admissible slot (id read, view checked).
        const int b = min(BOUND_PER_CHUNK, cand);
        cut(b, nullptr);
        const int n = min((int)misc[M_COUNT], b);
        const size_t at = qi * a.chunk_max + ci;
        if (t < n) a.part[at * BOUND_PER_CHUNK + t] = buf[t];
        if (t == 0) a.part_n[at] = (uint32_t)n;
        return;
    }
    cut(cand, gthr);
    id_reads = wave_sum(id_reads);
    if (lane == 0 && a.scanned) atomicAdd(a.scanned + 2, (unsigned long long)id_reads);
    id_reads = 0;

    if (nch > 1) {
        // Publish this chunk's list; the last chunk of the query to finish merges them all.
        // The list and its length are written with agent-scope atomic stores, which are coherent
        // across the XCDs' caches, and every store has completed (s_waitcnt) before the chunk
        // counts itself done; the last chunk reads them with agent-scope atomic loads. No
        // agent-scope fence: on this GPU one writes back and invalidates the whole L2 of the
        // XCD, stalling every workgroup there (a batch of split queries issues thousands).
        const size_t slot0 = qi * a.chunk_max;
        const int n = (int)misc[M_COUNT];
        unsigned long long* mine = reinterpret_cast<unsigned long long*>(a.part + (slot0 + ci) * (size_t)cand);
        for (int i = t; i < n; i += SCAN_THREADS)
            __hip_atomic_store(mine + i, ((unsigned long long)buf[i].y << 32) | buf[i].x, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        if (t == 0) __hip_atomic_store(a.part_n + slot0 + ci, (uint32_t)n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_s_waitcnt(0);  // every store of this thread acknowledged
        __syncthreads();
        if (t == 0)
            misc[M_LAST] = __hip_atomic_fetch_add(a.qdone + qi, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
                                   nch - 1
                               ? 1u
                               : 0u;
        __syncthreads();
        if (!misc[M_LAST]) return;
        // Stream the other chunks' entries through the buffer (own entries are in it already).
        for (uint32_t c = 0; c < nch; c++) {
            if (c == ci) continue;
            const int m = (int)min((uint32_t)cand, __hip_atomic_load(a.part_n + slot0 + c, __ATOMIC_RELAXED,
                                                                     __HIP_MEMORY_SCOPE_AGENT));
            merge_list(a.part + (slot0 + c) * (size_t)cand, m, false, true);
        }
        if ((int)misc[M_COUNT] > cand) scan_compact(buf, tmp, (int)misc[M_COUNT], cand, hist, misc, nullptr);
        if (t == 0) {  // ready for the next launch
            __hip_atomic_store(a.qdone + qi, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(gthr, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    if constexpr (MODE == SCAN_FINISH || MODE == SCAN_TIGHTEN) {
        if (ext_ok) {
            merge_list(a.seed + qi * (size_t)cand, (int)min((uint32_t)cand, a.seed_n[qi]), false);
            merge_list(a.ext + qi * (size_t)a.ext_cap, (int)a.ext_n[qi], true);
            if ((int)misc[M_COUNT] > cand) scan_compact(buf, tmp, (int)misc[M_COUNT], cand, hist, misc, nullptr);
            id_reads = wave_sum(id_reads);
            if (lane == 0 && a.scanned) atomicAdd(a.scanned + 2, (unsigned long long)id_reads);
        } else if (MODE == SCAN_FINISH && t == 0 && a.scanned) {
            atomicAdd(a.scanned + 3, 1ull);  // queries rescanned
        }
        __syncthreads();
        if (t == 0) a.ext_n[qi] = 0;  // ready for the next phase / launch
        if constexpr (MODE == SCAN_TIGHTEN) {
            if (!ext_ok) {  // lost: later phases append nothing for it
                if (t == 0) {
                    a.seed_n[qi] |= SEED_LOST;
                    a.seed_thr[qi] = ~0ull;
                }
                return;
            }
        }
    }
    if constexpr (MODE == SCAN_SEED || MODE == SCAN_TIGHTEN) {
        // The kept list and, when it is full, its worst entry: cand admitted slots compare at
        // least as high, so a later slot at or below it cannot enter the final list.
        const int n = (int)misc[M_COUNT];
        uint2* dst = a.seed + qi * (size_t)cand;
        unsigned long long lo = ~0ull;
        for (int i = t; i < n; i += SCAN_THREADS) {
            dst[i] = buf[i];
            lo = min(lo, pack_pair(buf[i]));
        }
        lo = wave_min_u64(lo);
        __syncthreads();
        unsigned long long* red = reinterpret_cast<unsigned long long*>(hist);
        if (lane == 0) red[wave] = lo;
        __syncthreads();
        if (t == 0) {
            for (int w = 1; w < SCAN_THREADS / 64; w++) lo = min(lo, red[w]);
            a.seed_n[qi] = (uint32_t)n;
            a.seed_thr[qi] = n >= cand ? lo : 0ull;
        }
        return;
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
                   uint32_t forced_chunk_blocks, uint32_t min_chunk_blocks) {
    // A chunk holds at least 24 * cand slots (and 64 blocks): a split query's last chunk merges
    // up to cand entries from each of the others alone, so fewer, larger chunks balance that
    // serial merge against the parallel scan (a lone 113.5M query: ~35 chunks).
    const double chmin = std::max({64.0, std::ceil(24.0 * cand / BLOCK), (double)min_chunk_blocks});
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
        "1 slot/thread + prefetch, cut at 1.5 cand", "1 slot/thread, cut at 1.25 cand",
        "cluster-major MFMA, split chains, integer bound epilogue", "cluster-major MFMA, one chain per column half, f32 epilogue"};
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

// ------------------------------------------------------------ cluster-major scan -----
// In a large batch a stored slot is probed by many queries (tens, at 10,000 queries over 16,384
// clusters), and the per-query scan reads its code once for each. The cluster-major pass reads
// a block chunk of one segment (a cluster's base or delta list) once for a tile of up to 32 of
// the queries that probe it and scores the tile with int8 MFMA: for 16 slots x 16 queries x 4
// groups per v_mfma_i32_16x16x64_i8, A holds the one-hot bytes of each slot's nibble per group
// and B the query's table entries minus 128, so the product sums T[g][nibble_g] - 128 over the
// 64 groups: score_code (src/lut.rs:58-61) minus 64 * 128, exactly (|sum| <= 8192 in i32).
//
// The pass needs each query's admission threshold up front. The seed pass (k_scan_t in
// SCAN_SEED mode) scans the query's nearest probes (whole clusters, at least seed_slots slots)
// and keeps their best cand entries; the cand-th of those is a lower bound of the final
// cand-th best, so a later slot at or below it never enters the result, and the cluster pass
// appends only (est, slot) pairs above it. The finish pass merges the seed list and the
// appended list through the same TopK buffer as the per-query scan (the union holds every
// final entry, and the order (est desc, slot asc) is total, so the kept set is the same), or,
// if the appended list overflowed its capacity, rescans the query in full.

typedef int v4i __attribute__((ext_vector_type(4)));

struct CmArgs {
    const uint4* codes;
    const float* mul;
    const uint32_t* ids;
    const uint8_t* tables;
    const float4* qab;
    const uint4* items;          // (first slot, slots, first seg_q entry, queries)
    uint32_t items_cap;
    const uint32_t* seg_ioff;
    const uint32_t* seg_q;
    const float* seg_cent;
    uint32_t nseg;
    const unsigned long long* thr;
    ViewArgs view;
    uint2* ext;
    uint32_t* ext_n;
    uint32_t ext_cap;
    uint32_t* work;              // phases counters
    unsigned long long* scanned;
    unsigned long long* pairs;   // (query, slot) pairs scored, the MFMA pass's own count
    unsigned long long* staged;  // pairs its integer bound passed to the exact estimate
    uint32_t phase, phases;      // this launch runs work items [total * phase / phases, ...)
};

__device__ __forceinline__ void cm_segment(const uint32_t* block_off, const uint32_t* dblock_off, uint32_t dbase,
                                           uint32_t k, uint32_t sg, uint32_t& b0, uint32_t& nb) {
    if (sg < k) {
        b0 = block_off[sg];
        nb = block_off[sg + 1] - b0;
    } else {
        const uint32_t d0 = dblock_off[sg - k];
        b0 = dbase + d0;
        nb = dblock_off[sg - k + 1] - d0;
    }
}

__global__ void k_cm_zero(uint32_t* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ __launch_bounds__(256) void k_pack_stats(const unsigned long long* cnt, const uint32_t* pn, size_t nq,
                                                    const uint32_t* ovf, unsigned long long* out) {
    __shared__ unsigned long long part[4];
    unsigned long long sum = 0;
    for (size_t i = threadIdx.x; i < nq; i += 256) sum += pn[i];
    for (int d = 32; d >= 1; d >>= 1) sum += __shfl_xor(sum, d, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = sum;
    __syncthreads();
    if (threadIdx.x < 4) out[threadIdx.x] = cnt[threadIdx.x];
    if (threadIdx.x == 0) {
        out[4] = part[0] + part[1] + part[2] + part[3];
        out[5] = *ovf;
        out[6] = cnt[4];
        out[7] = cnt[5];
    }
}

void launch_pack_stats(const unsigned long long* cnt, const uint32_t* pn, size_t nq, const uint32_t* ovf,
                       unsigned long long* out, hipStream_t st) {
    hipLaunchKernelGGL(k_pack_stats, dim3(1), dim3(256), 0, st, cnt, pn, nq, ovf, out);
}

void launch_zero_words(uint32_t* p, uint32_t n, hipStream_t st) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_cm_zero, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
}

// One wave per query: the seed pass's probe count (whole probes, in probe order, until
// seed_slots stored slots) and the per-segment counts of the later probes.
__global__ __launch_bounds__(256) void k_cm_prep(ScanArgs a, CmScratch cm) {
    const size_t q = (size_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (q >= cm.nq) return;
    const int pn = (int)a.probe_n[q];
    uint32_t run = 0, lim = (uint32_t)pn;
    bool found = false;
    for (int base = 0; base < pn && !found; base += 64) {
        const int p = base + lane;
        uint32_t nb = 0;
        if (p < pn) {
            const uint32_t c = a.probe_c[q * a.probe_stride + p];
            nb = a.block_off[c + 1] - a.block_off[c];
            if (a.dblock_off) nb += a.dblock_off[c + 1] - a.dblock_off[c];
        }
        const uint32_t incl = wave_incl_scan(nb) + run;
        const uint64_t hit = __ballot(p < pn && (uint64_t)incl * BLOCK >= cm.seed_slots);
        if (hit) {
            lim = (uint32_t)(base + __ffsll((long long)hit));  // through the first probe that reaches it
            found = true;
        }
        run = __shfl(incl, 63, 64);
    }
    if (lane == 0) cm.probe_lim[q] = lim;
    for (int p = (int)lim + lane; p < pn; p += 64) {
        const uint32_t c = a.probe_c[q * a.probe_stride + p];
        if (a.block_off[c + 1] > a.block_off[c]) atomicAdd(cm.seg_count + c, 1u);
        if (a.dblock_off && a.dblock_off[c + 1] > a.dblock_off[c]) atomicAdd(cm.seg_count + cm.k + c, 1u);
    }
}

// One workgroup: prefix sums of the per-segment query counts and work items; the counts are
// cleared for the scatter, which uses them as cursors.
__global__ __launch_bounds__(SEL_THREADS) void k_cm_offsets(ScanArgs a, CmScratch cm) {
    __shared__ uint32_t scratch[16];
    const int t = threadIdx.x;
    const uint32_t n = cm.nseg, per = (n + SEL_THREADS - 1) / SEL_THREADS;
    const uint32_t lo = min(n, t * per), hi = min(n, lo + per);
    uint32_t sq = 0, si = 0;
    for (uint32_t sg = lo; sg < hi; sg++) {
        uint32_t b0, nb;
        cm_segment(a.block_off, a.dblock_off, a.dbase, cm.k, sg, b0, nb);
        const uint32_t qc = cm.seg_count[sg];
        sq += qc;
        si += ((qc + 31) / 32) * ((nb + cm.chunk_blocks - 1) / cm.chunk_blocks);
    }
    uint32_t tq, ti;
    uint32_t oq = block_excl_scan(sq, scratch, &tq);
    uint32_t oi = block_excl_scan(si, scratch + 8, &ti);
    for (uint32_t sg = lo; sg < hi; sg++) {
        uint32_t b0, nb;
        cm_segment(a.block_off, a.dblock_off, a.dbase, cm.k, sg, b0, nb);
        const uint32_t qc = cm.seg_count[sg];
        cm.seg_qoff[sg] = oq;
        cm.seg_ioff[sg] = oi;
        oq += qc;
        oi += ((qc + 31) / 32) * ((nb + cm.chunk_blocks - 1) / cm.chunk_blocks);
        cm.seg_count[sg] = 0;
    }
    if (t == 0) {
        cm.seg_qoff[n] = tq;
        cm.seg_ioff[n] = ti;
    }
    if (t < (int)cm.phases) cm.work[t] = 0;
}

// One wave per query: lists each later probe's query and centroid score under its segment.
__global__ __launch_bounds__(256) void k_cm_scatter(ScanArgs a, CmScratch cm) {
    const size_t q = (size_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (q >= cm.nq) return;
    const int pn = (int)a.probe_n[q];
    for (int p = (int)cm.probe_lim[q] + lane; p < pn; p += 64) {
        const uint32_t c = a.probe_c[q * a.probe_stride + p];
        const float cent = a.probe_s[q * a.probe_stride + p];
        if (a.block_off[c + 1] > a.block_off[c]) {
            const uint32_t pos = cm.seg_qoff[c] + atomicAdd(cm.seg_count + c, 1u);
            cm.seg_q[pos] = (uint32_t)q;
            cm.seg_cent[pos] = cent;
        }
        if (a.dblock_off && a.dblock_off[c + 1] > a.dblock_off[c]) {
            const uint32_t sg = cm.k + c;
            const uint32_t pos = cm.seg_qoff[sg] + atomicAdd(cm.seg_count + sg, 1u);
            cm.seg_q[pos] = (uint32_t)q;
            cm.seg_cent[pos] = cent;
        }
    }
}

// One wave per segment: the segment's work items, 32-query tiles fastest (the tiles of one
// block chunk run close together and share its reads in L2).
__global__ __launch_bounds__(256) void k_cm_items(ScanArgs a, CmScratch cm) {
    const uint32_t sg = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (sg >= cm.nseg) return;
    uint32_t b0, nb;
    cm_segment(a.block_off, a.dblock_off, a.dbase, cm.k, sg, b0, nb);
    const uint32_t q0 = cm.seg_qoff[sg], qc = cm.seg_qoff[sg + 1] - q0;
    const uint32_t i0 = cm.seg_ioff[sg], n = cm.seg_ioff[sg + 1] - i0;
    if (n == 0) return;
    const uint32_t tiles = (qc + 31) / 32;
    for (uint32_t i = lane; i < n; i += 64) {
        const uint32_t tile = i % tiles, chunk = i / tiles;
        const uint32_t blo = chunk * cm.chunk_blocks, bhi = min(nb, blo + cm.chunk_blocks);
        if (i0 + i < cm.items_cap)
            cm.items[i0 + i] = make_uint4((b0 + blo) * BLOCK, (bhi - blo) * BLOCK, q0 + tile * 32,
                                          min(32u, qc - tile * 32));
    }
}

// The per-query constants of one C column (query) of a tile half.
struct CmCol {
    float a, b, cent, thrf;
    unsigned long long thr;
    uint32_t q;
};

// Per-wave staging of appended entries, flushed once per work item (or when full): a past-the-
// threshold slot costs an LDS write instead of a global atomic round trip inside the tile loop.
constexpr int CM_STAGE = 256;
// 16-slot tiles whose codes a wave has in flight ahead of the one it scores.
constexpr int CM_PREFETCH = 4;

// Orders this wave's LDS accesses (lanes exchange data through the wave's staging area).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ONEHOT: where the one-hot A operands come from: CM_OH_LDS a 16-entry LDS table (2 VALU +
// 1 LDS read per step); any other value takes the table on even steps and shifts and selects
// (about 7 VALU) on odd ones, halving the LDS reads.
constexpr int CM_OH_LDS = 1;
template <int ONEHOT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_scan_cm(CmArgs a) {
    __shared__ uint4 onehot[16];
    __shared__ uint2 stage_e[4][CM_STAGE];       // (key, ~slot)
    __shared__ uint8_t stage_c[4][CM_STAGE];     // tile column (query) 0..31
    __shared__ uint32_t stage_cnt[4][32], stage_base[4][32], stage_q[4][32];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    if (t < 16) {
        uint32_t w[4] = {0u, 0u, 0u, 0u};
        w[t >> 2] = 1u << (8 * (t & 3));
        onehot[t] = make_uint4(w[0], w[1], w[2], w[3]);
    }
    __syncthreads();
    const int r = lane >> 4, j = lane & 15;
    // The host sizes the descriptor array by a bound of the work items; never read past it.
    const uint32_t total = min(a.seg_ioff[a.nseg], a.items_cap);
    const uint32_t i_lo = (uint32_t)((uint64_t)total * a.phase / a.phases);
    const uint32_t i_hi = (uint32_t)((uint64_t)total * (a.phase + 1) / a.phases);
    uint32_t sn = 0;  // staged entries (the same in every lane)
    // Appends the staged entries to their queries' lists: one global atomic per column.
    auto flush = [&]() {
        if (sn == 0) return;
        if (lane < 32) stage_cnt[wave][lane] = 0;
        wave_lds_sync();
        uint32_t loc[CM_STAGE / 64];
#pragma unroll
        for (int k = 0; k < CM_STAGE / 64; k++) {
            const uint32_t i = lane + 64 * k;
            loc[k] = i < sn ? atomicAdd(&stage_cnt[wave][stage_c[wave][i]], 1u) : 0u;
        }
        wave_lds_sync();
        if (lane < 32) {
            const uint32_t cnt = stage_cnt[wave][lane];
            stage_base[wave][lane] = cnt ? atomicAdd(a.ext_n + stage_q[wave][lane], cnt) : 0u;
        }
        wave_lds_sync();
#pragma unroll
        for (int k = 0; k < CM_STAGE / 64; k++) {
            const uint32_t i = lane + 64 * k;
            if (i < sn) {
                const uint32_t c = stage_c[wave][i];
                const uint32_t at = stage_base[wave][c] + loc[k];
                if (at < a.ext_cap) a.ext[(size_t)stage_q[wave][c] * a.ext_cap + at] = stage_e[wave][i];
            }
        }
        wave_lds_sync();
        sn = 0;
    };
    // The next work item is claimed, and its descriptor read, while the current one runs.
    auto claim = [&]() {
        uint32_t v = 0;
        if (lane == 0) v = i_lo + atomicAdd(a.work + a.phase, 1u);
        return (uint32_t)__builtin_amdgcn_readlane((int)v, 0);
    };
    uint32_t nidx = claim();
    uint4 nitem = nidx < i_hi ? a.items[nidx] : make_uint4(0u, 0u, 0u, 0u);
    for (;;) {
        const uint32_t idx = nidx;
        if (idx >= i_hi) break;
        const uint4 item = nitem;
        nidx = claim();
        if (nidx < i_hi) nitem = a.items[nidx];
        const uint32_t s_lo = item.x, s_hi = item.x + item.y, qpos0 = item.z, nvalid = item.w;
        const bool two = nvalid > 16;
        // B operands: lane (j, r) holds query j's table entries of groups 4 st + r, minus 128.
        uint4 B0[16], B1[16];
        CmCol col[2];
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const uint32_t n = 16 * h + j;
            const bool valid = n < nvalid;
            const uint32_t pos = qpos0 + (valid ? n : 0);
            const uint32_t q = a.seg_q[pos];
            const float4 qab = a.qab[q];
            col[h].a = qab.x;
            col[h].b = qab.y;
            col[h].cent = a.seg_cent[pos];
            col[h].q = q;
            // ~0: nothing passes (a padding column, or a query whose list was lost).
            col[h].thr = valid ? a.thr[q] : ~0ull;
            col[h].thrf = col[h].thr == ~0ull ? INFINITY
                          : col[h].thr == 0   ? -INFINITY
                                              : key2f((uint32_t)(col[h].thr >> 32));
            if (h == 1 && !two) continue;
            const uint4* tp = reinterpret_cast<const uint4*>(a.tables + (size_t)q * 1024);
#pragma unroll
            for (int st = 0; st < 16; st++) {
                uint4 v = tp[4 * st + r];
                v.x ^= 0x80808080u;
                v.y ^= 0x80808080u;
                v.z ^= 0x80808080u;
                v.w ^= 0x80808080u;
                if (h == 0) B0[st] = v;
                else B1[st] = v;
            }
        }
        if (lane == 0 && a.scanned) {
            atomicAdd(a.scanned, (unsigned long long)item.y * nvalid);
            atomicAdd(a.pairs, (unsigned long long)item.y * nvalid);
        }
        if (lane < 16) {
            stage_q[wave][lane] = col[0].q;
            stage_q[wave][lane + 16] = col[1].q;
        }
        wave_lds_sync();

        // Lane (j, r) of C holds slots 4 r .. 4 r + 3 of the 16-slot tile for query column j
        // (column 16 h + j of the tile). Entries past the threshold are staged unchecked: the
        // merge passes drop padding and hidden rows.
        auto epilogue = [&](const v4i& acc, const CmCol& cc, uint32_t slot0, const float4& mu, uint32_t colx) {
            const float m4[4] = {mu.x, mu.y, mu.z, mu.w};
            uint32_t keys[4] = {0u, 0u, 0u, 0u}, mask = 0;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const uint32_t sc = (uint32_t)acc[i];
                float est = cc.cent + m4[i] * (cc.a * (float)sc + cc.b);
                est = est + 0.0f;
                if (est >= cc.thrf) {
                    const uint32_t slot = slot0 + 4 * r + i;
                    const uint32_t key = f2key(est);
                    if ((((unsigned long long)key << 32) | (unsigned long long)(~slot)) > cc.thr) {
                        mask |= 1u << i;
                        keys[i] = key;
                    }
                }
            }
            if (!__ballot(mask != 0)) return;
            const uint32_t c = (uint32_t)__popc(mask);
            const uint32_t incl = wave_incl_scan(c);
            const uint32_t tot = __shfl(incl, 63, 64);
            if (sn + tot > (uint32_t)CM_STAGE) flush();
            uint32_t at = sn + incl - c;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                if ((mask >> i) & 1u) {
                    stage_e[wave][at] = make_uint2(keys[i], ~(slot0 + 4 * r + i));
                    stage_c[wave][at] = (uint8_t)colx;
                    at++;
                }
            }
            sn += tot;
        };
        auto tile16 = [&](const uint2& cw, const float4& mu, uint32_t slot0, auto two_c) {
            constexpr bool TWO = decltype(two_c)::value;
            const uint32_t w[2] = {cw.x, cw.y};
            // The accumulators start at 64 * 128, so they end at score_code itself.
            v4i acc0 = {64 * 128, 64 * 128, 64 * 128, 64 * 128}, acc1 = acc0;
            // The A operand of step st: the one-hot bytes of group 4 st + r of the slot, nibble st
            // of the lane's 8-byte chunk of the code.
            auto operand = [&](int st) -> v4i {
                const uint32_t nib = (w[st >> 3] >> (4 * (st & 7))) & 15u;
                if (ONEHOT == CM_OH_LDS || (st & 1) == 0) {
                    const uint4 oh = onehot[nib];
                    return v4i{(int)oh.x, (int)oh.y, (int)oh.z, (int)oh.w};
                } else {
                    const unsigned long long v = 1ull << ((nib & 7u) * 8u);
                    const bool hi8 = nib >= 8u;
                    return v4i{hi8 ? 0 : (int)(uint32_t)v, hi8 ? 0 : (int)(uint32_t)(v >> 32),
                               hi8 ? (int)(uint32_t)v : 0, hi8 ? (int)(uint32_t)(v >> 32) : 0};
                }
            };
            // LDS reads run PIPE steps ahead of the MFMAs that consume them.
            constexpr int PIPE = 4;
            v4i pipe[PIPE];
#pragma unroll
            for (int st = 0; st < PIPE; st++) pipe[st] = operand(st);
#pragma unroll
            for (int st = 0; st < 16; st++) {
                const v4i av = pipe[st % PIPE];
                if (st + PIPE < 16) pipe[st % PIPE] = operand(st + PIPE);
                const v4i b0v = {(int)B0[st].x, (int)B0[st].y, (int)B0[st].z, (int)B0[st].w};
                acc0 = __builtin_amdgcn_mfma_i32_16x16x64_i8(av, b0v, acc0, 0, 0, 0);
                if constexpr (TWO) {
                    const v4i b1v = {(int)B1[st].x, (int)B1[st].y, (int)B1[st].z, (int)B1[st].w};
                    acc1 = __builtin_amdgcn_mfma_i32_16x16x64_i8(av, b1v, acc1, 0, 0, 0);
                }
            }
            if constexpr (ONEHOT == CM_OH_LDS) {
                // Keep the table reads PIPE steps ahead: the reads and address arithmetic of the
                // prologue first, then per step one read, its address, and the step's MFMAs.
                __builtin_amdgcn_sched_group_barrier(0x002, 2 * PIPE, 0);  // VALU
                __builtin_amdgcn_sched_group_barrier(0x100, PIPE, 0);      // DS read
#pragma unroll
                for (int st = 0; st < 16; st++) {
                    if (st + PIPE < 16) {
                        __builtin_amdgcn_sched_group_barrier(0x002, 2, 0);
                        __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
                    }
                    __builtin_amdgcn_sched_group_barrier(0x008, TWO ? 2 : 1, 0);  // MFMA
                }
            }
            epilogue(acc0, col[0], slot0, mu, (uint32_t)j);
            if constexpr (TWO) epilogue(acc1, col[1], slot0, mu, 16u + j);
        };
        // 16-slot tiles of the chunk in order, the codes and multipliers of the next CM_PREFETCH
        // tiles in flight (lane (j, r): chunk r of slot j's code, mul of slots 4 r .. 4 r + 3).
        constexpr int PF = ONEHOT == CM_OH_LDS ? CM_PREFETCH : CM_PREFETCH - 2;  // registers: no spills
        const uint32_t ntiles = (s_hi - s_lo) / 16;
        uint2 cr[PF];
        float4 mr[PF];
        auto load = [&](uint32_t ti, uint2& c, float4& m) {
            const uint32_t slot0 = s_lo + 16 * ti;
            c = reinterpret_cast<const uint2*>(a.codes + 2 * (size_t)(slot0 + j))[r];
            m = *reinterpret_cast<const float4*>(a.mul + slot0 + 4 * r);
        };
#pragma unroll
        for (int p = 0; p < PF; p++)
            if ((uint32_t)p < ntiles) load(p, cr[p], mr[p]);
        for (uint32_t t0 = 0; t0 < ntiles; t0 += PF) {
#pragma unroll
            for (int p = 0; p < PF; p++) {
                const uint32_t ti = t0 + p;
                if (ti < ntiles) {
                    const uint2 c = cr[p];
                    const float4 m = mr[p];
                    if (ti + PF < ntiles) load(ti + PF, cr[p], mr[p]);
                    if (two) tile16(c, m, s_lo + 16 * ti, std::true_type{});
                    else tile16(c, m, s_lo + 16 * ti, std::false_type{});
                }
            }
        }
        flush();  // before the next item replaces the column queries
    }
}

// The cluster-major MFMA pass, second form (scan variant 10). Against k_scan_cm:
//  - A dependent int8 MFMA waits out the full latency of the one before it (about three issue
//    intervals), so a 16-query tile accumulates its 16 steps in two chains (even and odd steps)
//    summed at the end, and a 32-query tile runs its two column halves as two chains.
//  - The one-hot operands come from the 16-entry table (one LDS row, so reads never conflict)
//    at an offset taken from the code with one masked byte select.
//  - The epilogue does no float work per (query, slot) pair. Each lane compares its four scores
//    with an integer bound tau for its column: a score below tau gives an estimate below the
//    query's threshold. est = cent + m (a sc + b) rises with sc (a > 0, m >= 0); with m between
//    the lane's least and greatest multiplier mn <= m <= mx and T = thr - cent - err (err bounds
//    the f32 rounding of the estimate and of T: 2^-21 (|cent| + |thr| + mx (4 * 16320 a + 3 |b|))
//    is eight times its first-order bound), est >= thr needs m (a sc + b) >= T, hence
//    a sc + b >= T / mx when T > 0 and a sc + b >= T / mn otherwise; tau is that bound on sc
//    less 2 (the error of computing it is far below one unit). Scores at or above tau are staged
//    as (slot, score, column); the flush computes their estimates in the f32 steps of the
//    per-query scan and appends those past the threshold, so the appended lists are exactly
//    k_scan_cm's.
constexpr int CM2_STAGE = 512;

// Bits 4..7 of byte k of x as a byte offset (0, 16, ..., 240): one masked byte select (SDWA).
__device__ __forceinline__ uint32_t nib_off(uint32_t x, int k) {
    const uint32_t m = 0xF0u;
    uint32_t r;
    switch (k) {
        case 0: return x & m;
        case 1: asm("v_and_b32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:BYTE_1 src1_sel:DWORD"
                    : "=v"(r) : "v"(x), "s"(m));
                return r;
        case 2: asm("v_and_b32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:BYTE_2 src1_sel:DWORD"
                    : "=v"(r) : "v"(x), "s"(m));
                return r;
        default: asm("v_and_b32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:BYTE_3 src1_sel:DWORD"
                     : "=v"(r) : "v"(x), "s"(m));
                 return r;
    }
}

struct CmColC {  // a tile column's query constants, for the flush
    float a, b, cent, thrf;
    unsigned long long thr;
    uint32_t q, pad;
};

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_scan_cm2(CmArgs a) {
    __shared__ uint4 onehot[16];              // 256 B: one LDS row, conflict-free reads
    __shared__ uint2 stage[4][CM2_STAGE];     // (slot, score << 5 | column)
    __shared__ CmColC cols[4][32];
    __shared__ uint32_t ccnt[4][32], cbase[4][32];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    if (t < 16) {
        uint32_t w[4] = {0u, 0u, 0u, 0u};
        w[t >> 2] = 1u << (8 * (t & 3));
        onehot[t] = make_uint4(w[0], w[1], w[2], w[3]);
    }
    __syncthreads();
    const int r = lane >> 4, j = lane & 15;
    // The host sizes the descriptor array by a bound of the work items; never read past it.
    const uint32_t total = min(a.seg_ioff[a.nseg], a.items_cap);
    const uint32_t i_lo = (uint32_t)((uint64_t)total * a.phase / a.phases);
    const uint32_t i_hi = (uint32_t)((uint64_t)total * (a.phase + 1) / a.phases);
    uint32_t sn = 0;  // staged entries (the same in every lane)
    // Exact estimates of the staged entries; those past their column's threshold are appended,
    // one global atomic per column and chunk of 64.
    auto flush = [&]() {
        if (lane == 0 && sn && a.staged) atomicAdd(a.staged, (unsigned long long)sn);
        for (uint32_t c0 = 0; c0 < sn; c0 += 64) {
            if (lane < 32) ccnt[wave][lane] = 0;
            wave_lds_sync();
            const uint32_t i = c0 + lane;
            bool pass = false;
            uint32_t key = 0, slot = 0, col = 0, loc = 0;
            if (i < sn) {
                const uint2 e = stage[wave][i];
                slot = e.x;
                col = e.y & 31u;
                const uint32_t sc = e.y >> 5;
                const CmColC cc = cols[wave][col];
                const float m = a.mul[slot];
                float est = cc.cent + m * (cc.a * (float)sc + cc.b);
                est = est + 0.0f;
                if (est >= cc.thrf) {
                    key = f2key(est);
                    pass = (((unsigned long long)key << 32) | (unsigned long long)(~slot)) > cc.thr;
                }
                if (pass) loc = atomicAdd(&ccnt[wave][col], 1u);
            }
            wave_lds_sync();
            if (lane < 32) {
                const uint32_t cnt = ccnt[wave][lane];
                cbase[wave][lane] = cnt ? atomicAdd(a.ext_n + cols[wave][lane].q, cnt) : 0u;
            }
            wave_lds_sync();
            if (pass) {
                const uint32_t at = cbase[wave][col] + loc;
                if (at < a.ext_cap) a.ext[(size_t)cols[wave][col].q * a.ext_cap + at] = make_uint2(key, ~slot);
            }
            wave_lds_sync();
        }
        sn = 0;
    };
    // The next work item is claimed, and its descriptor read, while the current one runs.
    auto claim = [&]() {
        uint32_t v = 0;
        if (lane == 0) v = i_lo + atomicAdd(a.work + a.phase, 1u);
        return (uint32_t)__builtin_amdgcn_readlane((int)v, 0);
    };
    uint32_t nidx = claim();
    uint4 nitem = nidx < i_hi ? a.items[nidx] : make_uint4(0u, 0u, 0u, 0u);
    for (;;) {
        const uint32_t idx = nidx;
        if (idx >= i_hi) break;
        const uint4 item = nitem;
        nidx = claim();
        if (nidx < i_hi) nitem = a.items[nidx];
        // The descriptor is the same in every lane: kept in scalar registers.
        const uint32_t s_lo = __builtin_amdgcn_readfirstlane(item.x);
        const uint32_t s_n = __builtin_amdgcn_readfirstlane(item.y);
        const uint32_t qpos0 = __builtin_amdgcn_readfirstlane(item.z);
        const uint32_t nvalid = __builtin_amdgcn_readfirstlane(item.w);
        const bool two = nvalid > 16;
        // B operands: lane (j, r) holds query j's table entries of groups 4 st + r, minus 128.
        uint4 B0[16], B1[16];
        // Per column half: T = thr - cent, the rounding-margin terms, -b and 1 / a.
        float cT[2], cS[2], cK[2], cnb[2], cia[2];
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const uint32_t n = 16 * h + j;
            const bool valid = n < nvalid;
            const uint32_t pos = qpos0 + (valid ? n : 0);
            const uint32_t q = a.seg_q[pos];
            const float4 qab = a.qab[q];
            const float cent = a.seg_cent[pos];
            // ~0: nothing passes (a padding column, or a query whose list was lost).
            const unsigned long long thr = valid ? a.thr[q] : ~0ull;
            const float thrf = thr == ~0ull ? INFINITY : thr == 0 ? -INFINITY : key2f((uint32_t)(thr >> 32));
            if (lane < 16) {
                CmColC cc;
                cc.a = qab.x;
                cc.b = qab.y;
                cc.cent = cent;
                cc.thrf = thrf;
                cc.thr = thr;
                cc.q = q;
                cc.pad = 0;
                cols[wave][n] = cc;
            }
            const bool finite = thr != ~0ull && thr != 0;
            cT[h] = thr == ~0ull ? INFINITY : thr == 0 ? -INFINITY : thrf - cent;
            // 1e-30: the margin is never exactly zero.
            cS[h] = finite ? fabsf(cent) + fabsf(thrf) + 1.0e-30f : 0.0f;
            cK[h] = 65280.0f * qab.x + 3.0f * fabsf(qab.y);
            cnb[h] = -qab.y;
            cia[h] = __builtin_amdgcn_rcpf(qab.x);
            if (h == 1 && !two) continue;
            const uint4* tp = reinterpret_cast<const uint4*>(a.tables + (size_t)q * 1024);
#pragma unroll
            for (int st = 0; st < 16; st++) {
                uint4 v = tp[4 * st + r];
                v.x ^= 0x80808080u;
                v.y ^= 0x80808080u;
                v.z ^= 0x80808080u;
                v.w ^= 0x80808080u;
                if (h == 0) B0[st] = v;
                else B1[st] = v;
            }
        }
        if (lane == 0 && a.scanned) {
            atomicAdd(a.scanned, (unsigned long long)s_n * nvalid);
            atomicAdd(a.pairs, (unsigned long long)s_n * nvalid);
        }
        wave_lds_sync();

        // Lane (j, r) of C holds slots 4 r .. 4 r + 3 of the 16-slot tile for query column j
        // (column 16 h + j of the tile).
        // The bound is compared in f32 (scores are exact there): a NaN bound (0 x inf, a zero
        // multiplier against a margin-adjusted threshold of exactly 0) stages everything.
        auto stage_half = [&](const v4i& acc, int h, uint32_t slot0, float mx, float rmx, float rmn) {
            if (sn > (uint32_t)(CM2_STAGE - 256)) flush();  // a half adds at most 256 entries
            const float tpv = fmaf(-0x1p-21f, fmaf(mx, cK[h], cS[h]), cT[h]);
            const float inv = tpv > 0.0f ? rmx : rmn;
            const float tf = fmaf(fmaf(tpv, inv, cnb[h]), cia[h], -2.0f);
            const uint32_t colx = (uint32_t)(16 * h + j), sb = slot0 + 4 * r;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const bool p = !((float)acc[i] < tf);
                const unsigned long long m = __ballot(p);
                if (m) {
                    if (p) {
                        const uint32_t at =
                            sn + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                        stage[wave][at] = make_uint2(sb + i, ((uint32_t)acc[i] << 5) | colx);
                    }
                    sn += (uint32_t)__popcll(m);
                }
            }
        };
        auto tile16 = [&](const uint2& cw, const float4& mu, uint32_t slot0, auto two_c) {
            constexpr bool TWO = decltype(two_c)::value;
            // The even chain starts at 64 * 128, so the sum of the chains is score_code itself.
            v4i e0 = {64 * 128, 64 * 128, 64 * 128, 64 * 128}, o0 = {0, 0, 0, 0}, e1 = e0;
            // Operand of step st: the one-hot bytes of group 4 st + r of the slot, nibble st of
            // the lane's 8-byte chunk of the code. Its table offset (nibble x 16) is bits 4..7 of
            // a code byte: of byte st / 2 itself for the high nibble, of the word shifted left by
            // 4 for the low one (one masked byte select per operand).
            const uint32_t w[2] = {cw.x, cw.y}, w4[2] = {cw.x << 4, cw.y << 4};
            auto operand = [&](int st) -> v4i {
                const uint32_t off = nib_off((st & 1) ? w[st >> 3] : w4[st >> 3], (st >> 1) & 3);
                const uint4 oh = *reinterpret_cast<const uint4*>(reinterpret_cast<const uint8_t*>(onehot) + off);
                return v4i{(int)oh.x, (int)oh.y, (int)oh.z, (int)oh.w};
            };
            constexpr int PIPE = 4;
            v4i pipe[PIPE];
#pragma unroll
            for (int st = 0; st < PIPE; st++) pipe[st] = operand(st);
#pragma unroll
            for (int st = 0; st < 16; st++) {
                const v4i av = pipe[st % PIPE];
                if (st + PIPE < 16) pipe[st % PIPE] = operand(st + PIPE);
                const v4i b0v = {(int)B0[st].x, (int)B0[st].y, (int)B0[st].z, (int)B0[st].w};
                if constexpr (TWO) {
                    // Two halves: already two chains.
                    const v4i b1v = {(int)B1[st].x, (int)B1[st].y, (int)B1[st].z, (int)B1[st].w};
                    e0 = __builtin_amdgcn_mfma_i32_16x16x64_i8(av, b0v, e0, 0, 0, 0);
                    e1 = __builtin_amdgcn_mfma_i32_16x16x64_i8(av, b1v, e1, 0, 0, 0);
                } else if (st & 1) {
                    o0 = __builtin_amdgcn_mfma_i32_16x16x64_i8(av, b0v, o0, 0, 0, 0);
                } else {
                    e0 = __builtin_amdgcn_mfma_i32_16x16x64_i8(av, b0v, e0, 0, 0, 0);
                }
            }
            // Multipliers are >= +0, where the f32 order is the order of the bit patterns.
            const uint32_t ux = __float_as_uint(mu.x), uy = __float_as_uint(mu.y), uz = __float_as_uint(mu.z),
                           uw = __float_as_uint(mu.w);
            const float mx = __uint_as_float(max(max(ux, uy), max(uz, uw)));
            const float mn = __uint_as_float(min(min(ux, uy), min(uz, uw)));
            const float rmx = __builtin_amdgcn_rcpf(mx), rmn = __builtin_amdgcn_rcpf(mn);
            if constexpr (TWO) {
                stage_half(e0, 0, slot0, mx, rmx, rmn);
                stage_half(e1, 1, slot0, mx, rmx, rmn);
            } else {
                stage_half(e0 + o0, 0, slot0, mx, rmx, rmn);
            }
        };
        // 16-slot tiles of the chunk in order, the codes and multipliers of the next CM_PREFETCH
        // tiles in flight (lane (j, r): chunk r of slot j's code, mul of slots 4 r .. 4 r + 3).
        // The loads take a wave-uniform base and a 32-bit lane offset; a tile's slots are
        // refilled with the tile PF ahead once the tile has consumed them (loads past the last
        // tile re-read it, in bounds, so no condition keeps later loads from staying in flight).
        constexpr int PF = CM_PREFETCH;
        const uint32_t ntiles = s_n / 16;
        const uint8_t* cbase = reinterpret_cast<const uint8_t*>(a.codes) + 32 * (size_t)s_lo;
        const uint8_t* mbase = reinterpret_cast<const uint8_t*>(a.mul) + 4 * (size_t)s_lo;
        const uint32_t clane = 32 * (uint32_t)j + 8 * (uint32_t)r, mlane = 16 * (uint32_t)r;
        const uint32_t tlast = ntiles > 0 ? ntiles - 1 : 0;
        auto load_c = [&](uint32_t ti) {
            return *reinterpret_cast<const uint2*>(cbase + 512 * min(ti, tlast) + clane);
        };
        auto load_m = [&](uint32_t ti) {
            return *reinterpret_cast<const float4*>(mbase + 64 * min(ti, tlast) + mlane);
        };
        uint2 cr[PF];
        float4 mr[PF];
#pragma unroll
        for (int p = 0; p < PF; p++) {
            cr[p] = load_c((uint32_t)p);
            mr[p] = load_m((uint32_t)p);
        }
        // Whole groups of PF tiles run with no condition around the loads (so the waits for a
        // tile's data count the later loads as still in flight), the last < PF tiles after.
        const uint32_t nfull = ntiles & ~(uint32_t)(PF - 1);
        auto group = [&](auto two_c) {
            for (uint32_t t0 = 0; t0 < nfull; t0 += PF) {
#pragma unroll
                for (int p = 0; p < PF; p++) {
                    const uint32_t ti = t0 + p;
                    tile16(cr[p], mr[p], s_lo + 16 * ti, two_c);
                    cr[p] = load_c(ti + PF);
                    mr[p] = load_m(ti + PF);
                }
            }
#pragma unroll
            for (int p = 0; p < PF - 1; p++)
                if (nfull + p < ntiles) tile16(cr[p], mr[p], s_lo + 16 * (nfull + p), two_c);
        };
        if (two) group(std::true_type{});
        else group(std::false_type{});
        flush();  // before the next item replaces the column constants
    }
}

void launch_scan_cm(const ScanArgs& a_in, const CmScratch& cm_in, size_t nq, int variant, hipStream_t st) {
    if (nq == 0) return;
    CmScratch cm = cm_in;
    cm.nq = nq;
    ScanArgs a = a_in;
    a.chunk_max = 1;
    a.chunk_blocks = 0xFFFFFFFFu;
    a.seg_stride = scan_segments(a);
    a.trigger = std::max(a.cand + 1, a.cand * 6 / 4);
    a.probe_lim = cm.probe_lim;
    const size_t lds = scan_layout(a.cand, a.seg_stride, 2, a.trigger).total;
    hipLaunchKernelGGL(k_cm_zero, dim3((cm.nseg + 255) / 256), dim3(256), 0, st, cm.seg_count, cm.nseg);
    const unsigned qgrid = (unsigned)((nq + 3) / 4);
    hipLaunchKernelGGL(k_cm_prep, dim3(qgrid), dim3(256), 0, st, a, cm);
    hipLaunchKernelGGL(k_cm_offsets, dim3(1), dim3(SEL_THREADS), 0, st, a, cm);
    hipLaunchKernelGGL(k_cm_scatter, dim3(qgrid), dim3(256), 0, st, a, cm);
    hipLaunchKernelGGL(k_cm_items, dim3((cm.nseg + 3) / 4), dim3(256), 0, st, a, cm);
    // The seed pass streams each query's nearest probes: 4 slots per thread keep more of its
    // short scan in flight.
    const size_t lds_seed = scan_layout(a.cand, a.seg_stride, 4, a.trigger).total;
    if (lds_seed > 64 * 1024)
        (void)hipFuncSetAttribute((const void*)k_scan_t<64, false, 4, false, 0, SCAN_SEED>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_seed);
    if (lds > 64 * 1024)
        (void)hipFuncSetAttribute((const void*)k_scan_t<64, false, 2, false, 0, SCAN_FINISH>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL((k_scan_t<64, false, 4, false, 0, SCAN_SEED>), dim3((unsigned)nq), dim3(SCAN_THREADS),
                       lds_seed, st, a);

    CmArgs c{};
    c.codes = a.codes;
    c.mul = a.mul;
    c.ids = a.ids;
    c.tables = a.tables;
    c.qab = a.qab;
    c.items = cm.items;
    c.items_cap = cm.items_cap;
    c.seg_ioff = cm.seg_ioff;
    c.seg_q = cm.seg_q;
    c.seg_cent = cm.seg_cent;
    c.nseg = cm.nseg;
    c.thr = a.seed_thr;
    c.view = a.view;
    c.ext = a.ext;
    c.ext_n = a.ext_n;
    c.ext_cap = a.ext_cap;
    c.work = cm.work;
    c.scanned = a.scanned;
    c.pairs = a.scanned ? a.scanned + 4 : nullptr;
    c.staged = a.scanned ? a.scanned + 5 : nullptr;
    c.phases = cm.phases;
    int dev = 0, cus = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (cus <= 0) cus = 256;
    if (lds > 64 * 1024)
        (void)hipFuncSetAttribute((const void*)k_scan_t<64, false, 2, false, 0, SCAN_TIGHTEN>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    // Phases of the work list; between them each query's kept list absorbs what was appended
    // and its threshold rises, so later phases append less.
    for (uint32_t ph = 0; ph < cm.phases; ph++) {
        c.phase = ph;
        const bool timed = cm.ev && 2 * ph + 1 < cm.nev;
        if (timed) (void)hipEventRecord(cm.ev[2 * ph], st);
        if (variant == SCAN_VARIANT_CM)
            hipLaunchKernelGGL(k_scan_cm2, dim3((unsigned)cus * 4), dim3(256), 0, st, c);
        else
            hipLaunchKernelGGL(k_scan_cm<CM_OH_LDS>, dim3((unsigned)cus * 4), dim3(256), 0, st, c);
        if (timed) (void)hipEventRecord(cm.ev[2 * ph + 1], st);
        if (ph + 1 < cm.phases)
            hipLaunchKernelGGL((k_scan_t<64, false, 2, false, 0, SCAN_TIGHTEN>), dim3((unsigned)nq),
                               dim3(SCAN_THREADS), lds, st, a);
    }
    hipLaunchKernelGGL((k_scan_t<64, false, 2, false, 0, SCAN_FINISH>), dim3((unsigned)nq), dim3(SCAN_THREADS), lds,
                       st, a);
}

// ------------------------------------------------------------------ small batches ------
// A few queries alone cannot fill the GPU through per-query workgroups that keep their own
// candidate lists: a chunk's workgroup admits and cuts thousands of entries under a low
// threshold, and the last chunk merges every chunk's list. For small batches the scan runs
// in four steps instead: (1) SCAN_BOUND: every 32-block chunk of a query's probes keeps its
// best 16 admissible entries; (2) k_scan_bound_select: the query's threshold is the
// (cand + 1)-th best of their union, a subset of the admissible entries, so it lies strictly
// below the cand-th best of all of them; (3) SCAN_APPEND: every slot above the threshold is
// appended to the query's list, which therefore holds the top cand; (4) SCAN_FINISH merges the
// list (reading ids and the view there, as after the cluster pass) or, when the list
// overflowed, rescans the query.
constexpr int BOUND_MAX = 8192;  // union entries one query's select reads (the first chunks')

__global__ __launch_bounds__(SEL_THREADS) void k_scan_bound_select(ScanArgs a) {
    extern __shared__ __align__(16) uint8_t lds[];
    uint2* ent = reinterpret_cast<uint2*>(lds);
    __shared__ uint32_t hist[256], sh[4], cnt;
    const size_t qi = blockIdx.x;
    const int t = threadIdx.x;
    if (t == 0) cnt = 0;
    __syncthreads();
    const uint32_t chunks = min(a.chunk_max, (uint32_t)(BOUND_MAX / BOUND_PER_CHUNK));
    for (uint32_t c = t; c < chunks; c += SEL_THREADS) {
        const size_t at = qi * a.chunk_max + c;
        const uint32_t m = a.part_n[at];
        if (m) {
            const uint32_t o = atomicAdd(&cnt, m);
            for (uint32_t i = 0; i < m; i++) ent[o + i] = a.part[at * BOUND_PER_CHUNK + i];
        }
    }
    __syncthreads();
    const int n = (int)cnt;
    const int k = a.cand + 1;
    unsigned long long thr = 0;
    if (n >= k) thr = pack_pair(block_kth_largest(ent, n, k, hist, sh));
    if (t == 0) {
        a.seed_thr[qi] = thr;
        a.seed_n[qi] = 0;
    }
}

template <int G, bool L2>
static void launch_small_t(const ScanArgs& a_in, size_t nq, uint32_t bound_chunks, uint32_t append_chunk_blocks,
                           uint32_t append_chunks, hipStream_t st) {
    if (nq == 0) return;
    ScanArgs a = a_in;
    a.seg_stride = scan_segments(a);
    a.trigger = std::max(a.cand + 1, a.cand * 6 / 4);
    const size_t lds4 = scan_layout(a.cand, a.seg_stride, 4, a.trigger).total;
    if (lds4 > 64 * 1024) {
        (void)hipFuncSetAttribute((const void*)k_scan_t<G, L2, 4, false, 0, SCAN_BOUND>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds4);
        (void)hipFuncSetAttribute((const void*)k_scan_t<G, L2, 4, false, 0, SCAN_APPEND>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds4);
    }
    if (lds4 > 64 * 1024)
        (void)hipFuncSetAttribute((const void*)k_scan_t<G, L2, 4, false, 0, SCAN_FINISH>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds4);
    ScanArgs b = a;  // one round of 4 slots per thread = 32 blocks per chunk
    b.chunk_blocks = 32;
    b.chunk_max = bound_chunks;
    hipLaunchKernelGGL((k_scan_t<G, L2, 4, false, 0, SCAN_BOUND>), dim3((unsigned)(nq * b.chunk_max)),
                       dim3(SCAN_THREADS), lds4, st, b);
    const size_t sel_lds = (size_t)std::min<uint32_t>(bound_chunks * BOUND_PER_CHUNK, BOUND_MAX) * sizeof(uint2);
    (void)hipFuncSetAttribute((const void*)k_scan_bound_select, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)sel_lds);
    hipLaunchKernelGGL(k_scan_bound_select, dim3((unsigned)nq), dim3(SEL_THREADS), sel_lds, st, b);
    ScanArgs p = a;
    p.chunk_blocks = append_chunk_blocks;
    p.chunk_max = append_chunks;
    hipLaunchKernelGGL((k_scan_t<G, L2, 4, false, 0, SCAN_APPEND>), dim3((unsigned)(nq * p.chunk_max)),
                       dim3(SCAN_THREADS), lds4, st, p);
    ScanArgs f = a;
    f.chunk_max = 1;
    f.chunk_blocks = 0xFFFFFFFFu;
    hipLaunchKernelGGL((k_scan_t<G, L2, 4, false, 0, SCAN_FINISH>), dim3((unsigned)nq), dim3(SCAN_THREADS), lds4,
                       st, f);
}

void launch_scan_small(const ScanArgs& a, size_t nq, bool l2, uint32_t bound_chunks, uint32_t append_chunk_blocks,
                       uint32_t append_chunks, hipStream_t st) {
    if (l2)
        launch_small_t<32, true>(a, nq, bound_chunks, append_chunk_blocks, append_chunks, st);
    else
        launch_small_t<64, false>(a, nq, bound_chunks, append_chunk_blocks, append_chunks, st);
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
