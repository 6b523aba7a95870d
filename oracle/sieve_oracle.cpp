// sieve_oracle.cpp — TEST INFRASTRUCTURE, NOT PRODUCT CODE.
//
// A CPU restatement of the reference's x86 (non-NEON) search path, used only to check results:
// by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg. Nothing in the
// product (a-game-of-vectors_amd/) links, loads or calls it.
//
// Every function follows the Rust it cites (paths are into the sieve crate,
// 18Abhinav07/A-Game-Of-Vectors) with Rust's floating-point rules: IEEE f32, no FMA
// contraction (built with -ffp-contract=off), left-to-right sums, `Sum for f32` seeded with
// -0.0, round() half away from zero, `as u8/i8` saturating.
//
// Parity status of each part is recorded in the header of oracle/oracle.py.

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fcntl.h>
#include <limits>
#include <memory>
#include <thread>
#include <unordered_map>
#include <string>
#include <unistd.h>
#include <vector>

#include "../include/sieve_gpu.h"

namespace {

constexpr int DIM = 768;
constexpr int PDIM = 256;
constexpr int BLOCK = 32;
constexpr uint32_t PAD = 0xFFFFFFFFu;
constexpr float EPS0 = 1.9f;  // src/rabitq.rs:28

// ---------------------------------------------------------------- Rng -----------------
// splitmix64, src/ivf.rs:120-135.
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    size_t below(size_t n) { return (size_t)(next() % (uint64_t)n); }
};

// Rust `Iterator::sum::<f32>()`: fold from -0.0, left to right.
inline float rsum_seed() { return -0.0f; }

// ---------------------------------------------------------------- TopK ----------------
// Amortised top-k, src/topk.rs:4-62. Ties at the k-th boundary are broken by nth_element
// here and by select_nth_unstable in the reference: results agree up to ties.
template <class S>
struct TopK {
    size_t k;
    std::vector<std::pair<S, uint32_t>> items;
    bool has_thr = false;
    S thr{};
    explicit TopK(size_t k_) : k(k_) { items.reserve(2 * std::max<size_t>(k_, 1)); }
    bool rejects(S score) const { return has_thr && score <= thr; }
    void push(S score, uint32_t id) {  // src/topk.rs:27-36
        if (has_thr && score <= thr) return;
        items.emplace_back(score, id);
        if (items.size() >= 2 * k) compact();
    }
    void compact() {  // src/topk.rs:38-44
        if (items.size() > k) {
            std::nth_element(items.begin(), items.begin() + (k - 1), items.end(),
                             [](const auto& a, const auto& b) { return b.first < a.first; });
            items.resize(k);
            thr = items[k - 1].first;
            has_thr = true;
        }
    }
    std::vector<std::pair<S, uint32_t>> into_sorted() {  // src/topk.rs:47-61
        compact();
        std::sort(items.begin(), items.end(), [](const auto& a, const auto& b) {
            if (a.first != b.first) return b.first < a.first;
            return a.second < b.second;
        });
        if (items.size() > k) items.resize(k);
        return items;
    }
};

// ---------------------------------------------------------------- blas ----------------
// gemm_abt, non-macOS fallback (src/blas.rs:43-54): each c[i][j] a sequential f32 sum.
void gemm_abt(const float* a, const float* b, size_t m, size_t n, size_t k, float* c, int threads) {
    auto rows = [&](size_t r0, size_t r1) {
        for (size_t i = r0; i < r1; i++) {
            const float* ar = a + i * k;
            for (size_t j = 0; j < n; j++) {
                const float* br = b + j * k;
                float acc = rsum_seed();
                for (size_t t = 0; t < k; t++) acc = acc + ar[t] * br[t];
                c[i * n + j] = acc;
            }
        }
    };
    if (threads <= 1 || m < 2) {
        rows(0, m);
        return;
    }
    std::vector<std::thread> th;
    size_t per = (m + threads - 1) / threads;
    for (int t = 0; t < threads; t++) {
        size_t r0 = std::min(m, t * per), r1 = std::min(m, (t + 1) * per);
        if (r0 < r1) th.emplace_back(rows, r0, r1);
    }
    for (auto& x : th) x.join();
}

// ---------------------------------------------------------------- lut -----------------
inline uint8_t nibble(const uint8_t* code, int g) {  // src/lut.rs:28-30
    return (code[g >> 1] >> ((g & 1) * 4)) & 0x0F;
}

// Lut::new, src/lut.rs:33-56. tables is G x 16.
void lut_new(const float* q, int G, uint8_t* tables, float* scale_out, float* bias_out) {
    std::vector<float> f((size_t)G * 16);
    for (int g = 0; g < G; g++)
        for (int c = 0; c < 16; c++) {
            float v = rsum_seed();
            for (int j = 0; j < 4; j++)
                if ((c >> j) & 1) v = v + q[4 * g + j];
            f[g * 16 + c] = v;
        }
    std::vector<float> mins(G);
    for (int g = 0; g < G; g++) {
        float m = std::numeric_limits<float>::max();
        for (int c = 0; c < 16; c++) m = std::fmin(m, f[g * 16 + c]);
        mins[g] = m;
    }
    float max_range = 0.0f;
    for (int g = 0; g < G; g++) {
        float mx = std::numeric_limits<float>::lowest();
        for (int c = 0; c < 16; c++) mx = std::fmax(mx, f[g * 16 + c]);
        max_range = std::fmax(max_range, mx - mins[g]);
    }
    float scale = max_range > 0.0f ? 255.0f / max_range : 1.0f;
    for (int g = 0; g < G; g++)
        for (int c = 0; c < 16; c++) {
            float v = std::round((f[g * 16 + c] - mins[g]) * scale);
            v = std::min(std::max(v, 0.0f), 255.0f);
            tables[g * 16 + c] = (uint8_t)v;
        }
    float bias = rsum_seed();
    for (int g = 0; g < G; g++) bias = bias + mins[g];
    *scale_out = scale;
    *bias_out = bias;
}

uint16_t score_code(const uint8_t* tables, int G, const uint8_t* code) {  // src/lut.rs:59-61
    uint16_t s = 0;
    for (int g = 0; g < G; g++) s = (uint16_t)(s + tables[g * 16 + nibble(code, g)]);
    return s;
}

// pack_block, src/lut.rs:73-83: codes is n x (G/2) bytes, n <= 32, out G x 16 bytes.
void pack_block(const uint8_t* codes, int n, int G, uint8_t* out) {
    int cb = G / 2;
    auto get = [&](int k, int g) -> uint8_t { return k < n ? nibble(codes + (size_t)k * cb, g) : 0; };
    for (int g = 0; g < G; g++)
        for (int k = 0; k < 16; k++) out[g * 16 + k] = (uint8_t)(get(k, g) | (get(k + 16, g) << 4));
}

void unpack_slot(const uint8_t* block, int G, int k, uint8_t* code) {  // src/lut.rs:85-94
    std::memset(code, 0, G / 2);
    for (int g = 0; g < G; g++) {
        uint8_t b = block[g * 16 + (k & 15)];
        uint8_t nib = k < 16 ? (b & 0x0F) : (b >> 4);
        code[g >> 1] |= (uint8_t)(nib << ((g & 1) * 4));
    }
}

// scan_block_scalar, src/lut.rs:117-127, over nblocks blocks (scan, src/lut.rs:96-103).
void scan(const uint8_t* tables, int G, const uint8_t* blocks, size_t nblocks, uint16_t* out) {
    for (size_t b = 0; b < nblocks; b++) {
        const uint8_t* blk = blocks + b * (size_t)G * 16;
        uint16_t* o = out + b * BLOCK;
        for (int k = 0; k < BLOCK; k++) o[k] = 0;
        for (int g = 0; g < G; g++) {
            const uint8_t* t = tables + g * 16;
            for (int k = 0; k < 16; k++) {
                uint8_t x = blk[g * 16 + k];
                o[k] = (uint16_t)(o[k] + t[x & 0x0F]);
                o[k + 16] = (uint16_t)(o[k + 16] + t[x >> 4]);
            }
        }
    }
}

// ---------------------------------------------------------------- rabitq --------------
// Rotation::with_dim signs, src/rabitq.rs:40-44.
void rotation_signs(uint64_t seed, int d, float* signs) {
    Rng r(seed);
    for (int i = 0; i < 3 * d; i++) signs[i] = (r.next() & 1) == 1 ? 1.0f : -1.0f;
}

void fwht(float* v, int n) {  // src/rabitq.rs:72-84
    for (int h = 1; h < n; h *= 2)
        for (int blk = 0; blk < n; blk += 2 * h)
            for (int i = 0; i < h; i++) {
                float x = v[blk + i], y = v[blk + h + i];
                v[blk + i] = x + y;
                v[blk + h + i] = x - y;
            }
}

void rotation_apply(const float* signs, int d, float* v) {  // src/rabitq.rs:59-69
    for (int r = 0; r < 3; r++) {
        for (int i = 0; i < d; i++) v[i] = v[i] * signs[r * d + i];
        fwht(v, d);
    }
    float norm = std::pow((float)d, -1.5f);
    for (int i = 0; i < d; i++) v[i] = v[i] * norm;
}

// encode, src/rabitq.rs:86-108. Writes d/8 code bytes and [mul, err].
void rabitq_encode(const float* xr, const float* cr, int d, uint8_t* code, float* mul_err) {
    std::memset(code, 0, d / 8);
    float nr2 = 0.0f, l1 = 0.0f;
    for (int i = 0; i < d; i++) {
        float r = xr[i] - cr[i];
        nr2 = nr2 + r * r;
        l1 = l1 + std::fabs(r);
        if (r > 0.0f) code[i >> 3] |= (uint8_t)(1u << (i & 7));
    }
    if (l1 < 1e-12f) {
        mul_err[0] = 0.0f;
        mul_err[1] = 0.0f;
        return;
    }
    float df = (float)d;
    float nr = std::sqrt(nr2);
    float od = l1 / (nr * std::sqrt(df));
    float inner = (1.0f - od * od) / (od * od * (df - 1.0f));
    inner = std::fmax(inner, 0.0f);
    float err = nr * EPS0 * std::sqrt(inner);
    mul_err[0] = nr2 / l1;
    mul_err[1] = err;
}

struct QueryState {  // rabitq::Query, src/rabitq.rs:112-131
    std::vector<uint8_t> tables;
    float scale, bias, a, b;
};

void query_new(const float* signs, int d, const float* q, QueryState& qs) {  // src/rabitq.rs:119-125
    std::vector<float> qr(q, q + d);
    rotation_apply(signs, d, qr.data());
    int G = d / 4;
    qs.tables.resize((size_t)G * 16);
    lut_new(qr.data(), G, qs.tables.data(), &qs.scale, &qs.bias);
    float sum = rsum_seed();
    for (int i = 0; i < d; i++) sum = sum + qr[i];
    qs.a = 2.0f / qs.scale;
    qs.b = 2.0f * qs.bias - sum;
}

inline float estimate(const QueryState& qs, float cent, float mul, uint16_t sc) {  // src/rabitq.rs:127-130
    return cent + mul * (qs.a * (float)sc + qs.b);
}

// ---------------------------------------------------------------- quant ---------------
float normalize(float* v, int n) {  // src/quant.rs:5-13
    float s = rsum_seed();
    for (int i = 0; i < n; i++) s = s + v[i] * v[i];
    float nn = std::sqrt(s);
    if (nn > 0.0f) {
        float inv = 1.0f / nn;
        for (int i = 0; i < n; i++) v[i] = v[i] * inv;
    }
    return nn;
}

void sign_code(const float* p, int n, uint8_t* out) {  // src/quant.rs:24-34
    std::memset(out, 0, n / 8);
    for (int i = 0; i < n; i++)
        if (p[i] > 0.0f) out[i >> 3] |= (uint8_t)(1u << (i & 7));
}

float dot_f32(const float* a, const float* b, int n) {  // src/quant.rs:86-100
    float acc[16];
    for (float& x : acc) x = 0.0f;
    int full = n / 16 * 16;
    for (int i = 0; i < full; i += 16)
        for (int j = 0; j < 16; j++) acc[j] = acc[j] + a[i + j] * b[i + j];
    float tail = rsum_seed();
    for (int i = full; i < n; i++) tail = tail + a[i] * b[i];
    float s = rsum_seed();
    for (float x : acc) s = s + x;
    return s + tail;
}

float dot_i8(const float* w, const int8_t* x, int n) {  // src/quant.rs:70-84
    float acc[16];
    for (float& v : acc) v = 0.0f;
    int full = n / 16 * 16;
    for (int i = 0; i < full; i += 16)
        for (int j = 0; j < 16; j++) acc[j] = acc[j] + w[i + j] * (float)x[i + j];
    float tail = rsum_seed();
    for (int i = full; i < n; i++) tail = tail + w[i] * (float)x[i];
    float s = rsum_seed();
    for (float v : acc) s = s + v;
    return s + tail;
}

inline int8_t sat_i8(float v) {  // Rust `f32 as i8`: saturating, NaN -> 0
    if (v != v) return 0;
    if (v >= 127.0f) return 127;
    if (v <= -128.0f) return -128;
    return (int8_t)v;
}

float query_i8(const float* w, int n, int8_t* out) {  // src/quant.rs:104-111
    float m = 1e-12f;
    for (int i = 0; i < n; i++) m = std::fmax(m, std::fabs(w[i]));
    float s = 127.0f / m;
    for (int i = 0; i < n; i++) out[i] = sat_i8(std::round(w[i] * s));
    return s;
}

int32_t dot_i8i8(const int8_t* a, const int8_t* b, int n) {  // src/quant.rs:123-125
    int32_t s = 0;
    for (int i = 0; i < n; i++) s += (int32_t)a[i] * (int32_t)b[i];
    return s;
}

// ---------------------------------------------------------------- tier2 ---------------
constexpr float T2_T = 0.9816f;  // src/tier2.rs:10-11
constexpr float T2_LEVELS[4] = {-1.510f, -0.4528f, 0.4528f, 1.510f};

void tier2_fit(const float* rows, size_t nrows, float* sigma) {  // src/tier2.rs:23-31
    float n = (float)std::max<size_t>(nrows, 1);
    std::vector<double> ss(PDIM, 0.0);
    for (size_t r = 0; r < nrows; r++)
        for (int i = 0; i < PDIM; i++) {
            float x = rows[r * PDIM + i];
            ss[i] += (double)(x * x);
        }
    for (int i = 0; i < PDIM; i++) sigma[i] = std::fmax((float)std::sqrt(ss[i] / (double)n), 1e-6f);
}

inline uint8_t tier2_level(float x, float s) {  // src/tier2.rs:30-34
    float t = T2_T * s;
    if (x < -t) return 0;
    if (x < 0.0f) return 1;
    if (x < t) return 2;
    return 3;
}

float tier2_encode(const float* sigma, const float* prefix, uint8_t* out) {  // src/tier2.rs:36-47
    std::memset(out, 0, PDIM / 4);
    float nn = 0.0f;
    for (int i = 0; i < PDIM; i++) {
        uint8_t l = tier2_level(prefix[i], sigma[i]);
        out[i >> 2] |= (uint8_t)(l << ((i & 3) * 2));
        float v = T2_LEVELS[l] * sigma[i];
        nn = nn + v * v;
    }
    return std::sqrt(nn);
}

void tier2_tables(const float* sigma, const float* q, float* tables) {  // src/tier2.rs:52-66
    for (int g = 0; g < PDIM / 4; g++) {
        float half[2][16];
        for (int h = 0; h < 2; h++)
            for (int b = 0; b < 16; b++) {
                float s = rsum_seed();
                for (int j = 0; j < 2; j++) {
                    int i = 4 * g + 2 * h + j;
                    s = s + q[i] * T2_LEVELS[(b >> (2 * j)) & 3] * sigma[i];
                }
                half[h][b] = s;
            }
        for (int b = 0; b < 256; b++) tables[g * 256 + b] = half[0][b & 15] + half[1][b >> 4];
    }
}

float tier2_score(const float* tables, const uint8_t* code, float norm) {  // src/tier2.rs:69-75
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int g = 0; g < PDIM / 4; g++) acc[g & 3] = acc[g & 3] + tables[g * 256 + code[g]];
    return (acc[0] + acc[1] + acc[2] + acc[3]) / norm;
}

// ---------------------------------------------------------------- ivf -----------------
struct Delta;

struct View {  // src/ivf.rs:462-496: tombstones, allowlist, live delta segment
    const uint64_t* deleted = nullptr;
    uint64_t deleted_words = 0;
    const uint64_t* allow = nullptr;
    uint64_t allow_words = 0;
    const Delta* delta = nullptr;
    static bool bit(const uint64_t* set, uint64_t words, uint32_t id) {  // src/ivf.rs:474-477
        uint64_t w = id >> 6;
        return w < words && ((set[w] >> (id & 63)) & 1) == 1;
    }
    bool hidden(uint32_t id) const {  // src/ivf.rs:480-483
        return (deleted && bit(deleted, deleted_words, id)) || (allow && !bit(allow, allow_words, id));
    }
};

struct Index {
    uint32_t k = 0;
    uint64_t n = 0, nblocks = 0;
    const float* centroids = nullptr;
    const uint64_t* block_off = nullptr;
    const uint8_t* blocks = nullptr;
    const uint32_t* ids = nullptr;
    const float* factors = nullptr;
    const float* rotation = nullptr;
    const int8_t* rerank = nullptr;
    const float* quant = nullptr;
    const float* norm256 = nullptr;
    const float* t2_sigma = nullptr;
    const uint8_t* t2_codes = nullptr;
    const float* t2_norm = nullptr;

    size_t slots(uint32_t c) const { return (size_t)(block_off[c + 1] - block_off[c]) * BLOCK; }  // src/ivf.rs:615-618
};

// One cluster's delta list in the base block layout (src/live.rs:37-66).
struct DeltaList {
    std::vector<uint8_t> blocks;
    std::vector<float> factors;
    std::vector<uint32_t> ids;
    std::vector<uint8_t> tail;  // raw codes of the last block, 32 B each
    size_t len = 0;
    void push(const uint8_t* code, const float* f, uint32_t id) {  // DeltaList::push, src/live.rs:49-65
        size_t k = len % BLOCK;
        if (k == 0) {
            blocks.resize(blocks.size() + 1024, 0);
            factors.resize(factors.size() + BLOCK * 2, 0.0f);
            ids.resize(ids.size() + BLOCK, PAD);
            tail.clear();
        }
        tail.insert(tail.end(), code, code + 32);
        pack_block(tail.data(), (int)(tail.size() / 32), PDIM / 4, blocks.data() + blocks.size() - 1024);
        size_t slot = ids.size() - BLOCK + k;
        ids[slot] = id;
        factors[slot * 2] = f[0];
        factors[slot * 2 + 1] = f[1];
        len++;
    }
};

// Rows added since the build (src/live.rs:68-117): delta lists, rows, prefix norms, tier-2
// codes and external ids; row ids count up from `base` (the corpus size).
struct Delta {
    uint64_t base = 0;
    std::vector<DeltaList> lists;
    std::vector<int8_t> rows;
    std::vector<float> norm256;
    std::vector<uint8_t> t2;
    std::vector<float> t2_norm;
    std::vector<uint32_t> ext;
    bool local(uint32_t id, size_t* i) const {
        if (id < base || id - base >= ext.size()) return false;
        *i = id - base;
        return true;
    }
    const DeltaList* list(uint32_t c) const {
        return c < lists.size() && lists[c].len > 0 ? &lists[c] : nullptr;
    }
};

// View::row and the prefix-norm / tier-2 lookups (src/ivf.rs:485-495, src/live.rs:84-103).
inline const int8_t* view_row(const Index& ix, const View& v, uint32_t id) {
    size_t i;
    if (v.delta && v.delta->local(id, &i)) return v.delta->rows.data() + i * DIM;
    return ix.rerank + (size_t)id * DIM;
}
inline float view_norm256(const Index& ix, const View& v, uint32_t id) {
    size_t i;
    if (v.delta && v.delta->local(id, &i)) return v.delta->norm256[i];
    return ix.norm256[id];
}
inline void view_tier2(const Index& ix, const View& v, uint32_t id, const uint8_t** code, float* norm) {
    size_t i;
    if (v.delta && !v.delta->t2_norm.empty() && v.delta->local(id, &i)) {
        *code = v.delta->t2.data() + i * 64;
        *norm = v.delta->t2_norm[i];
        return;
    }
    *code = ix.t2_codes + (size_t)id * 64;
    *norm = ix.t2_norm[id];
}

struct Probe {
    float score;
    uint32_t cluster;
};
struct Cand {
    float est;
    uint32_t row;
    float err;
};

inline bool has_margin(const sieve_params& p) { return !(p.margin != p.margin); }

// Index::route, src/ivf.rs:587-613.
std::vector<Probe> route(const Index& ix, const float* q256, const sieve_params& p, std::vector<float>& cent) {
    cent.resize(ix.k);
    gemm_abt(q256, ix.centroids, 1, ix.k, PDIM, cent.data(), 1);
    TopK<float> top(std::min<size_t>(p.nprobe, ix.k));
    for (uint32_t c = 0; c < ix.k; c++) top.push(cent[c], c);
    auto sorted = top.into_sorted();
    std::vector<Probe> probe;
    for (auto& x : sorted) probe.push_back({x.first, x.second});
    if (has_margin(p) && !probe.empty()) {
        float lim = probe[0].score - p.margin;
        std::vector<Probe> kept;
        for (auto& x : probe)
            if (x.score >= lim) kept.push_back(x);
        probe.swap(kept);
    }
    if (p.budget > 0) {
        uint64_t used = 0;
        size_t keep = 0;
        for (auto& x : probe) {
            used += ix.slots(x.cluster);
            if (used <= p.budget) keep++;
            else break;
        }
        probe.resize(std::min(probe.size(), std::max<size_t>(keep, 1)));
    }
    return probe;
}

// Index::candidates, src/ivf.rs:643-729 (RaBitQ and global-sign-code branches).
std::vector<Cand> candidates(const Index& ix, const float* q256, const std::vector<Probe>& probe,
                             const sieve_params& p, const View& v, uint64_t* scanned) {
    std::vector<Cand> cand;
    std::vector<uint16_t> scores;
    if (ix.rotation && ix.factors) {
        QueryState rq;
        query_new(ix.rotation, PDIM, q256, rq);
        TopK<float> top(p.cand);
        std::vector<std::pair<uint32_t, float>> meta;
        auto scan_list = [&](float cent, const uint8_t* blocks, size_t nb, const float* f, const uint32_t* ids) {
            size_t ns = nb * BLOCK;
            scores.resize(ns);
            scan(rq.tables.data(), PDIM / 4, blocks, nb, scores.data());
            *scanned += ns;
            for (size_t j = 0; j < ns; j++) {
                float est = estimate(rq, cent, f[2 * j], scores[j]);
                if (top.rejects(est) || ids[j] == PAD || v.hidden(ids[j])) continue;
                top.push(est, (uint32_t)meta.size());
                meta.emplace_back(ids[j], f[2 * j + 1]);
            }
        };
        for (const Probe& pr : probe) {
            uint64_t b0 = ix.block_off[pr.cluster], b1 = ix.block_off[pr.cluster + 1];
            scan_list(pr.score, ix.blocks + b0 * 1024, b1 - b0, ix.factors + b0 * BLOCK * 2, ix.ids + b0 * BLOCK);
            // The cluster's delta list right after its base list (src/ivf.rs:682-684).
            if (const DeltaList* d = v.delta ? v.delta->list(pr.cluster) : nullptr)
                scan_list(pr.score, d->blocks.data(), d->blocks.size() / 1024, d->factors.data(), d->ids.data());
        }
        for (auto& x : top.into_sorted()) cand.push_back({x.first, meta[x.second].first, meta[x.second].second});
    } else {
        std::vector<uint8_t> tables(PDIM / 4 * 16);
        float scale, bias;
        lut_new(q256, PDIM / 4, tables.data(), &scale, &bias);
        TopK<uint16_t> top(p.cand);
        for (const Probe& pr : probe) {
            uint64_t b0 = ix.block_off[pr.cluster], b1 = ix.block_off[pr.cluster + 1];
            size_t ns = (size_t)(b1 - b0) * BLOCK;
            scores.resize(ns);
            scan(tables.data(), PDIM / 4, ix.blocks + b0 * 1024, b1 - b0, scores.data());
            *scanned += ns;
            const uint32_t* ids = ix.ids + b0 * BLOCK;
            for (size_t j = 0; j < ns; j++) {
                if (top.rejects(scores[j]) || ids[j] == PAD || v.hidden(ids[j])) continue;
                top.push(scores[j], ids[j]);
            }
        }
        for (auto& x : top.into_sorted())
            cand.push_back({(float)x.first / scale + bias, x.second, std::numeric_limits<float>::infinity()});
    }
    // Sort by (row asc, est desc) and keep the first of each row (src/ivf.rs:725-726).
    std::sort(cand.begin(), cand.end(), [](const Cand& a, const Cand& b) {
        if (a.row != b.row) return a.row < b.row;
        return b.est < a.est;
    });
    std::vector<Cand> out;
    for (const Cand& c : cand)
        if (out.empty() || out.back().row != c.row) out.push_back(c);
    return out;
}

// prune_threshold, src/ivf.rs:540-550.
float prune_threshold(const std::vector<Cand>& cand, size_t m, float w) {
    if (m == 0 || cand.size() < m) return -std::numeric_limits<float>::infinity();
    std::vector<float> lo;
    for (const Cand& c : cand) lo.push_back(c.est - w * c.err);
    std::nth_element(lo.begin(), lo.begin() + (m - 1), lo.end(), [](float a, float b) { return b < a; });
    return lo[m - 1];
}

struct RerankQuery {  // the int8 query of a rerank pass (src/ivf.rs:753-755)
    std::vector<int8_t> wq;
    float inv;
};

RerankQuery rerank_query(const Index& ix, const float* q, int n) {
    std::vector<float> w(n);
    for (int i = 0; i < n; i++) w[i] = q[i] / ix.quant[i];
    RerankQuery r;
    r.wq.resize(n);
    r.inv = 1.0f / query_i8(w.data(), n, r.wq.data());
    return r;
}

// read_and_rerank, src/ivf.rs:818-832 (all rows are resident, so base and delta rows are
// one store; the pushes keep its order: base rows first).
void read_and_rerank(const Index& ix, const View& v, const float* q768, const std::vector<uint32_t>& ids,
                     TopK<float>& out) {
    RerankQuery r = rerank_query(ix, q768, DIM);
    for (int pass = 0; pass < 2; pass++)  // corpus rows first, then delta rows
        for (uint32_t id : ids)
            if ((id < ix.n) == (pass == 0))
                out.push((float)dot_i8i8(r.wq.data(), view_row(ix, v, id), DIM) * r.inv, id);
}

// Index::rerank, src/ivf.rs:731-815.
std::vector<std::pair<float, uint32_t>> rerank(const Index& ix, const float* q768, const float* q256,
                                               std::vector<Cand> cand, const sieve_params& p, const View& v,
                                               uint64_t* reranked) {
    if (p.prune_m > 0) {
        float tau = prune_threshold(cand, p.prune_m, p.prune_w);
        std::vector<Cand> kept;
        for (const Cand& c : cand)
            if (c.est + p.prune_w * c.err >= tau) kept.push_back(c);
        cand.swap(kept);
    }
    *reranked += cand.size();
    TopK<float> out(p.k);
    switch (p.rerank) {
        case SIEVE_RERANK_NONE:
            for (const Cand& c : cand) out.push(c.est, c.row);
            break;
        case SIEVE_RERANK_FULL: {
            RerankQuery r = rerank_query(ix, q768, DIM);
            for (const Cand& c : cand)
                out.push((float)dot_i8i8(r.wq.data(), view_row(ix, v, c.row), DIM) * r.inv, c.row);
            break;
        }
        case SIEVE_RERANK_SSD: {
            TopK<float> best(p.rerank_r);
            for (const Cand& c : cand) best.push(c.est, c.row);
            std::vector<uint32_t> ids;
            for (auto& x : best.into_sorted()) ids.push_back(x.second);
            read_and_rerank(ix, v, q768, ids, out);
            break;
        }
        default: {
            if (p.rerank == SIEVE_RERANK_TIERED && ix.t2_sigma) {
                std::vector<float> tables(64 * 256);
                tier2_tables(ix.t2_sigma, q256, tables.data());
                TopK<float> pre(p.rerank_r);
                for (const Cand& c : cand) {
                    const uint8_t* code;
                    float norm;
                    view_tier2(ix, v, c.row, &code, &norm);
                    pre.push(tier2_score(tables.data(), code, norm), c.row);
                }
                std::vector<uint32_t> ids;
                for (auto& x : pre.into_sorted()) ids.push_back(x.second);
                read_and_rerank(ix, v, q768, ids, out);
                break;
            }
            // Prefix, or Tiered without a 2-bit tier.
            RerankQuery r = rerank_query(ix, q256, PDIM);
            TopK<float> pre(p.rerank == SIEVE_RERANK_TIERED ? p.rerank_r : p.k);
            for (const Cand& c : cand) {
                float norm = view_norm256(ix, v, c.row);
                pre.push((float)dot_i8i8(r.wq.data(), view_row(ix, v, c.row), PDIM) * r.inv / norm, c.row);
            }
            if (p.rerank == SIEVE_RERANK_PREFIX) {
                out = pre;
            } else {
                RerankQuery f = rerank_query(ix, q768, DIM);
                for (auto& x : pre.into_sorted())
                    out.push((float)dot_i8i8(f.wq.data(), view_row(ix, v, x.second), DIM) * f.inv, x.second);
            }
        }
    }
    return out.into_sorted();
}

}  // namespace

// ======================================================================================
// extern "C" surface for ctypes (oracle/oracle.py)
// ======================================================================================
extern "C" {

uint64_t or_rng_next(uint64_t* state) {
    Rng r(*state);
    uint64_t v = r.next();
    *state = r.s;
    return v;
}

void or_gemm_abt(const float* a, const float* b, size_t m, size_t n, size_t k, float* c, int threads) {
    gemm_abt(a, b, m, n, k, c, threads);
}

void or_lut_new(const float* q, int G, uint8_t* tables, float* scale, float* bias) { lut_new(q, G, tables, scale, bias); }
uint16_t or_score_code(const uint8_t* tables, int G, const uint8_t* code) { return score_code(tables, G, code); }
float or_float_score(const float* q, int d, const uint8_t* code) {  // src/lut.rs:68-71
    float s = rsum_seed();
    for (int i = 0; i < d; i++)
        if ((code[i >> 3] >> (i & 7)) & 1) s = s + q[i];
    return s;
}
void or_pack_block(const uint8_t* codes, int n, int G, uint8_t* out) { pack_block(codes, n, G, out); }
void or_unpack_slot(const uint8_t* block, int G, int k, uint8_t* code) { unpack_slot(block, G, k, code); }
void or_scan(const uint8_t* tables, int G, const uint8_t* blocks, size_t nblocks, uint16_t* out) {
    scan(tables, G, blocks, nblocks, out);
}
uint32_t or_hamming(const uint8_t* a, const uint8_t* b, int nbytes) {  // src/lut.rs:157-166
    uint32_t s = 0;
    for (int i = 0; i < nbytes; i++) s += (uint32_t)__builtin_popcount((unsigned)(a[i] ^ b[i]));
    return s;
}

void or_rotation_signs(uint64_t seed, int d, float* signs) { rotation_signs(seed, d, signs); }
void or_fwht(float* v, int n) { fwht(v, n); }
void or_rotation_apply(const float* signs, int d, float* v) { rotation_apply(signs, d, v); }
void or_rabitq_encode(const float* xr, const float* cr, int d, uint8_t* code, float* mul_err) {
    rabitq_encode(xr, cr, d, code, mul_err);
}
// rabitq::Query::new: tables (d/4 x 16), then [scale, bias, a, b] in out4.
void or_query_new(const float* signs, int d, const float* q, uint8_t* tables, float* out4) {
    QueryState qs;
    query_new(signs, d, q, qs);
    std::memcpy(tables, qs.tables.data(), qs.tables.size());
    out4[0] = qs.scale;
    out4[1] = qs.bias;
    out4[2] = qs.a;
    out4[3] = qs.b;
}
float or_estimate(float cent, float mul, uint16_t sc, float a, float b) { return cent + mul * (a * (float)sc + b); }

float or_normalize(float* v, int n) { return normalize(v, n); }
void or_sign_code(const float* p, int n, uint8_t* out) { sign_code(p, n, out); }
void or_int8_fit(const float* rows, size_t nrows, float* scale) {  // Int8Quantizer::fit, src/quant.rs:40-50
    std::vector<float> maxabs(DIM, 1e-6f);
    for (size_t r = 0; r < nrows; r++)
        for (int i = 0; i < DIM; i++) maxabs[i] = std::fmax(maxabs[i], std::fabs(rows[r * DIM + i]));
    for (int i = 0; i < DIM; i++) scale[i] = 127.0f / maxabs[i];
}
void or_int8_encode(const float* v, const float* scale, int d, int8_t* out) {  // src/quant.rs:52-56
    for (int i = 0; i < d; i++) {
        float x = std::round(v[i] * scale[i]);
        x = std::min(std::max(x, -127.0f), 127.0f);
        out[i] = (int8_t)x;
    }
}
float or_dot_f32(const float* a, const float* b, int n) { return dot_f32(a, b, n); }
float or_dot_i8(const float* w, const int8_t* x, int n) { return dot_i8(w, x, n); }
float or_query_i8(const float* w, int n, int8_t* out) { return query_i8(w, n, out); }
int32_t or_dot_i8i8(const int8_t* a, const int8_t* b, int n) { return dot_i8i8(a, b, n); }

// TopK over (score, id) pairs pushed in order: best first, (score desc, id asc).
size_t or_topk_f32(const float* scores, const uint32_t* ids, size_t n, size_t k, float* out_s, uint32_t* out_i) {
    TopK<float> t(k);
    for (size_t i = 0; i < n; i++) t.push(scores[i], ids[i]);
    auto s = t.into_sorted();
    for (size_t i = 0; i < s.size(); i++) {
        out_s[i] = s[i].first;
        out_i[i] = s[i].second;
    }
    return s.size();
}
size_t or_topk_u16(const uint16_t* scores, const uint32_t* ids, size_t n, size_t k, uint16_t* out_s, uint32_t* out_i) {
    TopK<uint16_t> t(k);
    for (size_t i = 0; i < n; i++) t.push(scores[i], ids[i]);
    auto s = t.into_sorted();
    for (size_t i = 0; i < s.size(); i++) {
        out_s[i] = s[i].first;
        out_i[i] = s[i].second;
    }
    return s.size();
}

void or_tier2_fit(const float* rows, size_t nrows, float* sigma) { tier2_fit(rows, nrows, sigma); }
float or_tier2_encode(const float* sigma, const float* prefix, uint8_t* out) { return tier2_encode(sigma, prefix, out); }
void or_tier2_tables(const float* sigma, const float* q, float* tables) { tier2_tables(sigma, q, tables); }
float or_tier2_score(const float* tables, const uint8_t* code, float norm) { return tier2_score(tables, code, norm); }

float or_prune_threshold(const float* est, const float* err, size_t n, size_t m, float w) {
    std::vector<Cand> c(n);
    for (size_t i = 0; i < n; i++) c[i] = {est[i], (uint32_t)i, err[i]};
    return prune_threshold(c, m, w);
}

// ---- the index -------------------------------------------------------------------------
void* or_index_create(const sieve_index_desc* d) {
    Index* ix = new Index();
    ix->k = d->k;
    ix->n = d->n;
    ix->nblocks = d->nblocks;
    ix->centroids = d->centroids;
    ix->block_off = d->block_off;
    ix->blocks = d->blocks;
    ix->ids = d->ids;
    ix->factors = d->factors;
    ix->rotation = d->rotation;
    ix->rerank = d->rerank;
    ix->quant = d->quant;
    ix->norm256 = d->norm256;
    ix->t2_sigma = d->tier2_sigma;
    ix->t2_codes = d->tier2_codes;
    ix->t2_norm = d->tier2_norm;
    return ix;
}
void or_index_free(void* ix) { delete (Index*)ix; }

static View make_view(const sieve_view* v, size_t qi) {
    View out;
    if (!v) return out;
    out.deleted = v->deleted;
    out.deleted_words = v->deleted_words;
    if (v->allow && (v->allow_stride == 0 || !v->allow_flags || v->allow_flags[qi])) {
        out.allow = v->allow + (v->allow_stride ? qi * v->allow_stride : 0);
        out.allow_words = v->allow_words;
    }
    return out;
}

// route for one query; returns the number of probes written.
size_t or_route(const void* h, const float* q256, const sieve_params* p, float* scores, uint32_t* clusters) {
    const Index& ix = *(const Index*)h;
    std::vector<float> cent;
    auto probe = route(ix, q256, *p, cent);
    for (size_t i = 0; i < probe.size(); i++) {
        scores[i] = probe[i].score;
        clusters[i] = probe[i].cluster;
    }
    return probe.size();
}

// candidates for one query over a given probe list; returns the number written.
size_t or_candidates(const void* h, const float* q256, const float* pscores, const uint32_t* pclusters, size_t np,
                     const sieve_params* p, const sieve_view* v, size_t qi, float* est, uint32_t* rows, float* errs,
                     uint64_t* scanned) {
    const Index& ix = *(const Index*)h;
    std::vector<Probe> probe(np);
    for (size_t i = 0; i < np; i++) probe[i] = {pscores[i], pclusters[i]};
    uint64_t sc = 0;
    auto c = candidates(ix, q256, probe, *p, make_view(v, qi), &sc);
    if (scanned) *scanned += sc;
    for (size_t i = 0; i < c.size(); i++) {
        est[i] = c[i].est;
        rows[i] = c[i].row;
        errs[i] = c[i].err;
    }
    return c.size();
}

size_t or_rerank(const void* h, const float* q768, const float* q256, const float* est, const uint32_t* rows,
                 const float* errs, size_t nc, const sieve_params* p, float* out_s, uint32_t* out_r, uint64_t* reranked) {
    const Index& ix = *(const Index*)h;
    std::vector<Cand> c(nc);
    for (size_t i = 0; i < nc; i++) c[i] = {est[i], rows[i], errs[i]};
    uint64_t rr = 0;
    auto res = rerank(ix, q768, q256, c, *p, View{}, &rr);
    if (reranked) *reranked += rr;
    for (size_t i = 0; i < res.size(); i++) {
        out_s[i] = res[i].first;
        out_r[i] = res[i].second;
    }
    return res.size();
}

// Index::search_view for one query (src/ivf.rs:627-636).
static size_t search_one(const Index& ix, const float* q768, const float* q256, const sieve_params& p, const View& v,
                         float* out_s, uint32_t* out_r, sieve_timing* t) {
    std::vector<float> cent;
    auto probe = route(ix, q256, p, cent);
    uint64_t scanned = 0, reranked = 0;
    auto c = candidates(ix, q256, probe, p, v, &scanned);
    auto res = rerank(ix, q768, q256, c, p, v, &reranked);
    if (t) {
        t->probed += probe.size();
        t->scanned += scanned;
        t->reranked += reranked;
    }
    for (size_t i = 0; i < res.size(); i++) {
        out_s[i] = res[i].first;
        out_r[i] = res[i].second;
    }
    return res.size();
}

// nq queries on `threads` workers, each taking whole queries from a shared counter, as the
// reference's ivf_run (src/bin/search.rs:310-332). Outputs are nq x p->k, counts per query.
void or_search_batch(const void* h, const float* q768, const float* q256, size_t nq, const sieve_params* p,
                     const sieve_view* v, int threads, float* out_s, uint32_t* out_r, uint32_t* counts,
                     sieve_timing* t) {
    const Index& ix = *(const Index*)h;
    std::atomic<size_t> next{0};
    std::vector<sieve_timing> per(std::max(threads, 1));
    for (auto& x : per) std::memset(&x, 0, sizeof x);
    auto work = [&](int w) {
        for (;;) {
            size_t i = next.fetch_add(1);
            if (i >= nq) break;
            counts[i] = (uint32_t)search_one(ix, q768 + i * DIM, q256 + i * PDIM, *p, make_view(v, i),
                                             out_s + i * p->k, out_r + i * p->k, &per[w]);
        }
    };
    if (threads <= 1) {
        work(0);
    } else {
        std::vector<std::thread> th;
        for (int w = 0; w < threads; w++) th.emplace_back(work, w);
        for (auto& x : th) x.join();
    }
    if (t) {
        for (auto& x : per) {
            t->probed += x.probed;
            t->scanned += x.scanned;
            t->reranked += x.reranked;
        }
    }
}

// Single-thread latency of each query in ns (src/bin/search.rs:730-738).
void or_search_latency(const void* h, const float* q768, const float* q256, size_t nq, const sieve_params* p,
                       uint64_t* lat_ns) {
    const Index& ix = *(const Index*)h;
    std::vector<float> s(p->k);
    std::vector<uint32_t> r(p->k);
    View v;
    for (size_t i = 0; i < nq; i++) {
        auto t0 = std::chrono::steady_clock::now();
        search_one(ix, q768 + i * DIM, q256 + i * PDIM, *p, v, s.data(), r.data(), nullptr);
        lat_ns[i] = (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0).count();
    }
}

// ---- build-side restatements, for checking the product's index writer ----------------
// assign_top, src/ivf.rs:137-158: the m best centroids of each row, best first.
void or_assign_top(const float* rows, size_t n, const float* cents, size_t k, size_t m, uint32_t* out, int threads) {
    const size_t B = 8192;
    std::vector<float> s;
    for (size_t r0 = 0; r0 < n; r0 += B) {
        size_t rn = std::min(B, n - r0);
        s.resize(rn * k);
        gemm_abt(rows + r0 * PDIM, cents, rn, k, PDIM, s.data(), threads);
        for (size_t i = 0; i < rn; i++) {
            const float* row = s.data() + i * k;
            if (m == 1) {
                size_t best = 0;
                float bv = std::numeric_limits<float>::lowest();
                for (size_t j = 0; j < k; j++)
                    if (row[j] > bv) {
                        bv = row[j];
                        best = j;
                    }
                out[(r0 + i)] = (uint32_t)best;
                continue;
            }
            TopK<float> t(m);
            for (size_t j = 0; j < k; j++) t.push(row[j], (uint32_t)j);
            auto so = t.into_sorted();
            for (size_t j = 0; j < m && j < so.size(); j++) out[(r0 + i) * m + j] = so[j].second;
        }
    }
}

// soar_secondary, src/ivf.rs:164-186.
uint32_t or_soar_secondary(const float* x, const float* cents, const uint32_t* cand, size_t ncand, float lambda) {
    const float* c1 = cents + (size_t)cand[0] * PDIM;
    float x_c1 = dot_f32(x, c1, PDIM);
    float r2 = std::fmax(2.0f - 2.0f * x_c1, 1e-12f);
    float r_x = 1.0f - x_c1;
    float best = std::numeric_limits<float>::max();
    uint32_t bc = cand[1];
    for (size_t i = 1; i < ncand; i++) {
        const float* cv = cents + (size_t)cand[i] * PDIM;
        float x_c = dot_f32(x, cv, PDIM);
        float r_c = x_c - dot_f32(c1, cv, PDIM);
        float along = r_x - r_c;
        float loss = (2.0f - 2.0f * x_c) + lambda * along * along / r2;
        if (loss < best) {
            best = loss;
            bc = cand[i];
        }
    }
    return bc;
}

// kmeans, src/ivf.rs:188-225, with balance None (no split_large): initial centroids drawn
// from Rng(seed), then per iteration assign (the first best centroid), the members' sums in
// sample order from +0.0, quant::normalize, and a fresh sample row for each empty cluster in
// cluster order.
void or_kmeans(const float* sample, size_t n, size_t k, size_t iters, uint64_t seed, float* cents, int threads) {
    Rng rng(seed);
    for (size_t c = 0; c < k; c++) std::memcpy(cents + c * PDIM, sample + rng.below(n) * PDIM, PDIM * sizeof(float));
    std::vector<uint32_t> a(n);
    std::vector<float> sums(k * PDIM);
    std::vector<size_t> counts(k);
    for (size_t it = 0; it < iters; it++) {
        or_assign_top(sample, n, cents, k, 1, a.data(), threads);
        std::fill(sums.begin(), sums.end(), 0.0f);
        std::fill(counts.begin(), counts.end(), 0);
        for (size_t i = 0; i < n; i++) {
            counts[a[i]]++;
            float* s = sums.data() + (size_t)a[i] * PDIM;
            const float* x = sample + i * PDIM;
            for (int d = 0; d < PDIM; d++) s[d] = s[d] + x[d];
        }
        for (size_t c = 0; c < k; c++) {
            float* dst = cents + c * PDIM;
            if (counts[c] == 0) {
                std::memcpy(dst, sample + rng.below(n) * PDIM, PDIM * sizeof(float));
            } else {
                std::memcpy(dst, sums.data() + c * PDIM, PDIM * sizeof(float));
                normalize(dst, PDIM);
            }
        }
    }
}

// The (primary, secondary) clusters of each row as ivf::build assigns them (src/ivf.rs:313-321):
// assign_top(16) and soar_secondary, or (assign, PAD) when lambda is NaN.
void or_build_pairs(const float* rows, size_t n, const float* cents, size_t k, float lambda, uint32_t* pairs,
                    int threads) {
    const bool soar = lambda == lambda;
    const size_t m = soar ? std::min<size_t>(16, k) : 1;
    std::vector<uint32_t> top(n * 16, 0u);  // a list shorter than 16 reads cluster 0 past its end
    std::vector<uint32_t> t(n * m);
    or_assign_top(rows, n, cents, k, m, t.data(), threads);
    for (size_t i = 0; i < n; i++) {
        for (size_t j = 0; j < m; j++) top[i * 16 + j] = t[i * m + j];
        pairs[2 * i] = top[i * 16];
        pairs[2 * i + 1] = soar ? or_soar_secondary(rows + i * PDIM, cents, top.data() + i * 16, 16, lambda) : PAD;
    }
}

// ---- L2 path over 100-d int8 rows (src/bigann.rs) -----------------------------------
// Index arrays: leaves k x 100 f32, leaf norms k, block_off k+1, blocks nblocks x 512 B,
// factors nblocks x 32 x [mul, |x|^2], ids nblocks x 32, rotation 3 x 128 signs,
// base n x 100 i8.
struct L2Index {
    uint32_t k;
    const float* leaves;
    const float* leaf_norm;
    const uint64_t* block_off;
    const uint8_t* blocks;
    const float* factors;
    const uint32_t* ids;
    const float* rotation;
    const int8_t* base;
    // Multi-bit refinement (src/bigann.rs:38-113): `bits` (1 = none) and, per slot in slot order,
    // refine_bytes(bits) bytes: the low bits - 1 planes of the code, then the scale f (LE f32).
    int bits = 1;
    const uint8_t* extra = nullptr;
};

// ---------------------------------------------------------------- 1B build ------------
constexpr int L2_D = 100;

// norms, src/bigann.rs:135-137: sum of x * x per centroid, left to right from -0.0.
void or_l2_norms(const float* cents, size_t k, float* out) {
    for (size_t c = 0; c < k; c++) {
        float s = rsum_seed();
        for (int i = 0; i < L2_D; i++) s = s + cents[c * L2_D + i] * cents[c * L2_D + i];
        out[c] = s;
    }
}

// top_l2, src/bigann.rs:117-133: the m best centroids of each row by 2 <x, c> - |c|^2 (the
// x86 gemm_abt dot), best first, ties by id (TopK::into_sorted).
void or_top_l2(const float* rows, size_t n, const float* cents, const float* cnorm, size_t k, size_t m, uint32_t* out_id,
               float* out_s, int threads) {
    const size_t B = 4096;
    std::vector<float> s;
    for (size_t r0 = 0; r0 < n; r0 += B) {
        const size_t rn = std::min(B, n - r0);
        s.resize(rn * k);
        gemm_abt(rows + r0 * L2_D, cents, rn, k, L2_D, s.data(), threads);
        for (size_t i = 0; i < rn; i++) {
            TopK<float> t(m);
            for (size_t j = 0; j < k; j++) t.push(2.0f * s[i * k + j] - cnorm[j], (uint32_t)j);
            auto so = t.into_sorted();
            for (size_t j = 0; j < m; j++) {
                out_id[(r0 + i) * m + j] = j < so.size() ? so[j].second : 0u;
                out_s[(r0 + i) * m + j] = j < so.size() ? so[j].first : 0.0f;
            }
        }
    }
}

// kmeans_l2, src/bigann.rs:139-161: Lloyd in L2; f64 sums in sample order, means (s / m) as f32;
// an empty cluster takes a fresh sample row.
void or_kmeans_l2(const float* sample, size_t n, size_t k, size_t iters, uint64_t seed, float* cents, int threads) {
    Rng rng(seed);
    for (size_t c = 0; c < k; c++) std::memcpy(cents + c * L2_D, sample + rng.below(n) * L2_D, L2_D * sizeof(float));
    std::vector<float> cn(k), sc(n);
    std::vector<uint32_t> a(n);
    std::vector<double> sums(k * L2_D);
    std::vector<size_t> counts(k);
    for (size_t it = 0; it < iters; it++) {
        or_l2_norms(cents, k, cn.data());
        or_top_l2(sample, n, cents, cn.data(), k, 1, a.data(), sc.data(), threads);
        std::fill(sums.begin(), sums.end(), 0.0);
        std::fill(counts.begin(), counts.end(), 0);
        for (size_t i = 0; i < n; i++) {
            counts[a[i]]++;
            for (int d = 0; d < L2_D; d++) sums[(size_t)a[i] * L2_D + d] += (double)sample[i * L2_D + d];
        }
        for (size_t c = 0; c < k; c++) {
            float* dst = cents + c * L2_D;
            if (counts[c] == 0) {
                std::memcpy(dst, sample + rng.below(n) * L2_D, L2_D * sizeof(float));
            } else {
                for (int d = 0; d < L2_D; d++) dst[d] = (float)(sums[c * L2_D + d] / (double)counts[c]);
            }
        }
    }
}

// train, src/bigann.rs:175-227, without the file writes: the Rng(42) sample of base rows
// (indices sorted), the coarse rows at an even stride through it, k1 coarse cells by
// kmeans_l2 (seed 7), every sample row to its nearest cell, and k2 leaves per cell by
// kmeans_l2 over the cell's rows (seed = the cell; an empty cell's leaves are its centroid).
void or_l2_train(const int8_t* base, size_t n, size_t k1, size_t k2, size_t sample_n, size_t iters, float* coarse,
                 float* leaves, int threads) {
    Rng rng(42);
    std::vector<size_t> idx(sample_n);
    for (auto& i : idx) i = rng.below(n);
    std::sort(idx.begin(), idx.end());
    std::vector<float> sample(sample_n * L2_D);
    for (size_t r = 0; r < sample_n; r++)
        for (int d = 0; d < L2_D; d++) sample[r * L2_D + d] = (float)base[idx[r] * L2_D + d];
    const size_t coarse_n = std::min(sample_n, k1 * 512);
    const size_t stride = sample_n / coarse_n;
    std::vector<float> crow(coarse_n * L2_D);
    for (size_t r = 0; r < coarse_n; r++)
        std::memcpy(crow.data() + r * L2_D, sample.data() + r * stride * L2_D, L2_D * sizeof(float));
    or_kmeans_l2(crow.data(), coarse_n, k1, iters, 7, coarse, threads);
    std::vector<float> cn(k1), sc(sample_n);
    std::vector<uint32_t> a(sample_n);
    or_l2_norms(coarse, k1, cn.data());
    or_top_l2(sample.data(), sample_n, coarse, cn.data(), k1, 1, a.data(), sc.data(), threads);
    std::vector<std::vector<uint32_t>> members(k1);
    for (size_t i = 0; i < sample_n; i++) members[a[i]].push_back((uint32_t)i);
    for (size_t c = 0; c < k1; c++) {
        std::vector<float> rows;
        if (members[c].empty()) {
            rows.assign(coarse + c * L2_D, coarse + (c + 1) * L2_D);
        } else {
            rows.resize(members[c].size() * L2_D);
            for (size_t j = 0; j < members[c].size(); j++)
                std::memcpy(rows.data() + j * L2_D, sample.data() + (size_t)members[c][j] * L2_D, L2_D * sizeof(float));
        }
        or_kmeans_l2(rows.data(), rows.size() / L2_D, k2, iters, (uint64_t)c, leaves + c * k2 * L2_D, threads);
    }
}

// Model::assign, src/bigann.rs:277-302: each row's leaf is the best child of its 4 nearest
// coarse cells, the cells taken in ascending order and a later one winning only if strictly
// better.
void or_l2_assign(const float* rows, size_t n, const float* coarse, size_t k1, const float* leaves, size_t k2,
                  uint32_t* out, int threads) {
    const size_t TOP = 4;
    std::vector<float> cn(k1), ln(k1 * k2), ts(n * TOP);
    std::vector<uint32_t> top(n * TOP);
    or_l2_norms(coarse, k1, cn.data());
    or_l2_norms(leaves, k1 * k2, ln.data());
    or_top_l2(rows, n, coarse, cn.data(), k1, TOP, top.data(), ts.data(), threads);
    for (size_t i = 0; i < n; i++) {
        uint32_t cells[TOP];
        std::copy(top.begin() + i * TOP, top.begin() + (i + 1) * TOP, cells);
        std::sort(cells, cells + TOP);
        uint32_t best = 0;
        float bs = std::numeric_limits<float>::lowest();
        for (size_t t = 0; t < TOP; t++) {
            const uint32_t c = cells[t];
            uint32_t j;
            float s;
            or_top_l2(rows + i * L2_D, 1, leaves + (size_t)c * k2 * L2_D, ln.data() + (size_t)c * k2, k2, 1, &j, &s, 1);
            if (s > bs) {
                bs = s;
                best = (uint32_t)(c * k2 + j);
            }
        }
        out[i] = best;
    }
}

// ---------------------------------------------------------------- refinement ----------
constexpr int RF_PD = 128, RF_CODE = RF_PD / 8, RF_MAX_BITS = 3;

// refine_bytes, src/bigann.rs:110-113.
size_t refine_bytes(int bits) { return bits > 1 ? (size_t)(bits - 1) * RF_CODE + 4 : 0; }

// refine_encode, src/bigann.rs:70-108: the uniform `bits`-bit code u of the rotated residual r
// (levels u - h, h = (2^bits - 1) / 2) at the step of 24 tried that maximises cos(r, u - h), its
// top bit forced to the sign; writes the low bits - 1 planes (plane p = bit p of u, dim i at
// byte i / 8, bit i % 8) and returns f = |r|^2 / <r, u - h> (0 when that is not positive).
float refine_encode(const float* r, int bits, uint8_t* planes) {
    std::memset(planes, 0, (RF_MAX_BITS - 1) * RF_CODE);
    const uint32_t top = 1u << (bits - 1);
    const float h = (float)(2 * top - 1) / 2.0f;
    float amax = 0.0f;
    for (int i = 0; i < RF_PD; i++) amax = std::fmax(amax, std::fabs(r[i]));
    if (amax < 1e-12f) return 0.0f;
    auto level = [&](float v, float delta) -> uint32_t {
        float x = std::floor(v / delta + h + 0.5f);
        x = x < 0.0f ? 0.0f : (x > (float)(2 * top - 1) ? (float)(2 * top - 1) : x);  // f32::clamp
        const uint32_t u = x != x ? 0u : (uint32_t)x;                                  // `as u32`
        return v > 0.0f ? (u | top) : (u & (top - 1));
    };
    float best = -std::numeric_limits<float>::max(), best_delta = amax / h;
    for (int st = 0; st < 24; st++) {
        const float delta = amax / h * (0.3f + 0.035f * (float)st);
        float dot = 0.0f, n2 = 0.0f;
        for (int i = 0; i < RF_PD; i++) {
            const float w = (float)level(r[i], delta) - h;
            dot = dot + r[i] * w;
            n2 = n2 + w * w;
        }
        const float cosv = dot / std::sqrt(n2);
        if (cosv > best) {
            best = cosv;
            best_delta = delta;
        }
    }
    float rr = 0.0f, dot = 0.0f;
    for (int i = 0; i < RF_PD; i++) {
        const uint32_t u = level(r[i], best_delta);
        for (int p = 0; p < bits - 1; p++) planes[p * RF_CODE + (i >> 3)] |= (uint8_t)(((u >> p) & 1u) << (i & 7));
        rr = rr + r[i] * r[i];
        dot = dot + r[i] * ((float)u - h);
    }
    return dot > 0.0f ? rr / dot : 0.0f;
}

// Refiner, src/bigann.rs:565-602: per code byte k and byte value v, the sum of the rotated query
// over the dims whose bits are set in v (t[v] = t[v & (v - 1)] + q[8 k + lowest set bit]), and
// off = h * sum(q).
struct Refiner {
    std::vector<float> tab;  // RF_CODE x 256
    float off = 0.0f;
    int bits = 1;
    Refiner(const float* qr, int bits_) : tab((size_t)RF_CODE * 256, 0.0f), bits(bits_) {
        for (int k = 0; k < RF_CODE; k++) {
            float* t = tab.data() + k * 256;
            t[0] = 0.0f;
            for (int v = 1; v < 256; v++) t[v] = t[v & (v - 1)] + qr[8 * k + __builtin_ctz((unsigned)v)];
        }
        const float h = (float)((1u << bits) - 1) / 2.0f;
        float sum = rsum_seed();
        for (int i = 0; i < RF_PD; i++) sum = sum + qr[i];
        off = h * sum;
    }
    float plane(const uint8_t* p) const {
        float s = rsum_seed();
        for (int k = 0; k < RF_CODE; k++) s = s + tab[(size_t)k * 256 + p[k]];
        return s;
    }
    // <q, r> ~ f * (<q, u> - h * sum q), u = 2^(bits - 1) sign + sum_p 2^p plane_p.
    float estimate(const uint8_t* sign, const uint8_t* planes, float f) const {
        float u = (float)(1u << (bits - 1)) * plane(sign);
        for (int p = 0; p < bits - 1; p++) u = u + (float)(1u << p) * plane(planes + p * RF_CODE);
        return f * (u - off);
    }
};

void or_l2_leaf_norms(const float* leaves, size_t k, float* out) {  // norms, src/bigann.rs:135-137
    for (size_t c = 0; c < k; c++) {
        float s = rsum_seed();
        for (int i = 0; i < 100; i++) s = s + leaves[c * 100 + i] * leaves[c * 100 + i];
        out[c] = s;
    }
}

void* or_l2_create(uint32_t k, const float* leaves, const float* leaf_norm, const uint64_t* block_off,
                   const uint8_t* blocks, const float* factors, const uint32_t* ids, const float* rotation,
                   const int8_t* base) {
    return new L2Index{k, leaves, leaf_norm, block_off, blocks, factors, ids, rotation, base};
}
void or_l2_free(void* h) { delete (L2Index*)h; }

// bigann::Index::search without refinement (src/bigann.rs:652-747): route, candidates, rerank.
// Writes probes (count returned in *nprobe_out), candidates and the final top-k.
size_t or_l2_search_refine(const void* h, const int8_t* q, uint32_t nprobe, uint32_t cand, uint32_t k, float margin,
                           uint32_t refine, uint32_t* probe_out, uint32_t* nprobe_out, uint32_t* cand_ids,
                           float* cand_est, uint32_t* ncand_out, float* out_s, uint32_t* out_r);
size_t or_l2_search(const void* h, const int8_t* q, uint32_t nprobe, uint32_t cand, uint32_t k, float margin,
                    uint32_t* probe_out, uint32_t* nprobe_out, uint32_t* cand_ids, float* cand_est,
                    uint32_t* ncand_out, float* out_s, uint32_t* out_r) {
    return or_l2_search_refine(h, q, nprobe, cand, k, margin, 0, probe_out, nprobe_out, cand_ids, cand_est, ncand_out,
                               out_s, out_r);
}

// bigann::Index::search (src/bigann.rs:652-747) with the refinement parameter: on an index with
// bits > 1 and refine > 0, the 1-bit shortlist of max(refine, cand) gates which slots are
// re-estimated from all planes (at the moment each is scanned, in probe order, as
// src/bigann.rs:699-729 does), and the best cand by the refined estimate are reranked.
size_t or_l2_search_refine(const void* h, const int8_t* q, uint32_t nprobe, uint32_t cand, uint32_t k, float margin,
                           uint32_t refine, uint32_t* probe_out, uint32_t* nprobe_out, uint32_t* cand_ids,
                           float* cand_est, uint32_t* ncand_out, float* out_s, uint32_t* out_r) {
    const L2Index& ix = *(const L2Index*)h;
    const int D = 100, PD = 128, G = 32;
    // route, src/bigann.rs:659-678
    std::vector<float> qf(D);
    for (int i = 0; i < D; i++) qf[i] = (float)q[i];
    std::vector<float> dots(ix.k);
    gemm_abt(qf.data(), ix.leaves, 1, ix.k, D, dots.data(), 1);
    TopK<float> top(nprobe);
    for (uint32_t c = 0; c < ix.k; c++) top.push(2.0f * dots[c] - ix.leaf_norm[c], c);
    auto probe = top.into_sorted();
    if (margin > 0.0f && !probe.empty()) {
        float qq = rsum_seed();
        for (int i = 0; i < D; i++) qq = qq + qf[i] * qf[i];
        float lim = (qq - probe[0].first) * (1.0f + margin);
        std::vector<std::pair<float, uint32_t>> kept;
        for (auto& x : probe)
            if (qq - x.first <= lim) kept.push_back(x);
        probe.swap(kept);
    }
    *nprobe_out = (uint32_t)probe.size();
    for (size_t i = 0; i < probe.size(); i++) probe_out[i] = probe[i].second;
    // candidates, src/bigann.rs:683-734
    std::vector<float> qp(PD, 0.0f);
    for (int i = 0; i < D; i++) qp[i] = (float)q[i];
    QueryState rq;
    query_new(ix.rotation, PD, qp.data(), rq);
    const bool refining = ix.bits > 1 && refine > 0;
    std::unique_ptr<Refiner> rf;
    if (refining) {
        std::vector<float> qr(qp);
        rotation_apply(ix.rotation, PD, qr.data());  // pad_rotate, src/bigann.rs:249-254
        rf.reset(new Refiner(qr.data(), ix.bits));
    }
    const size_t extra = refine_bytes(ix.bits);
    std::vector<std::pair<uint32_t, float>> info;  // (id, |x|^2); est lives in the TopK
    std::vector<float> info_est;
    // Without refinement `ct` is the candidate list; with it, the 1-bit shortlist that gates
    // which slots are re-estimated, and `best` keeps the top cand by the refined estimate.
    TopK<float> ct(refining ? std::max(refine, cand) : cand);
    TopK<float> best(cand);
    std::vector<uint16_t> scores;
    uint8_t sign[RF_CODE];
    for (auto& pr : probe) {
        uint32_t c = pr.second;
        float cent = dots[c];
        uint64_t b0 = ix.block_off[c], b1 = ix.block_off[c + 1];
        size_t ns = (size_t)(b1 - b0) * BLOCK;
        scores.resize(ns);
        scan(rq.tables.data(), G, ix.blocks + b0 * 512, b1 - b0, scores.data());
        const float* f = ix.factors + b0 * BLOCK * 2;
        const uint32_t* ids = ix.ids + b0 * BLOCK;
        for (size_t j = 0; j < ns; j++) {
            if (ids[j] == PAD) continue;
            float mul = f[2 * j], nx = f[2 * j + 1];
            float est = 2.0f * (cent + mul * (rq.a * (float)scores[j] + rq.b)) - nx;
            if (!ct.has_thr || est > ct.thr) {
                ct.push(est, (uint32_t)info.size());
                if (refining) {
                    unpack_slot(ix.blocks + (b0 + j / BLOCK) * 512, G, (int)(j % BLOCK), sign);
                    const uint8_t* e = ix.extra + (b0 * BLOCK + j) * extra;
                    float fs;
                    std::memcpy(&fs, e + extra - 4, 4);
                    est = 2.0f * (cent + rf->estimate(sign, e, fs)) - nx;
                    best.push(est, (uint32_t)info.size());
                }
                info.emplace_back(ids[j], nx);
                info_est.push_back(est);
            }
        }
    }
    auto cs = refining ? best.into_sorted() : ct.into_sorted();
    *ncand_out = (uint32_t)cs.size();
    // rerank, src/bigann.rs:736-747
    TopK<float> out(k);
    for (size_t i = 0; i < cs.size(); i++) {
        uint32_t id = info[cs[i].second].first;
        float nx = info[cs[i].second].second;
        cand_ids[i] = id;
        cand_est[i] = info_est[cs[i].second];
        const int8_t* x = ix.base + (size_t)id * D;
        int32_t dot = 0;
        for (int j = 0; j < D; j++) dot += (int32_t)q[j] * (int32_t)x[j];
        out.push(2.0f * (float)dot - nx, id);
    }
    auto res = out.into_sorted();
    for (size_t i = 0; i < res.size(); i++) {
        out_s[i] = res[i].first;
        out_r[i] = res[i].second;
    }
    return res.size();
}

// Attaches a multi-bit refinement section (per slot, refine_bytes(bits) bytes) to an L2 index.
void or_l2_set_refine(void* h, int bits, const uint8_t* extra) {
    L2Index& ix = *(L2Index*)h;
    ix.bits = bits;
    ix.extra = extra;
}

void or_refine_encode(const float* r, int bits, uint8_t* planes, float* f) { *f = refine_encode(r, bits, planes); }

// Refiner::estimate for one rotated query and one slot (sign code, planes, f).
float or_refine_estimate(const float* qr, int bits, const uint8_t* sign, const uint8_t* planes, float f) {
    return Refiner(qr, bits).estimate(sign, planes, f);
}

// hot_leaves, src/bigann.rs:537-563: the leaves to keep resident under `budget` bytes, hottest
// per byte first: how often the queries route to a leaf (nprobe, cut by margin) over its record's
// bytes (record_bytes[c]); leaves no query probes are left out. Returns the count written.
size_t or_l2_hot_leaves(const void* h, const int8_t* q, size_t nq, uint32_t nprobe, float margin,
                        const uint64_t* record_bytes, uint64_t budget, uint32_t* out) {
    const L2Index& ix = *(const L2Index*)h;
    std::vector<uint32_t> hits(ix.k, 0), probe(std::max<uint32_t>(nprobe, 1)), ci(1);
    std::vector<float> ce(1), s(10);
    std::vector<uint32_t> r(10);
    uint32_t np = 0, nc = 0;
    for (size_t i = 0; i < nq; i++) {
        or_l2_search_refine(h, q + i * 100, nprobe, 1, 10, margin, 0, probe.data(), &np, ci.data(), ce.data(), &nc,
                            s.data(), r.data());
        for (uint32_t j = 0; j < np; j++) hits[probe[j]]++;
    }
    auto heat = [&](uint32_t c) { return (double)hits[c] / (double)std::max<uint64_t>(record_bytes[c], 1); };
    std::vector<uint32_t> order;
    for (uint32_t c = 0; c < ix.k; c++)
        if (hits[c] > 0) order.push_back(c);
    // slice::sort_by is stable: equal heat keeps leaf order
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return heat(b) < heat(a); });
    uint64_t used = 0;
    size_t n = 0;
    for (uint32_t c : order) {
        used += record_bytes[c];
        if (used > budget) break;
        out[n++] = c;
    }
    return n;
}

// or_l2_search over nq queries (nq x 100 i8), whole queries per worker thread as the reference's
// search bench runs them (src/bin/search.rs:310-332); top-k into out_s / out_r (nq x k).
void or_l2_search_batch(const void* h, const int8_t* q, size_t nq, uint32_t nprobe, uint32_t cand, uint32_t k,
                        float margin, uint32_t refine, int threads, float* out_s, uint32_t* out_r, uint32_t* counts) {
    std::atomic<size_t> next{0};
    auto work = [&]() {
        std::vector<uint32_t> probe(std::max<uint32_t>(nprobe, 1)), ci(std::max<uint32_t>(cand, 1));
        std::vector<float> ce(std::max<uint32_t>(cand, 1));
        uint32_t np = 0, nc = 0;
        for (;;) {
            size_t i = next.fetch_add(1);
            if (i >= nq) break;
            counts[i] = (uint32_t)or_l2_search_refine(h, q + i * 100, nprobe, cand, k, margin, refine, probe.data(), &np,
                                                      ci.data(), ce.data(), &nc, out_s + i * k, out_r + i * k);
        }
    };
    if (threads <= 1) {
        work();
    } else {
        std::vector<std::thread> th;
        for (int w = 0; w < threads; w++) th.emplace_back(work);
        for (auto& x : th) x.join();
    }
}

// Single-thread latency of each L2 query in ns (src/bin/search.rs:730-738).
void or_l2_search_latency(const void* h, const int8_t* q, size_t nq, uint32_t nprobe, uint32_t cand, uint32_t k,
                          float margin, uint32_t refine, uint64_t* lat_ns) {
    std::vector<uint32_t> probe(std::max<uint32_t>(nprobe, 1)), ci(std::max<uint32_t>(cand, 1)), r(k);
    std::vector<float> ce(std::max<uint32_t>(cand, 1)), s(k);
    uint32_t np = 0, nc = 0;
    for (size_t i = 0; i < nq; i++) {
        auto t0 = std::chrono::steady_clock::now();
        or_l2_search_refine(h, q + i * 100, nprobe, cand, k, margin, refine, probe.data(), &np, ci.data(), ce.data(), &nc,
                            s.data(), r.data());
        lat_ns[i] = (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0).count();
    }
}

}  // extern "C"

// ---- live updates, src/live.rs ------------------------------------------------------
// Live: the immutable index plus a Delta (inserted rows in per-cluster delta lists), a
// tombstone bitset over row ids and the external-id map. One writer at a time.
struct OLive {
    const Index* ix;
    bool soar;
    float lambda;
    std::vector<float> cents_rot;  // k x 256, Rotation::apply of each centroid (src/live.rs:172-180)
    Delta delta;
    std::vector<uint64_t> deleted;
    std::unordered_map<uint32_t, uint32_t> by_ext;
    std::string wal;  // the log (src/live.rs:161), "" without one
    int wal_fd = -1;
    ~OLive() {
        if (wal_fd >= 0) ::close(wal_fd);
    }
};

// Live::log (src/live.rs:251-258): the records, then fdatasync.
static bool live_log(OLive& l, const std::vector<uint8_t>& rec) {
    if (l.wal_fd < 0) return true;
    size_t done = 0;
    while (done < rec.size()) {
        ssize_t w = ::write(l.wal_fd, rec.data() + done, rec.size() - done);
        if (w <= 0) return false;
        done += (size_t)w;
    }
    return ::fdatasync(l.wal_fd) == 0;
}

// OP_INSERT / OP_DELETE and the record sizes, src/live.rs:131-135.
constexpr uint8_t OP_INSERT = 1, OP_DELETE = 2;
constexpr size_t INSERT_BYTES = 1 + 4 + 4 + DIM, DELETE_BYTES = 1 + 4;

static void le32_put(std::vector<uint8_t>& v, uint32_t x) {
    for (int i = 0; i < 4; i++) v.push_back((uint8_t)(x >> (8 * i)));
}
static uint32_t le32_get(const uint8_t* p) {
    return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24;
}

static bool live_bit(const std::vector<uint64_t>& set, uint32_t id) {  // src/live.rs:119-121
    size_t w = id >> 6;
    return w < set.size() && ((set[w] >> (id & 63)) & 1) == 1;
}
static void live_set_bit(std::vector<uint64_t>& set, uint32_t id) {  // src/live.rs:123-129
    size_t w = id >> 6;
    if (set.size() <= w) set.resize(w + 1, 0);
    set[w] |= 1ull << (id & 63);
}
// Live::current, src/live.rs:260-266: the row an external id resolves to.
static bool live_current(const OLive& l, uint32_t e, uint32_t* id) {
    auto it = l.by_ext.find(e);
    if (it != l.by_ext.end()) {
        *id = it->second;
        return true;
    }
    if ((uint64_t)e < l.ix->n && !live_bit(l.deleted, e)) {
        *id = e;
        return true;
    }
    return false;
}
static uint32_t live_ext_of(const OLive& l, uint32_t id) {  // src/live.rs:268-273
    size_t i;
    return l.delta.local(id, &i) ? l.delta.ext[i] : id;
}

// Live::apply_insert, src/live.rs:317-374: encode as ivf::build encodes a row, then append.
static void live_apply_insert(OLive& l, const uint32_t* ext, const int8_t* rows, const float* norms, size_t n) {
    if (n == 0) return;
    const Index& ix = *l.ix;
    std::vector<float> xs(n * PDIM);
    for (size_t j = 0; j < n; j++) {
        float* x = xs.data() + j * PDIM;
        for (int i = 0; i < PDIM; i++) x[i] = (float)rows[j * DIM + i] / ix.quant[i];
        normalize(x, PDIM);
    }
    std::vector<uint32_t> pairs(n * 2, PAD);
    if (l.soar) {
        const size_t m = 16;  // SOAR_CAND, src/live.rs:35
        std::vector<uint32_t> top(n * m, 0);
        or_assign_top(xs.data(), n, ix.centroids, ix.k, m, top.data(), 1);
        for (size_t j = 0; j < n; j++) {
            pairs[2 * j] = top[j * m];
            pairs[2 * j + 1] = or_soar_secondary(xs.data() + j * PDIM, ix.centroids, top.data() + j * m, m, l.lambda);
        }
    } else {
        std::vector<uint32_t> top(n);
        or_assign_top(xs.data(), n, ix.centroids, ix.k, 1, top.data(), 1);
        for (size_t j = 0; j < n; j++) pairs[2 * j] = top[j];
    }
    std::vector<uint8_t> codes(n * 2 * 32, 0);
    std::vector<float> fs(n * 2 * 2, 0.0f);
    std::vector<float> xr(PDIM);
    for (size_t j = 0; j < n; j++) {
        std::copy(xs.begin() + j * PDIM, xs.begin() + (j + 1) * PDIM, xr.begin());
        rotation_apply(ix.rotation, PDIM, xr.data());
        for (int h = 0; h < 2; h++) {
            uint32_t c = pairs[2 * j + h];
            if (c == PAD) continue;
            rabitq_encode(xr.data(), l.cents_rot.data() + (size_t)c * PDIM, PDIM, codes.data() + (2 * j + h) * 32,
                          fs.data() + (2 * j + h) * 2);
        }
    }
    if (ix.t2_sigma) {
        size_t at = l.delta.t2.size();
        l.delta.t2.resize(at + n * 64, 0);
        for (size_t j = 0; j < n; j++)
            l.delta.t2_norm.push_back(tier2_encode(ix.t2_sigma, xs.data() + j * PDIM, l.delta.t2.data() + at + j * 64));
    }
    for (size_t j = 0; j < n; j++) {
        uint32_t old;
        if (live_current(l, ext[j], &old)) live_set_bit(l.deleted, old);
        uint32_t id = (uint32_t)(l.delta.base + l.delta.ext.size());
        l.delta.ext.push_back(ext[j]);
        l.delta.rows.insert(l.delta.rows.end(), rows + j * DIM, rows + (j + 1) * DIM);
        l.delta.norm256.push_back(norms[j]);
        l.by_ext[ext[j]] = id;
        for (int h = 0; h < 2; h++) {
            uint32_t c = pairs[2 * j + h];
            if (c == PAD) continue;
            l.delta.lists[c].push(codes.data() + (2 * j + h) * 32, fs.data() + (2 * j + h) * 2, id);
        }
    }
}

extern "C" {

void* or_live_create(const void* index, int soar, float lambda) {
    const Index& ix = *(const Index*)index;
    OLive* l = new OLive();
    l->ix = &ix;
    l->soar = soar != 0;
    l->lambda = lambda;
    l->cents_rot.assign(ix.centroids, ix.centroids + (size_t)ix.k * PDIM);
    for (uint32_t c = 0; c < ix.k; c++) rotation_apply(ix.rotation, PDIM, l->cents_rot.data() + (size_t)c * PDIM);
    l->delta.base = ix.n;
    l->delta.lists.resize(ix.k);
    return l;
}
void or_live_free(void* h) { delete (OLive*)h; }

// Live::insert, src/live.rs:275-294 (no log): int8 rows by quant.encode, prefix norms of the
// float vectors, then apply_insert.
int or_live_insert(void* h, const uint32_t* ext, const float* vecs, size_t n) {
    OLive& l = *(OLive*)h;
    std::vector<int8_t> rows(n * DIM);
    std::vector<float> norms(n);
    for (size_t j = 0; j < n; j++) {
        or_int8_encode(vecs + j * DIM, l.ix->quant, DIM, rows.data() + j * DIM);
        float s = rsum_seed();
        for (int i = 0; i < PDIM; i++) s = s + vecs[j * DIM + i] * vecs[j * DIM + i];
        norms[j] = std::sqrt(s);
    }
    std::vector<uint8_t> rec;
    for (size_t j = 0; j < n; j++) {  // src/live.rs:284-290
        rec.push_back(OP_INSERT);
        le32_put(rec, ext[j]);
        uint32_t nb;
        std::memcpy(&nb, &norms[j], 4);
        le32_put(rec, nb);
        rec.insert(rec.end(), reinterpret_cast<const uint8_t*>(rows.data() + j * DIM),
                   reinterpret_cast<const uint8_t*>(rows.data() + (j + 1) * DIM));
    }
    if (!live_log(l, rec)) return -1;
    live_apply_insert(l, ext, rows.data(), norms.data(), n);
    return 0;
}

static void live_apply_delete(OLive& l, const uint32_t* ext, size_t n) {  // src/live.rs:308-315
    for (size_t j = 0; j < n; j++) {
        uint32_t id;
        if (live_current(l, ext[j], &id)) {
            live_set_bit(l.deleted, id);
            l.by_ext.erase(ext[j]);
        }
    }
}

// Live::delete, src/live.rs:296-306: log, then apply_delete.
int or_live_delete(void* h, const uint32_t* ext, size_t n) {
    OLive& l = *(OLive*)h;
    std::vector<uint8_t> rec;
    for (size_t j = 0; j < n; j++) {
        rec.push_back(OP_DELETE);
        le32_put(rec, ext[j]);
    }
    if (!live_log(l, rec)) return -1;
    live_apply_delete(l, ext, n);
    return 0;
}

uint32_t or_live_current(const void* h, uint32_t ext) {
    uint32_t id;
    return live_current(*(const OLive*)h, ext, &id) ? id : PAD;
}
size_t or_live_rows(const void* h) { return ((const OLive*)h)->delta.ext.size(); }

// Delta row j: int8 row, prefix norm, tier-2 code and norm (if any), external id.
void or_live_row(const void* h, size_t j, int8_t* row, float* norm, uint8_t* t2, float* t2n, uint32_t* ext) {
    const Delta& d = ((const OLive*)h)->delta;
    std::memcpy(row, d.rows.data() + j * DIM, DIM);
    *norm = d.norm256[j];
    if (!d.t2_norm.empty()) {
        std::memcpy(t2, d.t2.data() + j * 64, 64);
        *t2n = d.t2_norm[j];
    }
    *ext = d.ext[j];
}

// Cluster c's delta list: returns its slot count; copies raw codes (unpacked from the block
// layout), [mul, err] and row ids of its first `cap` slots.
size_t or_live_list(const void* h, uint32_t c, size_t cap, uint8_t* codes, float* factors, uint32_t* ids) {
    const DeltaList& dl = ((const OLive*)h)->delta.lists[c];
    for (size_t j = 0; j < dl.len && j < cap; j++) {
        unpack_slot(dl.blocks.data() + j / BLOCK * 1024, PDIM / 4, (int)(j % BLOCK), codes + j * 32);
        factors[2 * j] = dl.factors[2 * j];
        factors[2 * j + 1] = dl.factors[2 * j + 1];
        ids[j] = dl.ids[j];
    }
    return dl.len;
}

size_t or_live_deleted(const void* h, uint64_t* words, size_t cap) {
    const auto& d = ((const OLive*)h)->deleted;
    for (size_t i = 0; i < d.size() && i < cap; i++) words[i] = d[i];
    return d.size();
}

// Live::replay (src/live.rs:209-249) and the log opened for appending (src/live.rs:197-205).
// stats: inserts, deletes, torn bytes, and the byte offset of a bad op. Returns 0, -1 for an
// unknown op (nothing after it applied, the file untouched), -2 for an I/O error.
int or_live_open_wal(void* h, const char* path, uint64_t* stats) {
    OLive& l = *(OLive*)h;
    std::vector<uint8_t> buf;
    if (FILE* f = std::fopen(path, "rb")) {
        uint8_t tmp[65536];
        size_t r;
        while ((r = std::fread(tmp, 1, sizeof tmp, f)) > 0) buf.insert(buf.end(), tmp, tmp + r);
        std::fclose(f);
    }
    stats[0] = stats[1] = stats[2] = stats[3] = 0;
    size_t pos = 0;
    std::vector<uint32_t> ext;
    std::vector<int8_t> rows;
    std::vector<float> norms;
    while (pos < buf.size()) {
        size_t need;
        if (buf[pos] == OP_INSERT)
            need = INSERT_BYTES;
        else if (buf[pos] == OP_DELETE)
            need = DELETE_BYTES;
        else {
            stats[3] = pos;
            return -1;
        }
        if (pos + need > buf.size()) break;
        const uint8_t* r = buf.data() + pos;
        if (r[0] == OP_INSERT) {
            ext.push_back(le32_get(r + 1));
            uint32_t nb = le32_get(r + 5);
            float nf;
            std::memcpy(&nf, &nb, 4);
            norms.push_back(nf);
            rows.insert(rows.end(), reinterpret_cast<const int8_t*>(r + 9), reinterpret_cast<const int8_t*>(r + 9) + DIM);
            stats[0]++;
        } else {
            live_apply_insert(l, ext.data(), rows.data(), norms.data(), ext.size());
            ext.clear();
            rows.clear();
            norms.clear();
            uint32_t e = le32_get(r + 1);
            live_apply_delete(l, &e, 1);
            stats[1]++;
        }
        pos += need;
    }
    live_apply_insert(l, ext.data(), rows.data(), norms.data(), ext.size());
    stats[2] = buf.size() - pos;
    if (stats[2] && ::truncate(path, (off_t)pos) != 0) return -2;
    l.wal = path;
    l.wal_fd = ::open(path, O_WRONLY | O_CREAT | O_APPEND, 0644);
    return l.wal_fd >= 0 ? 0 : -2;
}

// Live::truncate_wal, src/live.rs:470-479.
int or_live_truncate_wal(void* h) {
    OLive& l = *(OLive*)h;
    if (l.wal_fd < 0) return 0;
    return ::ftruncate(l.wal_fd, 0) == 0 && ::fsync(l.wal_fd) == 0 ? 0 : -2;
}

// The row segment of a compacted index, Live::open (src/live.rs:183-195): rows, prefix norms,
// tier-2 codes and norms (t2 NULL: none), external ids, tombstone words; every delta row whose
// tombstone is clear is its external id's row.
void or_live_load_segment(void* h, const int8_t* rows, const float* norm256, const uint8_t* t2, const float* t2n,
                          const uint32_t* ext, size_t nrows, const uint64_t* deleted, size_t words) {
    OLive& l = *(OLive*)h;
    Delta& d = l.delta;
    d.rows.assign(rows, rows + nrows * DIM);
    d.norm256.assign(norm256, norm256 + nrows);
    if (t2) {
        d.t2.assign(t2, t2 + nrows * 64);
        d.t2_norm.assign(t2n, t2n + nrows);
    } else {
        d.t2.clear();
        d.t2_norm.clear();
    }
    d.ext.assign(ext, ext + nrows);
    l.deleted.assign(deleted, deleted + words);
    l.by_ext.clear();
    for (size_t i = 0; i < nrows; i++) {
        uint32_t id = (uint32_t)(l.ix->n + i);
        if (!live_bit(l.deleted, id)) l.by_ext[ext[i]] = id;
    }
}

// Live::compact's merge (src/live.rs:402-448): per cluster the base list's then the delta list's
// slots, PAD and tombstoned ones left out, packed into blocks. Writes the new lists if they fit
// in cap_blocks; returns their block count. stats: slots kept, slots dropped.
size_t or_live_compact(const void* h, size_t cap_blocks, uint64_t* block_off, uint8_t* blocks, float* factors,
                       uint32_t* ids, uint64_t* stats) {
    const OLive& l = *(const OLive*)h;
    const Index& ix = *l.ix;
    const size_t BB = 1024;  // BLOCK_BYTES of the 256-d codes
    struct Slot {
        uint8_t code[32];
        float f[2];
        uint32_t id;
    };
    std::vector<Slot> live;
    size_t nblocks = 0;
    uint64_t slots = 0, dropped = 0;
    for (uint32_t c = 0; c < ix.k; c++) {
        live.clear();
        const size_t b0 = ix.block_off[c], b1 = ix.block_off[c + 1];
        const DeltaList& dl = l.delta.lists[c];
        struct Src {
            const uint8_t* bl;
            const float* fs;
            const uint32_t* is;
            size_t n;
        } srcs[2] = {{ix.blocks + b0 * BB, ix.factors + b0 * BLOCK * 2, ix.ids + b0 * BLOCK, (b1 - b0) * BLOCK},
                     {dl.blocks.data(), dl.factors.data(), dl.ids.data(), dl.len ? dl.ids.size() : 0}};
        for (const Src& sr : srcs)
            for (size_t j = 0; j < sr.n; j++) {
                uint32_t id = sr.is[j];
                if (id == PAD) continue;
                if (live_bit(l.deleted, id)) {
                    dropped++;
                    continue;
                }
                Slot x;
                unpack_slot(sr.bl + j / BLOCK * BB, PDIM / 4, (int)(j % BLOCK), x.code);
                x.f[0] = sr.fs[j * 2];
                x.f[1] = sr.fs[j * 2 + 1];
                x.id = id;
                live.push_back(x);
            }
        if (block_off && nblocks <= cap_blocks) block_off[c] = nblocks;
        for (size_t o = 0; o < live.size(); o += BLOCK) {
            size_t m = std::min<size_t>(BLOCK, live.size() - o);
            if (nblocks < cap_blocks) {
                std::vector<uint8_t> codes(m * 32);
                for (size_t q = 0; q < m; q++) std::memcpy(codes.data() + q * 32, live[o + q].code, 32);
                pack_block(codes.data(), (int)m, PDIM / 4, blocks + nblocks * BB);
                for (size_t q = 0; q < (size_t)BLOCK; q++) {
                    factors[(nblocks * BLOCK + q) * 2] = q < m ? live[o + q].f[0] : 0.0f;
                    factors[(nblocks * BLOCK + q) * 2 + 1] = q < m ? live[o + q].f[1] : 0.0f;
                    ids[nblocks * BLOCK + q] = q < m ? live[o + q].id : PAD;
                }
            }
            nblocks++;
        }
        slots += live.size();
    }
    if (block_off && nblocks <= cap_blocks) block_off[ix.k] = nblocks;
    if (stats) {
        stats[0] = slots;
        stats[1] = dropped;
    }
    return nblocks;
}

// Live::search for nq queries (src/live.rs:381-385): search_view over tombstones and the delta
// segment (and an optional per-query allowlist), results as external ids.
void or_live_search_batch(const void* h, const float* q768, const float* q256, size_t nq, const sieve_params* p,
                          const sieve_view* v, int threads, float* out_s, uint32_t* out_r, uint32_t* counts) {
    const OLive& l = *(const OLive*)h;
    const Index& ix = *l.ix;
    std::atomic<size_t> next{0};
    auto work = [&]() {
        for (;;) {
            size_t i = next.fetch_add(1);
            if (i >= nq) break;
            View vw = make_view(v, i);
            vw.deleted = l.deleted.data();
            vw.deleted_words = l.deleted.size();
            vw.delta = &l.delta;
            size_t n = search_one(ix, q768 + i * DIM, q256 + i * PDIM, *p, vw, out_s + i * p->k, out_r + i * p->k,
                                  nullptr);
            for (size_t j = 0; j < n; j++) out_r[i * p->k + j] = live_ext_of(l, out_r[i * p->k + j]);
            counts[i] = (uint32_t)n;
        }
    };
    if (threads <= 1) {
        work();
    } else {
        std::vector<std::thread> th;
        for (int w = 0; w < threads; w++) th.emplace_back(work);
        for (auto& x : th) x.join();
    }
}

}  // extern "C"

extern "C" {
// DeltaList::push of n raw codes (32 B), factors and ids into a fresh list; copies out its
// blocks (nblocks x 1024), factors and ids (nblocks x 32 each). Returns nblocks.
size_t or_delta_list_fixture(size_t n, const uint8_t* codes, const float* factors, const uint32_t* ids,
                             size_t cap_blocks, uint8_t* blocks, float* out_f, uint32_t* out_ids) {
    DeltaList dl;
    for (size_t i = 0; i < n; i++) dl.push(codes + i * 32, factors + 2 * i, ids[i]);
    size_t nb = dl.blocks.size() / 1024;
    if (nb <= cap_blocks) {
        std::memcpy(blocks, dl.blocks.data(), dl.blocks.size());
        std::memcpy(out_f, dl.factors.data(), dl.factors.size() * 4);
        std::memcpy(out_ids, dl.ids.data(), dl.ids.size() * 4);
    }
    return nb;
}
}  // extern "C"

// ---- metadata filters, src/filter.rs --------------------------------------------------
struct OFilter {  // Filter, src/filter.rs:18-39
    std::vector<uint64_t> allow;
    std::vector<uint32_t> rows;         // allowed rows, ascending
    std::vector<uint32_t> per_cluster;  // allowed slots per cluster
};

extern "C" {

constexpr uint64_t BRUTE_MAX = 250000;  // src/filter.rs:55
constexpr double POST_MIN = 0.3;        // src/filter.rs:57
constexpr uint64_t EXPAND_TARGET = 200000;  // src/filter.rs:59

// Filter::new, src/filter.rs:27-38.
void* or_filter_create(const void* index, const uint64_t* allow, size_t words) {
    const Index& ix = *(const Index*)index;
    OFilter* f = new OFilter();
    f->allow.assign(allow, allow + words);
    for (size_t w = 0; w < words; w++)
        for (int i = 0; i < 64; i++)
            if ((allow[w] >> i) & 1) f->rows.push_back((uint32_t)(w * 64 + i));
    f->per_cluster.resize(ix.k);
    for (uint32_t c = 0; c < ix.k; c++) {
        uint32_t cnt = 0;
        for (uint64_t s = ix.block_off[c] * BLOCK; s < ix.block_off[c + 1] * BLOCK; s++) {
            uint32_t id = ix.ids[s];
            if (id != PAD && View::bit(allow, words, id)) cnt++;
        }
        f->per_cluster[c] = cnt;
    }
    return f;
}
void or_filter_free(void* f) { delete (OFilter*)f; }
size_t or_filter_rows(const void* f) { return ((const OFilter*)f)->rows.size(); }
void or_filter_per_cluster(const void* f, uint32_t* out) {
    const auto& p = ((const OFilter*)f)->per_cluster;
    std::copy(p.begin(), p.end(), out);
}

// Plan::resolve, src/filter.rs:60-69: plan 0 post, 1 expand (target), 2 brute, 3 auto.
uint32_t or_plan_resolve(uint32_t plan, uint64_t rows, uint64_t n) {
    if (plan != 3) return plan;
    if (rows <= BRUTE_MAX) return 2;
    if ((double)rows >= POST_MIN * (double)n) return 0;
    return 1;
}

// filter::search, src/filter.rs:72-110, for one query; returns the resolved plan.
uint32_t or_filter_search(const void* index, const void* filter, uint32_t plan, uint64_t target, const float* q768,
                          const float* q256, const sieve_params* p, float* out_s, uint32_t* out_r, uint32_t* count) {
    const Index& ix = *(const Index*)index;
    const OFilter& f = *(const OFilter*)filter;
    View v;
    v.allow = f.allow.data();
    v.allow_words = f.allow.size();
    uint32_t rp = or_plan_resolve(plan, f.rows.size(), ix.n);
    if (plan == 3 && rp == 1) target = EXPAND_TARGET;
    std::vector<std::pair<float, uint32_t>> res;
    if (rp == 0) {
        size_t n = search_one(ix, q768, q256, *p, v, out_s, out_r, nullptr);
        *count = (uint32_t)n;
        return rp;
    } else if (rp == 1) {
        sieve_params all = *p;
        all.nprobe = ix.k;
        all.margin = std::numeric_limits<float>::quiet_NaN();
        all.budget = 0;
        std::vector<float> cent;
        auto order = route(ix, q256, all, cent);
        std::vector<Probe> probe;
        uint64_t covered = 0;
        for (const Probe& x : order) {
            if (f.per_cluster[x.cluster] == 0) continue;
            if (covered >= target) break;  // take_while(covered < target)
            covered += f.per_cluster[x.cluster];
            probe.push_back(x);
        }
        uint64_t sc = 0, rr = 0;
        auto c = candidates(ix, q256, probe, *p, v, &sc);
        res = rerank(ix, q768, q256, c, *p, v, &rr);
    } else {
        RerankQuery r = rerank_query(ix, q768, DIM);
        TopK<float> top(p->k);
        for (uint32_t row : f.rows) top.push((float)dot_i8i8(r.wq.data(), ix.rerank + (size_t)row * DIM, DIM) * r.inv, row);
        res = top.into_sorted();
    }
    for (size_t i = 0; i < res.size(); i++) {
        out_s[i] = res[i].first;
        out_r[i] = res[i].second;
    }
    *count = (uint32_t)res.size();
    return rp;
}

}  // extern "C"
