// Durability of the live state: the write-ahead log (Live::log and Live::replay,
// src/live.rs:209-258), the row segment a compacted index carries (Live::open,
// src/live.rs:166-207) and compaction (Live::compact, src/live.rs:395-468; Live::truncate_wal,
// :470-479).
//
// Compaction runs on the device: one workgroup per cluster counts the live slots of the base
// list and the delta list (PAD and tombstoned slots dropped), a scan turns the counts into the
// new block offsets, a second pass writes every kept slot to its place in base-then-delta order
// (the reference's order), and the codes are packed into the reference's block layout. The host
// writes the files: blocks.bin, factors.f32, ids.u32, block_off.u64, postings.bin and
// postings.off (postings::pack, src/postings.rs:89-114), meta.json, and the row segment.
#include <dirent.h>
#include <errno.h>
#include <fcntl.h>
#include <limits.h>
#include <stdlib.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>

#include <hipcub/hipcub.hpp>

#include "live.hpp"

using namespace sieve;

namespace {

constexpr int CT = 256;  // threads per cluster in the compaction passes

__device__ __forceinline__ bool slot_kept(uint32_t id, const uint64_t* __restrict__ deleted, uint64_t words) {
    if (id == PAD) return false;
    const uint64_t w = id >> 6;
    return !(w < words && ((deleted[w] >> (id & 63)) & 1ull));
}

// The slot at position i of cluster c's list: base slots first, then the delta slots.
struct ClusterRange {
    size_t b0, nbase, d0, n;
    __device__ size_t slot(size_t i) const { return i < nbase ? b0 + i : d0 + (i - nbase); }
};

__device__ __forceinline__ ClusterRange cluster_range(uint32_t c, const uint32_t* __restrict__ block_off,
                                                      const uint32_t* __restrict__ dblock_off, size_t dbase_slot) {
    ClusterRange r;
    r.b0 = (size_t)block_off[c] * BLOCK;
    r.nbase = (size_t)(block_off[c + 1] - block_off[c]) * BLOCK;
    r.d0 = dbase_slot + (size_t)dblock_off[c] * BLOCK;
    r.n = r.nbase + (size_t)(dblock_off[c + 1] - dblock_off[c]) * BLOCK;
    return r;
}

// counts[c] = kept slots of cluster c; dropped[c] = its tombstoned ones.
__global__ __launch_bounds__(CT) void k_compact_count(const uint32_t* __restrict__ block_off,
                                                      const uint32_t* __restrict__ dblock_off, size_t dbase_slot,
                                                      const uint32_t* __restrict__ ids,
                                                      const uint64_t* __restrict__ deleted, uint64_t words,
                                                      uint32_t* __restrict__ counts, uint32_t* __restrict__ dropped) {
    const uint32_t c = blockIdx.x;
    const ClusterRange r = cluster_range(c, block_off, dblock_off, dbase_slot);
    uint32_t kept = 0, gone = 0;
    for (size_t i = threadIdx.x; i < r.n; i += CT) {
        const uint32_t id = ids[r.slot(i)];
        if (id == PAD) continue;
        if (slot_kept(id, deleted, words))
            kept++;
        else
            gone++;
    }
    __shared__ uint32_t sk[CT], sg[CT];
    sk[threadIdx.x] = kept;
    sg[threadIdx.x] = gone;
    __syncthreads();
    for (int o = CT / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            sk[threadIdx.x] += sk[threadIdx.x + o];
            sg[threadIdx.x] += sg[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        counts[c] = sk[0];
        dropped[c] = sg[0];
    }
}

__global__ void k_blocks_of_counts(const uint32_t* __restrict__ counts, uint32_t k, uint32_t* __restrict__ nb) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c <= k) nb[c] = c < k ? (counts[c] + BLOCK - 1) / BLOCK : 0u;
}

// Output slots: zero codes and factors, PAD ids (the padding of each cluster's last block).
__global__ void k_compact_fill(uint32_t* __restrict__ raw, float* __restrict__ fac, uint32_t* __restrict__ ids,
                               size_t slots) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= slots) return;
    for (int w = 0; w < 8; w++) raw[8 * i + w] = 0u;
    fac[2 * i] = 0.0f;
    fac[2 * i + 1] = 0.0f;
    ids[i] = PAD;
}

// Every kept slot of cluster c to new_off[c] * 32 + its rank among the kept slots, in list order:
// raw codes (bit i = dimension i), [mul, err], row id.
__global__ __launch_bounds__(CT) void k_compact_scatter(const uint32_t* __restrict__ block_off,
                                                        const uint32_t* __restrict__ dblock_off, size_t dbase_slot,
                                                        const uint32_t* __restrict__ ids,
                                                        const uint64_t* __restrict__ deleted, uint64_t words,
                                                        const uint4* __restrict__ codes,
                                                        const float* __restrict__ mul, const float* __restrict__ err,
                                                        const uint32_t* __restrict__ new_off,
                                                        uint32_t* __restrict__ oraw, float* __restrict__ ofac,
                                                        uint32_t* __restrict__ oids) {
    const uint32_t c = blockIdx.x;
    const ClusterRange r = cluster_range(c, block_off, dblock_off, dbase_slot);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __shared__ uint32_t wsum[CT / 64];
    size_t at = (size_t)new_off[c] * BLOCK;
    for (size_t i0 = 0; i0 < r.n; i0 += CT) {
        const size_t i = i0 + threadIdx.x;
        size_t s = 0;
        bool keep = false;
        if (i < r.n) {
            s = r.slot(i);
            keep = slot_kept(ids[s], deleted, words);
        }
        const unsigned long long bal = __ballot(keep);
        const uint32_t before = (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[wv] = (uint32_t)__popcll(bal);
        __syncthreads();
        uint32_t off = 0, total = 0;
        for (int w = 0; w < CT / 64; w++) {
            if (w < wv) off += wsum[w];
            total += wsum[w];
        }
        if (keep) {
            const size_t d = at + off + before;
            uint32_t ch[8], raw[8];
            const uint4 a = codes[2 * s], b = codes[2 * s + 1];
            ch[0] = a.x, ch[1] = a.y, ch[2] = a.z, ch[3] = a.w, ch[4] = b.x, ch[5] = b.y, ch[6] = b.z, ch[7] = b.w;
            code_from_chunked(ch, raw);
            for (int w = 0; w < 8; w++) oraw[8 * d + w] = raw[w];
            ofac[2 * d] = mul[s];
            ofac[2 * d + 1] = err[s];
            oids[d] = ids[s];
        }
        at += total;
        __syncthreads();
    }
}

// lut::pack_block (src/lut.rs:72-82) of every 32-slot block: byte kk of group g holds the
// nibble g of slot kk in its low half and of slot kk + 16 in its high half.
__global__ void k_pack_blocks(const uint8_t* __restrict__ raw, size_t nblocks, uint8_t* __restrict__ out) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nblocks * BLOCK_BYTES) return;
    const size_t b = t / BLOCK_BYTES;
    const int g = (int)(t % BLOCK_BYTES) / 16, kk = (int)(t % 16);
    auto nib = [&](int slot) -> uint32_t {
        const uint8_t byte = raw[(b * BLOCK + slot) * 32 + (g >> 1)];
        return (byte >> ((g & 1) * 4)) & 15u;
    };
    out[t] = (uint8_t)(nib(kk) | nib(kk + 16) << 4);
}

inline unsigned grid(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

// ---- files ------------------------------------------------------------------------------

bool exists(const std::string& p) {
    struct stat st;
    return ::stat(p.c_str(), &st) == 0;
}

std::string real_path(const std::string& p) {
    char buf[PATH_MAX];
    if (!::realpath(p.c_str(), buf)) return std::string();
    return std::string(buf);
}

void mkdirs(const std::string& p) {
    std::string cur;
    size_t i = 0;
    while (i <= p.size()) {
        const size_t j = p.find('/', i);
        cur = p.substr(0, j == std::string::npos ? p.size() : j);
        if (!cur.empty() && !exists(cur))
            SIEVE_REQUIRE(::mkdir(cur.c_str(), 0755) == 0 || errno == EEXIST, SIEVE_EIO, "cannot create " + cur);
        if (j == std::string::npos) break;
        i = j + 1;
    }
}

std::string read_text(const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "rb");
    SIEVE_REQUIRE(f, SIEVE_EIO, "cannot open " + path);
    std::string s;
    char buf[65536];
    size_t r;
    while ((r = std::fread(buf, 1, sizeof buf, f)) > 0) s.append(buf, r);
    std::fclose(f);
    return s;
}

// Reads a whole file of T (store::read_vec, src/store.rs:37-44).
template <class T>
std::vector<T> read_vec(const std::string& path) {
    Mapped m;
    map_file(path, m, true);
    SIEVE_REQUIRE(m.len % sizeof(T) == 0, SIEVE_EIO, path + " is not a whole number of items");
    std::vector<T> v(m.len / sizeof(T));
    if (m.len) std::memcpy(v.data(), m.p, m.len);
    return v;
}

struct File {
    FILE* f = nullptr;
    std::string path;
    explicit File(const std::string& p) : path(p) {
        f = std::fopen(p.c_str(), "wb");
        SIEVE_REQUIRE(f, SIEVE_EIO, "cannot create " + p);
        std::setvbuf(f, nullptr, _IOFBF, 8 << 20);
    }
    void write(const void* p, size_t n) {
        if (n) SIEVE_REQUIRE(std::fwrite(p, 1, n, f) == n, SIEVE_EIO, "cannot write " + path);
    }
    void finish() {
        const int rc = std::fclose(f);
        f = nullptr;
        SIEVE_REQUIRE(rc == 0, SIEVE_EIO, "cannot write " + path);
    }
    ~File() {
        if (f) std::fclose(f);
    }
};

void write_file(const std::string& path, const void* p, size_t n) {
    File f(path);
    f.write(p, n);
    f.finish();
}

// Device array -> file.
template <class T>
void write_device(const std::string& path, const T* d, size_t n) {
    std::vector<T> h(n);
    if (n) SIEVE_HIP(hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost));
    write_file(path, h.data(), n * sizeof(T));
}

// ---- meta.json --------------------------------------------------------------------------
// The top level of a JSON object as (key, raw value text) pairs: enough to read soar_lambda and
// to add the compaction's keys while keeping every other member as it was.

size_t skip_ws(const std::string& s, size_t i) {
    while (i < s.size() && (s[i] == ' ' || s[i] == '\n' || s[i] == '\r' || s[i] == '\t')) i++;
    return i;
}

size_t skip_string(const std::string& s, size_t i) {  // s[i] == '"'
    for (i++; i < s.size(); i++) {
        if (s[i] == '\\')
            i++;
        else if (s[i] == '"')
            return i + 1;
    }
    throw Error{SIEVE_EIO, "meta.json: unterminated string"};
}

size_t skip_value(const std::string& s, size_t i) {
    if (i >= s.size()) throw Error{SIEVE_EIO, "meta.json: value expected"};
    if (s[i] == '"') return skip_string(s, i);
    if (s[i] == '{' || s[i] == '[') {
        int depth = 0;
        while (i < s.size()) {
            if (s[i] == '"') {
                i = skip_string(s, i);
                continue;
            }
            if (s[i] == '{' || s[i] == '[') depth++;
            if (s[i] == '}' || s[i] == ']') {
                if (--depth == 0) return i + 1;
            }
            i++;
        }
        throw Error{SIEVE_EIO, "meta.json: unbalanced brackets"};
    }
    while (i < s.size() && s[i] != ',' && s[i] != '}' && s[i] != ']' && s[i] != ' ' && s[i] != '\n' && s[i] != '\r' &&
           s[i] != '\t')
        i++;
    return i;
}

std::vector<std::pair<std::string, std::string>> json_members(const std::string& s) {
    std::vector<std::pair<std::string, std::string>> out;
    size_t i = skip_ws(s, 0);
    SIEVE_REQUIRE(i < s.size() && s[i] == '{', SIEVE_EIO, "meta.json: not an object");
    i = skip_ws(s, i + 1);
    if (i < s.size() && s[i] == '}') return out;
    for (;;) {
        SIEVE_REQUIRE(i < s.size() && s[i] == '"', SIEVE_EIO, "meta.json: key expected");
        const size_t k1 = skip_string(s, i);
        std::string key = s.substr(i + 1, k1 - i - 2);
        i = skip_ws(s, k1);
        SIEVE_REQUIRE(i < s.size() && s[i] == ':', SIEVE_EIO, "meta.json: ':' expected");
        i = skip_ws(s, i + 1);
        const size_t v1 = skip_value(s, i);
        out.emplace_back(key, s.substr(i, v1 - i));
        i = skip_ws(s, v1);
        SIEVE_REQUIRE(i < s.size(), SIEVE_EIO, "meta.json: truncated");
        if (s[i] == '}') break;
        SIEVE_REQUIRE(s[i] == ',', SIEVE_EIO, "meta.json: ',' expected");
        i = skip_ws(s, i + 1);
    }
    return out;
}

// meta["soar_lambda"].as_f64() (src/live.rs:170-171): NaN for null, a missing key or a
// non-number.
float meta_soar_lambda(const std::string& dir) {
    const std::string p = dir + "/meta.json";
    SIEVE_REQUIRE(exists(p), SIEVE_EIO, "index meta.json: cannot open " + p);
    for (const auto& kv : json_members(read_text(p)))
        if (kv.first == "soar_lambda") {
            const char* b = kv.second.c_str();
            char* e = nullptr;
            const double v = std::strtod(b, &e);
            if (e != b && *e == 0) return (float)v;
            return std::numeric_limits<float>::quiet_NaN();
        }
    return std::numeric_limits<float>::quiet_NaN();
}

std::string json_string(const std::string& s) {
    std::string o = "\"";
    for (char ch : s) {
        if (ch == '"' || ch == '\\') {
            o += '\\';
            o += ch;
        } else if ((unsigned char)ch < 0x20) {
            char buf[8];
            std::snprintf(buf, sizeof buf, "\\u%04x", (unsigned)(unsigned char)ch);
            o += buf;
        } else {
            o += ch;
        }
    }
    return o + "\"";
}

std::string json_float(float v) {
    if (v != v) return "null";
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.9g", (double)v);
    std::string s(buf);
    if (s.find_first_of(".eE") == std::string::npos) s += ".0";
    return s;
}

// Members sorted by key, pretty-printed with two-space indentation (serde_json's
// to_string_pretty of a map).
std::string json_object(std::vector<std::pair<std::string, std::string>> m) {
    std::stable_sort(m.begin(), m.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    std::string o = "{";
    for (size_t i = 0; i < m.size(); i++) {
        o += i ? ",\n  " : "\n  ";
        o += json_string(m[i].first) + ": " + m[i].second;
    }
    return o + (m.empty() ? "}" : "\n}");
}

void set_member(std::vector<std::pair<std::string, std::string>>& m, const std::string& key, const std::string& v) {
    for (auto& kv : m)
        if (kv.first == key) {
            kv.second = v;
            return;
        }
    m.emplace_back(key, v);
}

// ---- the log ------------------------------------------------------------------------------

// Live::replay (src/live.rs:209-249): applies the log's records in order, consecutive inserts as
// one batch, and cuts a torn final record off the file.
void replay(sieve_live* l, const std::string& path, sieve_replay_stats* st) {
    const auto t0 = std::chrono::steady_clock::now();
    Mapped m;
    map_file(path, m, false);  // a missing log is an empty one
    const uint8_t* buf = static_cast<const uint8_t*>(m.p);
    const size_t len = m.len;
    sieve_replay_stats stats{};
    std::vector<uint32_t> ext;
    std::vector<int8_t> rows;
    std::vector<float> norms;
    const hipStream_t s = l->work.st;
    auto flush = [&] {
        if (ext.empty()) return;
        const size_t n = ext.size();
        SIEVE_REQUIRE(l->nrows + n <= l->cap_rows, SIEVE_ELIMIT,
                      "replay: the log holds more delta rows than max_rows (" + std::to_string(l->cap_rows) + ")");
        SIEVE_HIP(hipMemcpyAsync(l->rows + l->nrows * DIM, rows.data(), n * DIM, hipMemcpyHostToDevice, s));
        SIEVE_HIP(hipMemcpyAsync(l->norm256 + l->nrows, norms.data(), n * 4, hipMemcpyHostToDevice, s));
        SIEVE_HIP(hipStreamSynchronize(s));
        live_apply_rows(l, ext.data(), n);
        ext.clear();
        rows.clear();
        norms.clear();
    };
    size_t pos = 0;
    while (pos < len) {
        size_t need;
        if (buf[pos] == WAL_OP_INSERT)
            need = WAL_INSERT_BYTES;
        else if (buf[pos] == WAL_OP_DELETE)
            need = WAL_DELETE_BYTES;
        else
            throw Error{SIEVE_EIO, path + ": bad op " + std::to_string(buf[pos]) + " at byte " + std::to_string(pos)};
        if (pos + need > len) break;
        const uint8_t* r = buf + pos;
        if (r[0] == WAL_OP_INSERT) {
            ext.push_back(get_le32(r + 1));
            const uint32_t nb = get_le32(r + 5);
            float nf;
            std::memcpy(&nf, &nb, 4);
            norms.push_back(nf);
            rows.insert(rows.end(), reinterpret_cast<const int8_t*>(r + 9), reinterpret_cast<const int8_t*>(r + 9) + DIM);
            stats.inserts++;
            // Batches are applied in order either way; a bounded one keeps the staging small.
            if (ext.size() == 65536) flush();
        } else {
            flush();
            const uint32_t e = get_le32(r + 1);
            live_apply_delete(l, &e, 1);
            stats.deletes++;
        }
        pos += need;
    }
    flush();
    stats.torn = len - pos;
    if (stats.torn) SIEVE_REQUIRE(::truncate(path.c_str(), (off_t)pos) == 0, SIEVE_EIO, "cannot truncate " + path);
    stats.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (st) *st = stats;
}

// The row segment of a compacted index (src/live.rs:183-195): rows, prefix norms, tier-2 codes,
// external ids and tombstones; a delta row whose tombstone is clear is its external id's row.
void load_segment(sieve_live* l, const std::string& dir) {
    sieve_index* ix = l->ix;
    const std::string I = dir + "/";
    const auto rows = read_vec<int8_t>(I + "live.rows.i8");
    const auto norm = read_vec<float>(I + "live.norm256.f32");
    const auto t2 = read_vec<uint8_t>(I + "live.t2.bin");
    const auto t2n = read_vec<float>(I + "live.t2.norm");
    const auto ext = read_vec<uint32_t>(I + "live.ext.u32");
    const auto del = read_vec<uint64_t>(I + "live.deleted.u64");
    const size_t r = ext.size();
    SIEVE_REQUIRE(r == ix->seg_rows, SIEVE_EINVAL,
                  "live.ext.u32 holds " + std::to_string(r) + " rows; the index's lists reference " +
                      std::to_string(ix->seg_rows));
    SIEVE_REQUIRE(r <= l->cap_rows, SIEVE_ELIMIT,
                  "the row segment holds " + std::to_string(r) + " rows; max_rows is " + std::to_string(l->cap_rows));
    SIEVE_REQUIRE(rows.size() == r * DIM && norm.size() == r, SIEVE_EIO, "live.rows.i8 / live.norm256.f32 size");
    if (ix->t2_codes)
        SIEVE_REQUIRE(t2.size() == r * 64 && t2n.size() == r, SIEVE_EIO, "live.t2.bin / live.t2.norm size");
    SIEVE_REQUIRE(del.size() <= l->deleted_words, SIEVE_EIO, "live.deleted.u64 covers more rows than the index");
    const hipStream_t s = l->work.st;
    if (r) {
        SIEVE_HIP(hipMemcpyAsync(l->rows, rows.data(), r * DIM, hipMemcpyHostToDevice, s));
        SIEVE_HIP(hipMemcpyAsync(l->norm256, norm.data(), r * 4, hipMemcpyHostToDevice, s));
        SIEVE_HIP(hipMemcpyAsync(l->ext, ext.data(), r * 4, hipMemcpyHostToDevice, s));
        if (ix->t2_codes) {
            SIEVE_HIP(hipMemcpyAsync(l->t2, t2.data(), r * 64, hipMemcpyHostToDevice, s));
            SIEVE_HIP(hipMemcpyAsync(l->t2n, t2n.data(), r * 4, hipMemcpyHostToDevice, s));
        }
    }
    std::copy(del.begin(), del.end(), l->deleted_host.begin());
    l->ndeleted = 0;
    for (uint64_t w : del) l->ndeleted += (uint64_t)__builtin_popcountll(w);
    if (!del.empty()) SIEVE_HIP(hipMemcpyAsync(l->deleted, del.data(), del.size() * 8, hipMemcpyHostToDevice, s));
    SIEVE_HIP(hipStreamSynchronize(s));
    for (size_t i = 0; i < r; i++) {
        const uint32_t id = (uint32_t)(l->n + i);
        if (!l->bit_host(id)) l->by_ext[ext[i]] = id;
    }
    l->nrows = r;
    l->seg_rows = r;
}

// ---- compaction -----------------------------------------------------------------------------

void compact(const sieve_live* l, const std::string& out, sieve_compact_stats* stp) {
    const auto t0 = std::chrono::steady_clock::now();
    const sieve_index* ix = l->ix;
    const std::string src = l->index_dir.empty() ? std::string() : real_path(l->index_dir);
    SIEVE_REQUIRE(l->index_dir.empty() || !src.empty(), SIEVE_EIO, "cannot resolve " + l->index_dir);
    SIEVE_REQUIRE(src.empty() || real_path(out) != src, SIEVE_EINVAL, "compaction must write a new directory");
    mkdirs(out);
    const uint32_t k = ix->k;
    hipStream_t s = nullptr;
    SIEVE_HIP(hipStreamCreate(&s));
    struct Cleanup {
        hipStream_t s;
        std::vector<void*> ps;
        ~Cleanup() {
            for (void* p : ps) (void)hipFree(p);
            (void)hipStreamDestroy(s);
        }
    } cl{s, {}};
    auto dalloc = [&](size_t bytes) {
        void* p = nullptr;
        SIEVE_HIP(hipMalloc(&p, std::max<size_t>(bytes, 16)));
        cl.ps.push_back(p);
        return p;
    };
    uint32_t* counts = static_cast<uint32_t*>(dalloc((k + 1) * 4));
    uint32_t* dropped = static_cast<uint32_t*>(dalloc((k + 1) * 4));
    uint32_t* nb = static_cast<uint32_t*>(dalloc((k + 1) * 4));
    uint32_t* new_off = static_cast<uint32_t*>(dalloc((k + 1) * 4));
    const size_t dbase_slot = ix->nslots;
    SIEVE_HIP(hipMemsetAsync(counts, 0, (k + 1) * 4, s));
    SIEVE_HIP(hipMemsetAsync(dropped, 0, (k + 1) * 4, s));
    hipLaunchKernelGGL(k_compact_count, dim3(k), dim3(CT), 0, s, ix->block_off, l->dblock_off, dbase_slot, ix->ids,
                       l->deleted, l->deleted_words, counts, dropped);
    hipLaunchKernelGGL(k_blocks_of_counts, dim3(grid(k + 1, 256)), dim3(256), 0, s, counts, k, nb);
    size_t tmp_bytes = 0;
    SIEVE_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, nb, new_off, (int)(k + 1), s));
    void* tmp = dalloc(tmp_bytes);
    SIEVE_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, nb, new_off, (int)(k + 1), s));
    SIEVE_HIP(hipGetLastError());
    std::vector<uint32_t> h_off(k + 1), h_cnt(k), h_drop(k);
    SIEVE_HIP(hipMemcpyAsync(h_off.data(), new_off, (k + 1) * 4, hipMemcpyDeviceToHost, s));
    SIEVE_HIP(hipMemcpyAsync(h_cnt.data(), counts, k * 4, hipMemcpyDeviceToHost, s));
    SIEVE_HIP(hipMemcpyAsync(h_drop.data(), dropped, k * 4, hipMemcpyDeviceToHost, s));
    SIEVE_HIP(hipStreamSynchronize(s));
    const size_t nblocks = h_off[k], slots_out = nblocks * BLOCK;
    uint32_t* oraw = static_cast<uint32_t*>(dalloc(slots_out * 32));
    float* ofac = static_cast<float*>(dalloc(slots_out * 8));
    uint32_t* oids = static_cast<uint32_t*>(dalloc(slots_out * 4));
    uint8_t* oblk = static_cast<uint8_t*>(dalloc(nblocks * BLOCK_BYTES));
    if (slots_out)
        hipLaunchKernelGGL(k_compact_fill, dim3(grid(slots_out, 256)), dim3(256), 0, s, oraw, ofac, oids, slots_out);
    hipLaunchKernelGGL(k_compact_scatter, dim3(k), dim3(CT), 0, s, ix->block_off, l->dblock_off, dbase_slot, ix->ids,
                       l->deleted, l->deleted_words, ix->codes, ix->mul, ix->err, new_off, oraw, ofac, oids);
    if (nblocks)
        hipLaunchKernelGGL(k_pack_blocks, dim3(grid(nblocks * BLOCK_BYTES, 256)), dim3(256), 0, s,
                           reinterpret_cast<const uint8_t*>(oraw), nblocks, oblk);
    SIEVE_HIP(hipGetLastError());
    std::vector<uint8_t> blocks(nblocks * BLOCK_BYTES);
    std::vector<float> fac(slots_out * 2);
    std::vector<uint32_t> ids(slots_out);
    SIEVE_HIP(hipMemcpyAsync(blocks.data(), oblk, blocks.size(), hipMemcpyDeviceToHost, s));
    SIEVE_HIP(hipMemcpyAsync(fac.data(), ofac, fac.size() * 4, hipMemcpyDeviceToHost, s));
    SIEVE_HIP(hipMemcpyAsync(ids.data(), oids, ids.size() * 4, hipMemcpyDeviceToHost, s));
    SIEVE_HIP(hipStreamSynchronize(s));

    const std::string O = out + "/";
    write_file(O + "blocks.bin", blocks.data(), blocks.size());
    write_file(O + "factors.f32", fac.data(), fac.size() * 4);
    write_file(O + "ids.u32", ids.data(), ids.size() * 4);
    std::vector<uint64_t> off64(h_off.begin(), h_off.end());
    write_file(O + "block_off.u64", off64.data(), off64.size() * 8);
    // Centroids, rotation and the corpus-indexed tier files: shared with the source directory by
    // symlink, as the reference does; an index created from arrays has no directory, so they are
    // written from the device.
    const char* shared[] = {"centroids.f32", "rotation.f32", "tier2.sigma", "tier2.bin", "tier2.norm"};
    if (!src.empty()) {
        for (const char* f : shared) {
            const std::string from = src + "/" + f, to = O + f;
            if (!exists(from)) continue;
            (void)::unlink(to.c_str());
            const std::string target = real_path(from);
            SIEVE_REQUIRE(!target.empty() && ::symlink(target.c_str(), to.c_str()) == 0, SIEVE_EIO,
                          "cannot link " + to);
        }
    } else {
        write_device(O + "centroids.f32", ix->cents, (size_t)k * PDIM);
        write_device(O + "rotation.f32", ix->rotation, 3 * PDIM);
        if (ix->t2_codes) {
            write_device(O + "tier2.sigma", ix->t2_sigma, PDIM);
            write_device(O + "tier2.bin", ix->t2_codes, (size_t)ix->n * 64);
            write_device(O + "tier2.norm", ix->t2_norm, (size_t)ix->n);
        }
    }
    std::vector<std::pair<std::string, std::string>> meta;
    if (!src.empty() && exists(src + "/meta.json")) {
        meta = json_members(read_text(src + "/meta.json"));
    } else {
        set_member(meta, "k", std::to_string(k));
        set_member(meta, "rabitq", "true");
        set_member(meta, "rows", std::to_string(ix->n));
        set_member(meta, "soar_lambda", l->soar ? json_float(l->lambda) : "null");
    }
    set_member(meta, "compacted_from", src.empty() ? "null" : json_string(src));
    set_member(meta, "live_rows", std::to_string(l->nrows));
    set_member(meta, "deleted", std::to_string(l->ndeleted));
    const std::string js = json_object(meta);
    write_file(O + "meta.json", js.data(), js.size());
    // The row segment (src/live.rs:460-465). The tombstones as the reference keeps them: words up
    // to the last one with a bit set.
    const size_t r = l->nrows;
    write_device(O + "live.rows.i8", l->rows, r * DIM);
    write_device(O + "live.norm256.f32", l->norm256, r);
    write_device(O + "live.t2.bin", l->t2, ix->t2_codes ? r * 64 : 0);
    write_device(O + "live.t2.norm", l->t2n, ix->t2_codes ? r : 0);
    write_device(O + "live.ext.u32", l->ext, r);
    size_t words = l->deleted_host.size();
    while (words && l->deleted_host[words - 1] == 0) words--;
    write_file(O + "live.deleted.u64", l->deleted_host.data(), words * 8);
    // postings::pack (src/postings.rs:89-114): per cluster its blocks, factors and ids, the record
    // padded to 4 KiB.
    {
        constexpr uint64_t ALIGN = 4096, RECORD = BLOCK_BYTES + BLOCK * 8 + BLOCK * 4;
        File f(O + "postings.bin");
        std::vector<uint64_t> poff(k + 1);
        const std::vector<uint8_t> zeros(ALIGN, 0);
        uint64_t pos = 0;
        for (uint32_t c = 0; c < k; c++) {
            const size_t b0 = h_off[c], b1 = h_off[c + 1];
            poff[c] = pos;
            f.write(blocks.data() + b0 * BLOCK_BYTES, (b1 - b0) * BLOCK_BYTES);
            f.write(fac.data() + b0 * BLOCK * 2, (b1 - b0) * BLOCK * 8);
            f.write(ids.data() + b0 * BLOCK, (b1 - b0) * BLOCK * 4);
            pos += (b1 - b0) * RECORD;
            const uint64_t pad = (pos + ALIGN - 1) / ALIGN * ALIGN - pos;
            f.write(zeros.data(), pad);
            pos += pad;
        }
        poff[k] = pos;
        f.finish();
        write_file(O + "postings.off", poff.data(), poff.size() * 8);
    }
    sieve_compact_stats st{};
    for (uint32_t c = 0; c < k; c++) {
        st.slots += h_cnt[c];
        st.dropped += h_drop[c];
    }
    st.blocks = nblocks;
    st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (stp) *stp = st;
}

}  // namespace

namespace sieve {

void live_log(sieve_live* l, const std::vector<uint8_t>& rec) {
    if (l->wal_fd < 0 || rec.empty()) return;
    size_t done = 0;
    while (done < rec.size()) {
        const ssize_t w = ::write(l->wal_fd, rec.data() + done, rec.size() - done);
        if (w < 0 && errno == EINTR) continue;
        SIEVE_REQUIRE(w > 0, SIEVE_EIO, "cannot write " + l->wal_path);
        done += (size_t)w;
    }
    SIEVE_REQUIRE(::fdatasync(l->wal_fd) == 0, SIEVE_EIO, "cannot sync " + l->wal_path);
}

}  // namespace sieve

extern "C" {

int sieve_live_open_dir(sieve_index* ix, const sieve_live_opts* o, const char* index_dir, const char* wal_path,
                        sieve_replay_stats* replay_out, sieve_live** out) {
    return guarded([&] {
        SIEVE_REQUIRE(ix && out, SIEVE_EINVAL, "null argument");
        sieve_live_opts opts = o ? *o : sieve_live_opts{};
        const std::string dir = index_dir ? index_dir : "";
        SIEVE_REQUIRE(!dir.empty() || ix->seg_rows == 0, SIEVE_EINVAL,
                      "the index's lists hold rows of its live segment: give the directory that holds live.*");
        if (!dir.empty()) opts.soar_lambda = meta_soar_lambda(dir);
        SIEVE_HIP(hipSetDevice(ix->device));
        std::unique_ptr<sieve_live> l(live_create(ix, opts));
        try {
            l->index_dir = dir;
            if (!dir.empty() && exists(dir + "/live.ext.u32")) load_segment(l.get(), dir);
            SIEVE_REQUIRE(l->seg_rows == ix->seg_rows, SIEVE_EINVAL,
                          "the index's lists reference a live segment that " + dir + " does not hold");
            if (wal_path) {
                l->wal_path = wal_path;
                replay(l.get(), l->wal_path, replay_out);
                l->wal_fd = ::open(wal_path, O_WRONLY | O_CREAT | O_APPEND | O_CLOEXEC, 0644);
                SIEVE_REQUIRE(l->wal_fd >= 0, SIEVE_EIO, "cannot open " + l->wal_path);
            } else if (replay_out) {
                *replay_out = sieve_replay_stats{};
            }
        } catch (...) {
            ix->live = nullptr;
            throw;
        }
        *out = l.release();
    });
}

int sieve_live_compact(const sieve_live* l, const char* out_dir, sieve_compact_stats* stats) {
    return guarded([&] {
        SIEVE_REQUIRE(l && out_dir && *out_dir, SIEVE_EINVAL, "null argument");
        SIEVE_HIP(hipSetDevice(l->ix->device));
        live_settle(l);
        compact(l, out_dir, stats);
    });
}

int sieve_live_truncate_wal(sieve_live* l) {
    return guarded([&] {
        SIEVE_REQUIRE(l, SIEVE_EINVAL, "null argument");
        if (l->wal_fd < 0) return;
        SIEVE_REQUIRE(::ftruncate(l->wal_fd, 0) == 0, SIEVE_EIO, "cannot truncate " + l->wal_path);
        SIEVE_REQUIRE(::fsync(l->wal_fd) == 0, SIEVE_EIO, "cannot sync " + l->wal_path);
    });
}

}  // extern "C"
