// The 1B L2 index in the reference's on-disk form, and its hot-leaf policy.
//
// sieve_l2_index_open reads what bigann::Index::open reads (src/bigann.rs:634-641):
//   Model::open (src/bigann.rs:257-266): centroids.f32 (the k leaves x 100 f32, cell major),
//     coarse.f32 and k2.u64 (the two-level model the builder assigns rows with; required as
//     there, not used by search), rotation.f32 (3 x 128 signs), bits.u64 (refinement width,
//     absent = 1, src/bigann.rs:229-234);
//   Postings::open_ext (src/postings.rs:209-229): postings.off (k + 1 record offsets, equal in
//     length to block_off.u64), block_off.u64, postings.bin: record c at postings.off[c], 4 KiB
//     aligned, holds for its nb blocks the 512-byte code blocks, nb x 32 x [mul, |x|^2] f32, nb x
//     32 u32 row ids and nb x 32 x refine_bytes(bits) extra bytes (src/postings.rs:26-35, 53-75);
//   Base::open (src/bigann.rs:476-489): base.00, base.01, ... in base_dir, each with a
//     base.NN.done marker, 100,000,000 rows of 100 int8 each (the last may be shorter).
// The sections are gathered record by record into the library's arrays through pinned
// staging, the base rows are streamed into HBM, and the index is then created as from arrays.
//
// sieve_l2_hot_leaves is hot_leaves (src/bigann.rs:537-563): the leaves worth keeping resident
// under a byte budget, by how often a sample of queries routes to each per byte of its record.
#include <algorithm>
#include <vector>

#include "host.hpp"

using namespace sieve;

namespace {

constexpr int L2D = 100;
constexpr uint64_t CHUNK_ROWS = 100000000ull;  // src/bigann.rs:57-58
constexpr uint64_t RECORD_ALIGN = 4096;         // src/postings.rs:38

// Bytes of one block's slice of a record with refinement width bits (block_record_ext,
// src/postings.rs:31-34, with 512-byte code blocks).
uint64_t record_block_bytes(int bits) { return 512 + 32 * 8 + 32 * 4 + 32 * (uint64_t)SIEVE_REFINE_BYTES(bits); }

template <class T>
std::vector<T> read_vec(const std::string& path, bool required, bool* found = nullptr) {
    Mapped m;
    const bool ok = map_file(path, m, required);
    if (found) *found = ok;
    std::vector<T> v;
    if (!ok) return v;
    SIEVE_REQUIRE(m.len % sizeof(T) == 0, SIEVE_EINVAL, path + " is not a whole number of items");
    v.resize(m.len / sizeof(T));
    if (m.len) std::memcpy(v.data(), m.p, m.len);
    return v;
}

bool file_exists(const std::string& p) {
    struct stat st;
    return ::stat(p.c_str(), &st) == 0;
}

// Host-to-device copies through one pinned buffer.
struct Stager {
    uint8_t* buf = nullptr;
    size_t cap = 0;
    explicit Stager(size_t bytes) : cap(bytes) {
        SIEVE_HIP(hipHostMalloc((void**)&buf, cap, hipHostMallocDefault));
    }
    ~Stager() {
        if (buf) (void)hipHostFree(buf);
    }
    void copy(void* dst, const void* src, size_t bytes) {
        for (size_t o = 0; o < bytes; o += cap) {
            const size_t m = std::min(cap, bytes - o);
            std::memcpy(buf, static_cast<const uint8_t*>(src) + o, m);
            SIEVE_HIP(hipMemcpy(static_cast<uint8_t*>(dst) + o, buf, m, hipMemcpyHostToDevice));
        }
    }
};

// A device array the reader fills and frees unless handed over.
struct DevArray {
    void* p = nullptr;
    ~DevArray() {
        if (p) (void)hipFree(p);
    }
    template <class T>
    T* alloc(size_t n) {
        p = dmalloc<T>(n, nullptr);
        return static_cast<T*>(p);
    }
};

}  // namespace

extern "C" {

int sieve_l2_index_open(const char* index_dir, const char* base_dir, sieve_l2_index** out) {
    return guarded([&] {
        SIEVE_REQUIRE(index_dir && base_dir && out, SIEVE_EINVAL, "null argument");
        const std::string I(index_dir), B(base_dir);
        // Model::open
        std::vector<float> leaves = read_vec<float>(I + "/centroids.f32", true);
        std::vector<float> coarse = read_vec<float>(I + "/coarse.f32", true);
        std::vector<uint64_t> k2 = read_vec<uint64_t>(I + "/k2.u64", true);
        std::vector<float> rot = read_vec<float>(I + "/rotation.f32", true);
        SIEVE_REQUIRE(leaves.size() % L2D == 0 && !leaves.empty(), SIEVE_EINVAL, "centroids.f32 is not k x 100 f32");
        SIEVE_REQUIRE(coarse.size() % L2D == 0, SIEVE_EINVAL, "coarse.f32 is not k1 x 100 f32");
        SIEVE_REQUIRE(k2.size() == 1 && k2[0] >= 1, SIEVE_EINVAL, "k2.u64 must hold one positive count");
        SIEVE_REQUIRE(rot.size() == 3 * 128, SIEVE_EINVAL, "rotation.f32 is not 128-dim");
        bool has_bits = false;
        std::vector<uint64_t> bitsv = read_vec<uint64_t>(I + "/bits.u64", false, &has_bits);
        const uint64_t bits = has_bits ? (bitsv.empty() ? 0 : bitsv[0]) : 1;
        SIEVE_REQUIRE(bits >= 1 && bits <= SIEVE_REFINE_MAX_BITS, SIEVE_EINVAL,
                      "bits.u64 = " + std::to_string(bits) + ", expected 1 ..= 3");
        const uint32_t k = (uint32_t)(leaves.size() / L2D);
        // Postings::open_ext
        std::vector<uint64_t> off = read_vec<uint64_t>(I + "/postings.off", true);
        std::vector<uint64_t> bo = read_vec<uint64_t>(I + "/block_off.u64", true);
        SIEVE_REQUIRE(off.size() == bo.size(), SIEVE_EINVAL, "postings.off does not match block_off.u64");
        SIEVE_REQUIRE(bo.size() == (size_t)k + 1 && bo[0] == 0, SIEVE_EINVAL,
                      "block_off.u64 does not have k + 1 entries from 0");
        Mapped post;
        map_file(I + "/postings.bin", post, true);
        const int ib = (int)bits;
        const uint64_t rb = SIEVE_REFINE_BYTES(ib), rec = record_block_bytes(ib);
        const uint64_t nblocks = bo[k], nslots = nblocks * BLOCK;
        SIEVE_REQUIRE(nslots < (1ull << 32) - 1, SIEVE_ELIMIT, "more than 2^32 - 1 slots");
        for (uint32_t c = 0; c < k; c++) {
            SIEVE_REQUIRE(bo[c + 1] >= bo[c], SIEVE_EINVAL, "block_off.u64 is not monotone");
            SIEVE_REQUIRE(off[c] % RECORD_ALIGN == 0 && off[c] + (bo[c + 1] - bo[c]) * rec <= post.len, SIEVE_EINVAL,
                          "postings record " + std::to_string(c) + " lies outside postings.bin or is not aligned");
        }
        // Base::open: base.NN for every NN with a .done marker, in order.
        std::vector<std::unique_ptr<Mapped>> chunks;
        uint64_t n = 0;
        for (int c = 0; c < 100; c++) {
            char name[32];
            std::snprintf(name, sizeof name, "/base.%02d", c);
            if (!file_exists(B + name + std::string(".done"))) break;
            chunks.emplace_back(new Mapped());
            map_file(B + name, *chunks.back(), true);
            SIEVE_REQUIRE(chunks.back()->len % L2D == 0 && chunks.back()->len / L2D <= CHUNK_ROWS, SIEVE_EINVAL,
                          B + name + " is not whole 100-byte rows of one chunk");
            n += chunks.back()->len / L2D;
        }
        SIEVE_REQUIRE(!chunks.empty(), SIEVE_EIO, "no base.NN chunks in " + B);
        // Row id r lives in chunk r / CHUNK_ROWS (src/bigann.rs:58, :509): every chunk but the last
        // is full.
        for (size_t c = 0; c + 1 < chunks.size(); c++)
            SIEVE_REQUIRE(chunks[c]->len / L2D == CHUNK_ROWS, SIEVE_EINVAL,
                          "base chunk " + std::to_string(c) + " is short but not the last");
        DevArray dblocks, dfactors, dids, dextra, dbase;
        uint8_t* blocks = dblocks.alloc<uint8_t>(nblocks * 512);
        float* factors = dfactors.alloc<float>(nslots * 2);
        uint32_t* ids = dids.alloc<uint32_t>(nslots);
        uint8_t* extra = rb ? dextra.alloc<uint8_t>(nslots * rb) : nullptr;
        int8_t* base = dbase.alloc<int8_t>(n * L2D);
        {
            // The records' sections, leaf by leaf, into the four arrays.
            Stager sg(size_t(64) << 20);
            const uint8_t* P = static_cast<const uint8_t*>(post.p);
            for (uint32_t c = 0; c < k; c++) {
                const uint64_t nb = bo[c + 1] - bo[c];
                if (!nb) continue;
                const uint8_t* r = P + off[c];
                sg.copy(blocks + bo[c] * 512, r, nb * 512);
                sg.copy(factors + bo[c] * BLOCK * 2, r + nb * 512, nb * BLOCK * 8);
                sg.copy(ids + bo[c] * BLOCK, r + nb * (512 + BLOCK * 8), nb * BLOCK * 4);
                if (rb) sg.copy(extra + bo[c] * BLOCK * rb, r + nb * (512 + BLOCK * 12), nb * BLOCK * rb);
            }
            uint64_t row = 0;
            for (auto& ch : chunks) {
                sg.copy(base + row * L2D, ch->p, ch->len);
                row += ch->len / L2D;
            }
        }
        sieve_l2_desc d{};
        d.k = k;
        d.n = n;
        d.nblocks = nblocks;
        d.leaves = leaves.data();
        d.block_off = bo.data();
        d.blocks = blocks;
        d.ids = ids;
        d.factors = factors;
        d.rotation = rot.data();
        d.base = base;
        d.borrow_base = 1;
        d.refine_bits = (uint32_t)bits;
        d.refine = extra;
        sieve_l2_index* ix = create_l2_index(&d);
        ix->base_owned = true;  // the rows read here belong to the index
        dbase.p = nullptr;
        *out = ix;
    });
}

int sieve_l2_hot_leaves(const sieve_l2_index* ix, sieve_stream* s, const int8_t* q, size_t nq, uint32_t nprobe,
                        float margin, uint64_t budget, uint32_t* leaves, uint32_t* count) {
    return guarded([&] {
        SIEVE_REQUIRE(ix && s && count && (nq == 0 || (q && leaves)), SIEVE_EINVAL, "null argument");
        SIEVE_REQUIRE(nprobe >= 1 && nprobe <= (uint32_t)MAX_PROBE, SIEVE_ELIMIT,
                      "nprobe must be in [1, " + std::to_string(MAX_PROBE) + "]");
        *count = 0;
        std::vector<uint64_t> hits(ix->k, 0);
        const int pstride = (int)std::min<uint32_t>(nprobe, ix->k);
        std::vector<uint32_t> hc, hn;
        // Route in sub-batches (each a full routing of its queries), counting probes per leaf.
        const size_t sub = 4096;
        for (size_t q0 = 0; q0 < nq; q0 += sub) {
            const size_t m = std::min(sub, nq - q0);
            const int8_t* dq = dev_in(s, B_Q768, q + q0 * L2D, m * L2D);
            L2Prep prep;
            float* ps;
            uint32_t *pc, *pn;
            l2_route(ix, s, dq, m, nprobe, margin, false, prep, &ps, &pc, &pn);
            hc.resize(m * pstride);
            hn.resize(m);
            SIEVE_HIP(hipMemcpyAsync(hc.data(), pc, hc.size() * 4, hipMemcpyDeviceToHost, s->st));
            SIEVE_HIP(hipMemcpyAsync(hn.data(), pn, hn.size() * 4, hipMemcpyDeviceToHost, s->st));
            SIEVE_HIP(hipStreamSynchronize(s->st));
            for (size_t i = 0; i < m; i++)
                for (uint32_t j = 0; j < hn[i]; j++) hits[hc[i * pstride + j]]++;
        }
        // heat = probes per record byte; a record is its blocks' sections, padded to 4 KiB.
        const uint64_t rec = record_block_bytes(ix->bits);
        auto bytes = [&](uint32_t c) {
            const uint64_t nb = ix->hblock_off[c + 1] - ix->hblock_off[c];
            return (nb * rec + RECORD_ALIGN - 1) / RECORD_ALIGN * RECORD_ALIGN;
        };
        auto heat = [&](uint32_t c) { return (double)hits[c] / (double)std::max<uint64_t>(bytes(c), 1); };
        std::vector<uint32_t> order;
        for (uint32_t c = 0; c < ix->k; c++)
            if (hits[c] > 0) order.push_back(c);
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return heat(b) < heat(a); });
        uint64_t used = 0;
        uint32_t m = 0;
        for (uint32_t c : order) {
            used += bytes(c);
            if (used > budget) break;
            leaves[m++] = c;
        }
        *count = m;
    });
}

}  // extern "C"
