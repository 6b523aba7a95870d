// The device live state (Live, src/live.rs:151-163) shared by the update path (live.hip) and
// its durability half (livefs.hip: the write-ahead log, the row segment of a compacted index and
// compaction).
#pragma once

#include <unistd.h>

#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "host.hpp"

struct sieve_live {
    sieve_index* ix = nullptr;
    bool soar = true;
    float lambda = 1.0f;
    uint64_t n = 0;            // corpus rows
    uint64_t cap_rows = 0;     // delta rows the device holds
    uint64_t cap_slots = 0;    // arena slots after the base slots
    uint64_t nrows = 0, nlog = 0, ndeleted = 0;
    mutable uint64_t arena_blocks = 0;  // exact, or an upper bound while a publish is pending
    uint64_t seg_rows = 0;     // delta rows read from a compacted index's row segment (live.*)
    size_t device_bytes = 0;
    float* cents_rot = nullptr;
    int8_t* rows = nullptr;
    float* norm256 = nullptr;
    uint8_t* t2 = nullptr;
    float* t2n = nullptr;
    uint32_t* ext = nullptr;
    uint64_t* deleted = nullptr;
    uint64_t deleted_words = 0;
    uint4* lcode = nullptr;
    float* lmul = nullptr;
    float* lerr = nullptr;
    uint32_t* lrow = nullptr;
    uint32_t* lcl = nullptr;
    uint32_t* dblock_off = nullptr;  // k + 1
    // writer-side state on the host
    std::unordered_map<uint32_t, uint32_t> by_ext;
    std::vector<uint64_t> deleted_host;
    std::string index_dir;  // the directory the index was opened from ("" for an index from arrays)
    std::string wal_path;   // the write-ahead log ("" without one)
    int wal_fd = -1;
    sieve_stream work;  // scratch for the encoder and the arena rebuild
    DevBuf xs, zq, pairs, sort_keys, sort_vals, sort_tmp, counts, nb, cstart, words_i, words_v;
    // Staged writes: tombstone words changed and rows applied since the last publish.
    std::vector<uint64_t> pend;
    bool arena_dirty = false;
    // A server's writes publish on a stream of their own, ordered after the batches in flight by
    // device-side waits instead of a drain (live_publish_async). At most one such publish is
    // pending; host-side readers settle it first (live_settle).
    mutable std::mutex pub_mu;
    hipStream_t pub = nullptr;
    hipEvent_t pub_ev = nullptr;
    mutable bool pub_pending = false;
    mutable bool blocks_from_pub = false;
    uint32_t* blocks_pinned = nullptr;  // the arena's block count, read back by the publish

    ~sieve_live() {
        if (pub) {
            (void)hipStreamSynchronize(pub);
            (void)hipStreamDestroy(pub);
        }
        if (pub_ev) (void)hipEventDestroy(pub_ev);
        if (blocks_pinned) (void)hipHostFree(blocks_pinned);
        void* ps[] = {cents_rot, rows, norm256, t2, t2n, ext, deleted, lcode, lmul, lerr, lrow, lcl, dblock_off};
        for (void* p : ps)
            if (p) (void)hipFree(p);
        for (DevBuf* b : {&xs, &zq, &pairs, &sort_keys, &sort_vals, &sort_tmp, &counts, &nb, &cstart, &words_i,
                          &words_v})
            b->release();
        if (wal_fd >= 0) (void)::close(wal_fd);
    }

    sieve::LiveArgs args() const {
        sieve::LiveArgs a;
        a.dblock_off = dblock_off;
        a.dbase = (uint32_t)ix->nblocks;
        a.dblocks = (uint32_t)arena_blocks;
        a.n_base = (uint32_t)n;
        a.drows = rows;
        a.dnorm256 = norm256;
        a.dt2 = ix->t2_codes ? t2 : nullptr;
        a.dt2n = ix->t2_codes ? t2n : nullptr;
        a.dext = ext;
        a.deleted = deleted;
        a.deleted_words = deleted_words;
        return a;
    }

    // Live::current (src/live.rs:260-266).
    bool current(uint32_t e, uint32_t* id) const {
        auto it = by_ext.find(e);
        if (it != by_ext.end()) {
            *id = it->second;
            return true;
        }
        if ((uint64_t)e < n && !bit_host(e)) {
            *id = e;
            return true;
        }
        return false;
    }
    bool bit_host(uint32_t id) const {
        const size_t w = id >> 6;
        return w < deleted_host.size() && ((deleted_host[w] >> (id & 63)) & 1) == 1;
    }
};

namespace sieve {

// Log records (src/live.rs:131-135): the op, the external id (u32, little-endian), and for an
// insert the prefix norm (f32, little-endian) and the int8 row.
constexpr uint8_t WAL_OP_INSERT = 1, WAL_OP_DELETE = 2;
constexpr size_t WAL_INSERT_BYTES = 1 + 4 + 4 + DIM, WAL_DELETE_BYTES = 1 + 4;

inline void put_le32(uint8_t* p, uint32_t v) {
    for (int i = 0; i < 4; i++) p[i] = (uint8_t)(v >> (8 * i));
}
inline uint32_t get_le32(const uint8_t* p) {
    return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24;
}
inline void wal_insert_record(uint32_t ext, float norm, const int8_t* row, uint8_t* out) {
    uint32_t nb;
    std::memcpy(&nb, &norm, 4);
    out[0] = WAL_OP_INSERT;
    put_le32(out + 1, ext);
    put_le32(out + 5, nb);
    std::memcpy(out + 9, row, DIM);
}
inline void wal_delete_record(uint32_t ext, uint8_t* out) {
    out[0] = WAL_OP_DELETE;
    put_le32(out + 1, ext);
}

// The live state of `ix` with room for o.max_rows delta rows (Live::open without a segment or
// a log), attached to the index.
sieve_live* live_create(sieve_index* ix, const sieve_live_opts& o);
// Live::apply_insert (src/live.rs:317-374) of n rows already quantised into the delta arrays at
// row l->nrows on: ext_host are their external ids (host memory). Without publish the tombstones
// and the arena wait for the next live_publish / live_publish_async.
void live_apply_rows(sieve_live* l, const uint32_t* ext_host, size_t n, bool publish = true);
// Live::apply_delete (src/live.rs:308-315); ext_host in host memory.
void live_apply_delete(sieve_live* l, const uint32_t* ext_host, size_t n, bool publish = true);
// Publishes the staged writes on the work stream (tombstone words, then the arena rebuilt from
// the slot log if rows were applied) and waits for them.
void live_publish(sieve_live* l);
// Uploads the tombstone words whose indexes `changed` lists from the host copy.
void live_push_words(sieve_live* l, hipStream_t st, const std::vector<uint64_t>& changed);
// Live::log (src/live.rs:251-258): appends the records to the write-ahead log, then waits for
// the data to be durable. No-op without a log.
void live_log(sieve_live* l, const std::vector<uint8_t>& rec);

}  // namespace sieve
