// Serving loop: continuous batching of single-query requests over one resident index.
//
// Requests are submitted one at a time from any host thread and queue in host memory. A
// batcher thread forms a device batch from everything pending (up to max_batch queries) as
// soon as a slot is free, and launches the whole pipeline (prep, route, scan, rerank) for it
// on the slot's own stream, so several batches are on the device at once and overlap. A slot
// holds the device scratch of one batch; the number of slots is bounded by the HBM budget,
// which is how memory gates admission: a request waits in the host queue until a slot, i.e.
// device memory, frees up. A slot's scratch is reserved at creation for the worst case of every
// path a batch can take (sizing mode of the pipeline), so batches never allocate. A completer thread retires batches in launch order and hands each
// request its own results.
//
// Each query is scanned and reranked by its own workgroups, so a request's results do not
// depend on which other requests share its batch.
//
// Filters and live writes (the serving config with filters and live updates): a filtered
// request is batched only with queued requests carrying the same filter and plan and runs
// filter::search on the device; with a live state attached, searches read its delta lists and
// tombstones, and inserts / deletes queue with the searches. A write is a barrier in
// submission order, kept on the device rather than by draining it: the batcher encodes the
// write while earlier batches run, publishes its tombstones and delta lists on the live state's
// stream after those batches' events, and makes every later batch's stream wait for the
// publish, so each search sees exactly the writes submitted before it.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "host.hpp"

using namespace sieve;

namespace {

using Clock = std::chrono::steady_clock;

enum Kind { K_SEARCH = 0, K_INSERT = 1, K_DELETE = 2 };

struct Request {
    int kind = K_SEARCH;
    const sieve_filter* filter = nullptr;
    uint32_t plan = SIEVE_PLAN_AUTO;
    std::vector<uint32_t> ext;  // writes
    std::vector<float> vecs;
    std::vector<float> q768, q256;
    std::vector<float> scores;
    std::vector<uint32_t> rows;
    uint32_t count = 0;
    int status = 0;
    std::string error;
    bool done = false;
    Clock::time_point t_submit, t_launch, t_done;
    uint32_t batch_size = 0;
    uint64_t batch_id = 0;
};

struct HostPinned {
    void* p = nullptr;
    ~HostPinned() {
        if (p) (void)hipHostFree(p);
    }
};

struct BatchSlot {
    std::unique_ptr<sieve_stream> s;
    HostPinned q768, q256, scores, rows, counts, ovf;
    float* dq768 = nullptr;
    float* dq256 = nullptr;
    float* dscores = nullptr;
    uint32_t* drows = nullptr;
    uint32_t* dcounts = nullptr;
    hipEvent_t done = nullptr;
    std::vector<uint64_t> tickets;
    uint64_t batch_id = 0;
    bool busy = false;
    ~BatchSlot() {
        void* ps[] = {dq768, dq256, dscores, drows, dcounts};
        for (void* p : ps)
            if (p) (void)hipFree(p);
        if (done) (void)hipEventDestroy(done);
    }
};

}  // namespace

struct sieve_server {
    const sieve_index* ix = nullptr;
    sieve_live* live = nullptr;
    LiveArgs la;
    hipEvent_t pub_ev = nullptr;  // the live state's publish event (owned by it), once writes ran
    sieve_params p{};
    sieve_server_opts o{};
    std::vector<std::unique_ptr<BatchSlot>> slots;
    std::mutex mu;
    std::condition_variable cv_batcher, cv_completer, cv_done;
    std::deque<uint64_t> pending;
    std::unordered_map<uint64_t, Request> reqs;
    std::deque<BatchSlot*> inflight;
    int launching = 0;  // batches taken from `pending` and not yet in `inflight`
    uint64_t next_ticket = 1, next_batch = 1;
    bool stop = false;
    sieve_server_stats stats{};
    std::thread batcher, completer;
    int device = 0;
};

namespace {

// Applies one write request (the batcher holds no batch in flight), then refreshes the delta
// segment later batches read (its block count grows with inserts).
// Stages one write (encoded into the slot log, logged, host state updated); its tombstones and
// arena wait for the publish that follows the run of writes.
void apply_write(sieve_server* sv, Request& r) {
    const int rc = guarded([&] {
        SIEVE_HIP(hipSetDevice(sv->device));
        if (r.kind == K_INSERT)
            live_insert(sv->live, r.ext.data(), r.vecs.data(), r.ext.size(), false);
        else
            live_delete(sv->live, r.ext.data(), r.ext.size(), false);
    });
    r.status = rc;
    if (rc != 0) r.error = sieve_last_error();
}

void batcher_loop(sieve_server* sv) {
    (void)hipSetDevice(sv->device);
    const size_t K = sv->p.k;
    std::vector<Request*> src;
    for (;;) {
        BatchSlot* slot = nullptr;
        std::vector<uint64_t> take;
        const sieve_filter* filter = nullptr;
        uint32_t plan = SIEVE_PLAN_AUTO;
        {
            std::unique_lock<std::mutex> lk(sv->mu);
            auto free_slot = [&]() -> BatchSlot* {
                for (auto& x : sv->slots)
                    if (!x->busy) return x.get();
                return nullptr;
            };
            auto ready = [&] {
                if (sv->pending.empty()) return false;
                const Request& h = sv->reqs[sv->pending.front()];
                if (h.kind != K_SEARCH) return true;  // a write waits for earlier batches on the device
                return free_slot() != nullptr;
            };
            sv->cv_batcher.wait(lk, [&] { return sv->stop || ready(); });
            if (sv->stop) return;
            const uint64_t head = sv->pending.front();
            Request& hr = sv->reqs[head];
            if (hr.kind != K_SEARCH) {
                // The writes queued back to back at the head are staged in order beside the batches
                // in flight (they write only rows, slot-log entries and host state those batches
                // do not read), then published together on the live state's stream once the
                // batches in flight have finished there: a device-side wait, not a drain. Later
                // batches wait for the publish on the device.
                std::vector<Request*> writes;
                while (!sv->pending.empty() && sv->reqs[sv->pending.front()].kind != K_SEARCH) {
                    Request& w = sv->reqs[sv->pending.front()];
                    sv->pending.pop_front();
                    w.t_launch = Clock::now();
                    writes.push_back(&w);  // node-based map: stable until waited out
                }
                // Only this thread records the slots' events, so these stay the batches' own.
                std::vector<hipEvent_t> waits;
                for (const BatchSlot* x : sv->inflight) waits.push_back(x->done);
                sv->launching++;
                lk.unlock();
                for (Request* w : writes) apply_write(sv, *w);
                std::string perr;
                const int prc = guarded([&] {
                    SIEVE_HIP(hipSetDevice(sv->device));
                    sv->pub_ev = live_publish_async(sv->live, waits.data(), waits.size());
                });
                if (prc != 0) perr = sieve_last_error();
                sv->la = live_args(sv->live);
                lk.lock();
                sv->launching--;
                for (Request* w : writes) {
                    if (prc != 0 && w->status == 0) {
                        w->status = prc;
                        w->error = perr;
                    }
                    w->done = true;
                    w->t_done = Clock::now();
                }
                sv->stats.writes += writes.size();
                sv->cv_done.notify_all();
                continue;
            }
            slot = free_slot();
            filter = hr.filter;
            plan = hr.plan;
            // Requests sharing the head's filter and plan, in order, up to the next write.
            std::deque<uint64_t> rest;
            while (!sv->pending.empty()) {
                const uint64_t tk = sv->pending.front();
                const Request& r = sv->reqs[tk];
                if (r.kind != K_SEARCH) break;
                sv->pending.pop_front();
                if (take.size() < sv->o.max_batch && r.filter == filter && r.plan == plan)
                    take.push_back(tk);
                else
                    rest.push_back(tk);
            }
            while (!rest.empty()) {
                sv->pending.push_front(rest.back());
                rest.pop_back();
            }
            const size_t m = take.size();
            slot->busy = true;
            slot->batch_id = sv->next_batch++;
            sv->launching++;
            auto now = Clock::now();
            src.clear();
            for (size_t i = 0; i < m; i++) {
                Request& r = sv->reqs[take[i]];
                r.t_launch = now;
                r.batch_size = (uint32_t)m;
                r.batch_id = slot->batch_id;
                src.push_back(&r);  // node-based map: stable until the request is waited out
            }
            sv->stats.batches++;
            sv->stats.queries += m;
        }
        for (size_t i = 0; i < src.size(); i++) {
            std::memcpy(static_cast<float*>(slot->q768.p) + i * DIM, src[i]->q768.data(), DIM * 4);
            std::memcpy(static_cast<float*>(slot->q256.p) + i * PDIM, src[i]->q256.data(), PDIM * 4);
        }
        const size_t nq = take.size();
        sieve_stream* s = slot->s.get();
        int status = 0;
        std::string err;
        try {
            SIEVE_HIP(hipMemcpyAsync(slot->dq768, slot->q768.p, nq * DIM * 4, hipMemcpyHostToDevice, s->st));
            SIEVE_HIP(hipMemcpyAsync(slot->dq256, slot->q256.p, nq * PDIM * 4, hipMemcpyHostToDevice, s->st));
            if (sv->pub_ev) SIEVE_HIP(hipStreamWaitEvent(s->st, sv->pub_ev, 0));  // the last writes' publish
            const LiveArgs* lv = sv->live ? &sv->la : nullptr;
            if (filter)
                enqueue_filter_search(sv->ix, filter, plan, 0, s, slot->dq256, slot->dq768, nq, &sv->p,
                                      slot->dscores, slot->drows, slot->dcounts, lv);
            else
                enqueue_search(sv->ix, s, slot->dq256, slot->dq768, nq, &sv->p, ViewArgs{}, slot->dscores,
                               slot->drows, slot->dcounts, lv);
            SIEVE_HIP(hipMemcpyAsync(slot->scores.p, slot->dscores, nq * K * 4, hipMemcpyDeviceToHost, s->st));
            SIEVE_HIP(hipMemcpyAsync(slot->rows.p, slot->drows, nq * K * 4, hipMemcpyDeviceToHost, s->st));
            SIEVE_HIP(hipMemcpyAsync(slot->counts.p, slot->dcounts, nq * 4, hipMemcpyDeviceToHost, s->st));
            SIEVE_HIP(hipMemcpyAsync(slot->ovf.p, s->b[B_OVF].p, 4, hipMemcpyDeviceToHost, s->st));
            SIEVE_HIP(hipEventRecord(slot->done, s->st));
        } catch (const Error& e) {
            status = e.code;
            err = e.msg;
        }
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->launching--;
        slot->tickets = std::move(take);
        if (status != 0) {
            for (uint64_t tk : slot->tickets) {
                Request& r = sv->reqs[tk];
                r.status = status;
                r.error = err;
                r.done = true;
                r.t_done = Clock::now();
            }
            slot->tickets.clear();
            slot->busy = false;
            sv->cv_done.notify_all();
            sv->cv_batcher.notify_one();
            continue;
        }
        sv->inflight.push_back(slot);
        sv->cv_completer.notify_one();
    }
}

void completer_loop(sieve_server* sv) {
    (void)hipSetDevice(sv->device);
    const size_t K = sv->p.k;
    for (;;) {
        BatchSlot* slot = nullptr;
        {
            std::unique_lock<std::mutex> lk(sv->mu);
            sv->cv_completer.wait(lk, [&] { return !sv->inflight.empty() || sv->stop; });
            if (sv->inflight.empty()) return;  // stopping and drained
            slot = sv->inflight.front();
        }
        hipError_t e = hipEventSynchronize(slot->done);
        const uint32_t ovf = *static_cast<uint32_t*>(slot->ovf.p);
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->inflight.pop_front();
        const auto now = Clock::now();
        for (size_t i = 0; i < slot->tickets.size(); i++) {
            Request& r = sv->reqs[slot->tickets[i]];
            if (e != hipSuccess) {
                r.status = SIEVE_EHIP;
                r.error = hipGetErrorString(e);
            } else if (ovf) {
                r.status = SIEVE_ELIMIT;
                r.error = "expand plan: the cut needs more clusters than its probe limit";
            } else {
                const uint32_t n = static_cast<uint32_t*>(slot->counts.p)[i];
                r.count = n;
                r.scores.assign(static_cast<float*>(slot->scores.p) + i * K, static_cast<float*>(slot->scores.p) + i * K + n);
                r.rows.assign(static_cast<uint32_t*>(slot->rows.p) + i * K, static_cast<uint32_t*>(slot->rows.p) + i * K + n);
            }
            r.done = true;
            r.t_done = now;
        }
        slot->tickets.clear();
        slot->busy = false;
        sv->cv_done.notify_all();
        sv->cv_batcher.notify_one();
    }
}

template <class T>
void pinned(HostPinned& h, size_t n) {
    SIEVE_HIP(hipHostMalloc(&h.p, std::max<size_t>(n * sizeof(T), 16), hipHostMallocDefault));
}

// Device bytes of a slot's own staging (inputs and outputs of a batch of mb queries).
size_t slot_io_bytes(const sieve_params& p, size_t mb) { return mb * (DIM + PDIM) * 4 + mb * (p.k * 8 + 4) + 16; }

// The scratch a slot's stream needs for every path a batch of 1 .. mb queries can take: the plain
// search (small-batch, per-query or cluster-major scan by batch size), Post, Expand and Brute
// filters, with the live delta lists at their largest. Recorded per scratch slot in `need`
// (B_NSLOTS entries) by running the pipeline in sizing mode, which enqueues nothing.
void slot_scratch_need(const sieve_server* sv, size_t mb, size_t* need) {
    sieve_stream z;
    z.sizing = need;
    LiveArgs lmax;
    const LiveArgs* lv = nullptr;
    if (sv->live) {
        lmax = live_args_max(sv->live);
        lv = &lmax;
    }
    const float* q = reinterpret_cast<const float*>(uintptr_t(256));
    float* os = reinterpret_cast<float*>(uintptr_t(256));
    uint32_t* ou = reinterpret_cast<uint32_t*>(uintptr_t(256));
    for (size_t m = 1; m <= mb; m++) {
        enqueue_search(sv->ix, &z, q, q, m, &sv->p, ViewArgs{}, os, ou, ou, lv);
        try {
            size_filter_search(sv->ix, &z, m, &sv->p, lv);
        } catch (const Error&) {
            // a plan these parameters cannot run fails its requests with that error instead
        }
    }
    z.sizing = nullptr;
}

}  // namespace

extern "C" {}  // namespace

extern "C" {

int sieve_server_create(const sieve_index* ix, const sieve_params* p, const sieve_server_opts* o, sieve_server** out) {
    return guarded([&] {
        SIEVE_REQUIRE(ix && p && out, SIEVE_EINVAL, "null argument");
        check_params(p, true);
        std::unique_ptr<sieve_server> sv(new sieve_server());
        SIEVE_HIP(hipGetDevice(&sv->device));
        sv->ix = ix;
        sv->p = *p;
        sv->o = o ? *o : sieve_server_opts{};
        sv->live = sv->o.live;
        if (sv->live) {
            SIEVE_REQUIRE(ix->live == sv->live, SIEVE_EINVAL, "opts.live is not this index's live state");
            live_settle(sv->live);  // a publish of an earlier server's writes
            sv->la = live_args(sv->live);
        }
        if (sv->o.max_batch == 0) sv->o.max_batch = 1024;
        size_t budget = sv->o.hbm_budget;
        if (budget == 0) {
            size_t fr = 0, tot = 0;
            SIEVE_HIP(hipMemGetInfo(&fr, &tot));
            budget = fr / 4;
        }
        // Admission by HBM: a slot holds the worst-case scratch of every path a batch can take,
        // reserved here, so no batch allocates device memory once the server runs.
        std::vector<size_t> need(B_NSLOTS, 0);
        slot_scratch_need(sv.get(), sv->o.max_batch, need.data());
        size_t per = slot_io_bytes(*p, sv->o.max_batch);
        for (size_t b : need) per += b;
        size_t nslots = budget / per;
        if (sv->o.max_inflight) nslots = std::min<size_t>(nslots, sv->o.max_inflight);
        SIEVE_REQUIRE(nslots >= 1, SIEVE_ENOMEM,
                      "HBM budget below one batch slot (" + std::to_string(per) + " bytes)");
        nslots = std::min<size_t>(nslots, 16);
        sv->o.max_inflight = (uint32_t)nslots;
        const size_t mb = sv->o.max_batch, K = p->k;
        for (size_t i = 0; i < nslots; i++) {
            std::unique_ptr<BatchSlot> sl(new BatchSlot());
            sl->s.reset(new sieve_stream());
            sl->s->device = sv->device;
            SIEVE_HIP(hipStreamCreateWithFlags(&sl->s->st, hipStreamNonBlocking));
            ensure_events(sl->s.get());
            SIEVE_HIP(hipEventCreateWithFlags(&sl->done, hipEventDisableTiming));
            pinned<float>(sl->q768, mb * DIM);
            pinned<float>(sl->q256, mb * PDIM);
            pinned<float>(sl->scores, mb * K);
            pinned<uint32_t>(sl->rows, mb * K);
            pinned<uint32_t>(sl->counts, mb);
            pinned<uint32_t>(sl->ovf, 1);
            sl->dq768 = dmalloc<float>(mb * DIM, nullptr);
            sl->dq256 = dmalloc<float>(mb * PDIM, nullptr);
            sl->dscores = dmalloc<float>(mb * K, nullptr);
            sl->drows = dmalloc<uint32_t>(mb * K, nullptr);
            sl->dcounts = dmalloc<uint32_t>(mb, nullptr);
            // The reserved scratch, zeroed (counters that kernels leave at zero start zeroed).
            sieve_stream* s = sl->s.get();
            for (int b = 0; b < B_NSLOTS; b++)
                if (need[b]) grow_zeroed<uint8_t>(s->b[b], need[b], s->st);
            SIEVE_HIP(hipStreamSynchronize(s->st));
            sv->slots.push_back(std::move(sl));
        }
        sv->stats.slots = (uint32_t)nslots;
        sv->stats.slot_bytes = per;
        sieve_server* raw = sv.get();
        raw->batcher = std::thread(batcher_loop, raw);
        raw->completer = std::thread(completer_loop, raw);
        *out = sv.release();
    });
}

int sieve_server_submit(sieve_server* sv, const float* q768, const float* q256, uint64_t* ticket) {
    return guarded([&] {
        SIEVE_REQUIRE(sv && q768 && q256 && ticket, SIEVE_EINVAL, "null argument");
        Request r;
        r.q768.assign(q768, q768 + DIM);
        r.q256.assign(q256, q256 + PDIM);
        r.t_submit = Clock::now();
        std::lock_guard<std::mutex> lk(sv->mu);
        SIEVE_REQUIRE(!sv->stop, SIEVE_EINVAL, "server is stopping");
        uint64_t tk = sv->next_ticket++;
        sv->reqs.emplace(tk, std::move(r));
        sv->pending.push_back(tk);
        *ticket = tk;
        sv->cv_batcher.notify_one();
    });
}

namespace {
int submit_request(sieve_server* sv, Request&& r, uint64_t* ticket) {
    r.t_submit = Clock::now();
    std::lock_guard<std::mutex> lk(sv->mu);
    SIEVE_REQUIRE(!sv->stop, SIEVE_EINVAL, "server is stopping");
    uint64_t tk = sv->next_ticket++;
    sv->reqs.emplace(tk, std::move(r));
    sv->pending.push_back(tk);
    *ticket = tk;
    sv->cv_batcher.notify_one();
    return 0;
}
}  // namespace

int sieve_server_submit_filtered(sieve_server* sv, const float* q768, const float* q256, const sieve_filter* filter,
                                 uint32_t plan, uint64_t* ticket) {
    return guarded([&] {
        SIEVE_REQUIRE(sv && q768 && q256 && filter && ticket, SIEVE_EINVAL, "null argument");
        SIEVE_REQUIRE(plan <= SIEVE_PLAN_AUTO, SIEVE_EINVAL, "unknown plan");
        Request r;
        r.filter = filter;
        r.plan = plan;
        r.q768.assign(q768, q768 + DIM);
        r.q256.assign(q256, q256 + PDIM);
        submit_request(sv, std::move(r), ticket);
    });
}

int sieve_server_insert(sieve_server* sv, const uint32_t* ext, const float* vecs, size_t n, uint64_t* ticket) {
    return guarded([&] {
        SIEVE_REQUIRE(sv && ticket && (n == 0 || (ext && vecs)), SIEVE_EINVAL, "null argument");
        SIEVE_REQUIRE(sv->live, SIEVE_EINVAL, "the server has no live state");
        Request r;
        r.kind = K_INSERT;
        r.ext.assign(ext, ext + n);
        r.vecs.assign(vecs, vecs + n * DIM);
        submit_request(sv, std::move(r), ticket);
    });
}

int sieve_server_delete(sieve_server* sv, const uint32_t* ext, size_t n, uint64_t* ticket) {
    return guarded([&] {
        SIEVE_REQUIRE(sv && ticket && (n == 0 || ext), SIEVE_EINVAL, "null argument");
        SIEVE_REQUIRE(sv->live, SIEVE_EINVAL, "the server has no live state");
        Request r;
        r.kind = K_DELETE;
        r.ext.assign(ext, ext + n);
        submit_request(sv, std::move(r), ticket);
    });
}

int sieve_server_wait(sieve_server* sv, uint64_t ticket, float* scores, uint32_t* rows, uint32_t* count,
                      sieve_request_info* info) {
    int rc = 0;
    std::string err;
    int g = guarded([&] {
        SIEVE_REQUIRE(sv && count, SIEVE_EINVAL, "null argument");
        std::unique_lock<std::mutex> lk(sv->mu);
        auto it = sv->reqs.find(ticket);
        SIEVE_REQUIRE(it != sv->reqs.end(), SIEVE_EINVAL, "unknown ticket " + std::to_string(ticket));
        // Submits may rehash the map while this thread waits: keep the element (whose address
        // is stable), not the iterator, and find the ticket again to erase it.
        Request& r = it->second;
        sv->cv_done.wait(lk, [&] { return r.done; });
        *count = r.count;
        if (scores) std::memcpy(scores, r.scores.data(), r.count * 4);
        if (rows) std::memcpy(rows, r.rows.data(), r.count * 4);
        if (info) {
            auto ms = [](Clock::duration d) { return std::chrono::duration<double, std::milli>(d).count(); };
            info->queue_ms = ms(r.t_launch - r.t_submit);
            info->total_ms = ms(r.t_done - r.t_submit);
            info->batch_size = r.batch_size;
            info->batch_id = r.batch_id;
        }
        rc = r.status;
        err = r.error;
        sv->reqs.erase(ticket);
    });
    if (g != 0) return g;
    if (rc != 0) {
        set_error(err);
        return rc;
    }
    return 0;
}

int sieve_server_get_stats(sieve_server* sv, sieve_server_stats* out) {
    return guarded([&] {
        SIEVE_REQUIRE(sv && out, SIEVE_EINVAL, "null argument");
        std::lock_guard<std::mutex> lk(sv->mu);
        *out = sv->stats;
        out->pending = sv->pending.size();
        out->inflight = sv->inflight.size();
    });
}

void sieve_server_destroy(sieve_server* sv) {
    if (!sv) return;
    {
        std::unique_lock<std::mutex> lk(sv->mu);
        // let queued work finish: stop admitting, drain pending and in-flight batches
        sv->cv_done.wait(lk, [&] { return sv->pending.empty() && sv->inflight.empty() && sv->launching == 0; });
        sv->stop = true;
    }
    sv->cv_batcher.notify_all();
    sv->cv_completer.notify_all();
    if (sv->batcher.joinable()) sv->batcher.join();
    if (sv->completer.joinable()) sv->completer.join();
    for (auto& sl : sv->slots) (void)hipStreamSynchronize(sl->s->st);
    if (sv->live) (void)guarded([&] { live_settle(sv->live); });
    delete sv;
}

}  // extern "C"
