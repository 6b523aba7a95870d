// sieve_search_batch_noalloc: the batched search with every scratch buffer carved from one
// fixed device array, for callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The array `scratch` (NOALLOC_SCRATCH_BYTES) lives in this file's device code, so each GPU
// the library runs on holds its own copy until the process ends. Host code never passes the
// array's name to a kernel: that is a host-side stand-in, not a device address. At set-up
// (the first sieve_stream_create on a device, with that device current) the array's address on
// that device is resolved with hipGetSymbolAddress and kept per device; every buffer a call's
// kernels use is an offset into that address.
//
// A call sizes its scratch with the pipeline's sizing mode (the same code path as the enqueue,
// recording instead of allocating), lays the buffers out in the array at 256-byte offsets, and
// rejects a call whose buffers do not fit. Buffers that the kernels need zeroed at the start
// (split-scan counters, appended-list lengths) are cleared by a kernel on every call, since the
// array keeps the previous call's contents. Nothing is allocated, freed or synchronised: the
// search is enqueued on the caller's stream and its results are ready when that stream is.
// All calls on one GPU share the array, so two calls in flight on one device at once (on any
// streams) race; the caller serialises them (one stream, or events).
#include <mutex>

#include "host.hpp"

using namespace sieve;

namespace {

constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
constexpr int NOALLOC_MAX_DEVICES = 64;

static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

std::mutex g_scratch_mu;
uint8_t* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};
hipError_t g_scratch_err[NOALLOC_MAX_DEVICES] = {};

// Resolves the device address of `scratch` on device dev (the current device), once per device;
// returns it, or null with the failure kept for the error message.
uint8_t* resolve_scratch(int dev) {
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e == hipSuccess && p) {
            g_scratch_ptr[dev] = static_cast<uint8_t*>(p);
        } else {
            (void)hipGetLastError();
            g_scratch_err[dev] = e == hipSuccess ? hipErrorInvalidSymbol : e;
        }
    }
    return g_scratch_ptr[dev];
}

uint8_t* scratch_on(int dev) {
    SIEVE_REQUIRE(dev >= 0 && dev < NOALLOC_MAX_DEVICES, SIEVE_EINVAL, "device id out of range");
    uint8_t* p = resolve_scratch(dev);
    SIEVE_REQUIRE(p, SIEVE_EHIP,
                  std::string("hipGetSymbolAddress of the fixed scratch array: ") +
                      hipGetErrorString(g_scratch_err[dev]));
    return p;
}

// A stream object whose scratch slots point into the fixed array and whose stream and events
// are the caller's: it owns nothing, so it hands everything back before its destructor runs.
struct BorrowedStream {
    sieve_stream t;
    BorrowedStream(const sieve_stream* s) {
        t.st = s->st;
        t.device = s->device;
        for (size_t e = 0; e < sizeof t.ev / sizeof t.ev[0]; e++) t.ev[e] = s->ev[e];
    }
    ~BorrowedStream() {
        for (auto& b : t.b) {
            b.p = nullptr;
            b.cap = 0;
        }
        for (auto& e : t.ev) e = nullptr;
        t.st = nullptr;
    }
};

void require_device(const void* p, const char* what) {
    SIEVE_REQUIRE(p && on_device(p), SIEVE_EINVAL, std::string(what) + " must be device memory");
}

}  // namespace

namespace sieve {
// Set-up of the fixed scratch array on device dev (current): called by sieve_stream_create, so
// the array is resolved (and its code object placed on the device) before any call that must
// not allocate. A failure is reported by sieve_search_batch_noalloc, not here.
void noalloc_setup(int dev) { (void)resolve_scratch(dev); }
}  // namespace sieve

extern "C" {

int sieve_search_batch_noalloc(const sieve_index* ix, sieve_stream* s, const float* q768, const float* q256,
                               size_t nq, const sieve_params* p, const sieve_view* v, float* scores, uint32_t* rows,
                               uint32_t* counts) {
    return guarded([&] {
        SIEVE_REQUIRE(ix && s, SIEVE_EINVAL, "null argument");
        check_params(p, true);
        if (nq == 0) return;
        int dev = 0;
        SIEVE_HIP(hipGetDevice(&dev));
        SIEVE_REQUIRE(dev == s->device && dev == ix->device, SIEVE_EINVAL,
                      "the index, the stream and the current device differ");
        require_device(q768, "q768");
        require_device(q256, "q256");
        require_device(scores, "scores");
        require_device(rows, "rows");
        require_device(counts, "counts");
        ViewArgs view{};
        if (v) {
            if (v->deleted) require_device(v->deleted, "view.deleted");
            if (v->allow) require_device(v->allow, "view.allow");
            if (v->allow_flags) require_device(v->allow_flags, "view.allow_flags");
            SIEVE_REQUIRE(!v->allow || !v->allow_stride || v->allow_stride >= v->allow_words, SIEVE_EINVAL,
                          "allow_stride smaller than allow_words");
            view.deleted = v->deleted;
            view.deleted_words = v->deleted ? v->deleted_words : 0;
            view.allow = v->allow;
            view.allow_words = v->allow ? v->allow_words : 0;
            view.allow_stride = v->allow ? v->allow_stride : 0;
            view.allow_flags = v->allow ? v->allow_flags : nullptr;
        }
        // Sizes of this call's buffers, by the enqueue itself in sizing mode.
        size_t need[B_NSLOTS] = {};
        uint8_t zeroed[B_NSLOTS] = {};
        {
            sieve_stream z;
            z.sizing = need;
            z.sizing_zeroed = zeroed;
            enqueue_search(ix, &z, q256, q768, nq, p, view, scores, rows, counts);
            z.sizing = nullptr;
            z.sizing_zeroed = nullptr;
        }
        size_t off[B_NSLOTS] = {}, total = 0;
        for (int i = 0; i < B_NSLOTS; i++) {
            if (!need[i]) continue;
            off[i] = total;
            total += (need[i] + 255) & ~size_t(255);
        }
        SIEVE_REQUIRE(total <= NOALLOC_SCRATCH_BYTES, SIEVE_ELIMIT,
                      "this batch needs " + std::to_string(total) + " bytes of scratch, more than the fixed " +
                          std::to_string(NOALLOC_SCRATCH_BYTES) + " (use fewer queries per call)");
        uint8_t* base = scratch_on(dev);
        BorrowedStream bs(s);
        for (int i = 0; i < B_NSLOTS; i++) {
            if (!need[i]) continue;
            bs.t.b[i].p = base + off[i];
            bs.t.b[i].cap = need[i];
            if (zeroed[i]) launch_zero_words(static_cast<uint32_t*>(bs.t.b[i].p), (uint32_t)((need[i] + 3) / 4), s->st);
        }
        enqueue_search(ix, &bs.t, q256, q768, nq, p, view, scores, rows, counts);
        SIEVE_HIP(hipGetLastError());
    });
}

}  // extern "C"
