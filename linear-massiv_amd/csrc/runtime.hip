// Context, error reporting, staging/scratch memory and kernel-family timing.
#include <cstring>
#include <vector>

#include "common.hpp"

namespace lmx {

namespace {
thread_local std::string g_last_error = "no error";
constexpr int kMaxDevices = 64;
Context *g_ctx[kMaxDevices] = {};
std::mutex g_ctx_mu;

const char *kFamilies[] = {"gemm", "sgemm", "getrf_panel", "trsm", "laswp", "potrf_diag",
                           "syrk", "getrs", "refine", "getrf", "potrf", "batch", "gemm64", "sgemm64", "ssyrk",
                           "bgemm", "bgemm3", "gemv"};
constexpr int kNumFamilies = sizeof(kFamilies) / sizeof(kFamilies[0]);
struct EventPair {
    hipEvent_t a, b;
    double work;
};
std::mutex g_time_mu;
bool g_timing = false;
std::vector<EventPair> g_pairs[kNumFamilies];
std::vector<hipEvent_t> g_event_pool;

int family_slot(const char *f) {
    for (int i = 0; i < kNumFamilies; ++i)
        if (std::strcmp(f, kFamilies[i]) == 0) return i;
    return -1;
}

hipEvent_t take_event() {
    if (!g_event_pool.empty()) {
        hipEvent_t e = g_event_pool.back();
        g_event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    LMX_HIP(hipEventCreate(&e));
    return e;
}

Context *make_context(int dev) {
    auto *c = new Context();
    c->device = dev;
    LMX_HIP(hipSetDevice(dev));
    // Default priority, non-blocking with respect to the legacy stream.
    LMX_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    hipDeviceProp_t prop;
    LMX_HIP(hipGetDeviceProperties(&prop, dev));
    c->cu_count = prop.multiProcessorCount;
    // The default pool returns freed memory at every synchronisation; then the
    // first allocations after the factorisation's info read (a host
    // synchronisation) took 3 ms to map memory again.
    hipMemPoolProps props = {};
    props.allocType = hipMemAllocationTypePinned;
    props.location.type = hipMemLocationTypeDevice;
    props.location.id = dev;
    LMX_HIP(hipMemPoolCreate(&c->pool, &props));
    uint64_t keep = UINT64_MAX;
    LMX_HIP(hipMemPoolSetAttribute(c->pool, hipMemPoolAttrReleaseThreshold, &keep));
    return c;
}
}  // namespace

void throw_hip(hipError_t e, const char *what, const char *file, int line) {
    char buf[512];
    std::snprintf(buf, sizeof buf, "HIP error %d (%s) in %s at %s:%d", (int)e,
                  hipGetErrorString(e), what, file, line);
    throw Error(e == hipErrorOutOfMemory ? LMX_ERR_NOMEM : LMX_ERR_DEVICE, buf);
}

void set_last_error(const std::string &m) { g_last_error = m; }

Context &ctx() {
    int dev = 0;
    LMX_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw Error(LMX_ERR_DEVICE, "device id out of range");
    std::lock_guard<std::mutex> g(g_ctx_mu);
    if (!g_ctx[dev]) g_ctx[dev] = make_context(dev);
    return *g_ctx[dev];
}

Context *context_of(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> g(g_ctx_mu);
    return g_ctx[dev];
}

// NULL is HIP's default (legacy) stream, which orders against every other
// blocking stream -- the stream torch uses unless told otherwise.
hipStream_t resolve_stream(void *s) { return static_cast<hipStream_t>(s); }

void *staging(Context &c, int slot, size_t bytes) {
    if (bytes == 0) bytes = 256;
    if (c.staging_bytes[slot] < bytes) {
        if (c.staging[slot]) {
            LMX_HIP(hipStreamSynchronize(c.stream));
            LMX_HIP(hipFree(c.staging[slot]));
            c.staging[slot] = nullptr;
            c.staging_bytes[slot] = 0;
        }
        size_t sz = bytes + bytes / 8;
        LMX_HIP(hipMalloc(&c.staging[slot], sz));
        c.staging_bytes[slot] = sz;
    }
    return c.staging[slot];
}

// Scratch sizes are rounded up to 1/8 of their power of two, so that a
// freed block serves later requests of about the same size; a cached block
// is taken for a request of at least 4/5 of its size.
namespace {
constexpr size_t kScratchCacheBytes = size_t(8) << 30;

size_t scratch_round(size_t bytes) {
    size_t p2 = 256;
    while (p2 < bytes) p2 <<= 1;
    const size_t g = std::max<size_t>(256, p2 / 8);
    return (bytes + g - 1) / g * g;
}

// Frees every cached block (stream-ordered on the stream it was last used
// on); the caller holds scratch_mu.
void release_cached(Context &c) {
    for (auto &e : c.scratch_free)
        for (auto &b : e.second) (void)hipFreeAsync(b.p, e.first);
    c.scratch_free.clear();
    c.scratch_cached = 0;
}
}  // namespace

Scratch::Scratch(size_t req, hipStream_t st) : s(st), bytes(scratch_round(req == 0 ? 256 : req)) {
    Context &c = ctx();
    // Under stream capture the allocation must be the graph's own (a cached
    // block would be handed out again while the graph still uses it).
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    LMX_HIP(hipStreamIsCapturing(s, &cs));
    if (cs == hipStreamCaptureStatusNone) device = c.device;
    if (device >= 0) {
        std::lock_guard<std::mutex> g(c.scratch_mu);
        for (auto &e : c.scratch_free) {
            if (e.first != s) continue;
            auto &v = e.second;
            int best = -1;
            for (int i = 0; i < (int)v.size(); ++i)
                if (v[i].bytes >= bytes && v[i].bytes / 5 * 4 <= bytes && (best < 0 || v[i].bytes < v[best].bytes))
                    best = i;
            if (best >= 0) {
                p = v[best].p;
                bytes = v[best].bytes;
                c.scratch_cached -= bytes;
                v[best] = v.back();
                v.pop_back();
                return;
            }
            break;
        }
    }
    hipError_t e = hipMallocFromPoolAsync(&p, bytes, c.pool, s);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        {
            std::lock_guard<std::mutex> g(c.scratch_mu);
            release_cached(c);
        }
        e = hipMallocFromPoolAsync(&p, bytes, c.pool, s);
    }
    if (e != hipSuccess) {
        p = nullptr;
        throw_hip(e, "hipMallocFromPoolAsync", __FILE__, __LINE__);
    }
}

Scratch::~Scratch() {
    if (!p) return;
    Context *c = device >= 0 ? context_of(device) : nullptr;
    if (c) {
        std::lock_guard<std::mutex> g(c->scratch_mu);
        if (c->scratch_cached + bytes <= kScratchCacheBytes) {
            for (auto &e : c->scratch_free)
                if (e.first == s) {
                    e.second.push_back({p, bytes});
                    c->scratch_cached += bytes;
                    return;
                }
            c->scratch_free.push_back({s, {{p, bytes}}});
            c->scratch_cached += bytes;
            return;
        }
    }
    (void)hipFreeAsync(p, s);
}

AuxStream::AuxStream() {
    Context &c = ctx();
    device = c.device;
    {
        std::lock_guard<std::mutex> g(c.scratch_mu);
        if (!c.aux_free.empty()) {
            aux = c.aux_free.back().aux;
            ready = c.aux_free.back().ready;
            done = c.aux_free.back().done;
            c.aux_free.pop_back();
            return;
        }
    }
    // Default priority, non-blocking with respect to the legacy stream.
    LMX_HIP(hipStreamCreateWithFlags(&aux, hipStreamNonBlocking));
    LMX_HIP(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
    LMX_HIP(hipEventCreateWithFlags(&done, hipEventDisableTiming));
}

AuxStream::~AuxStream() {
    Context *c = context_of(device);
    if (c && aux && ready && done) {
        std::lock_guard<std::mutex> g(c->scratch_mu);
        c->aux_free.push_back({aux, ready, done});
        return;
    }
    if (ready) (void)hipEventDestroy(ready);
    if (done) (void)hipEventDestroy(done);
    if (aux) (void)hipStreamDestroy(aux);
}

bool timing_enabled() { return g_timing; }

TimedRegion::TimedRegion(const char *family, hipStream_t st, double w) : s(st), work(w) {
    if (!g_timing) return;
    std::lock_guard<std::mutex> g(g_time_mu);
    slot = family_slot(family);
    if (slot < 0) return;
    e0 = take_event();
    e1 = take_event();
    LMX_HIP(hipEventRecord(e0, s));
}

TimedRegion::~TimedRegion() {
    if (slot < 0) return;
    std::lock_guard<std::mutex> g(g_time_mu);
    if (hipEventRecord(e1, s) == hipSuccess) g_pairs[slot].push_back({e0, e1, work});
}

}  // namespace lmx

using namespace lmx;

extern "C" {

int lmx_init(int device) {
    return guarded([&] {
        int n = 0;
        LMX_HIP(hipGetDeviceCount(&n));
        if (device < 0 || device >= n) fail_arg("lmx_init: no such device");
        LMX_HIP(hipSetDevice(device));
        (void)ctx();
        return LMX_OK;
    });
}

int lmx_finalize(void) {
    return guarded([&] {
        std::lock_guard<std::mutex> g(g_ctx_mu);
        bool any = false;
        for (int d = 0; d < kMaxDevices; ++d) any = any || g_ctx[d] != nullptr;
        if (!any) return LMX_OK;  // nothing initialised: no HIP call at all
        int cur = 0;
        LMX_HIP(hipGetDevice(&cur));
        for (int d = 0; d < kMaxDevices; ++d) {
            Context *c = g_ctx[d];
            if (!c) continue;
            LMX_HIP(hipSetDevice(d));
            LMX_HIP(hipStreamSynchronize(c->stream));
            for (int i = 0; i < 4; ++i)
                if (c->staging[i]) LMX_HIP(hipFree(c->staging[i]));
            LMX_HIP(hipStreamDestroy(c->stream));
            // scratch may still be in use on callers' streams until now
            LMX_HIP(hipDeviceSynchronize());
            // every stream is idle: the blocks go back to the pool on the
            // legacy stream (the stream a block last served may be gone)
            for (auto &e : c->scratch_free)
                for (auto &b : e.second) LMX_HIP(hipFreeAsync(b.p, nullptr));
            LMX_HIP(hipDeviceSynchronize());
            c->scratch_free.clear();
            c->scratch_cached = 0;
            for (auto &a : c->aux_free) {
                LMX_HIP(hipEventDestroy(a.ready));
                LMX_HIP(hipEventDestroy(a.done));
                LMX_HIP(hipStreamDestroy(a.aux));
            }
            c->aux_free.clear();
            if (c->pool) LMX_HIP(hipMemPoolDestroy(c->pool));
            delete c;
            g_ctx[d] = nullptr;
        }
        LMX_HIP(hipSetDevice(cur));
        return LMX_OK;
    });
}

const char *lmx_last_error(void) { return g_last_error.c_str(); }

int lmx_version(void) { return 100; }

int lmx_kernel_timing(int enable) {
    std::lock_guard<std::mutex> g(g_time_mu);
    g_timing = enable != 0;
    return LMX_OK;
}

int lmx_kernel_timing_reset(void) {
    return guarded([&] {
        std::lock_guard<std::mutex> g(g_time_mu);
        for (auto &v : g_pairs)
            for (auto &p : v) {
                LMX_HIP(hipEventSynchronize(p.b));
                g_event_pool.push_back(p.a);
                g_event_pool.push_back(p.b);
            }
        for (auto &v : g_pairs) v.clear();
        return LMX_OK;
    });
}

int lmx_kernel_timing_get(const char *family, int64_t *count, double *total_ms, double *total_work) {
    return guarded([&] {
        std::lock_guard<std::mutex> g(g_time_mu);
        int s = family_slot(family ? family : "");
        if (s < 0) fail_arg("unknown kernel family");
        double tot = 0, work = 0;
        for (auto &p : g_pairs[s]) {
            LMX_HIP(hipEventSynchronize(p.b));
            float ms = 0;
            LMX_HIP(hipEventElapsedTime(&ms, p.a, p.b));
            tot += ms;
            work += p.work;
        }
        if (count) *count = (int64_t)g_pairs[s].size();
        if (total_ms) *total_ms = tot;
        if (total_work) *total_work = work;
        return LMX_OK;
    });
}

}  // extern "C"
