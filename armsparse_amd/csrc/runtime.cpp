// Library runtime: status strings, HIP error capture, stream-ordered
// workspace, optional event timing of kernel launches.
#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "common.hpp"

namespace armsp {

namespace {
thread_local int g_last_hip_error = 0;

std::mutex g_pool_mu;
bool g_pool_ready[64] = {};
hipMemPool_t g_pools[64] = {};

// Workspace for streams under HIP graph capture, where nothing may allocate:
// one reserved arena per device (armsp_workspace_reserve), carved from the
// bottom as a stack (an entry point may hold two workspaces at once, freed in
// reverse order).  A captured graph keeps the arena's addresses, so an arena
// is never freed: a larger reservation retires the old one.
std::mutex g_arena_mu;
char* g_arena[64] = {};
size_t g_arena_bytes[64] = {};
size_t g_arena_used[64] = {};
std::vector<std::pair<size_t, size_t>> g_arena_marks[64];  // [begin, end) of the live blocks, newest last
size_t g_ws_high[64] = {};              // the largest workspace a call on the device has asked for
std::vector<void*> g_retired;

struct TimingRecord {
    std::string family;
    hipEvent_t start, stop;
};
std::mutex g_timing_mu;
bool g_timing = false;
std::vector<TimingRecord*> g_records;

struct FamilyTotal {
    std::string family;
    double ms;
    int64_t launches;
};
std::vector<FamilyTotal> g_totals;

void resolve_records_locked() {
    for (TimingRecord* r : g_records) {
        if (r->stop) {
            hipEventSynchronize(r->stop);
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, r->start, r->stop) == hipSuccess) {
                bool found = false;
                for (auto& t : g_totals)
                    if (t.family == r->family) {
                        t.ms += ms;
                        t.launches += 1;
                        found = true;
                    }
                if (!found) g_totals.push_back({r->family, (double)ms, 1});
            }
            hipEventDestroy(r->stop);
        }
        hipEventDestroy(r->start);
        delete r;
    }
    g_records.clear();
}
}  // namespace

// ARMSP_FAIL_ALLOC set to a non-empty value other than "0" (read once per
// process): every device allocation of the library fails as if the device were
// out of memory, before any HIP call — the failure paths can be exercised
// without a GPU.
bool alloc_injected() {
    static const bool on = [] {
        const char* e = std::getenv("ARMSP_FAIL_ALLOC");
        return e && *e && std::strcmp(e, "0") != 0;
    }();
    return on;
}

int hip_fail(hipError_t e) {
    g_last_hip_error = (int)e;
    if (e == hipErrorOutOfMemory) return ARMSP_ERROR_ALLOC;
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return ARMSP_ERROR_NO_DEVICE;
    return ARMSP_ERROR_HIP;
}

// Workspaces of calls outside graph capture come from a small cache per host
// thread and (device, stream): two slots, since an entry point holds at most two
// workspaces at once.  A slot keeps its buffer between calls, so a steady
// sequence of calls makes no allocator call at all; calls on one stream run in
// stream order, and each host thread has its own slots, so two threads that
// enqueue on the same stream never share a buffer.  A slot that is too small is
// released with hipFreeAsync on the same stream (after the work that used it)
// and replaced.  Buffers come from the library's own stream-ordered pool per
// device (the device's default pool, which other libraries in the process share,
// is left untouched).  A thread keeps entries for at most kWsStreams streams: the
// least recently used idle entry makes room, and its buffers, like all of a
// thread's buffers when the thread ends, are released after a device-wide
// synchronisation, since the stream they were used on may no longer exist.
namespace {
constexpr size_t kWsStreams = 8;

struct WsSlot {
    void* p = nullptr;
    size_t bytes = 0;
    bool busy = false;
};
struct WsEntry {
    int dev;
    hipStream_t s;
    WsSlot slot[2];
};

void ws_release_entry(WsEntry& e) {
    int cur = -1;
    const bool have_cur = hipGetDevice(&cur) == hipSuccess;
    bool switched = false;
    bool synced = false;
    for (auto& sl : e.slot) {
        if (!sl.p) continue;
        if (!switched && have_cur && cur != e.dev) switched = hipSetDevice(e.dev) == hipSuccess;
        if (!synced) synced = hipDeviceSynchronize() == hipSuccess;  // work of any stream that used it
        (void)hipFreeAsync(sl.p, nullptr);
        sl = WsSlot{};
    }
    if (switched) (void)hipSetDevice(cur);
    (void)hipGetLastError();
}

struct WsCache {
    std::vector<WsEntry> entries;  // least recently used first
    ~WsCache() {
        for (auto& e : entries) ws_release_entry(e);
    }
};
thread_local WsCache t_ws;

// The cache entry of (dev, s), most recently used last; nullptr when every entry
// holds a busy slot and none can make room (the caller then allocates uncached).
WsEntry* ws_entry(int dev, hipStream_t s) {
    auto& v = t_ws.entries;
    for (size_t i = 0; i < v.size(); ++i)
        if (v[i].dev == dev && v[i].s == s) {
            if (i + 1 != v.size()) std::rotate(v.begin() + (std::ptrdiff_t)i, v.begin() + (std::ptrdiff_t)i + 1, v.end());
            return &v.back();
        }
    if (v.size() >= kWsStreams) {
        size_t i = 0;
        while (i < v.size() && (v[i].slot[0].busy || v[i].slot[1].busy)) ++i;
        if (i == v.size()) return nullptr;
        ws_release_entry(v[i]);
        v.erase(v.begin() + (std::ptrdiff_t)i);
    }
    v.push_back(WsEntry{dev, s, {}});
    return &v.back();
}

int pool_alloc(void** p, size_t bytes, int dev, hipStream_t s) {
    hipMemPool_t pool;
    {
        std::lock_guard<std::mutex> g(g_pool_mu);
        if (!g_pool_ready[dev]) {
            hipMemPoolProps props;
            std::memset(&props, 0, sizeof(props));
            props.allocType = hipMemAllocationTypePinned;
            props.location.type = hipMemLocationTypeDevice;
            props.location.id = dev;
            ARMSP_HIP(hipMemPoolCreate(&g_pools[dev], &props));
            uint64_t thr = ~0ULL;
            ARMSP_HIP(hipMemPoolSetAttribute(g_pools[dev], hipMemPoolAttrReleaseThreshold, &thr));
            g_pool_ready[dev] = true;
        }
        pool = g_pools[dev];
    }
    ARMSP_HIP(hipMallocFromPoolAsync(p, bytes, pool, s));
    return ARMSP_SUCCESS;
}
}  // namespace

int ws_alloc(void** p, size_t bytes, hipStream_t s) {
    *p = nullptr;
    if (bytes == 0) return ARMSP_SUCCESS;
    if (alloc_injected()) return hip_fail(hipErrorOutOfMemory);
    int dev = 0;
    ARMSP_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return ARMSP_ERROR_NOT_SUPPORTED;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    ARMSP_HIP(hipStreamIsCapturing(s, &cs));
    if (cs != hipStreamCaptureStatusNone) {
        std::lock_guard<std::mutex> g(g_arena_mu);
        const size_t off = (g_arena_used[dev] + 255) & ~(size_t)255;
        if (!g_arena[dev] || off > g_arena_bytes[dev] || bytes > g_arena_bytes[dev] - off) return ARMSP_ERROR_ALLOC;
        g_arena_marks[dev].push_back({off, off + bytes});
        g_arena_used[dev] = off + bytes;
        *p = g_arena[dev] + off;
        return ARMSP_SUCCESS;
    }
    {
        std::lock_guard<std::mutex> g(g_arena_mu);
        if (bytes > g_ws_high[dev]) g_ws_high[dev] = bytes;
    }
    WsEntry* ep = ws_entry(dev, s);
    if (!ep) return pool_alloc(p, bytes, dev, s);
    WsEntry& e = *ep;
    WsSlot* fit = nullptr;
    WsSlot* idle = nullptr;
    for (auto& sl : e.slot) {
        if (sl.busy) continue;
        if (sl.bytes >= bytes && (!fit || sl.bytes < fit->bytes)) fit = &sl;
        if (!idle || sl.bytes < idle->bytes) idle = &sl;
    }
    if (fit) {
        fit->busy = true;
        *p = fit->p;
        return ARMSP_SUCCESS;
    }
    if (!idle) return pool_alloc(p, bytes, dev, s);  // both slots held: an uncached buffer
    const size_t want = (bytes + ((size_t)1 << 20) - 1) & ~(((size_t)1 << 20) - 1);  // whole MiB
    void* q = nullptr;
    const int st = pool_alloc(&q, want, dev, s);
    if (st) return st;
    if (idle->p) hipFreeAsync(idle->p, s);
    idle->p = q;
    idle->bytes = want;
    idle->busy = true;
    *p = q;
    return ARMSP_SUCCESS;
}

void ws_free(void* p, hipStream_t s) {
    if (!p) return;
    int dev = 0;
    if (hipGetDevice(&dev) == hipSuccess && dev >= 0 && dev < 64) {
        std::lock_guard<std::mutex> g(g_arena_mu);
        char* c = static_cast<char*>(p);
        if (g_arena[dev] && c >= g_arena[dev] && c < g_arena[dev] + g_arena_bytes[dev]) {
            auto& marks = g_arena_marks[dev];
            const size_t off = (size_t)(c - g_arena[dev]);
            for (size_t i = marks.size(); i-- > 0;)
                if (marks[i].first == off) {
                    marks.erase(marks.begin() + (std::ptrdiff_t)i);
                    break;
                }
            // space above the newest live block is free again (a hole below it waits for it)
            g_arena_used[dev] = marks.empty() ? 0 : marks.back().second;
            return;
        }
        for (auto& e : t_ws.entries)  // a cached slot goes back to its entry, whatever stream the caller names
            for (auto& sl : e.slot)
                if (sl.busy && sl.p == p) {
                    sl.busy = false;
                    return;
                }
    }
    hipFreeAsync(p, s);
}

int workspace_reserve(size_t bytes, size_t* reserved) {
    if (alloc_injected()) return hip_fail(hipErrorOutOfMemory);
    int dev = 0;
    ARMSP_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return ARMSP_ERROR_NOT_SUPPORTED;
    std::lock_guard<std::mutex> g(g_arena_mu);
    // room for two workspaces at once, each as large as any seen so far
    size_t want = bytes > 2 * g_ws_high[dev] + 256 ? bytes : 2 * g_ws_high[dev] + 256;
    want = (want + 255) & ~(size_t)255;
    if (want > g_arena_bytes[dev]) {
        if (!g_arena_marks[dev].empty()) return ARMSP_ERROR_NOT_SUPPORTED;  // a capture holds the arena
        void* a = nullptr;
        ARMSP_HIP(hipMalloc(&a, want));
        if (g_arena[dev]) g_retired.push_back(g_arena[dev]);
        g_arena[dev] = static_cast<char*>(a);
        g_arena_bytes[dev] = want;
        g_arena_used[dev] = 0;
    }
    if (reserved) *reserved = g_arena_bytes[dev];
    return ARMSP_SUCCESS;
}

namespace {
thread_local int g_timing_paused = 0;
}

bool timing_on() { return g_timing && !g_timing_paused; }

TimingPause::TimingPause() { ++g_timing_paused; }
TimingPause::~TimingPause() { --g_timing_paused; }

TimedLaunch::TimedLaunch(const char* family, hipStream_t s) {
    if (!timing_on()) return;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;  // no timing events inside a captured graph
    if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return;
    TimingRecord* r = new TimingRecord{family, nullptr, nullptr};
    if (hipEventCreate(&r->start) != hipSuccess) {
        delete r;
        return;
    }
    hipEventRecord(r->start, s);
    rec = r;
    stream_ = s;
}

void TimedLaunch::stop() {
    if (!rec) return;
    TimingRecord* r = static_cast<TimingRecord*>(rec);
    if (hipEventCreate(&r->stop) == hipSuccess) hipEventRecord(r->stop, stream_);
    std::lock_guard<std::mutex> g(g_timing_mu);
    g_records.push_back(r);
    rec = nullptr;
}

// compute units of the current device, looked up once per device
int device_cus() {
    static std::atomic<int> cached[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    int cus = cached[dev].load(std::memory_order_relaxed);
    if (cus <= 0) {
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        cached[dev].store(cus, std::memory_order_relaxed);
    }
    return cus;
}

}  // namespace armsp

extern "C" {

int armsp_version(void) { return ARMSP_VERSION; }

const char* armsp_status_string(int status) {
    switch (status) {
        case ARMSP_SUCCESS: return "success";
        case ARMSP_ERROR_INVALID_VALUE: return "invalid value";
        case ARMSP_ERROR_NOT_SUPPORTED: return "not supported";
        case ARMSP_ERROR_HIP: return "HIP runtime error";
        case ARMSP_ERROR_ALLOC: return "device allocation failed";
        case ARMSP_ERROR_NO_DEVICE: return "no usable gfx950 device";
        default: return "unknown status";
    }
}

int armsp_last_hip_error(void) { return armsp::g_last_hip_error; }

int armsp_device_available(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return 0;
    hipDeviceProp_t prop;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return 0;
    return std::string(prop.gcnArchName).rfind("gfx950", 0) == 0 ? 1 : 0;
}

int armsp_workspace_reserve(size_t bytes, size_t* reserved) { return armsp::workspace_reserve(bytes, reserved); }

int armsp_timing_enable(int on) {
    std::lock_guard<std::mutex> g(armsp::g_timing_mu);
    armsp::g_timing = on != 0;
    return ARMSP_SUCCESS;
}

int armsp_timing_reset(void) {
    std::lock_guard<std::mutex> g(armsp::g_timing_mu);
    armsp::resolve_records_locked();
    armsp::g_totals.clear();
    return ARMSP_SUCCESS;
}

int armsp_timing_read(const char* family, double* total_ms, int64_t* launches) {
    if (!family || !total_ms || !launches) return ARMSP_ERROR_INVALID_VALUE;
    std::lock_guard<std::mutex> g(armsp::g_timing_mu);
    armsp::resolve_records_locked();
    *total_ms = 0.0;
    *launches = 0;
    for (auto& t : armsp::g_totals)
        if (t.family == family) {
            *total_ms = t.ms;
            *launches = t.launches;
        }
    return ARMSP_SUCCESS;
}

}  // extern "C"
