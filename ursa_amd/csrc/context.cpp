// context.cpp — device selection, per-thread streams, error strings,
// allocation and the kernel-timing registry behind ursa_hip_profile_*.
#include <atomic>
#include <cstring>
#include <map>
#include <mutex>
#include <unordered_map>
#include <string>
#include <vector>

#include "common.hpp"

namespace ursa {

void cc_noalloc_setup();

namespace {
thread_local std::string t_last_error;
std::atomic<int> g_device{-1};
std::mutex g_init_mu;

struct ThreadStream {
    hipStream_t s = nullptr;
    int device = -1;
    // Streams are intentionally not destroyed at thread exit: the runtime may
    // already be torn down when thread-local destructors run at process exit.
};
thread_local ThreadStream t_stream;
// A second stream per thread for work that runs beside the main stream's
// (joined to it by events before the call returns); default priority.
struct SideStream {
    hipStream_t s = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    int device = -1;
};
thread_local SideStream t_side;
thread_local int t_current_device = -1;
thread_local hipStream_t t_user_stream = nullptr;
thread_local Arena* t_arena = nullptr;

struct PinnedScratch {
    void* p = nullptr;
    size_t bytes = 0;
    // Not freed at thread exit, for the same reason as the streams.
};
thread_local PinnedScratch t_pinned;

std::mutex g_arena_mu;
std::vector<std::pair<uintptr_t, uintptr_t>> g_arenas;

bool in_arena(const void* p) {
    const uintptr_t a = (uintptr_t)p;
    std::lock_guard<std::mutex> lk(g_arena_mu);
    for (auto& r : g_arenas)
        if (a >= r.first && a < r.second) return true;
    return false;
}

struct ProfRec {
    std::string name;
    hipEvent_t a, b;
    double bytes;
};
struct ProfAgg {
    double ms = 0.0;
    uint64_t launches = 0;
    double bytes = 0.0;
};
std::atomic<bool> g_prof_on{false};
std::atomic<bool> g_graphs_on{true};
std::mutex g_prof_mu;
std::vector<ProfRec> g_prof_pending;
std::map<std::string, ProfAgg> g_prof_agg;

void prof_drain_locked() {
    for (auto& r : g_prof_pending) {
        float ms = 0.f;
        if (hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            auto& g = g_prof_agg[r.name];
            g.ms += ms;
            g.launches += 1;
            g.bytes += r.bytes;
        }
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    g_prof_pending.clear();
}
}  // namespace

void fail(int code, const std::string& msg) { throw Error(code, msg); }

void hip_fail(hipError_t e, const char* expr, const char* file, int line) {
    int code = (e == hipErrorOutOfMemory) ? URSA_HIP_ENOMEM : URSA_HIP_EDEVICE;
    std::string f = file;
    auto p = f.find_last_of('/');
    if (p != std::string::npos) f = f.substr(p + 1);
    throw Error(code, std::string(hipGetErrorString(e)) + " in " + expr + " (" + f + ":" +
                          std::to_string(line) + ")");
}

void set_last_error(const char* msg) { t_last_error = msg ? msg : ""; }

int ensure_device() {
    int dev = g_device.load();
    if (dev < 0) {
        std::lock_guard<std::mutex> lk(g_init_mu);
        dev = g_device.load();
        if (dev < 0) {
            int count = 0;
            HIP_OK(hipGetDeviceCount(&count));
            URSA_REQUIRE(count > 0, URSA_HIP_EDEVICE, "no HIP device visible");
            dev = 0;
            hipDeviceProp_t prop;
            HIP_OK(hipGetDeviceProperties(&prop, dev));
            URSA_REQUIRE(std::strncmp(prop.gcnArchName, "gfx950", 6) == 0, URSA_HIP_EDEVICE,
                         std::string("this build targets gfx950 (MI355X); device 0 is ") + prop.gcnArchName);
            g_device.store(dev);
        }
    }
    if (t_current_device != dev) {
        HIP_OK(hipSetDevice(dev));
        t_current_device = dev;
    }
    return dev;
}

hipStream_t stream() {
    int dev = ensure_device();
    if (t_user_stream) return t_user_stream;
    if (!t_stream.s || t_stream.device != dev) {
        HIP_OK(hipStreamCreateWithFlags(&t_stream.s, hipStreamNonBlocking));
        t_stream.device = dev;
    }
    return t_stream.s;
}

void fork_side(hipStream_t main, hipStream_t* side) {
    const int dev = ensure_device();
    if (!t_side.s || t_side.device != dev) {
        HIP_OK(hipStreamCreateWithFlags(&t_side.s, hipStreamNonBlocking));
        HIP_OK(hipEventCreateWithFlags(&t_side.fork, hipEventDisableTiming));
        HIP_OK(hipEventCreateWithFlags(&t_side.join, hipEventDisableTiming));
        t_side.device = dev;
    }
    HIP_OK(hipEventRecord(t_side.fork, main));
    HIP_OK(hipStreamWaitEvent(t_side.s, t_side.fork, 0));
    *side = t_side.s;
}

void join_side(hipStream_t main) {
    HIP_OK(hipEventRecord(t_side.join, t_side.s));
    HIP_OK(hipStreamWaitEvent(main, t_side.join, 0));
}

// Device blocks freed outside a workspace arena are kept for reuse instead of
// going back to HIP: hipMalloc and hipFree cost tens to hundreds of
// microseconds each, and hipFree waits for the whole device. A block is kept
// only when it is freed on the stream it was allocated on, and it is handed
// out again only on that stream. Every kernel of the library runs on the
// calling thread's stream, so work enqueued there later is ordered after every
// earlier use of the block. Only the library's own per-thread streams qualify:
// they are never destroyed, so their handles are never reused, whereas a
// caller's stream may be. A call that keeps a block synchronizes its stream
// before returning (sync_kept_blocks), so calls still return with their
// results written, as they did when hipFree waited. A request takes the
// smallest idle block of at least its size and at most twice it (calls on
// graphs of different sizes reuse each other's blocks). Blocks above kKeepMax
// are never kept, the cache holds at most kCacheCap bytes, and a failed
// hipMalloc waits for the device and releases the whole cache before it
// retries.
namespace {
constexpr size_t kKeepMax = 512ull << 20;
constexpr size_t kCacheCap = 4ull << 30;
struct LiveBlock {
    size_t bytes;
    hipStream_t s;
};
std::mutex g_cache_mu;
std::unordered_map<void*, LiveBlock> g_live;                   // blocks in use that may be kept
std::multimap<std::pair<hipStream_t, size_t>, void*> g_kept;  // (stream, bytes) -> idle block
size_t g_kept_bytes = 0;
thread_local bool t_kept_since_sync = false;

size_t keep_round(size_t b) {
    return b < (1u << 20) ? (b + 4095) & ~size_t(4095) : (b + 65535) & ~size_t(65535);
}

void release_kept() {
    std::vector<void*> blocks;
    {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        for (auto& e : g_kept) blocks.push_back(e.second);
        g_kept.clear();
        g_kept_bytes = 0;
    }
    if (!blocks.empty()) {
        (void)hipDeviceSynchronize();
        for (void* p : blocks) (void)hipFree(p);
    }
}

void* hip_alloc_or_throw(size_t bytes) {
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        release_kept();
        e = hipMalloc(&p, bytes);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw Error(URSA_HIP_ENOMEM, "device allocation of " + std::to_string(bytes) + " bytes failed: " +
                                         hipGetErrorString(e));
    }
    return p;
}
}  // namespace

void sync_kept_blocks() {
    if (t_kept_since_sync) {  // blocks are only kept on the thread's own stream
        t_kept_since_sync = false;
        if (t_stream.s) HIP_OK(hipStreamSynchronize(t_stream.s));
    }
}

void* dev_alloc(size_t bytes) {
    ensure_device();
    if (Arena* a = t_arena) {
        const size_t at = (a->used + 255) & ~(size_t)255;
        if (at + bytes > a->cap)
            throw Error(URSA_HIP_ENOMEM, "workspace of " + std::to_string(bytes) +
                                             " bytes exceeds the query's admitted reservation (" +
                                             std::to_string(a->cap) + " bytes, " + std::to_string(at) + " used)");
        if (a->depth < 64) a->stack[a->depth++] = a->used;
        a->used = at + bytes;
        if (a->used > a->peak) a->peak = a->used;
        return a->base + at;
    }
    const size_t r = keep_round(bytes ? bytes : 1);
    hipStream_t s = (t_user_stream || r > kKeepMax) ? nullptr : stream();
    if (s) {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        auto it = g_kept.lower_bound({s, r});
        if (it != g_kept.end() && it->first.first == s && it->first.second <= 2 * r) {
            void* p = it->second;
            const size_t got = it->first.second;
            g_kept.erase(it);
            g_kept_bytes -= got;
            g_live[p] = LiveBlock{got, s};
            return p;
        }
    }
    void* p = hip_alloc_or_throw(s ? r : bytes);
    if (s) {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        g_live[p] = LiveBlock{r, s};
    }
    return p;
}

void dev_free(void* p) {
    if (!p) return;
    if (Arena* a = t_arena) {
        char* c = static_cast<char*>(p);
        if (c >= a->base && c < a->base + a->cap) {
            if (a->depth > 0) {
                const size_t prev = a->stack[a->depth - 1];
                if (c == a->base + ((prev + 255) & ~(size_t)255)) {
                    a->used = prev;
                    --a->depth;
                }
            }
            return;
        }
    }
    if (in_arena(p)) return;
    {
        hipStream_t cur = t_user_stream ? nullptr : t_stream.s;
        std::lock_guard<std::mutex> lk(g_cache_mu);
        auto it = g_live.find(p);
        if (it != g_live.end()) {
            const LiveBlock b = it->second;
            g_live.erase(it);
            if (cur && b.s == cur && g_kept_bytes + b.bytes <= kCacheCap) {
                g_kept.emplace(std::make_pair(b.s, b.bytes), p);
                g_kept_bytes += b.bytes;
                t_kept_since_sync = true;
                return;
            }
        }
    }
    (void)hipFree(p);
}

Arena* arena_current() { return t_arena; }
ScopedArena::ScopedArena(Arena* a) : prev_(t_arena) { t_arena = a; }
ScopedArena::~ScopedArena() { t_arena = prev_; }
NoArena::NoArena() : prev_(t_arena) { t_arena = nullptr; }
NoArena::~NoArena() { t_arena = prev_; }

void arena_register(const void* base, size_t bytes) {
    std::lock_guard<std::mutex> lk(g_arena_mu);
    g_arenas.emplace_back((uintptr_t)base, (uintptr_t)base + bytes);
}

void arena_unregister(const void* base) {
    std::lock_guard<std::mutex> lk(g_arena_mu);
    for (size_t i = 0; i < g_arenas.size(); ++i)
        if (g_arenas[i].first == (uintptr_t)base) {
            g_arenas.erase(g_arenas.begin() + i);
            return;
        }
}

void* pinned_scratch(size_t bytes) {
    ensure_device();
    if (t_pinned.bytes < bytes) {
        if (t_pinned.p) (void)hipHostFree(t_pinned.p);
        t_pinned.p = nullptr;
        t_pinned.bytes = 0;
        size_t want = 4096;
        while (want < bytes) want <<= 1;
        HIP_OK(hipHostMalloc(&t_pinned.p, want, hipHostMallocDefault));
        t_pinned.bytes = want;
    }
    return t_pinned.p;
}

namespace {
constexpr size_t kStageChunk = 4ull << 20;
constexpr size_t kStageDirect = 256ull << 10;
struct Staging {
    char* buf = nullptr;  // two chunks
    hipEvent_t ev[2] = {nullptr, nullptr};
    // Not freed at thread exit, for the same reason as the streams.
};
thread_local Staging t_stage;

bool is_pinned_host(const void* p) {
    hipPointerAttribute_t attr;
    hipError_t e = hipPointerGetAttributes(&attr, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return attr.type == hipMemoryTypeHost;
}
}  // namespace

void copy_to_host(void* dst, const void* src, size_t bytes, hipStream_t s) {
    if (bytes <= kStageDirect || is_pinned_host(dst)) {
        if (bytes) HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        return;
    }
    ensure_device();
    Staging& g = t_stage;
    if (!g.buf) {
        HIP_OK(hipHostMalloc((void**)&g.buf, 2 * kStageChunk, hipHostMallocDefault));
        for (auto& e : g.ev) HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    const char* from = static_cast<const char*>(src);
    char* to = static_cast<char*>(dst);
    const size_t chunks = (bytes + kStageChunk - 1) / kStageChunk;
    auto issue = [&](size_t c) {
        const size_t at = c * kStageChunk, len = std::min(kStageChunk, bytes - at);
        char* st = g.buf + (c & 1) * kStageChunk;
        HIP_OK(hipMemcpyAsync(st, from + at, len, hipMemcpyDeviceToHost, s));
        HIP_OK(hipEventRecord(g.ev[c & 1], s));
    };
    issue(0);
    for (size_t c = 0; c < chunks; ++c) {
        // Buffer (c+1)&1 held chunk c-1, which was copied out below in the
        // previous round, so the next DMA may overwrite it.
        if (c + 1 < chunks) issue(c + 1);
        HIP_OK(hipEventSynchronize(g.ev[c & 1]));
        const size_t at = c * kStageChunk, len = std::min(kStageChunk, bytes - at);
        std::memcpy(to + at, g.buf + (c & 1) * kStageChunk, len);
    }
    HIP_OK(hipStreamSynchronize(s));
}

bool is_device_ptr(const void* p) {
    if (!p) return false;
    hipPointerAttribute_t attr;
    hipError_t e = hipPointerGetAttributes(&attr, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return attr.type == hipMemoryTypeDevice;
}

bool prof_enabled() { return g_prof_on.load(std::memory_order_relaxed); }
bool graphs_enabled() { return g_graphs_on.load(std::memory_order_relaxed); }

void prof_push(const char* name, hipEvent_t a, hipEvent_t b, double bytes, hipStream_t) {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_prof_pending.push_back({name, a, b, bytes});
    if (g_prof_pending.size() > 4096) prof_drain_locked();
}

}  // namespace ursa

using namespace ursa;

extern "C" {

int ursa_hip_init(int device) {
    return guard([&] {
        int count = 0;
        HIP_OK(hipGetDeviceCount(&count));
        URSA_REQUIRE(device >= 0 && device < count, URSA_HIP_EINVAL, "device index out of range");
        hipDeviceProp_t prop;
        HIP_OK(hipGetDeviceProperties(&prop, device));
        URSA_REQUIRE(std::strncmp(prop.gcnArchName, "gfx950", 6) == 0, URSA_HIP_EDEVICE,
                     std::string("this build targets gfx950 (MI355X); device is ") + prop.gcnArchName);
        {
            std::lock_guard<std::mutex> lk(g_init_mu);
            g_device.store(device);
        }
        ensure_device();
        ursa::cc_noalloc_setup();  // device address of the no-allocation scratch
    });
}

int ursa_hip_shutdown(void) {
    return guard([&] {
        if (g_device.load() >= 0) {
            HIP_OK(hipDeviceSynchronize());
            release_kept();
        }
    });
}

const char* ursa_hip_last_error(void) { return t_last_error.c_str(); }

int ursa_hip_set_stream(void* s) {
    return guard([&] {
        ensure_device();
        t_user_stream = static_cast<hipStream_t>(s);
    });
}

int ursa_hip_synchronize(void) {
    return guard([&] { HIP_OK(hipStreamSynchronize(stream())); });
}

int ursa_hip_profile_enable(int on) {
    g_prof_on.store(on != 0);
    return URSA_HIP_OK;
}

int ursa_hip_graphs_enable(int on) {
    g_graphs_on.store(on != 0);
    return URSA_HIP_OK;
}

int ursa_hip_profile_reset(void) {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    prof_drain_locked();
    g_prof_agg.clear();
    return URSA_HIP_OK;
}

int ursa_hip_profile_query(const char* kernel, double* total_ms, uint64_t* launches, double* algo_bytes) {
    return guard([&] {
        URSA_REQUIRE(kernel != nullptr, URSA_HIP_EINVAL, "kernel name is NULL");
        std::lock_guard<std::mutex> lk(g_prof_mu);
        prof_drain_locked();
        auto it = g_prof_agg.find(kernel);
        ProfAgg a = it == g_prof_agg.end() ? ProfAgg{} : it->second;
        if (total_ms) *total_ms = a.ms;
        if (launches) *launches = a.launches;
        if (algo_bytes) *algo_bytes = a.bytes;
    });
}

int ursa_hip_device_alloc(uint64_t bytes, void** out) {
    return guard([&] {
        URSA_REQUIRE(out != nullptr, URSA_HIP_EINVAL, "out is NULL");
        ensure_device();
        // Caller-owned memory may be used on the caller's streams: never kept.
        *out = hip_alloc_or_throw(bytes ? bytes : 1);
    });
}

int ursa_hip_device_free(void* p) {
    return guard([&] {
        ensure_device();
        dev_free(p);
    });
}

int ursa_hip_memcpy_to_host(void* dst, const void* src_device, uint64_t bytes) {
    return guard([&] {
        hipStream_t s = stream();
        HIP_OK(hipMemcpyAsync(dst, src_device, bytes, hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
    });
}

int ursa_hip_memcpy_to_device(void* dst_device, const void* src, uint64_t bytes) {
    return guard([&] {
        hipStream_t s = stream();
        HIP_OK(hipMemcpyAsync(dst_device, src, bytes, hipMemcpyHostToDevice, s));
        HIP_OK(hipStreamSynchronize(s));
    });
}

uint64_t ursa_hip_profile_names(char* buf, uint64_t cap) {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    prof_drain_locked();
    std::string all;
    for (auto& kv : g_prof_agg) {
        all += kv.first;
        all += '\n';
    }
    if (buf && cap) {
        uint64_t k = std::min<uint64_t>(cap - 1, all.size());
        std::memcpy(buf, all.data(), k);
        buf[k] = 0;
    }
    return all.size();
}

}  // extern "C"
