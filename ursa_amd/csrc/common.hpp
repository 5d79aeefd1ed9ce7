// common.hpp — shared host-side plumbing of the gfx950 backend: error
// propagation into C status codes, per-thread streams, device buffers,
// host/device argument staging and optional per-kernel event timing.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <new>
#include <stdexcept>
#include <string>
#include <utility>

#include "../../include/ursa_hip.h"
#include "../../include/ursa_hip_ext.h"

namespace ursa {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void fail(int code, const std::string& msg);
[[noreturn]] void hip_fail(hipError_t e, const char* expr, const char* file, int line);

#define HIP_OK(expr)                                                              \
    do {                                                                          \
        hipError_t e_ = (expr);                                                   \
        if (e_ != hipSuccess) ::ursa::hip_fail(e_, #expr, __FILE__, __LINE__);    \
    } while (0)

#define URSA_REQUIRE(cond, code, msg)          \
    do {                                       \
        if (!(cond)) ::ursa::fail(code, msg);  \
    } while (0)

void set_last_error(const char* msg);

void sync_kept_blocks();

// Run f, translating exceptions into a status code and a thread-local message.
template <class F>
int guard(F&& f) {
    try {
        f();
        sync_kept_blocks();
        return URSA_HIP_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return URSA_HIP_ENOMEM;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return URSA_HIP_EINTERNAL;
    } catch (...) {
        set_last_error("unknown internal error");
        return URSA_HIP_EINTERNAL;
    }
}

// Device selected for this process (ursa_hip_init, default 0); makes it
// current on the calling thread.
int ensure_device();
// The calling thread's stream (non-blocking, default priority).
hipStream_t stream();
// Work beside `main`: fork_side returns the thread's side stream, ordered after
// everything enqueued on `main` so far; join_side orders `main` after
// everything enqueued on the side stream since. Memory the side work uses must
// be allocated before the fork and freed after the join.
void fork_side(hipStream_t main, hipStream_t* side);
void join_side(hipStream_t main);

void* dev_alloc(size_t bytes);
void dev_free(void* p);
// Synchronizes the calling thread's stream if a block was kept for reuse since
// the last call (see dev_free); guard calls it before an entry point returns.

// Workspace arena. While a ScopedArena is active on a thread, every dev_alloc
// on that thread is carved from the arena's range (256-byte aligned bump
// allocation, no HIP call) and dev_free of such a pointer does nothing; the
// range is reclaimed as a whole by its owner. The request server uses this to
// run queries inside the HBM reservation they were admitted with, so that the
// query path never calls hipMalloc/hipFree (hipFree waits for the device).
// Exceeding the range throws URSA_HIP_ENOMEM.
// Frees in reverse allocation order (the order DevBuf destructors run in)
// give their bytes back at once; other frees wait for the owner's reset.
struct Arena {
    char* base = nullptr;
    size_t cap = 0;
    size_t used = 0;
    size_t peak = 0;
    size_t stack[64];
    int depth = 0;
    void reset(char* b, size_t c) {
        base = b;
        cap = c;
        used = peak = 0;
        depth = 0;
    }
};
Arena* arena_current();
class ScopedArena {
   public:
    explicit ScopedArena(Arena* a);
    ~ScopedArena();
    ScopedArena(const ScopedArena&) = delete;
    ScopedArena& operator=(const ScopedArena&) = delete;

   private:
    Arena* prev_;
};
// Suspends the thread's arena: allocations that outlive the current call
// (topology caches built lazily) must come from hipMalloc.
class NoArena {
   public:
    NoArena();
    ~NoArena();
    NoArena(const NoArena&) = delete;
    NoArena& operator=(const NoArena&) = delete;

   private:
    Arena* prev_;
};
// Registers [base, base+bytes) as arena memory, so dev_free ignores pointers
// into it on any thread.
void arena_register(const void* base, size_t bytes);
void arena_unregister(const void* base);

// Per-thread pinned host scratch of at least `bytes` (kept for the life of
// the thread; grows by reallocation).
void* pinned_scratch(size_t bytes);

template <class T>
class DevBuf {
   public:
    DevBuf() = default;
    explicit DevBuf(size_t n) { alloc(n); }
    ~DevBuf() { reset(); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            reset();
            p_ = o.p_; n_ = o.n_;
            o.p_ = nullptr; o.n_ = 0;
        }
        return *this;
    }
    void alloc(size_t n) {
        reset();
        p_ = static_cast<T*>(dev_alloc((n ? n : 1) * sizeof(T)));
        n_ = n;
    }
    // Grow-only (contents not preserved).
    void ensure(size_t n) { if (!p_ || n_ < n) alloc(n); }
    void reset() {
        if (p_) dev_free(p_);
        p_ = nullptr;
        n_ = 0;
    }
    T* get() const { return p_; }
    size_t size() const { return n_; }
    size_t bytes() const { return n_ * sizeof(T); }
    explicit operator bool() const { return p_ != nullptr; }

   private:
    T* p_ = nullptr;
    size_t n_ = 0;
};

bool is_device_ptr(const void* p);

// Copies `bytes` from device memory to host memory at `dst` and synchronises
// stream s. Pageable destinations above a few hundred KB go through the
// calling thread's two pinned staging buffers: the DMA of one chunk overlaps
// the host copy of the previous one, and no process-wide staging path is
// shared with other threads' copies.
void copy_to_host(void* dst, const void* src, size_t bytes, hipStream_t s);

// A read-only argument: device pointers are used in place, host buffers
// are uploaded to a temporary device copy.
template <class T>
struct DevIn {
    const T* p = nullptr;
    DevBuf<T> own;
    DevIn(const T* user, size_t n, hipStream_t s) {
        if (!user || n == 0) { p = nullptr; return; }
        if (is_device_ptr(user)) { p = user; return; }
        own.alloc(n);
        HIP_OK(hipMemcpyAsync(own.get(), user, n * sizeof(T), hipMemcpyHostToDevice, s));
        p = own.get();
    }
};

// An output argument: device pointers are written in place; host buffers get
// a device staging copy that finish() copies back (and synchronises).
template <class T>
struct DevOut {
    T* user = nullptr;
    T* p = nullptr;
    size_t n = 0;
    bool dev = false;
    DevBuf<T> own;
    DevOut(T* u, size_t count) : user(u), n(count) {
        if (!user || n == 0) { p = nullptr; return; }
        dev = is_device_ptr(user);
        if (dev) { p = user; return; }
        own.alloc(n);
        p = own.get();
    }
    void finish(hipStream_t s) {
        if (user && n && !dev)
            copy_to_host(user, own.get(), n * sizeof(T), s);
        else
            HIP_OK(hipStreamSynchronize(s));
    }
};

// ---- per-kernel timing (ursa_hip_profile_*) ----
bool prof_enabled();
void prof_push(const char* name, hipEvent_t a, hipEvent_t b, double bytes, hipStream_t s);

// Launch through `launch()`; when profiling is on, bracket it with events on
// the launch stream and book `bytes` algorithmic bytes against `name`.
template <class L>
inline void timed(const char* name, double bytes, hipStream_t s, L&& launch) {
    if (!prof_enabled()) {
        launch();
        return;
    }
    hipEvent_t a, b;
    HIP_OK(hipEventCreate(&a));
    HIP_OK(hipEventCreate(&b));
    HIP_OK(hipEventRecord(a, s));
    launch();
    HIP_OK(hipEventRecord(b, s));
    prof_push(name, a, b, bytes, s);
}

// HIP graph replay of fixed launch sequences (ursa_hip_graphs_enable).
bool graphs_enabled();

// Records the launches `record()` makes on stream s the first time run() is
// called (stream capture, thread-local mode) and replays the instantiated
// graph on every call. Falls back to launching directly while graphs are
// disabled or kernel timing is on. The recorded launches must not allocate,
// synchronise or clear memory with hipMemsetAsync.
class CapturedLaunches {
   public:
    CapturedLaunches() = default;
    CapturedLaunches(const CapturedLaunches&) = delete;
    CapturedLaunches& operator=(const CapturedLaunches&) = delete;
    ~CapturedLaunches() {
        if (exec_) (void)hipGraphExecDestroy(exec_);
        if (graph_) (void)hipGraphDestroy(graph_);
    }
    template <class F>
    void run(hipStream_t s, F&& record) {
        if (!graphs_enabled() || prof_enabled()) {
            record();
            return;
        }
        if (!exec_) {
            HIP_OK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            try {
                record();
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(s, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            HIP_OK(hipStreamEndCapture(s, &graph_));
            HIP_OK(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
        }
        HIP_OK(hipGraphLaunch(exec_, s));
    }

   private:
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
};

inline unsigned blocks_for(uint64_t n, unsigned block, unsigned cap = 1u << 20) {
    uint64_t b = (n + block - 1) / block;
    if (b == 0) b = 1;
    if (b > cap) b = cap;
    return (unsigned)b;
}

inline int bits_for(uint64_t n) {  // bits needed to represent values < n
    int b = 0;
    while (b < 64 && (1ull << b) < n) ++b;
    return b == 0 ? 1 : b;
}

inline void check_launch() { HIP_OK(hipGetLastError()); }

}  // namespace ursa
