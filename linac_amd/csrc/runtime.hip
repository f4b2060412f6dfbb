// Per-device context, error reporting, scratch memory and host utilities.
#include <stdarg.h>
#include <stdlib.h>
#include <stdio.h>
#include <string.h>
#include <unistd.h>

#include <array>
#include <vector>

#include "runtime.hpp"

namespace linac {

static thread_local char g_error[2048] = "";
static pid_t g_owner_pid = 0;  // process that first touched the HIP runtime
static std::mutex g_ctx_mu;
static std::array<DeviceContext*, 64> g_ctx{};

int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof(g_error), fmt, ap);
    va_end(ap);
    return code;
}

// A HIP runtime initialised in a parent cannot be used from a fork()ed child
// (the driver state is not inherited).  Detect that before any HIP call and
// fail cleanly, the analogue of CUDA's "initialization error" in a forked
// child (tests/test_cuda_concurrency.py:18-27 of the reference).
static int check_process() {
    const pid_t me = getpid();
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    if (g_owner_pid == 0) {
        g_owner_pid = me;
        return LINAC_OK;
    }
    if (g_owner_pid != me)
        return fail(LINAC_ERR_FORKED,
                    "initialization error: the HIP runtime was initialised in process %d and this "
                    "process (%d) is a fork of it; start workers with the 'spawn' method",
                    (int)g_owner_pid, (int)me);
    return LINAC_OK;
}

int DeviceContext::get(ScratchSlot slot, size_t bytes, void** out) {
    Scratch& s = scratch[slot];
    if (bytes == 0) bytes = 256;
    if (s.bytes < bytes) {
        if (s.ptr) {
            LINAC_HIP(hipStreamSynchronize(stream));
            LINAC_HIP(hipFree(s.ptr));
            s.ptr = nullptr;
            s.bytes = 0;
        }
        const size_t want = (size_t)round_up((int64_t)bytes, (int64_t)2 << 20);
        hipError_t e = hipMalloc(&s.ptr, want);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            s.ptr = nullptr;
            return fail(LINAC_ERR_NOMEM, "device allocation of %zu bytes failed: %s", want,
                        hipGetErrorString(e));
        }
        s.bytes = want;
    }
    *out = s.ptr;
    return LINAC_OK;
}

int DeviceContext::release() {
    LINAC_HIP(hipStreamSynchronize(stream));
    for (auto& s : scratch) {
        if (s.ptr) LINAC_HIP(hipFree(s.ptr));
        s.ptr = nullptr;
        s.bytes = 0;
    }
    return LINAC_OK;
}

static std::vector<ShutdownHook>& shutdown_hooks() {
    static std::vector<ShutdownHook> hooks;
    return hooks;
}

void add_shutdown_hook(ShutdownHook fn) {
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    for (ShutdownHook h : shutdown_hooks())
        if (h == fn) return;
    shutdown_hooks().push_back(fn);
}

// Synchronises and frees everything the library holds on every device:
// cached engines and batches (hooks), scratch, the pinned page, the probe
// event and the stream.  Contexts are created again on the next call.
static int release_all(bool from_exit) {
    if (g_owner_pid != getpid()) return LINAC_OK;  // a forked child owns nothing
    std::vector<ShutdownHook> hooks;
    {
        std::lock_guard<std::mutex> lk(g_ctx_mu);
        hooks = shutdown_hooks();
    }
    bool any = false;
    {
        std::lock_guard<std::mutex> lk(g_ctx_mu);
        for (DeviceContext* c : g_ctx) any = any || c != nullptr;
    }
    if (!any) {
        if (!from_exit) thread_profiler().clear();
        return LINAC_OK;  // nothing to release (and no need to touch the runtime)
    }
    int rc = LINAC_OK;
    int cur = -1;
    const bool have_cur = hipGetDevice(&cur) == hipSuccess;
    for (size_t d = 0; d < g_ctx.size(); ++d) {
        DeviceContext* c = nullptr;
        {
            std::lock_guard<std::mutex> lk(g_ctx_mu);
            c = g_ctx[d];
            g_ctx[d] = nullptr;
        }
        if (!c) continue;
        {
            std::lock_guard<std::recursive_mutex> lk(c->mu);
            if (hipSetDevice(c->device) != hipSuccess) rc = LINAC_ERR_DEVICE;
            for (ShutdownHook h : hooks) h(c);
            if (c->release() != LINAC_OK) rc = LINAC_ERR_DEVICE;
            if (c->probe) (void)hipEventDestroy(c->probe);
            if (c->pinned) (void)hipHostFree(c->pinned);
            if (c->stream) (void)hipStreamDestroy(c->stream);
        }
        delete c;
    }
    // (at exit the calling thread's thread-local profiler is already gone)
    if (!from_exit) thread_profiler().clear();
    if (have_cur) (void)hipSetDevice(cur);
    return rc;
}

static void release_all_at_exit() { (void)release_all(true); }

int current_context(DeviceContext** out) {
    LINAC_TRY(check_process());
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(LINAC_ERR_DEVICE, "no usable GPU: %s", hipGetErrorString(e));
    }
    if (dev < 0 || dev >= (int)g_ctx.size()) return fail(LINAC_ERR_DEVICE, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    if (!g_ctx[dev]) {
        auto* c = new DeviceContext();
        c->device = dev;
        hipError_t se = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (se != hipSuccess) {
            delete c;
            return fail(LINAC_ERR_DEVICE, "hipStreamCreate: %s", hipGetErrorString(se));
        }
        void* pin = nullptr;
        se = hipHostMalloc(&pin, 4096, hipHostMallocDefault);
        if (se != hipSuccess) {
            (void)hipStreamDestroy(c->stream);
            delete c;
            return fail(LINAC_ERR_DEVICE, "hipHostMalloc: %s", hipGetErrorString(se));
        }
        c->pinned = (int64_t*)pin;
        se = hipEventCreateWithFlags(&c->probe, hipEventDisableTiming);
        if (se != hipSuccess) {
            (void)hipHostFree(pin);
            (void)hipStreamDestroy(c->stream);
            delete c;
            return fail(LINAC_ERR_DEVICE, "hipEventCreate: %s", hipGetErrorString(se));
        }
        g_ctx[dev] = c;
        // Registered after the HIP runtime is up, so it runs before the
        // runtime's own teardown (and before a profiler's) at process exit.
        static bool hooked = false;
        if (!hooked) {
            hooked = true;
            atexit(release_all_at_exit);
        }
    }
    *out = g_ctx[dev];
    return LINAC_OK;
}

int existing_context(DeviceContext** out) {
    LINAC_TRY(check_process());
    int dev = 0;
    LINAC_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= (int)g_ctx.size()) return fail(LINAC_ERR_DEVICE, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    if (!g_ctx[dev]) return fail(LINAC_ERR_INVALID, "device %d is not initialised: call linac_init(%d) first", dev, dev);
    *out = g_ctx[dev];
    return LINAC_OK;
}

int make_workspace(DeviceContext* ctx, int64_t N, int64_t M, ElimWorkspace* ws) {
    const int64_t cap = std::max<int64_t>(1, std::min(N, M));
    const size_t st_bytes = 256;
    const size_t piv_bytes = (size_t)round_up(cap * 8, 256);
    const size_t used_bytes = (size_t)round_up(N, 256);
    void* base = nullptr;
    LINAC_TRY(ctx->get(SLOT_STATE, st_bytes + 2 * piv_bytes + used_bytes, &base));
    char* b = (char*)base;
    ws->st = (ElimState*)b;
    ws->piv_rows = (int64_t*)(b + st_bytes);
    ws->piv_cols = (int64_t*)(b + st_bytes + piv_bytes);
    ws->used = (uint8_t*)(b + st_bytes + 2 * piv_bytes);
    ws->cap = cap;
    return LINAC_OK;
}

static uint64_t mulmod_host(uint64_t a, uint64_t b, uint64_t m) {
    return (uint64_t)(((unsigned __int128)a * b) % m);
}

static uint64_t powmod_host(uint64_t a, uint64_t e, uint64_t m) {
    uint64_t r = 1 % m;
    a %= m;
    while (e) {
        if (e & 1) r = mulmod_host(r, a, m);
        a = mulmod_host(a, a, m);
        e >>= 1;
    }
    return r;
}

// Deterministic Miller-Rabin: the first 12 prime bases decide every n < 3.3e24.
bool is_prime_u64(uint64_t n) {
    if (n < 2) return false;
    static const uint64_t small[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    for (uint64_t q : small) {
        if (n == q) return true;
        if (n % q == 0) return false;
    }
    uint64_t d = n - 1;
    int s = 0;
    while ((d & 1) == 0) { d >>= 1; ++s; }
    for (uint64_t a : small) {
        uint64_t x = powmod_host(a, d, n);
        if (x == 1 || x == n - 1) continue;
        bool comp = true;
        for (int i = 1; i < s; ++i) {
            x = mulmod_host(x, x, n);
            if (x == n - 1) { comp = false; break; }
        }
        if (comp) return false;
    }
    return true;
}

Profiler& thread_profiler() {
    static thread_local Profiler prof;
    return prof;
}

void Profiler::clear() {
    for (auto& x : samples) {
        (void)hipEventDestroy(x.a);
        (void)hipEventDestroy(x.b);
    }
    samples.clear();
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    t0 = t1 = nullptr;
}

Profiler::~Profiler() {
    if (g_owner_pid == getpid()) clear();
}

int Profiler::begin(hipStream_t s) {
    for (auto& x : samples) {
        (void)hipEventDestroy(x.a);
        (void)hipEventDestroy(x.b);
    }
    samples.clear();
    for (double& r : result) r = 0;
    if (!t0) LINAC_HIP(hipEventCreate(&t0));
    if (!t1) LINAC_HIP(hipEventCreate(&t1));
    LINAC_HIP(hipEventRecord(t0, s));
    return LINAC_OK;
}

int Profiler::sample_begin(hipStream_t s, int64_t tag, double bytes, double macs) {
    Sample x{nullptr, nullptr, bytes, macs, tag};
    LINAC_HIP(hipEventCreate(&x.a));
    LINAC_HIP(hipEventCreate(&x.b));
    LINAC_HIP(hipEventRecord(x.a, s));
    samples.push_back(x);
    return LINAC_OK;
}

int Profiler::sample_end(hipStream_t s) {
    LINAC_HIP(hipEventRecord(samples.back().b, s));
    return LINAC_OK;
}

int Profiler::finish(hipStream_t s) {
    LINAC_HIP(hipEventRecord(t1, s));
    LINAC_HIP(hipEventSynchronize(t1));
    float ms = 0;
    LINAC_HIP(hipEventElapsedTime(&ms, t0, t1));
    result[0] = ms;
    for (auto& x : samples) {
        LINAC_HIP(hipEventElapsedTime(&ms, x.a, x.b));
        result[1] += 1;
        result[2] += ms;
        result[3] += x.bytes;
        result[4] += x.macs;
    }
    return LINAC_OK;
}

Tuning& tuning() {
    static Tuning t;
    return t;
}

}  // namespace linac

using namespace linac;

extern "C" {

const char* linac_last_error(void) { return g_error; }

int linac_version(void) { return 10000; }

int linac_device_count(int* count) {
    if (!count) return fail(LINAC_ERR_INVALID, "count is NULL");
    *count = 0;
    LINAC_TRY(check_process());
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        n = 0;
    }
    *count = n;
    return LINAC_OK;
}

int linac_available(void) {
    int n = 0;
    if (linac_device_count(&n) != LINAC_OK) return 0;
    return n > 0 ? 1 : 0;
}

int linac_init(int device) {
    LINAC_TRY(check_process());
    LINAC_HIP(hipSetDevice(device));
    DeviceContext* ctx = nullptr;
    LINAC_TRY(current_context(&ctx));
    return noalloc_setup(device);
}

int linac_release_cache(void) {
    DeviceContext* ctx = nullptr;
    LINAC_TRY(current_context(&ctx));
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    return ctx->release();
}

int linac_shutdown(void) {
    LINAC_TRY(check_process());
    const int rc = release_all(false);
    return rc == LINAC_OK ? LINAC_OK : fail(rc, "releasing a device context failed");
}

int linac_is_prime(uint64_t n) { return is_prime_u64(n) ? 1 : 0; }

// The per-size-class slot named by key ("gemm_group.m" ...), with its range.
static bool sized_slot(Tuning& t, const char* key, int64_t** slot, int64_t* lo, int64_t* hi) {
    static const struct {
        const char* name;
        Tuning::Sized k;
        int64_t lo, hi;
    } sized[] = {{"gemm_group", Tuning::GEMM_GROUP, 1, 64},
                 {"inner_split", Tuning::INNER_SPLIT, 1, 32},
                 {"tile64", Tuning::TILE64, 1, 2},
                 {"fpanel_rows", Tuning::FPANEL_ROWS, 16, 32}};
    const char* dot = strchr(key, '.');
    if (!dot || strlen(dot) != 2) return false;
    const int c = dot[1] == 's' ? 0 : dot[1] == 'm' ? 1 : dot[1] == 'l' ? 2 : -1;
    if (c < 0) return false;
    for (auto& x : sized)
        if (strlen(x.name) == (size_t)(dot - key) && !strncmp(key, x.name, dot - key)) {
            *slot = &t.by_class[x.k][c];
            *lo = x.lo;
            *hi = x.hi;
            return true;
        }
    return false;
}

int linac_tune_set(const char* key, int64_t value) {
    if (!key) return fail(LINAC_ERR_INVALID, "tuning key is NULL");
    Tuning& t = tuning();
    struct Knob {
        const char* name;
        int64_t* slot;
        int64_t lo, hi;
    } knobs[] = {{"gf_blocked_min_rows", &t.gf_blocked_min_rows, 1, INT64_MAX},
                 {"gf_blocked_min_cols", &t.gf_blocked_min_cols, 1, INT64_MAX},
                 {"float_blocked_min_rows", &t.float_blocked_min_rows, 1, INT64_MAX},
                 {"float_blocked_min_cols", &t.float_blocked_min_cols, 1, INT64_MAX},
                 {"gemm_group", &t.gemm_group, 1, 64},
                 {"lookahead", &t.lookahead, 0, 1},
                 {"timeslice", &t.timeslice, 0, 1},
                 {"gemm_lds", &t.gemm_lds, 0, 1},
                 {"fast_panel", &t.fast_panel, 0, 2},
                 {"echelon", &t.echelon, 0, 1},
                 {"inner_compact", &t.inner_compact, 0, 1},
                 {"inner_threads", &t.inner_threads, 0, 1024},
                 {"inner_split", &t.inner_split, 0, 32},
                 {"tile64", &t.tile64, 1, 2},
                 {"u_passes", &t.u_passes, 0, 8},
                 {"fpanel_rows", &t.fpanel_rows, 16, 32}};
    for (auto& k : knobs)
        if (!strcmp(key, k.name)) {
            if (value < k.lo || value > k.hi)
                return fail(LINAC_ERR_INVALID, "%s = %lld outside [%lld, %lld]", key, (long long)value,
                            (long long)k.lo, (long long)k.hi);
            if (!strcmp(key, "inner_threads") && value != 0 && value != 256 && value != 512 && value != 1024)
                return fail(LINAC_ERR_INVALID, "inner_threads = %lld: 0, 256, 512 or 1024", (long long)value);
            if (!strcmp(key, "u_passes") && (value & (value - 1)) != 0)
                return fail(LINAC_ERR_INVALID, "u_passes = %lld: 0, 1, 2, 4 or 8", (long long)value);
            *k.slot = value;
            return LINAC_OK;
        }
    // per size class: "<knob>.s" / ".m" / ".l", 0 = the global value
    int64_t* slot = nullptr;
    int64_t lo = 0, hi = 0;
    if (sized_slot(t, key, &slot, &lo, &hi)) {
        if (value != 0 && (value < lo || value > hi))
            return fail(LINAC_ERR_INVALID, "%s = %lld outside {0} and [%lld, %lld]", key, (long long)value,
                        (long long)lo, (long long)hi);
        *slot = value;
        return LINAC_OK;
    }
    return fail(LINAC_ERR_INVALID, "unknown tuning key '%s'", key);
}

int linac_tune_get(const char* key, int64_t* value) {
    if (!key || !value) return fail(LINAC_ERR_INVALID, "NULL argument");
    const Tuning& t = tuning();
    const struct {
        const char* name;
        int64_t v;
    } knobs[] = {{"gf_blocked_min_rows", t.gf_blocked_min_rows}, {"gf_blocked_min_cols", t.gf_blocked_min_cols},
                 {"float_blocked_min_rows", t.float_blocked_min_rows},
                 {"float_blocked_min_cols", t.float_blocked_min_cols}, {"gemm_group", t.gemm_group},
                 {"lookahead", t.lookahead}, {"timeslice", t.timeslice}, {"gemm_lds", t.gemm_lds},
                 {"fast_panel", t.fast_panel}, {"echelon", t.echelon},
                 {"inner_compact", t.inner_compact}, {"inner_threads", t.inner_threads},
                 {"inner_split", t.inner_split},
                 {"tile64", t.tile64}, {"u_passes", t.u_passes}, {"fpanel_rows", t.fpanel_rows}};
    for (auto& k : knobs)
        if (!strcmp(key, k.name)) {
            *value = k.v;
            return LINAC_OK;
        }
    int64_t* slot = nullptr;
    int64_t lo = 0, hi = 0;
    if (sized_slot(tuning(), key, &slot, &lo, &hi)) {
        *value = *slot;
        return LINAC_OK;
    }
    return fail(LINAC_ERR_INVALID, "unknown tuning key '%s'", key);
}

int linac_device_name(char* out, int n) {
    if (!out || n <= 0) return fail(LINAC_ERR_INVALID, "bad output buffer");
    LINAC_TRY(check_process());
    int dev = 0;
    LINAC_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    LINAC_HIP(hipGetDeviceProperties(&prop, dev));
    snprintf(out, (size_t)n, "%s: %s", prop.gcnArchName, prop.name);
    return LINAC_OK;
}

int linac_profile_last(double* out, int n) {
    if (!out || n <= 0) return 0;
    const int m = n < 6 ? n : 6;
    for (int i = 0; i < m; ++i) out[i] = thread_profiler().result[i];
    return m;
}

}  // extern "C"
