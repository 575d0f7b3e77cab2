// C ABI plumbing: error string, device/memory/stream helpers and the
// per-device FFT plan cache.
#include <atomic>
#include <cstdarg>
#include <cstring>
#include <string>

#include "common.hpp"

namespace wct {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

static bool small_primes_only(int n)
{
    if (n < 1) return false;
    const int primes[] = {2, 3, 5, 7, 11, 13};
    for (int p : primes)
        while (n % p == 0) n /= p;
    return n == 1;
}

bool fft_supported_length(int n) { return small_primes_only(n); }

// Failure-path hooks: one armed entry point (by name) fails with a given
// status on its (after + 1)-th call from now; device allocations above
// the cap fail as out of memory.
static std::mutex g_hook_mutex;
static std::string g_inject_entry;
static int g_inject_after = 0, g_inject_status = 0;
static std::atomic<bool> g_inject_armed{false};
static std::atomic<size_t> g_alloc_cap{0};

int inject_check(const char* entry)
{
    if (!g_inject_armed.load(std::memory_order_relaxed)) return 0;
    std::lock_guard<std::mutex> lk(g_hook_mutex);
    if (!g_inject_armed.load() || g_inject_entry != entry) return 0;
    if (g_inject_after > 0) {
        --g_inject_after;
        return 0;
    }
    g_inject_armed = false;
    set_error("%s: injected failure (status %d)", entry, g_inject_status);
    return g_inject_status;
}

static int alloc_status(hipError_t e, size_t bytes)
{
    if (e == hipSuccess) return WCT_OK;
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        set_error("device allocation of %zu bytes: out of memory", bytes);
        return WCT_ERR_NOMEM;
    }
    set_error("device allocation of %zu bytes failed: %s", bytes, hipGetErrorString(e));
    return WCT_ERR_HIP;
}

static bool above_cap(size_t bytes)
{
    const size_t cap = g_alloc_cap.load(std::memory_order_relaxed);
    if (cap && bytes > cap) {
        set_error("device allocation of %zu bytes above the %zu-byte cap", bytes, cap);
        return true;
    }
    return false;
}

int dev_alloc(void** p, size_t bytes)
{
    *p = nullptr;
    if (above_cap(bytes)) return WCT_ERR_NOMEM;
    return alloc_status(hipMalloc(p, bytes ? bytes : 1), bytes);
}

int dev_alloc_async(void** p, size_t bytes, hipStream_t s)
{
    *p = nullptr;
    if (above_cap(bytes)) return WCT_ERR_NOMEM;
    return alloc_status(hipMallocAsync(p, bytes ? bytes : 1, s), bytes);
}

int lds_limit_bytes()
{
    static int lim = -1;
    if (lim < 0) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
            lim = v;
        else
            lim = 65536;
    }
    return lim;
}

// The radix schedule is sf_radix's (fft_lds.hpp), shared with the
// compile-time plans.
static void factorize(int n, FftPlan& pl, bool odd_first)
{
    int np = 0;
    while (np < kMaxPasses && sf_radix(n, np, odd_first)) {
        pl.rad[np] = sf_radix(n, np, odd_first);
        ++np;
    }
    pl.npass = np;
    pl.n = n;
}

template <typename T>
static void fill_twiddles(const FftPlan& pl, std::vector<T>& tab)
{
    // pass twiddles w^{r k} for k < Ns, r in [1, R), w = exp(-2 pi i/(Ns R))
    int Ns = 1;
    for (int p = 0; p < pl.npass; ++p) {
        const int R = pl.rad[p];
        const long M = (long)Ns * R;
        for (int k = 0; k < Ns; ++k)
            for (int r = 1; r < R; ++r) {
                const long q = ((long)r * k) % M;
                const double ang = -2.0 * M_PI * (double)q / (double)M;
                tab.push_back((T)cos(ang));
                tab.push_back((T)sin(ang));
            }
        Ns *= R;
    }
    // generic small-DFT tables exp(-2 pi i q / R), R = 7, 11, 13
    const int gr[] = {7, 11, 13};
    for (int R : gr)
        for (int q = 0; q < R; ++q) {
            const double ang = -2.0 * M_PI * q / R;
            tab.push_back((T)cos(ang));
            tab.push_back((T)sin(ang));
        }
}

using PlanKey = std::tuple<int, int, bool, bool>;   // device, n, double, odd-first schedule
static std::mutex g_plan_mutex;
static std::map<PlanKey, DevPlan>* g_plans = new std::map<PlanKey, DevPlan>();
using LineKey = std::tuple<int, int, int, bool>;
static std::map<LineKey, void*>* g_lines = new std::map<LineKey, void*>();

int get_plan(int n, bool is_double, const DevPlan** out, bool odd_first)
{
    WCT_REQUIRE(fft_supported_length(n), WCT_ERR_UNSUPPORTED,
                "FFT length %d has a prime factor above 13", n);
    int dev = 0;
    WCT_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_plan_mutex);
    auto key = PlanKey(dev, n, is_double, odd_first);
    auto it = g_plans->find(key);
    if (it != g_plans->end()) {
        *out = &it->second;
        return WCT_OK;
    }
    DevPlan dp;
    memset(&dp.plan, 0, sizeof(dp.plan));
    factorize(n, dp.plan, odd_first);
    int Ns = 1;
    size_t off = 0;
    for (int p = 0; p < dp.plan.npass; ++p) {
        dp.plan.twoff[p] = (int)off;
        off += (size_t)Ns * (dp.plan.rad[p] - 1);
        Ns *= dp.plan.rad[p];
    }
    dp.plan.dftoff = (int)off;
    dp.plan.twcount = (int)off + kGenericTabSize;
    if (is_double) {
        std::vector<double> tab;
        fill_twiddles<double>(dp.plan, tab);
        dp.tw_count = tab.size() / 2;
        WCT_TRY(dev_alloc(&dp.tw, tab.size() * sizeof(double)));
        WCT_HIP_CHECK(hipMemcpy(dp.tw, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
    } else {
        std::vector<float> tab;
        fill_twiddles<float>(dp.plan, tab);
        dp.tw_count = tab.size() / 2;
        WCT_TRY(dev_alloc(&dp.tw, tab.size() * sizeof(float)));
        WCT_HIP_CHECK(hipMemcpy(dp.tw, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    auto res = g_plans->emplace(key, dp);
    *out = &res.first->second;
    return WCT_OK;
}

int get_twiddle_line(int N, int count, bool is_double, const void** out)
{
    int dev = 0;
    WCT_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_plan_mutex);
    auto key = LineKey(dev, N, count, is_double);
    auto it = g_lines->find(key);
    if (it != g_lines->end()) {
        *out = it->second;
        return WCT_OK;
    }
    void* d = nullptr;
    if (is_double) {
        std::vector<double> tab(2 * (size_t)count);
        for (int k = 0; k < count; ++k) {
            const double ang = -2.0 * M_PI * (double)(k % N) / N;
            tab[2 * k] = cos(ang);
            tab[2 * k + 1] = sin(ang);
        }
        WCT_TRY(dev_alloc(&d, tab.size() * sizeof(double)));
        WCT_HIP_CHECK(hipMemcpy(d, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
    } else {
        std::vector<float> tab(2 * (size_t)count);
        for (int k = 0; k < count; ++k) {
            const double ang = -2.0 * M_PI * (double)(k % N) / N;
            tab[2 * k] = (float)cos(ang);
            tab[2 * k + 1] = (float)sin(ang);
        }
        WCT_TRY(dev_alloc(&d, tab.size() * sizeof(float)));
        WCT_HIP_CHECK(hipMemcpy(d, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    (*g_lines)[key] = d;
    *out = d;
    return WCT_OK;
}

int find_plan(int n, bool is_double, const DevPlan** out)
{
    int dev = 0;
    WCT_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_plan_mutex);
    auto it = g_plans->find(PlanKey(dev, n, is_double, false));
    WCT_REQUIRE(it != g_plans->end(), WCT_ERR_STATE, "no transform tables for length %d on device %d", n, dev);
    *out = &it->second;
    return WCT_OK;
}

int find_twiddle_line(int N, int count, bool is_double, const void** out)
{
    int dev = 0;
    WCT_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_plan_mutex);
    auto it = g_lines->find(LineKey(dev, N, count, is_double));
    WCT_REQUIRE(it != g_lines->end(), WCT_ERR_STATE, "no twiddle line for length %d on device %d", N, dev);
    *out = it->second;
    return WCT_OK;
}

}  // namespace wct

using namespace wct;

extern "C" {

const char* wct_last_error(void) { return g_err; }
int wct_abi_version(void) { return 5; }

int wct_device_count(int* n)
{
    WCT_REQUIRE(n, WCT_ERR_ARG, "null pointer");
    WCT_HIP_CHECK(hipGetDeviceCount(n));
    return WCT_OK;
}

int wct_set_device(int device)
{
    WCT_HIP_CHECK(hipSetDevice(device));
    return WCT_OK;
}

int wct_malloc(void** dptr, size_t bytes)
{
    WCT_ENTRY();
    WCT_REQUIRE(dptr, WCT_ERR_ARG, "null pointer");
    return dev_alloc(dptr, bytes);
}

int wct_free(void* dptr)
{
    if (dptr) WCT_HIP_CHECK(hipFree(dptr));
    return WCT_OK;
}

int wct_memcpy_htod(void* dst, const void* src, size_t bytes, void* stream)
{
    WCT_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, as_stream(stream)));
    WCT_HIP_CHECK(hipStreamSynchronize(as_stream(stream)));
    return WCT_OK;
}

int wct_memcpy_dtoh(void* dst, const void* src, size_t bytes, void* stream)
{
    WCT_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, as_stream(stream)));
    WCT_HIP_CHECK(hipStreamSynchronize(as_stream(stream)));
    return WCT_OK;
}

int wct_memcpy_dtod(void* dst, const void* src, size_t bytes, void* stream)
{
    WCT_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, as_stream(stream)));
    return WCT_OK;
}

int wct_memset(void* dptr, int value, size_t bytes, void* stream)
{
    WCT_HIP_CHECK(hipMemsetAsync(dptr, value, bytes, as_stream(stream)));
    return WCT_OK;
}

int wct_host_malloc(void** hptr, size_t bytes)
{
    WCT_REQUIRE(hptr, WCT_ERR_ARG, "null pointer");
    hipError_t e = hipHostMalloc(hptr, bytes ? bytes : 1, hipHostMallocDefault);
    if (e == hipErrorOutOfMemory) {
        set_error("hipHostMalloc(%zu) out of memory", bytes);
        return WCT_ERR_NOMEM;
    }
    WCT_HIP_CHECK(e);
    return WCT_OK;
}

int wct_host_free(void* hptr)
{
    if (hptr) WCT_HIP_CHECK(hipHostFree(hptr));
    return WCT_OK;
}

int wct_memcpy_htod_async(void* dst, const void* src, size_t bytes, void* stream)
{
    WCT_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, as_stream(stream)));
    return WCT_OK;
}

int wct_memcpy_dtoh_async(void* dst, const void* src, size_t bytes, void* stream)
{
    WCT_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, as_stream(stream)));
    return WCT_OK;
}

int wct_stream_create(void** stream)
{
    WCT_ENTRY();
    WCT_REQUIRE(stream, WCT_ERR_ARG, "null pointer");
    hipStream_t s;
    WCT_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (void*)s;
    return WCT_OK;
}

int wct_stream_destroy(void* stream)
{
    if (stream) WCT_HIP_CHECK(hipStreamDestroy(as_stream(stream)));
    return WCT_OK;
}

int wct_stream_sync(void* stream)
{
    WCT_HIP_CHECK(hipStreamSynchronize(as_stream(stream)));
    return WCT_OK;
}

int wct_debug_inject(const char* entry, int after, int status)
{
    std::lock_guard<std::mutex> lk(g_hook_mutex);
    if (!entry || !*entry || status == 0) {
        g_inject_armed = false;
        g_inject_entry.clear();
        return WCT_OK;
    }
    WCT_REQUIRE(after >= 0 && status < 0, WCT_ERR_ARG, "debug_inject: after %d must be >= 0 and status %d < 0",
                after, status);
    g_inject_entry = entry;
    g_inject_after = after;
    g_inject_status = status;
    g_inject_armed = true;
    return WCT_OK;
}

int wct_debug_alloc_cap(size_t bytes)
{
    g_alloc_cap = bytes;
    return WCT_OK;
}

int wct_dft_supported(int n) { return n >= 1 && n <= kMaxDftLength ? 1 : 0; }

int wct_dft_native_length(int n) { return fft_supported_length(n) ? 1 : 0; }

}  // extern "C"
