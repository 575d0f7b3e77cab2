// Shared host-side plumbing of the HIP library: status codes, the
// thread-local error string of the C ABI, and the per-device cache of
// FFT plans (radix schedule + twiddle tables in device memory).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/wct_hip.h"
#include "fft_lds.hpp"

namespace wct {

void set_error(const char* fmt, ...);

struct Status {
    int code;
};

#define WCT_HIP_CHECK(expr)                                                               \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) {                                                           \
            ::wct::set_error("%s:%d: %s failed: %s", __FILE__, __LINE__, #expr,           \
                             hipGetErrorString(e_));                                      \
            return WCT_ERR_HIP;                                                           \
        }                                                                                 \
    } while (0)

#define WCT_REQUIRE(cond, code, ...)                                                      \
    do {                                                                                  \
        if (!(cond)) {                                                                    \
            ::wct::set_error(__VA_ARGS__);                                                \
            return (code);                                                                \
        }                                                                                 \
    } while (0)

// ---- failure-path hooks ----------------------------------------------
// The status wct_debug_inject armed for this entry point's call, or 0.
int inject_check(const char* entry);
// First statement of every C-ABI entry point that does device work.
#define WCT_ENTRY()                                                                       \
    do {                                                                                  \
        if (const int rc_ = ::wct::inject_check(__func__)) return rc_;                    \
    } while (0)

// Device allocations of the library: WCT_ERR_NOMEM (with the error text)
// when HIP is out of memory or the request is above the cap that
// wct_debug_alloc_cap set; other HIP failures give WCT_ERR_HIP.
int dev_alloc(void** p, size_t bytes);
int dev_alloc_async(void** p, size_t bytes, hipStream_t s);

// A stream-ordered device temporary, released (hipFreeAsync on its
// stream) when it goes out of scope, on early returns included.
template <typename T>
struct DevTemp {
    T* p = nullptr;
    hipStream_t s;
    explicit DevTemp(hipStream_t stream) : s(stream) {}
    DevTemp(const DevTemp&) = delete;
    DevTemp& operator=(const DevTemp&) = delete;
    ~DevTemp()
    {
        if (p) (void)hipFreeAsync(p, s);
    }
    int alloc(size_t count) { return dev_alloc_async(reinterpret_cast<void**>(&p), count * sizeof(T), s); }
    operator T*() const { return p; }
};

#define WCT_TRY(expr)                                                                     \
    do {                                                                                  \
        if (const int rc_ = (expr)) return rc_;                                           \
    } while (0)

// Largest prime factor allowed by the LDS kernels.
bool fft_supported_length(int n);

// Largest transform length of the IDFT functions (other lengths than
// 2^a 3^b 5^c 7^d 11^e 13^f go through Bluestein's algorithm).
constexpr int kMaxDftLength = 1 << 20;

// Device limit of LDS bytes per workgroup (160 KB on gfx950).
int lds_limit_bytes();

// Raises a kernel's dynamic-LDS ceiling when a launch needs more than the
// default 64 KB.
template <typename K>
inline int allow_lds(K kernel, size_t bytes)
{
    if (bytes <= 65536) return WCT_OK;
    if (bytes > (size_t)lds_limit_bytes()) {
        set_error("launch needs %zu bytes of LDS, the device allows %d", bytes, lds_limit_bytes());
        return WCT_ERR_UNSUPPORTED;
    }
    WCT_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    return WCT_OK;
}

// A device-resident plan for one length and precision.
struct DevPlan {
    FftPlan plan;
    void* tw = nullptr;      // T2 table: pass twiddles then generic tables
    size_t tw_count = 0;
};

// Returns a cached plan (created on first use) for length n on the
// current device.  is_double selects double2 twiddles.
int get_plan(int n, bool is_double, const DevPlan** out, bool odd_first = false);

// Twiddles exp(-2 pi i k / N) for k in [0, count), as T2, cached.
int get_twiddle_line(int N, int count, bool is_double, const void** out);

// The cached plan / twiddle line when it exists on the current device,
// WCT_ERR_STATE when it does not (these never allocate).
int find_plan(int n, bool is_double, const DevPlan** out);
int find_twiddle_line(int N, int count, bool is_double, const void** out);

// Batched transforms used across translation units (dft.hip).
template <typename T>
int launch_rows(const void* in, void* out, int n, int nrows, long dist, int sign, T scale,
                hipStream_t s);
template <typename T>
int launch_cols_any(const void* in, void* out, int nrows, int ncols, int sign, T scale,
                    hipStream_t s);

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace wct
