// IDFT drop-in kernels: batched complex FFTs along the contiguous axis
// (rows) and along the strided axis (columns, through an LDS tile), and
// LDS-tiled transposes.  These implement the IDFT contract of
// iface/inc/WireCellIface/IDFT.h:9-127 with FftwDFT's conventions
// (aux/src/FftwDFT.cxx:88-283).
#include <map>
#include <mutex>
#include <tuple>

#include "common.hpp"
#include "gmem.hpp"

namespace wct {

constexpr int kBlock = 256;

// ---- rows: each slice of SLICE threads transforms one row in LDS ----
// LDS: rows of padded_len(n) entries, then the plan's twiddle table
// (TWG: no table in LDS, the passes read it through the caches, which
// for long rows leaves room for more workgroups per CU).
template <typename T, int SLICE, int NMAX, bool TWG = false>
__global__ __launch_bounds__(kBlock) void k_fft_rows(const typename C2<T>::type* __restrict__ in,
                                                     typename C2<T>::type* __restrict__ out,
                                                     int nrows, long dist, FftPlan pl,
                                                     const typename C2<T>::type* __restrict__ tw,
                                                     int sign, T scale)
{
    using T2 = typename C2<T>::type;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T2* smem = reinterpret_cast<T2*>(smem_raw);
    constexpr int RPB = kBlock / SLICE;
    const int slice = threadIdx.x / SLICE;
    const int tid = threadIdx.x % SLICE;
    const long row = (long)blockIdx.x * RPB + slice;
    const int n = pl.n;
    const int ld = padded_len(n);
    T2* buf = smem + (size_t)slice * ld;
    T2* twl = smem + (size_t)RPB * ld;
    if (!TWG) stage_twiddles(twl, tw, pl.twcount, threadIdx.x, kBlock);
    const bool active = row < nrows;
    if (active) {
        const auto src = G(in + row * dist);
        for (int i = tid; i < n; i += SLICE) buf[pidx(i)] = src[i];
    }
    __syncthreads();
    if (TWG)
        fft_lds<T2, SLICE, NMAX>(buf, pl, G(tw), sign, tid);
    else
        fft_lds<T2, SLICE, NMAX>(buf, pl, (const T2*)twl, sign, tid);
    if (active) {
        const auto dst = G(out + row * dist);
        for (int i = tid; i < n; i += SLICE) {
            T2 v = buf[pidx(i)];
            v.x *= scale;
            v.y *= scale;
            dst[i] = v;
        }
    }
}

// ---- rows of a length with a compile-time plan ----
// One row of N entries per workgroup of NT threads, unpadded in LDS, on
// the odd-first radix schedule (every LDS address a base plus an
// immediate); the schedule's twiddles (about N entries) are read through
// the caches so the row alone sets the LDS a workgroup takes.
template <int N, int NT>
__global__ __launch_bounds__(NT) void k_fft_rows_lin(const float2* __restrict__ in, float2* __restrict__ out,
                                                     long dist, const float2* __restrict__ tw, int sign, float scale)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float2* buf = reinterpret_cast<float2*>(smem_raw);
    const long row = blockIdx.x;
    const auto src = G(in + row * dist);
    for (int i = threadIdx.x; i < N; i += NT) buf[i] = src[i];
    __syncthreads();
    fft_lin<float2, NT, N>(buf, G(tw), sign, threadIdx.x);
    const auto dst = G(out + row * dist);
    for (int i = threadIdx.x; i < N; i += NT) {
        float2 v = buf[i];
        v.x *= scale;
        v.y *= scale;
        dst[i] = v;
    }
}

template <int N, int NT>
static int launch_rows_lin(const void* in, void* out, int nrows, long dist, int sign, float scale, hipStream_t s)
{
    const DevPlan* dl = nullptr;
    WCT_TRY(get_plan(N, false, &dl, true));
    const size_t lds = sizeof(float2) * N;
    WCT_TRY(allow_lds(k_fft_rows_lin<N, NT>, lds));
    hipLaunchKernelGGL((k_fft_rows_lin<N, NT>), dim3(nrows), dim3(NT), lds, s, (const float2*)in, (float2*)out, dist,
                       (const float2*)dl->tw, sign, scale);
    WCT_HIP_CHECK(hipGetLastError());
    return WCT_OK;
}

// ---- columns: a tile of TC columns, each transformed by its own slice ----
template <typename T, int TC, int NMAX>
__global__ __launch_bounds__(kBlock) void k_fft_cols(const typename C2<T>::type* __restrict__ in,
                                                     typename C2<T>::type* __restrict__ out,
                                                     int nrows, int ncols, long bstride, FftPlan pl,
                                                     const typename C2<T>::type* __restrict__ tw,
                                                     int sign, T scale)
{
    using T2 = typename C2<T>::type;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T2* smem = reinterpret_cast<T2*>(smem_raw);
    constexpr int SLICE = kBlock / TC;
    in += blockIdx.y * bstride;
    out += blockIdx.y * bstride;
    const int ld = padded_len(nrows) | 1;   // odd stride: the transposed store spreads over banks
    T2* twl = smem + (size_t)TC * ld;
    stage_twiddles(twl, tw, pl.twcount, threadIdx.x, kBlock);
    const long c0 = (long)blockIdx.x * TC;
    // coalesced load: TC consecutive columns of each row
    for (int idx = threadIdx.x; idx < nrows * TC; idx += kBlock) {
        const int cc = idx % TC, r = idx / TC;
        T2 v;
        v.x = 0;
        v.y = 0;
        if (c0 + cc < ncols) v = in[(long)r * ncols + c0 + cc];
        smem[cc * ld + pidx(r)] = v;
    }
    __syncthreads();
    const int slice = threadIdx.x / SLICE;
    const int tid = threadIdx.x % SLICE;
    fft_lds<T2, SLICE, NMAX>(smem + slice * ld, pl, twl, sign, tid);
    for (int idx = threadIdx.x; idx < nrows * TC; idx += kBlock) {
        const int cc = idx % TC, r = idx / TC;
        if (c0 + cc < ncols) {
            T2 v = smem[cc * ld + pidx(r)];
            v.x *= scale;
            v.y *= scale;
            out[(long)r * ncols + c0 + cc] = v;
        }
    }
}

// ---- transpose through a 32x32 LDS tile ----
template <typename V>
__global__ __launch_bounds__(kBlock) void k_transpose(const V* __restrict__ in, V* __restrict__ out,
                                                      int nrows, int ncols)
{
    __shared__ V tile[32][33];
    in += (long)blockIdx.z * nrows * ncols;
    out += (long)blockIdx.z * nrows * ncols;
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
    const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;   // 32 x 8
    for (int j = ty; j < 32; j += 8) {
        const int r = by + j, c = bx + tx;
        if (r < nrows && c < ncols) tile[j][tx] = in[(long)r * ncols + c];
    }
    __syncthreads();
    for (int j = ty; j < 32; j += 8) {
        const int r = bx + j, c = by + tx;   // out is ncols x nrows
        if (r < ncols && c < nrows) out[(long)r * nrows + c] = tile[tx][j];
    }
}

// ---- host dispatch ---------------------------------------------------

template <typename T>
static int launch_rows_lds(const void* in, void* out, int n, int nrows, long dist, int sign, T scale,
                           hipStream_t s)
{
    using T2 = typename C2<T>::type;
    const DevPlan* dp = nullptr;
    int rc = get_plan(n, sizeof(T) == 8, &dp);
    if (rc) return rc;
    const T2* tw = (const T2*)dp->tw;
    const T2* src = (const T2*)in;
    T2* dst = (T2*)out;
    constexpr bool dbl = sizeof(T) == 8;
    if constexpr (!dbl) {
        // lengths of common frames on compile-time plans: 6000 ticks (the
        // noise filters, the per-plane path), 6272 (the forward simulation's
        // padded 6000-tick window), 8000
        if (nrows >= 1) {
            if (n == 6000) return launch_rows_lin<6000, 512>(in, out, nrows, dist, sign, (float)scale, s);
            if (n == 6272) return launch_rows_lin<6272, 512>(in, out, nrows, dist, sign, (float)scale, s);
            if (n == 8000) return launch_rows_lin<8000, 512>(in, out, nrows, dist, sign, (float)scale, s);
        }
    }
#define WCT_ROWS(SL, NM, TWG)                                                                        \
    do {                                                                                             \
        const int rpb = kBlock / SL;                                                                 \
        dim3 grid((nrows + rpb - 1) / rpb);                                                          \
        size_t lds = ((size_t)rpb * padded_len(n) + (TWG ? 0 : dp->plan.twcount)) * sizeof(T2);      \
        int rc_ = allow_lds(k_fft_rows<T, SL, NM, TWG>, lds);                                        \
        if (rc_) return rc_;                                                                         \
        hipLaunchKernelGGL((k_fft_rows<T, SL, NM, TWG>), grid, dim3(kBlock), lds, s, src, dst, nrows, \
                           dist, dp->plan, tw, sign, scale);                                         \
    } while (0)
    // Rows above 1024 entries read the twiddles through the caches: their
    // tables (about n entries) would otherwise double the LDS a row takes.
    if (n <= 64) WCT_ROWS(16, 64, false);
    else if (n <= 256) WCT_ROWS(64, 256, false);
    else if (n <= 1024) WCT_ROWS(128, 1024, false);
    else if (n <= 4096) WCT_ROWS(256, 4096, true);
    else if (!dbl && n <= 8192) WCT_ROWS(256, 8192, true);
    else {
        set_error("transform length %d is above the LDS limit", n);
        return WCT_ERR_UNSUPPORTED;
    }
#undef WCT_ROWS
    WCT_HIP_CHECK(hipGetLastError());
    return WCT_OK;
}

template <typename T>
static int launch_cols(const void* in, void* out, int nrows, int ncols, int sign, T scale,
                       hipStream_t s, int nbatch = 1, long bstride = 0)
{
    using T2 = typename C2<T>::type;
    const DevPlan* dp = nullptr;
    int rc = get_plan(nrows, sizeof(T) == 8, &dp);
    if (rc) return rc;
    const T2* tw = (const T2*)dp->tw;
    const T2* src = (const T2*)in;
    T2* dst = (T2*)out;
#define WCT_COLS(TC, NM)                                                                         \
    do {                                                                                         \
        dim3 grid((ncols + TC - 1) / TC, nbatch);                                                \
        size_t lds = ((size_t)TC * (padded_len(nrows) | 1) + dp->plan.twcount) * sizeof(T2);     \
        int rc_ = allow_lds(k_fft_cols<T, TC, NM>, lds);                                         \
        if (rc_) return rc_;                                                                     \
        hipLaunchKernelGGL((k_fft_cols<T, TC, NM>), grid, dim3(kBlock), lds, s, src, dst, nrows, \
                           ncols, bstride, dp->plan, tw, sign, scale);                           \
    } while (0)
    // TC columns per tile, keeping the tile within 64 KB of LDS
    constexpr bool dbl = sizeof(T) == 8;
    if (nrows <= 256) WCT_COLS(8, 256);
    else if (!dbl && nrows <= 1024) WCT_COLS(4, 1024);
    else if (!dbl && nrows <= 2048) WCT_COLS(2, 2048);
    else if (!dbl && nrows <= 4096) WCT_COLS(1, 4096);
    else if (dbl && nrows <= 1024) WCT_COLS(2, 1024);
    else if (dbl && nrows <= 2048) WCT_COLS(1, 2048);
    else {
        set_error("column transform length %d is above the LDS limit", nrows);
        return WCT_ERR_UNSUPPORTED;
    }
#undef WCT_COLS
    WCT_HIP_CHECK(hipGetLastError());
    return WCT_OK;
}

template <typename V>
static int launch_transpose(const V* in, V* out, int nrows, int ncols, hipStream_t s)
{
    const size_t bytes = (size_t)nrows * ncols * sizeof(V);
    V* dst = out;
    DevTemp<V> tmp(s);
    if (in == out) {
        WCT_TRY(tmp.alloc((size_t)nrows * ncols));
        dst = tmp;
    }
    dim3 grid((ncols + 31) / 32, (nrows + 31) / 32);
    hipLaunchKernelGGL((k_transpose<V>), grid, dim3(kBlock), 0, s, in, dst, nrows, ncols);
    WCT_HIP_CHECK(hipGetLastError());
    if (tmp) WCT_HIP_CHECK(hipMemcpyAsync(out, tmp, bytes, hipMemcpyDeviceToDevice, s));
    return WCT_OK;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_fourstep_twiddle(typename C2<T>::type* x, int n1, int n2,
                                                             int n, long total, int sign)
{
    const long i = (long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= total) return;
    const long inner = i % ((long)n1 * n2);
    const int k1 = (int)(inner / n2), c = (int)(inner % n2);
    const long q = ((long)k1 * c) % n;
    double sn, cs;
    sincos(sign * 2.0 * M_PI * (double)q / (double)n, &sn, &cs);
    auto v = x[i];
    const T wr = (T)cs, wi = (T)sn;
    const T xr = v.x * wr - v.y * wi, xi = v.x * wi + v.y * wr;
    v.x = xr;
    v.y = xi;
    x[i] = v;
}

// Largest row length the single-pass LDS kernel takes.
template <typename T> constexpr int lds_row_limit() { return sizeof(T) == 8 ? 4096 : 8192; }

// ---- Bluestein (chirp-z) transforms for lengths with a prime factor
// above 13 ---------------------------------------------------------------
// With nk = (n^2 + k^2 - (k - n)^2) / 2,
//   X[k] = c_k sum_n (x_n c_n) conj(c_(k-n)),  c_n = exp(sign i pi n^2 / N),
// a circular convolution of length L >= 2N - 1 (L with prime factors 2, 3
// and 5), done with two length-L transforms and the transform of the
// chirp kernel b (b_m = conj(c_m) for |m| < N, circularly), which is
// built once per (N, L, sign) in FP64.  n^2 is reduced mod 2N before the
// angle is formed, so the chirp is accurate to the last bit of FP64.

__device__ __forceinline__ double2 chirp(long n, int N, int sign)
{
    const long q = (n * n) % (2L * N);
    double sn, cs;
    sincos((double)sign * M_PI * (double)q / (double)N, &sn, &cs);
    double2 r;
    r.x = cs;
    r.y = sn;
    return r;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_blue_pre(const typename C2<T>::type* __restrict__ in, long dist, int N,
                                                     int L, long total, int sign,
                                                     typename C2<T>::type* __restrict__ work)
{
    const long i = (long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= total) return;
    const long row = i / L;
    const int n = (int)(i % L);
    typename C2<T>::type v;
    v.x = 0;
    v.y = 0;
    if (n < N) {
        const auto x = in[row * dist + n];
        const double2 c = chirp(n, N, sign);
        v.x = (T)((double)x.x * c.x - (double)x.y * c.y);
        v.y = (T)((double)x.x * c.y + (double)x.y * c.x);
    }
    work[i] = v;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_blue_mul(typename C2<T>::type* __restrict__ work,
                                                     const typename C2<T>::type* __restrict__ bhat, int L,
                                                     long total)
{
    const long i = (long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= total) return;
    work[i] = cmul(work[i], bhat[i % L]);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_blue_post(const typename C2<T>::type* __restrict__ work, int N, int L,
                                                      long total, int sign, T scale,
                                                      typename C2<T>::type* __restrict__ out, long dist)
{
    const long i = (long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= total) return;
    const long row = i / N;
    const int k = (int)(i % N);
    const auto w = work[row * L + k];
    const double2 c = chirp(k, N, sign);
    typename C2<T>::type v;
    v.x = (T)(((double)w.x * c.x - (double)w.y * c.y) * (double)scale);
    v.y = (T)(((double)w.x * c.y + (double)w.y * c.x) * (double)scale);
    out[row * dist + k] = v;
}

__global__ __launch_bounds__(kBlock) void k_blue_kernel(int N, int L, int sign, double2* __restrict__ b)
{
    const int m = blockIdx.x * kBlock + threadIdx.x;
    if (m >= L) return;
    double2 v;
    v.x = 0;
    v.y = 0;
    const int d = m < N ? m : (L - m < N ? L - m : -1);
    if (d >= 0) {
        const double2 c = chirp(d, N, sign);
        v.x = c.x;
        v.y = -c.y;
    }
    b[m] = v;
}

__global__ __launch_bounds__(kBlock) void k_d2f(const double2* __restrict__ in, float2* __restrict__ out, int n)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        float2 v;
        v.x = (float)in[i].x;
        v.y = (float)in[i].y;
        out[i] = v;
    }
}

// Smallest length >= n with prime factors 2, 3 and 5.
static int smooth5_at_least(int n)
{
    for (int L = n;; ++L) {
        int m = L;
        for (int q : {2, 3, 5})
            while (m % q == 0) m /= q;
        if (m == 1) return L;
    }
}

using BlueKey = std::tuple<int, int, int, bool>;   // device, N, sign, double
static std::mutex g_blue_mutex;
static std::map<BlueKey, void*>* g_blue = new std::map<BlueKey, void*>();

template <typename T> int launch_rows(const void*, void*, int, int, long, int, T, hipStream_t);

// FFT_L of the chirp kernel for (N, sign), in T, cached per device.
template <typename T>
static int blue_kernel(int N, int L, int sign, const void** out, hipStream_t s)
{
    constexpr bool dbl = sizeof(T) == 8;
    int dev = 0;
    WCT_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_blue_mutex);
    const BlueKey key(dev, N, sign, dbl);
    auto it = g_blue->find(key);
    if (it != g_blue->end()) {
        *out = it->second;
        return WCT_OK;
    }
    // b is stream-ordered scratch when the table is kept in float
    DevTemp<double2> scratch(s);
    double2* b = nullptr;
    if (dbl) {
        WCT_TRY(dev_alloc((void**)&b, (size_t)L * sizeof(double2)));
    } else {
        WCT_TRY(scratch.alloc((size_t)L));
        b = scratch;
    }
    hipLaunchKernelGGL(k_blue_kernel, dim3((L + kBlock - 1) / kBlock), dim3(kBlock), 0, s, N, L, sign, b);
    WCT_HIP_CHECK(hipGetLastError());
    int rc = launch_rows<double>(b, b, L, 1, L, -1, 1.0, s);
    if (rc) {
        if (dbl) (void)hipFree(b);
        return rc;
    }
    void* res = b;
    if (!dbl) {
        float2* f = nullptr;
        WCT_TRY(dev_alloc((void**)&f, (size_t)L * sizeof(float2)));
        hipLaunchKernelGGL(k_d2f, dim3((L + kBlock - 1) / kBlock), dim3(kBlock), 0, s, b, f, L);
        WCT_HIP_CHECK(hipGetLastError());
        res = f;
    }
    WCT_HIP_CHECK(hipStreamSynchronize(s));
    (*g_blue)[key] = res;
    *out = res;
    return WCT_OK;
}

// Batched rows of any length N by Bluestein's algorithm.
template <typename T>
static int blue_rows(const void* in, void* out, int N, int nrows, long dist, int sign, T scale, hipStream_t s)
{
    using T2 = typename C2<T>::type;
    const int L = smooth5_at_least(2 * N - 1);
    const void* bhat = nullptr;
    int rc = blue_kernel<T>(N, L, sign, &bhat, s);
    if (rc) return rc;
    const long total = (long)nrows * L;
    DevTemp<T2> work(s);
    WCT_TRY(work.alloc(total));
    const dim3 g((total + kBlock - 1) / kBlock);
    hipLaunchKernelGGL((k_blue_pre<T>), g, dim3(kBlock), 0, s, (const T2*)in, dist, N, L, total, sign, work);
    WCT_HIP_CHECK(hipGetLastError());
    rc = launch_rows<T>(work, work, L, nrows, L, -1, T(1), s);
    if (rc) return rc;
    hipLaunchKernelGGL((k_blue_mul<T>), g, dim3(kBlock), 0, s, work, (const T2*)bhat, L, total);
    WCT_HIP_CHECK(hipGetLastError());
    rc = launch_rows<T>(work, work, L, nrows, L, +1, T(1) / T(L), s);
    if (rc) return rc;
    const long tout = (long)nrows * N;
    hipLaunchKernelGGL((k_blue_post<T>), dim3((tout + kBlock - 1) / kBlock), dim3(kBlock), 0, s, work, N, L, tout,
                       sign, scale, (T2*)out, dist);
    WCT_HIP_CHECK(hipGetLastError());
    return WCT_OK;
}

// Columns of any length (nrows) by transposing to rows and back.
template <typename T>
static int blue_cols(const void* in, void* out, int nrows, int ncols, int sign, T scale, hipStream_t s)
{
    using T2 = typename C2<T>::type;
    const size_t bytes = (size_t)nrows * ncols * sizeof(T2);
    DevTemp<T2> t(s);
    WCT_TRY(t.alloc(bytes / sizeof(T2)));
    dim3 g1((ncols + 31) / 32, (nrows + 31) / 32);
    hipLaunchKernelGGL((k_transpose<T2>), g1, dim3(kBlock), 0, s, (const T2*)in, t, nrows, ncols);
    WCT_HIP_CHECK(hipGetLastError());
    int rc = blue_rows<T>(t, t, nrows, ncols, nrows, sign, scale, s);
    if (rc) return rc;
    dim3 g2((nrows + 31) / 32, (ncols + 31) / 32);
    hipLaunchKernelGGL((k_transpose<T2>), g2, dim3(kBlock), 0, s, (const T2*)t, (T2*)out, ncols, nrows);
    WCT_HIP_CHECK(hipGetLastError());
    return WCT_OK;
}

// Batched row transforms of any supported length: one LDS pass when the
// row fits, otherwise the four-step split n = n1 * n2 (column
// transforms of length n1, twiddles, row transforms of length n2, then a
// transpose of each n1 x n2 block).  Rows are contiguous (dist == n).
template <typename T>
int launch_rows(const void* in, void* out, int n, int nrows, long dist, int sign, T scale,
                hipStream_t s)
{
    using T2 = typename C2<T>::type;
    WCT_REQUIRE(n >= 1 && n <= kMaxDftLength, WCT_ERR_UNSUPPORTED, "transform length %d not in [1, %d]", n,
                kMaxDftLength);
    if (!fft_supported_length(n)) return blue_rows<T>(in, out, n, nrows, dist, sign, scale, s);
    if (n <= lds_row_limit<T>()) return launch_rows_lds<T>(in, out, n, nrows, dist, sign, scale, s);
    WCT_REQUIRE(dist == n, WCT_ERR_UNSUPPORTED, "strided rows above the LDS limit");
    WCT_REQUIRE(fft_supported_length(n), WCT_ERR_UNSUPPORTED,
                "FFT length %d has a prime factor above 13", n);
    // n2: largest divisor within the row limit with n1 = n / n2 small enough for a column tile
    int n2 = 0;
    for (int d = lds_row_limit<T>() / 2; d >= 2; --d)
        if (n % d == 0 && n / d <= 2048) { n2 = d; break; }
    WCT_REQUIRE(n2 > 0, WCT_ERR_UNSUPPORTED, "no four-step split for length %d", n);
    const int n1 = n / n2;
    const long total = (long)nrows * n;
    T2* o = (T2*)out;
    int rc = launch_cols<T>(in, out, n1, n2, sign, T(1), s, nrows, (long)n);
    if (rc) return rc;
    hipLaunchKernelGGL((k_fourstep_twiddle<T>), dim3((total + kBlock - 1) / kBlock), dim3(kBlock), 0,
                       s, o, n1, n2, n, total, sign);
    WCT_HIP_CHECK(hipGetLastError());
    rc = launch_rows_lds<T>(out, out, n2, nrows * n1, n2, sign, scale, s);
    if (rc) return rc;
    DevTemp<T2> tmp(s);
    WCT_TRY(tmp.alloc(total));
    dim3 grid((n2 + 31) / 32, (n1 + 31) / 32, nrows);
    hipLaunchKernelGGL((k_transpose<T2>), grid, dim3(kBlock), 0, s, (const T2*)o, (T2*)tmp, n1, n2);
    WCT_HIP_CHECK(hipGetLastError());
    WCT_HIP_CHECK(hipMemcpyAsync(o, tmp, total * sizeof(T2), hipMemcpyDeviceToDevice, s));
    return WCT_OK;
}
template int launch_rows<float>(const void*, void*, int, int, long, int, float, hipStream_t);
template int launch_rows<double>(const void*, void*, int, int, long, int, double, hipStream_t);

template <typename T>
int launch_cols_any(const void* in, void* out, int nrows, int ncols, int sign, T scale, hipStream_t s)
{
    WCT_REQUIRE(nrows >= 1 && nrows <= kMaxDftLength, WCT_ERR_UNSUPPORTED, "transform length %d not in [1, %d]",
                nrows, kMaxDftLength);
    if (!fft_supported_length(nrows)) return blue_cols<T>(in, out, nrows, ncols, sign, scale, s);
    return launch_cols<T>(in, out, nrows, ncols, sign, scale, s);
}
template int launch_cols_any<float>(const void*, void*, int, int, int, float, hipStream_t);
template int launch_cols_any<double>(const void*, void*, int, int, int, double, hipStream_t);

template <typename T>
static int dft_1b(const T* in, T* out, int nrows, int ncols, int axis, int sign, hipStream_t s)
{
    WCT_REQUIRE(in && out && nrows > 0 && ncols > 0 && (axis == 0 || axis == 1), WCT_ERR_ARG,
                "dft 1b: bad arguments (nrows=%d ncols=%d axis=%d)", nrows, ncols, axis);
    if (axis == 1) {
        const T scale = sign > 0 ? T(1) / T(ncols) : T(1);
        return launch_rows<T>(in, out, ncols, nrows, ncols, sign, scale, s);
    }
    const T scale = sign > 0 ? T(1) / T(nrows) : T(1);
    return launch_cols_any<T>(in, out, nrows, ncols, sign, scale, s);
}

}  // namespace wct

using namespace wct;

extern "C" {

int wct_dft_fwd1d(const float* in, float* out, int size, void* stream)
{
    WCT_ENTRY();
    WCT_REQUIRE(in && out && size > 0, WCT_ERR_ARG, "fwd1d: bad arguments");
    return launch_rows<float>(in, out, size, 1, size, -1, 1.0f, as_stream(stream));
}

int wct_dft_inv1d(const float* in, float* out, int size, void* stream)
{
    WCT_ENTRY();
    WCT_REQUIRE(in && out && size > 0, WCT_ERR_ARG, "inv1d: bad arguments");
    return launch_rows<float>(in, out, size, 1, size, +1, 1.0f / (float)size, as_stream(stream));
}

int wct_dft_fwd1b(const float* in, float* out, int nrows, int ncols, int axis, void* stream)
{
    WCT_ENTRY();
    return dft_1b<float>(in, out, nrows, ncols, axis, -1, as_stream(stream));
}

int wct_dft_inv1b(const float* in, float* out, int nrows, int ncols, int axis, void* stream)
{
    WCT_ENTRY();
    return dft_1b<float>(in, out, nrows, ncols, axis, +1, as_stream(stream));
}

int wct_dft_fwd1b_f64(const double* in, double* out, int nrows, int ncols, int axis, void* stream)
{
    WCT_ENTRY();
    return dft_1b<double>(in, out, nrows, ncols, axis, -1, as_stream(stream));
}

int wct_dft_inv1b_f64(const double* in, double* out, int nrows, int ncols, int axis, void* stream)
{
    WCT_ENTRY();
    return dft_1b<double>(in, out, nrows, ncols, axis, +1, as_stream(stream));
}

// FftwDFT plans fftwf_plan_dft_2d(ncols, nrows): the buffer is an
// (R = ncols) x (C = nrows) row-major array for the transform.
int wct_dft_fwd2d(const float* in, float* out, int nrows, int ncols, void* stream)
{
    WCT_ENTRY();
    WCT_REQUIRE(in && out && nrows > 0 && ncols > 0, WCT_ERR_ARG, "fwd2d: bad arguments");
    const int R = ncols, C = nrows;
    hipStream_t s = as_stream(stream);
    int rc = launch_rows<float>(in, out, C, R, C, -1, 1.0f, s);
    if (rc) return rc;
    return launch_cols_any<float>(out, out, R, C, -1, 1.0f, s);
}

int wct_dft_inv2d(const float* in, float* out, int nrows, int ncols, void* stream)
{
    WCT_ENTRY();
    WCT_REQUIRE(in && out && nrows > 0 && ncols > 0, WCT_ERR_ARG, "inv2d: bad arguments");
    const int R = ncols, C = nrows;
    hipStream_t s = as_stream(stream);
    int rc = launch_rows<float>(in, out, C, R, C, +1, 1.0f / (float)C, s);
    if (rc) return rc;
    return launch_cols_any<float>(out, out, R, C, +1, 1.0f / (float)R, s);
}

int wct_dft_transpose_real(const float* in, float* out, int nrows, int ncols, void* stream)
{
    WCT_ENTRY();
    WCT_REQUIRE(in && out && nrows > 0 && ncols > 0, WCT_ERR_ARG, "transpose: bad arguments");
    return launch_transpose<float>(in, out, nrows, ncols, as_stream(stream));
}

int wct_dft_transpose_complex(const float* in, float* out, int nrows, int ncols, void* stream)
{
    WCT_ENTRY();
    WCT_REQUIRE(in && out && nrows > 0 && ncols > 0, WCT_ERR_ARG, "transpose: bad arguments");
    return launch_transpose<float2>((const float2*)in, (float2*)out, nrows, ncols,
                                    as_stream(stream));
}

}  // extern "C"
