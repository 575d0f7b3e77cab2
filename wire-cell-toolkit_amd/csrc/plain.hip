// The plain per-plane device path: OmnibusSigProc's control flow for one
// plane with one launch per step of the reference, FP32 throughout as the
// reference computes, buffers resident in device memory:
//   load_data (OmnibusSigProc.cxx:304-352)          k_load_real
//   decon_2D_init (:944-1051)
//     DftTools fwd_r2c along ticks                  k_real_to_complex + fwd1b(axis 1)
//     fwd along wires                               fwd1b(axis 0)
//     response spectrum, rebuilt per call (:988)    k_pad_response + fwd1b(1) + fwd1b(0)
//     c_data / R (:1001)                            k_cdiv
//     NaN/Inf fix-up and wire filter (:1004-1019)   k_fix_wire_filter
//     inverse along wires (:1022)                   inv1b(axis 0)
//     inv_c2r along ticks (:1025)                   k_mirror + inv1b(axis 1) + k_real_part
//     wire and tick rolls (:1028-1048)              k_roll
//     fwd_r2c of the result (:1049)                 k_real_to_complex + fwd1b(axis 1)
//   decon_2D_tightROI / tighterROI (:1077-1160)     k_mul_time_filter + k_mirror +
//                                                   inv1b(axis 1) + k_real_part
//   restore_baseline (:867-906)                     k_rows (one workgroup per wire)
//   cal_RMS + threshold (ROI_formation.cxx:364-456) k_rows
//   save_data (:383-418)                            k_save
// This is the denominator the batched fused kernels are measured against.
#include "common.hpp"
#include "rowstats.hpp"

namespace wct {

namespace {

constexpr int kEw = 256;   // threads per elementwise workgroup

inline dim3 ew_grid(long n)
{
    long b = (n + kEw - 1) / kEw;
    if (b > 1 << 20) b = 1 << 20;
    return dim3((unsigned)b);
}

__device__ __forceinline__ bool bad_at(const int* bp, const int* blo, const int* bhi, int w, int t, bool inclusive)
{
    if (!bp) return false;
    for (int q = bp[w]; q < bp[w + 1]; ++q) {
        const int lo = blo[q], hi = bhi[q];
        if (t >= lo && (inclusive ? t <= hi : t < hi)) return true;
    }
    return false;
}

__global__ void k_load_real(const float* adc, int ld, int nw, int nt, const int* bp, const int* blo,
                            const int* bhi, float* x)
{
    const long n = (long)nw * nt;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int w = (int)(i / nt), t = (int)(i % nt);
        float v = adc[(long)w * ld + t];
        if (bad_at(bp, blo, bhi, w, t, false)) v = 0.f;
        x[i] = v;
    }
}

__global__ void k_real_to_complex(const float* x, float2* s, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        s[i] = make_float2(x[i], 0.f);
}

__global__ void k_pad_response(const float* resp, int nr, int nw, int nt, float2* r)
{
    const long n = (long)nw * nt;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int w = (int)(i / nt);
        r[i] = make_float2(w < nr ? resp[i] : 0.f, 0.f);
    }
}

// std::complex<float> division in the scaled (Smith) form.
__device__ __forceinline__ float2 cdiv_f(float2 a, float2 b)
{
    if (fabsf(b.x) >= fabsf(b.y)) {
        if (b.x == 0.f && b.y == 0.f) return make_float2(a.x / 0.f, a.y / 0.f);
        const float r = b.y / b.x, d = b.x + b.y * r;
        return make_float2((a.x + a.y * r) / d, (a.y - a.x * r) / d);
    }
    const float r = b.x / b.y, d = b.y + b.x * r;
    return make_float2((a.x * r + a.y) / d, (a.y * r - a.x) / d);
}

__global__ void k_cdiv(float2* s, const float2* r, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        s[i] = cdiv_f(s[i], r[i]);
}

// |S| NaN -> -0, Inf -> 0, then the wire filter of wire-frequency row m.
__global__ void k_fix_wire_filter(float2* s, const float* fw, int nw, int nt)
{
    const long n = (long)nw * nt;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        float2 v = s[i];
        const float a = hypotf(v.x, v.y);
        if (isnan(a)) v = make_float2(-0.f, 0.f);
        else if (isinf(a)) v = make_float2(0.f, 0.f);
        const float f = fw[i / nt];
        s[i] = make_float2(v.x * f, v.y * f);
    }
}

// Spectrum::hermitian_mirror along each row (util/inc/WireCellUtil/
// Spectrum.h:56-89): bins above n/2 become conjugates of their mirrors,
// DC and Nyquist keep their real parts.
__global__ void k_mirror(float2* s, int nrows, int nt)
{
    const int half = nt / 2;
    const long n = (long)nrows * nt;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int t = (int)(i % nt);
        const long row = i - t;
        if (t == 0 || (nt % 2 == 0 && t == half)) {
            s[i].y = 0.f;
        } else if (t > half) {
            const float2 m = s[row + (nt - t)];
            s[i] = make_float2(m.x, -m.y);
        }
    }
}

__global__ void k_real_part(const float2* s, float* x, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        x[i] = s[i].x;
}

__global__ void k_roll(const float* x, float* y, int nw, int nt, int ws, int ts)
{
    const long n = (long)nw * nt;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int w = (int)(i / nt), t = (int)(i % nt);
        const int w2 = (int)(((long)w + ws) % nw), t2 = (int)(((long)t + ts) % nt);
        y[(long)w2 * nt + t2] = x[i];
    }
}

__global__ void k_mul_time_filter(const float2* c, const float* ft, float2* s, int nw, int nt)
{
    const long n = (long)nw * nt;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float f = ft[i % nt];
        const float2 v = c[i];
        s[i] = make_float2(v.x * f, v.y * f);
    }
}

// One workgroup per wire: restore_baseline of y (and of y2 when given),
// then, when rms is given, cal_RMS over the non-bad samples of y and the
// mask (y > th || y2 > th, th = f rms + 1).
template <int NT, int VPT>
__global__ void __launch_bounds__(NT) k_rows(float* y, float* y2, int nw, int nt, const int* bp, const int* blo,
                                             const int* bhi, float th_factor, float* rms_out,
                                             unsigned char* mask, int ld_mask)
{
    __shared__ RowScratch sh_;
    RowScratch* sh = &sh_;
    const int w = blockIdx.x;
    if (w >= nw) return;
    float val[VPT];
    uint64_t valid = 0, pm = 0;
    float* row = y + (long)w * nt;
#pragma unroll
    for (int i = 0; i < VPT; ++i) {
        const int t = threadIdx.x + i * NT;
        val[i] = t < nt ? row[t] : 0.f;
        if (t < nt) {
            valid |= 1ull << i;
            if (!bad_at(bp, blo, bhi, w, t, true)) pm |= 1ull << i;
        }
    }
    row_restore_baseline<NT, VPT>(val, valid, sh);
#pragma unroll
    for (int i = 0; i < VPT; ++i) {
        const int t = threadIdx.x + i * NT;
        if (t < nt) row[t] = val[i];
    }
    float val2[VPT];
    if (y2) {
        float* row2 = y2 + (long)w * nt;
#pragma unroll
        for (int i = 0; i < VPT; ++i) {
            const int t = threadIdx.x + i * NT;
            val2[i] = t < nt ? row2[t] : 0.f;
        }
        row_restore_baseline<NT, VPT>(val2, valid, sh);
#pragma unroll
        for (int i = 0; i < VPT; ++i) {
            const int t = threadIdx.x + i * NT;
            if (t < nt) row2[t] = val2[i];
        }
    }
    if (!rms_out) return;
    const float rms = row_cal_rms<NT, VPT>(val, pm, sh);
    const double th = (double)th_factor * (double)rms + 1.0;
    if (mask) {
        unsigned char* mk = mask + (long)w * ld_mask;
#pragma unroll
        for (int i = 0; i < VPT; ++i) {
            const int t = threadIdx.x + i * NT;
            if (t < nt) {
                unsigned m = 0;
                if ((pm >> i) & 1ull) {
                    if ((double)val[i] > th) m |= 1u;
                    if (y2 && (double)val2[i] > th) m |= 2u;
                }
                mk[t] = (unsigned char)m;
            }
        }
    }
    if (threadIdx.x == 0) rms_out[w] = rms;
}

__global__ void k_save(const float* y, float* out, int ld_out, int nw, int nt, const int* bp, const int* blo,
                       const int* bhi, int save_mode)
{
    const long n = (long)nw * nt;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int w = (int)(i / nt), t = (int)(i % nt);
        float v = y[i];
        if (save_mode) {
            if (!(v > 0.f)) v = 0.f;
            if (bad_at(bp, blo, bhi, w, t, false)) v = 0.f;
        }
        out[(long)w * ld_out + t] = v;
    }
}

template <int VPT>
int launch_rows_vpt(float* y, float* y2, int nw, int nt, const int* bp, const int* blo, const int* bhi,
                    float th, float* rms, unsigned char* mask, int ld_mask, hipStream_t s)
{
    hipLaunchKernelGGL((k_rows<256, VPT>), dim3(nw), dim3(256), 0, s, y, y2, nw, nt, bp, blo, bhi, th, rms, mask,
                       ld_mask);
    WCT_HIP_CHECK(hipGetLastError());
    return WCT_OK;
}

int launch_plain_rows(float* y, float* y2, int nw, int nt, const int* bp, const int* blo, const int* bhi, float th,
                      float* rms, unsigned char* mask, int ld_mask, hipStream_t s)
{
    const int v = (nt + 255) / 256;
    if (v <= 8) return launch_rows_vpt<8>(y, y2, nw, nt, bp, blo, bhi, th, rms, mask, ld_mask, s);
    if (v <= 16) return launch_rows_vpt<16>(y, y2, nw, nt, bp, blo, bhi, th, rms, mask, ld_mask, s);
    if (v <= 24) return launch_rows_vpt<24>(y, y2, nw, nt, bp, blo, bhi, th, rms, mask, ld_mask, s);
    if (v <= 32) return launch_rows_vpt<32>(y, y2, nw, nt, bp, blo, bhi, th, rms, mask, ld_mask, s);
    if (v <= 40) return launch_rows_vpt<40>(y, y2, nw, nt, bp, blo, bhi, th, rms, mask, ld_mask, s);
    if (v <= 48) return launch_rows_vpt<48>(y, y2, nw, nt, bp, blo, bhi, th, rms, mask, ld_mask, s);
    if (v <= 64) return launch_rows_vpt<64>(y, y2, nw, nt, bp, blo, bhi, th, rms, mask, ld_mask, s);
    set_error("plain path: rows of %d ticks exceed 16384", nt);
    return WCT_ERR_UNSUPPORTED;
}

struct PlainWs {
    float2 *S, *R, *C;
    float *X, *Y, *Y2;
};

size_t plain_ws_layout(int nw, int nt, char* base, PlainWs* ws)
{
    const size_t n = (size_t)nw * nt;
    const size_t cb = ((n * sizeof(float2)) + 255) & ~size_t(255);
    const size_t rb = ((n * sizeof(float)) + 255) & ~size_t(255);
    if (ws) {
        ws->S = (float2*)(base);
        ws->R = (float2*)(base + cb);
        ws->C = (float2*)(base + 2 * cb);
        ws->X = (float*)(base + 3 * cb);
        ws->Y = (float*)(base + 3 * cb + rb);
        ws->Y2 = (float*)(base + 3 * cb + 2 * rb);
    }
    return 3 * cb + 3 * rb;
}

#define PLAIN_TRY(expr)                 \
    do {                                \
        const int rc_ = (expr);         \
        if (rc_ != WCT_OK) return rc_;  \
    } while (0)

// inv_c2r along ticks: mirror, inverse c2c, real part.
int plain_c2r(float2* s, float* x, int nw, int nt, hipStream_t st)
{
    const long n = (long)nw * nt;
    hipLaunchKernelGGL(k_mirror, ew_grid(n), dim3(kEw), 0, st, s, nw, nt);
    WCT_HIP_CHECK(hipGetLastError());
    PLAIN_TRY(wct_dft_inv1b((const float*)s, (float*)s, nw, nt, 1, st));
    hipLaunchKernelGGL(k_real_part, ew_grid(n), dim3(kEw), 0, st, s, x, n);
    WCT_HIP_CHECK(hipGetLastError());
    return WCT_OK;
}

}  // namespace
}  // namespace wct

using namespace wct;

extern "C" {

size_t wct_plain_workspace_size(int nw, int nt)
{
    if (nw <= 0 || nt <= 0) return 0;
    return plain_ws_layout(nw, nt, nullptr, nullptr);
}

int wct_plain_decon_plane(const float* adc, int ld_adc, int nw, int nt, const float* resp, int nr,
                          const float* wire_filter, const float* tight_filter, const float* tighter_filter,
                          int wire_shift, int time_shift, float th_factor, const int* bad_ptr, const int* bad_lo,
                          const int* bad_hi, float* out, int ld_out, unsigned char* mask, int ld_mask, float* rms,
                          int save_mode, void* workspace, size_t ws_bytes, void* stream)
{
    WCT_ENTRY();
    WCT_REQUIRE(adc && resp && wire_filter && tight_filter && out && workspace, WCT_ERR_ARG,
                "plain decon: null pointer");
    WCT_REQUIRE(nw > 0 && nt > 0 && nr > 0 && nr <= nw && ld_adc >= nt && ld_out >= nt &&
                    (!mask || ld_mask >= nt),
                WCT_ERR_ARG, "plain decon: bad sizes (nw=%d nt=%d nr=%d)", nw, nt, nr);
    // any lengths: the IDFT functions take them (Bluestein above 13); the
    // row statistics hold up to 64 x 256 samples per row
    WCT_REQUIRE(nt <= 64 * 256, WCT_ERR_UNSUPPORTED, "plain decon: %d ticks above 16384", nt);
    WCT_REQUIRE(ws_bytes >= wct_plain_workspace_size(nw, nt), WCT_ERR_ARG, "plain decon: workspace too small");
    WCT_REQUIRE((bad_ptr == nullptr) == (bad_lo == nullptr) && (bad_lo == nullptr) == (bad_hi == nullptr),
                WCT_ERR_ARG, "plain decon: bad ranges need ptr, lo and hi");
    WCT_REQUIRE(!mask || rms, WCT_ERR_ARG, "plain decon: the mask needs the rms output");
    hipStream_t st = as_stream(stream);
    PlainWs ws;
    plain_ws_layout(nw, nt, (char*)workspace, &ws);
    const long n = (long)nw * nt;
    const dim3 g = ew_grid(n), b(kEw);

    // load_data, then decon_2D_init
    hipLaunchKernelGGL(k_load_real, g, b, 0, st, adc, ld_adc, nw, nt, bad_ptr, bad_lo, bad_hi, ws.X);
    hipLaunchKernelGGL(k_real_to_complex, g, b, 0, st, ws.X, ws.S, n);
    WCT_HIP_CHECK(hipGetLastError());
    PLAIN_TRY(wct_dft_fwd1b((const float*)ws.S, (float*)ws.S, nw, nt, 1, st));
    PLAIN_TRY(wct_dft_fwd1b((const float*)ws.S, (float*)ws.S, nw, nt, 0, st));
    hipLaunchKernelGGL(k_pad_response, g, b, 0, st, resp, nr, nw, nt, ws.R);
    WCT_HIP_CHECK(hipGetLastError());
    PLAIN_TRY(wct_dft_fwd1b((const float*)ws.R, (float*)ws.R, nw, nt, 1, st));
    PLAIN_TRY(wct_dft_fwd1b((const float*)ws.R, (float*)ws.R, nw, nt, 0, st));
    hipLaunchKernelGGL(k_cdiv, g, b, 0, st, ws.S, ws.R, n);
    hipLaunchKernelGGL(k_fix_wire_filter, g, b, 0, st, ws.S, wire_filter, nw, nt);
    WCT_HIP_CHECK(hipGetLastError());
    PLAIN_TRY(wct_dft_inv1b((const float*)ws.S, (float*)ws.S, nw, nt, 0, st));
    PLAIN_TRY(plain_c2r(ws.S, ws.X, nw, nt, st));
    const int ws_w = ((wire_shift % nw) + nw) % nw;
    const int ts = time_shift > 0 ? time_shift % nt : 0;
    hipLaunchKernelGGL(k_roll, g, b, 0, st, ws.X, ws.Y, nw, nt, ws_w, ts);
    hipLaunchKernelGGL(k_real_to_complex, g, b, 0, st, ws.Y, ws.C, n);
    WCT_HIP_CHECK(hipGetLastError());
    PLAIN_TRY(wct_dft_fwd1b((const float*)ws.C, (float*)ws.C, nw, nt, 1, st));

    // decon_2D_tighterROI (induction) and decon_2D_tightROI
    float* y2 = nullptr;
    if (tighter_filter) {
        hipLaunchKernelGGL(k_mul_time_filter, g, b, 0, st, ws.C, tighter_filter, ws.S, nw, nt);
        WCT_HIP_CHECK(hipGetLastError());
        PLAIN_TRY(plain_c2r(ws.S, ws.Y2, nw, nt, st));
        y2 = ws.Y2;
    }
    hipLaunchKernelGGL(k_mul_time_filter, g, b, 0, st, ws.C, tight_filter, ws.S, nw, nt);
    WCT_HIP_CHECK(hipGetLastError());
    PLAIN_TRY(plain_c2r(ws.S, ws.Y, nw, nt, st));

    // restore_baseline, cal_RMS and the threshold mask, then save_data
    PLAIN_TRY(launch_plain_rows(ws.Y, y2, nw, nt, bad_ptr, bad_lo, bad_hi, th_factor, rms, mask,
                                ld_mask, st));
    hipLaunchKernelGGL(k_save, g, b, 0, st, ws.Y, out, ld_out, nw, nt, bad_ptr, bad_lo, bad_hi, save_mode);
    WCT_HIP_CHECK(hipGetLastError());
    return WCT_OK;
}

}  // extern "C"
