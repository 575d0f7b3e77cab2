// Forward simulation convolution of Gen::ImpactTransform
// (gen/src/ImpactTransform.cxx:19-347) on the device.
//
// The reference packs impact group i (real part) and its mirror group
// 2 nd - i (imaginary part, wire order reversed) into one complex frame,
// takes its 2D transform, multiplies by the 2D transform of the group's
// response frame (2P + 1 non-zero wire rows), inverts the wire transform,
// sums the groups, inverts the time transform and returns the real part
// plus the row-reversed imaginary part.  Here the packing, the time
// transforms and the output combination are the same, and the wire
// direction is done without transforms: the product of wire transforms
// followed by the inverse wire transform is the circular convolution
// over wires with the 2P + 1 response rows, so in the (wire, time
// frequency) domain every output element is
//   acc(w, k) = sum_g sum_{j=-P..P} X_g(w - j mod nw, k) Rt_g(j, k)
// (Rt_g(j, .) the time transform of the group's response at wire offset
// j), summed over the nd + 1 packed frames in one pass.  That is
// 4 (2P + 1)(nd + 1) FMAs per element on the vector ALUs, against seven
// wire transforms per column and two extra frame passes for the
// transform route, and no response frame is ever stored.
//
// Kernels: k_imp_place (response rows, zero-padded or cut to nt ticks),
// k_imp_scatter (charges into the packed frames, float atomics),
// the batched row transform of dft.hip (time, forward and inverse),
// k_imp_wire_conv (the convolution above, LDS tile with a 2P halo),
// k_imp_combine (real + row-reversed imaginary), k_imp_window and
// k_imp_mul_spec (ImpactTransform::waveform and its long-range
// response).
#include "common.hpp"
#include "gmem.hpp"

struct wct_impact_s {
    int device;
    int num_group, nwires_resp, P, nresp;
    float* resp = nullptr;        // [num_group][nwires_resp][nresp]
    float2* frames = nullptr;     // (nd + 1) packed frames, nw x nt each
    size_t frames_cap = 0;        // complex entries
    float2* acc = nullptr;        // nw x nt
    size_t acc_cap = 0;
    float2* taps = nullptr;       // (nd + 1) x (2P + 1) x nt
    size_t taps_cap = 0;
};

namespace wct {

namespace {

constexpr int kImpBlock = 256;
constexpr int kMaxGroups = 64;
// Wire-convolution tile: 64 time-frequency bins (one per lane) by
// kConvWaves x kConvOut wires.
constexpr int kConvKt = 64;
constexpr int kConvOut = 8;
constexpr int kConvWaves = kImpBlock / kConvKt;
constexpr int kConvWires = kConvWaves * kConvOut;

struct GroupOffsets {
    long off[kMaxGroups + 1];
};

// FFTBestLength.cxx:12-62: the smallest tabulated length >= n, or n
// above the table; keep_odd_even picks the table of n's parity.
const long kBestAny[] = {33, 35, 40, 42, 44, 56, 64, 66, 70, 77, 80, 96, 98, 100, 112, 121, 128, 140, 143,
                         160, 168, 169, 200, 224, 256, 280, 320, 400, 448, 512, 560, 640, 686, 768, 800, 847,
                         896, 1024, 1120, 1280, 1331, 1372, 1400, 1536, 1600, 1792, 2048, 2560, 2662, 2744,
                         3200, 3584, 4096, 4116, 4480, 4802, 5120, 5324, 5376, 5544, 5600, 5632, 5929, 6174,
                         6272, 6292, 6655, 7168, 7840, 8192, 8232, 8960, 9317, 10240, 10648, 11979, 14336,
                         15972, 16384};
const long kBestEven[] = {40, 42, 44, 56, 64, 66, 70, 72, 80, 96, 98, 100, 112, 120, 128, 140, 160, 168,
                          200, 224, 256, 280, 320, 400, 448, 512, 560, 640, 686, 768, 800, 896, 1024, 1120,
                          1280, 1372, 1400, 1536, 1600, 1792, 2048, 2560, 2662, 2744, 3200, 3584, 4096, 4116,
                          4480, 4802, 5120, 5324, 5376, 5544, 5600, 5632, 6174, 6272, 6292, 6468, 7168, 7840,
                          8192, 8232, 8960, 10240, 10648, 11200, 11858, 11880, 14336, 15972, 16384};
const long kBestOdd[] = {33, 35, 39, 49, 55, 63, 65, 77, 81, 91, 99, 121, 143, 169, 175, 187, 225, 231,
                         245, 275, 297, 343, 363, 375, 385, 405, 441, 455, 495, 539, 605, 637, 693, 735,
                         847, 875, 891, 945, 1001, 1089, 1331, 1573, 1859, 1875, 2197, 2299, 2431, 2541,
                         2625, 2695, 3025, 3267, 3993, 4095, 4125, 4719, 4725, 5005, 5145, 5445, 5929, 6655,
                         7007, 7623, 7865, 9317, 9801, 11979, 14157, 14175, 15379, 16335};

template <size_t N>
long best_from(const long (&tab)[N], long n)
{
    if (n <= tab[N - 1])
        for (long e : tab)
            if (e >= n) return e;
    return n;
}

long fft_best_length(long n, bool keep_odd_even)
{
    if (!keep_odd_even) return best_from(kBestAny, n);
    return n % 2 == 0 ? best_from(kBestEven, n) : best_from(kBestOdd, n);
}

// Response rows of groups 0..nd at offsets -P..P, each the first
// min(nt, nresp) samples of the waveform, zero beyond (:158-168).
__global__ __launch_bounds__(kImpBlock) void k_imp_place(const float* __restrict__ resp, float2* __restrict__ taps,
                                                         int nrows, int nresp, int nt)
{
    const long i = (long)blockIdx.x * kImpBlock + threadIdx.x;
    if (i >= (long)nrows * nt) return;
    const int r = (int)(i / nt), t = (int)(i % nt);
    float2 v;
    v.x = t < nresp ? G(resp)[(long)r * nresp + t] : 0.f;
    v.y = 0.f;
    G(taps)[i] = v;
}

// Charges into the packed frames (:126-144, :226-231): group g < nd into
// frame g's real part at row ch - frame_start_ch, its mirror 2 nd - g into
// frame g's imaginary part at row frame_end_ch - 1 - ch, the central
// group nd into frame nd's real part.  Groups above 2 nd (an even group
// count) are not used by the reference either.  Entries outside the
// frame are skipped.
__global__ __launch_bounds__(kImpBlock) void k_imp_scatter(float2* __restrict__ frames, int nw, int nt, int ch0,
                                                           int tick0, int nd, int ngroups, GroupOffsets go,
                                                           const int* __restrict__ ch,
                                                           const int* __restrict__ tick,
                                                           const double* __restrict__ q)
{
    const long i = (long)blockIdx.x * kImpBlock + threadIdx.x;
    if (i >= go.off[ngroups]) return;
    int g = 0;
    while (g + 1 < ngroups && go.off[g + 1] <= i) ++g;
    if (g > 2 * nd) return;
    const int c = G(ch)[i];
    const int col = G(tick)[i] - tick0;
    const float v = (float)G(q)[i];
    int frame, row;
    bool imag = false;
    if (g <= nd) {
        frame = g;
        row = c - ch0;
    } else {
        frame = 2 * nd - g;
        row = ch0 + nw - 1 - c;
        imag = true;
    }
    if (row < 0 || row >= nw || col < 0 || col >= nt) return;
    float* p = reinterpret_cast<float*>(frames + ((long)frame * nw + row) * nt + col) + (imag ? 1 : 0);
    unsafeAtomicAdd(p, v);
}

// acc(w, k) = sum_g sum_{j=-P..P} X_g(w - j mod nw, k) Rt_g(j, k).
// One workgroup: kConvKt bins (lane = bin) by kConvWires wires; each
// wave kConvOut consecutive wires.  Per frame the tile's rows plus a halo
// of P rows each side are staged in LDS (rows of kConvKt bins).
__global__ __launch_bounds__(kImpBlock) void k_imp_wire_conv(const float2* __restrict__ frames,
                                                             const float2* __restrict__ taps,
                                                             float2* __restrict__ acc, int nw, int nt,
                                                             int nframes, int P)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float2* tile = reinterpret_cast<float2*>(smem_raw);
    const int lane = threadIdx.x % kConvKt, wave = threadIdx.x / kConvKt;
    const int k = blockIdx.x * kConvKt + lane;
    const int w0 = blockIdx.y * kConvWires;
    const int nrow = kConvWires + 2 * P;
    const int ntap = 2 * P + 1;
    const bool kin = k < nt;
    float2 a[kConvOut];
#pragma unroll
    for (int o = 0; o < kConvOut; ++o) a[o] = make_float2(0.f, 0.f);
    for (int f = 0; f < nframes; ++f) {
        const float2* X = frames + (long)f * nw * nt;
        __syncthreads();
        for (int r = wave; r < nrow; r += kConvWaves) {
            int w = w0 - P + r;
            w %= nw;
            if (w < 0) w += nw;
            float2 v = make_float2(0.f, 0.f);
            if (kin) v = G(X)[(long)w * nt + k];
            tile[r * kConvKt + lane] = v;
        }
        __syncthreads();
        const float2* T = taps + (long)f * ntap * nt;
        // output wire w0 + wave*kConvOut + o takes tile row
        // wave*kConvOut + o + P - j for offset j
        for (int jj = 0; jj < ntap; ++jj) {
            float2 t = make_float2(0.f, 0.f);
            if (kin) t = G(T)[(long)jj * nt + k];
            const int base = wave * kConvOut + 2 * P - jj;   // tile row for o = 0
#pragma unroll
            for (int o = 0; o < kConvOut; ++o) {
                const float2 x = tile[(base + o) * kConvKt + lane];
                a[o].x = fmaf(x.x, t.x, a[o].x);
                a[o].x = fmaf(-x.y, t.y, a[o].x);
                a[o].y = fmaf(x.x, t.y, a[o].y);
                a[o].y = fmaf(x.y, t.x, a[o].y);
            }
        }
    }
    if (!kin) return;
#pragma unroll
    for (int o = 0; o < kConvOut; ++o) {
        const int w = w0 + wave * kConvOut + o;
        if (w < nw) G(acc)[(long)w * nt + k] = a[o];
    }
}

// The same with P fixed at compile time: the thread's 2P + 1 taps of a
// frame sit in registers and each staged row is read from LDS once and
// used for every output it reaches (kConvOut + 2P reads instead of
// kConvOut (2P + 1)).
template <int PP>
__global__ __launch_bounds__(kImpBlock) void k_imp_wire_conv_p(const float2* __restrict__ frames,
                                                               const float2* __restrict__ taps,
                                                               float2* __restrict__ acc, int nw, int nt,
                                                               int nframes)
{
    constexpr int NTAP = 2 * PP + 1, NROW = kConvWires + 2 * PP, NS = kConvOut + 2 * PP;
    __shared__ float2 tile[NROW * kConvKt];
    const int lane = threadIdx.x % kConvKt, wave = threadIdx.x / kConvKt;
    const int k = blockIdx.x * kConvKt + lane;
    const int w0 = blockIdx.y * kConvWires;
    const bool kin = k < nt;
    float2 a[kConvOut];
#pragma unroll
    for (int o = 0; o < kConvOut; ++o) a[o] = make_float2(0.f, 0.f);
    for (int f = 0; f < nframes; ++f) {
        const float2* X = frames + (long)f * nw * nt;
        __syncthreads();
        for (int r = wave; r < NROW; r += kConvWaves) {
            int w = w0 - PP + r;
            w %= nw;
            if (w < 0) w += nw;
            float2 v = make_float2(0.f, 0.f);
            if (kin) v = G(X)[(long)w * nt + k];
            tile[r * kConvKt + lane] = v;
        }
        float2 t[NTAP];
        const float2* T = taps + (long)f * NTAP * nt;
#pragma unroll
        for (int jj = 0; jj < NTAP; ++jj) {
            t[jj] = make_float2(0.f, 0.f);
            if (kin) t[jj] = G(T)[(long)jj * nt + k];
        }
        __syncthreads();
        // output o takes staged row wave*kConvOut + s with tap o + 2P - s
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const float2 x = tile[(wave * kConvOut + s) * kConvKt + lane];
#pragma unroll
            for (int o = 0; o < kConvOut; ++o) {
                const int jj = o + 2 * PP - s;
                if (jj >= 0 && jj < NTAP) {
                    a[o].x = fmaf(x.x, t[jj].x, a[o].x);
                    a[o].x = fmaf(-x.y, t[jj].y, a[o].x);
                    a[o].y = fmaf(x.x, t[jj].y, a[o].y);
                    a[o].y = fmaf(x.y, t[jj].x, a[o].y);
                }
            }
        }
    }
    if (!kin) return;
#pragma unroll
    for (int o = 0; o < kConvOut; ++o) {
        const int w = w0 + wave * kConvOut + o;
        if (w < nw) G(acc)[(long)w * nt + k] = a[o];
    }
}

// m_decon_data = real(acc) + imag(acc) with rows reversed (:304-306).
__global__ __launch_bounds__(kImpBlock) void k_imp_combine(const float2* __restrict__ acc, float* __restrict__ out,
                                                           int nw, int nt)
{
    const long i = (long)blockIdx.x * kImpBlock + threadIdx.x;
    if (i >= (long)nw * nt) return;
    const int w = (int)(i / nt), t = (int)(i % nt);
    G(out)[i] = G(acc)[i].x + G(acc)[(long)(nw - 1 - w) * nt + t].y;
}

// waveform(iwire) rows (:319-325): samples [0, nsamples) of the readout,
// the frame's values on [frame_start_tick, frame_end_tick), zero
// elsewhere, written as real rows of length ld (complex when cplx).
__global__ __launch_bounds__(kImpBlock) void k_imp_window(const float* __restrict__ decon, void* __restrict__ out,
                                                          int nw, int nt, int tick0, int nsamples, int ld,
                                                          int cplx)
{
    const long i = (long)blockIdx.x * kImpBlock + threadIdx.x;
    if (i >= (long)nw * ld) return;
    const int w = (int)(i / ld), s = (int)(i % ld);
    const int t = s - tick0;
    const float v = (s < nsamples && t >= 0 && t < nt) ? G(decon)[(long)w * nt + t] : 0.f;
    if (cplx)
        G(reinterpret_cast<float2*>(out))[i] = make_float2(v, 0.f);
    else
        G(reinterpret_cast<float*>(out))[i] = v;
}

// Long-range response, zero-padded or cut to n samples, as a complex row.
__global__ __launch_bounds__(kImpBlock) void k_imp_long_row(const float* __restrict__ aux, int naux,
                                                            float2* __restrict__ row, int n)
{
    const int i = blockIdx.x * kImpBlock + threadIdx.x;
    if (i >= n) return;
    G(row)[i] = make_float2(i < naux ? G(aux)[i] : 0.f, 0.f);
}

// Every row times the long-range spectrum (:336-340).
__global__ __launch_bounds__(kImpBlock) void k_imp_mul_spec(float2* __restrict__ x, const float2* __restrict__ spec,
                                                            int nw, int n)
{
    const long i = (long)blockIdx.x * kImpBlock + threadIdx.x;
    if (i >= (long)nw * n) return;
    const float2 a = G(x)[i], b = G(spec)[i % n];
    G(x)[i] = make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Real part of the first nsamples of every row (inv_c2r, :341-342).
__global__ __launch_bounds__(kImpBlock) void k_imp_real_head(const float2* __restrict__ x, float* __restrict__ out,
                                                             int nw, int n, int nsamples)
{
    const long i = (long)blockIdx.x * kImpBlock + threadIdx.x;
    if (i >= (long)nw * nsamples) return;
    const int w = (int)(i / nsamples), s = (int)(i % nsamples);
    G(out)[i] = G(x)[(long)w * n + s].x;
}

inline unsigned blocks_for(long n) { return (unsigned)((n + kImpBlock - 1) / kImpBlock); }

int grow(float2** p, size_t* cap, size_t need)
{
    if (*cap >= need) return WCT_OK;
    if (*p) WCT_HIP_CHECK(hipFree(*p));
    *p = nullptr;
    *cap = 0;
    WCT_TRY(dev_alloc(reinterpret_cast<void**>(p), need * sizeof(float2)));
    *cap = need;
    return WCT_OK;
}

int check_frame(const wct_impact_s* h, const wct_impact_frame* fr)
{
    WCT_REQUIRE(fr && fr->nw > 0 && fr->nt > 0, WCT_ERR_ARG, "impact: empty frame");
    WCT_REQUIRE(fr->nw >= 2 * h->P + 1, WCT_ERR_ARG, "impact: frame of %d wires is narrower than the %d response rows",
                fr->nw, 2 * h->P + 1);
    WCT_REQUIRE(fr->frame_end_ch - fr->frame_start_ch == fr->nw && fr->frame_end_tick - fr->frame_start_tick == fr->nt,
                WCT_ERR_ARG, "impact: inconsistent frame (channels [%d, %d) for %d wires, ticks [%d, %d) for %d)",
                fr->frame_start_ch, fr->frame_end_ch, fr->nw, fr->frame_start_tick, fr->frame_end_tick, fr->nt);
    WCT_REQUIRE(fr->nt <= kMaxDftLength, WCT_ERR_UNSUPPORTED, "impact: %d ticks above the transform limit", fr->nt);
    int dev = 0;
    WCT_HIP_CHECK(hipGetDevice(&dev));
    WCT_REQUIRE(dev == h->device, WCT_ERR_STATE, "impact: handle made on device %d used on device %d", h->device, dev);
    return WCT_OK;
}

}  // namespace

}  // namespace wct

using namespace wct;

extern "C" {

int wct_impact_geometry(int num_group, int nwires_resp, int impact_lo, int impact_hi, int time_lo, int time_hi,
                        int waveform_pad, wct_impact_frame* out)
{
    WCT_REQUIRE(out && num_group >= 2 && nwires_resp >= 1 && waveform_pad >= 0 && impact_hi >= impact_lo &&
                    time_hi >= time_lo,
                WCT_ERR_ARG, "impact geometry: bad arguments");
    const int P = (int)std::floor((nwires_resp - 1) / 2.0 + 0.5);
    int start_ch = (int)std::floor(impact_lo * 1.0 / (num_group - 1)) - 1;
    int end_ch = (int)std::ceil(impact_hi * 1.0 / (num_group - 1)) + 2;
    if ((end_ch - start_ch) % 2 != 0) end_ch += 1;
    const int start_tick = time_lo - 1;
    int end_tick = time_hi + 2;
    if ((end_tick - start_tick) % 2 != 0) end_tick += 1;
    const long nwt = fft_best_length((long)end_ch - start_ch + 2 * P, true);
    const int npad_wire = (int)((nwt - end_ch + start_ch) / 2);
    const long ntt = fft_best_length((long)end_tick - start_tick + waveform_pad, false);
    const int npad_time = (int)(ntt - end_tick + start_tick);
    out->start_ch = start_ch;
    out->end_ch = end_ch;
    out->npad_wire = npad_wire;
    out->frame_start_ch = start_ch - npad_wire;
    out->frame_end_ch = end_ch + npad_wire;
    out->frame_start_tick = start_tick;
    out->frame_end_tick = end_tick + npad_time;
    out->nw = end_ch - start_ch + 2 * npad_wire;
    out->nt = end_tick + npad_time - start_tick;
    return WCT_OK;
}

long wct_fft_best_length(long n, int keep_odd_even) { return fft_best_length(n, keep_odd_even != 0); }

int wct_impact_create(wct_impact* out, int num_group, int nwires_resp, const float* resp, int nresp, void* stream)
{
    WCT_ENTRY();
    WCT_REQUIRE(out && resp && num_group >= 1 && num_group <= kMaxGroups && nwires_resp >= 1 &&
                    nwires_resp % 2 == 1 && nresp >= 1,
                WCT_ERR_ARG, "impact create: bad arguments (num_group %d in [1, %d], odd nwires_resp %d, nresp %d)",
                num_group, kMaxGroups, nwires_resp, nresp);
    auto* h = new wct_impact_s();
    WCT_HIP_CHECK(hipGetDevice(&h->device));
    h->num_group = num_group;
    h->nwires_resp = nwires_resp;
    h->P = (nwires_resp - 1) / 2;
    h->nresp = nresp;
    const size_t bytes = (size_t)num_group * nwires_resp * nresp * sizeof(float);
    int rc = dev_alloc(reinterpret_cast<void**>(&h->resp), bytes);
    if (!rc) {
        const hipError_t e = hipMemcpyAsync(h->resp, resp, bytes, hipMemcpyDefault, as_stream(stream));
        if (e != hipSuccess) {
            set_error("impact create: copy of the responses failed: %s", hipGetErrorString(e));
            rc = WCT_ERR_HIP;
        }
    }
    if (rc) {
        wct_impact_destroy(h);
        return rc;
    }
    *out = h;
    return WCT_OK;
}

int wct_impact_destroy(wct_impact h)
{
    if (!h) return WCT_OK;
    int rc = WCT_OK;
    for (void* p : {(void*)h->resp, (void*)h->frames, (void*)h->acc, (void*)h->taps})
        if (p && hipFree(p) != hipSuccess) rc = WCT_ERR_HIP;
    delete h;
    return rc;
}

int wct_impact_run(wct_impact h, const wct_impact_frame* fr, const long* group_offsets, const int* ch,
                   const int* tick, const double* q, float* out, void* stream)
{
    WCT_ENTRY();
    WCT_REQUIRE(h && fr && group_offsets && out, WCT_ERR_ARG, "impact run: bad arguments");
    WCT_TRY(check_frame(h, fr));
    const int G = h->num_group, nd = (G - 1) / 2, P = h->P, ntap = 2 * P + 1;
    const int nw = fr->nw, nt = fr->nt;
    GroupOffsets go{};
    for (int g = 0; g <= G; ++g) {
        go.off[g] = group_offsets[g];
        WCT_REQUIRE(go.off[g] >= 0 && (g == 0 || go.off[g] >= go.off[g - 1]), WCT_ERR_ARG,
                    "impact run: group offsets must be non-decreasing from 0");
    }
    WCT_REQUIRE(go.off[0] == 0, WCT_ERR_ARG, "impact run: group offsets must start at 0");
    const long ncharge = go.off[G];
    WCT_REQUIRE(ncharge == 0 || (ch && tick && q), WCT_ERR_ARG, "impact run: missing charge arrays");
    hipStream_t s = as_stream(stream);
    const int nframes = nd + 1;
    const size_t fsz = (size_t)nw * nt;
    WCT_TRY(grow(&h->frames, &h->frames_cap, nframes * fsz));
    WCT_TRY(grow(&h->acc, &h->acc_cap, fsz));
    WCT_TRY(grow(&h->taps, &h->taps_cap, (size_t)nframes * ntap * nt));

    // response rows of groups 0..nd and their time transforms
    hipLaunchKernelGGL(k_imp_place, dim3(blocks_for((long)nframes * ntap * nt)), dim3(kImpBlock), 0, s, h->resp,
                       h->taps, nframes * ntap, h->nresp, nt);
    WCT_HIP_CHECK(hipGetLastError());
    WCT_TRY(launch_rows<float>(h->taps, h->taps, nt, nframes * ntap, nt, -1, 1.0f, s));

    // packed charge frames and their time transforms
    WCT_HIP_CHECK(hipMemsetAsync(h->frames, 0, nframes * fsz * sizeof(float2), s));
    if (ncharge > 0) {
        hipLaunchKernelGGL(k_imp_scatter, dim3(blocks_for(ncharge)), dim3(kImpBlock), 0, s, h->frames, nw, nt,
                           fr->frame_start_ch, fr->frame_start_tick, nd, G, go, ch, tick, q);
        WCT_HIP_CHECK(hipGetLastError());
    }
    WCT_TRY(launch_rows<float>(h->frames, h->frames, nt, nframes * nw, nt, -1, 1.0f, s));

    // wire convolution summed over the frames
    dim3 grid((nt + kConvKt - 1) / kConvKt, (nw + kConvWires - 1) / kConvWires);
#define WCT_CONV_P(PP)                                                                                  \
    case PP:                                                                                            \
        hipLaunchKernelGGL(k_imp_wire_conv_p<PP>, grid, dim3(kImpBlock), 0, s, h->frames, h->taps, h->acc, \
                           nw, nt, nframes);                                                            \
        break
    switch (P) {
        WCT_CONV_P(1);
        WCT_CONV_P(2);
        WCT_CONV_P(3);
        WCT_CONV_P(4);
        WCT_CONV_P(5);
        WCT_CONV_P(6);
        WCT_CONV_P(7);
        WCT_CONV_P(8);
        WCT_CONV_P(9);
        WCT_CONV_P(10);
        WCT_CONV_P(11);
        WCT_CONV_P(12);
    default: {
        const size_t lds = (size_t)(kConvWires + 2 * P) * kConvKt * sizeof(float2);
        WCT_TRY(allow_lds(k_imp_wire_conv, lds));
        hipLaunchKernelGGL(k_imp_wire_conv, grid, dim3(kImpBlock), lds, s, h->frames, h->taps, h->acc, nw, nt,
                           nframes, P);
    }
    }
#undef WCT_CONV_P
    WCT_HIP_CHECK(hipGetLastError());

    // inverse time transform and the output combination
    WCT_TRY(launch_rows<float>(h->acc, h->acc, nt, nw, nt, +1, 1.0f / nt, s));
    hipLaunchKernelGGL(k_imp_combine, dim3(blocks_for((long)fsz)), dim3(kImpBlock), 0, s, h->acc, out, nw, nt);
    WCT_HIP_CHECK(hipGetLastError());
    return WCT_OK;
}

int wct_impact_waveforms(wct_impact h, const wct_impact_frame* fr, const float* decon, int nsamples,
                         const float* long_aux, int nlong, int long_pad, float* out, void* stream)
{
    WCT_ENTRY();
    WCT_REQUIRE(h && fr && decon && out && nsamples >= 1 && nlong >= 0 && long_pad >= 0 && (nlong == 0 || long_aux),
                WCT_ERR_ARG, "impact waveforms: bad arguments");
    WCT_TRY(check_frame(h, fr));
    hipStream_t s = as_stream(stream);
    const int nw = fr->nw;
    if (nlong == 0) {
        hipLaunchKernelGGL(k_imp_window, dim3(blocks_for((long)nw * nsamples)), dim3(kImpBlock), 0, s, decon,
                           (void*)out, nw, fr->nt, fr->frame_start_tick, nsamples, nsamples, 0);
        WCT_HIP_CHECK(hipGetLastError());
        return WCT_OK;
    }
    const long n = fft_best_length((long)nsamples + long_pad, false);
    WCT_REQUIRE(n <= kMaxDftLength, WCT_ERR_UNSUPPORTED, "impact waveforms: %ld samples above the transform limit", n);
    DevTemp<float2> rows(s), spec(s);
    WCT_TRY(rows.alloc((size_t)nw * n));
    WCT_TRY(spec.alloc((size_t)n));
    hipLaunchKernelGGL(k_imp_window, dim3(blocks_for((long)nw * n)), dim3(kImpBlock), 0, s, decon, (void*)rows.p, nw,
                       fr->nt, fr->frame_start_tick, nsamples, (int)n, 1);
    WCT_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_imp_long_row, dim3(blocks_for(n)), dim3(kImpBlock), 0, s, long_aux, nlong, spec.p, (int)n);
    WCT_HIP_CHECK(hipGetLastError());
    WCT_TRY(launch_rows<float>(spec.p, spec.p, (int)n, 1, n, -1, 1.0f, s));
    WCT_TRY(launch_rows<float>(rows.p, rows.p, (int)n, nw, n, -1, 1.0f, s));
    hipLaunchKernelGGL(k_imp_mul_spec, dim3(blocks_for((long)nw * n)), dim3(kImpBlock), 0, s, rows.p, spec.p, nw,
                       (int)n);
    WCT_HIP_CHECK(hipGetLastError());
    WCT_TRY(launch_rows<float>(rows.p, rows.p, (int)n, nw, n, +1, 1.0f / n, s));
    hipLaunchKernelGGL(k_imp_real_head, dim3(blocks_for((long)nw * nsamples)), dim3(kImpBlock), 0, s, rows.p, out,
                       nw, (int)n, nsamples);
    WCT_HIP_CHECK(hipGetLastError());
    return WCT_OK;
}

}  // extern "C"
