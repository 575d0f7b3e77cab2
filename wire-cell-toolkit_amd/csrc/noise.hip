// Per-channel noise filter of ProtoDUNE-SP (Protodune::OneChannelNoise::apply,
// sigproc/src/Protodune.cxx:753-918) for every channel of a frame at once.
//
// Stages (one stream): k_nf_load (two channels' samples into one complex
// row, the first as real part, the second as imaginary part), forward
// row transforms (dft.hip), k_nf_spec (one workgroup per channel pair:
// each channel's spectrum separated from the pair's, X_a(k) = (Z(k) +
// Z*(n-k)) / 2, X_b(k) = (Z(k) - Z*(n-k)) / 2i; the partial-waveform
// test, the RC-undershoot division, the collection plane's harmonic
// removal, the DC bin, and the Hermitian mirror of DftTools::inv_c2r
// written back as Y_a + i Y_b), inverse row transforms (whose real and
// imaginary parts are then the two channels' real waveforms), k_nf_time
// (one workgroup per channel: baseline by binned median, the adaptive
// baseline of partial channels, the signal-protected RMS test, the flag
// removal).  Two channels per transform halve the transform work of a
// real-input c2c transform per channel, which the reference does.
//
// The per-channel algorithms of sigproc/src/Microboone.cxx are sequential
// loops in the reference; here each is a workgroup-parallel pass over the
// channel's samples in LDS with the same arithmetic: float histograms for
// the binned percentiles (three percentiles of CalcRMSWithFlags from one
// histogram), direct 21-sample window sums for RawAdapativeBaselineAlg
// and the nearest filled/flagged/edge bin for its interpolation, the
// 16-sample dilation of SignalFilter.
#include "common.hpp"
#include "gmem.hpp"

namespace wct {

namespace {

constexpr int kNfBlock = 512;   // 8 waves per channel: 16 per CU at two channels' LDS
constexpr int kNfWaves = kNfBlock / 64;
constexpr float kNfFlag = 10000.0f;

struct NfShared {
    double d[kNfWaves];
    float f[kNfWaves];
    int i[kNfWaves];
    float res[4];
    int flag;
};

__device__ __forceinline__ double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ __forceinline__ int wave_sum(int v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ __forceinline__ float wave_min(float v)
{
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ float wave_max(float v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

// Block reductions in a fixed order (deterministic).
__device__ double block_sum(double v, NfShared& sh)
{
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) sh.d[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0;
    for (int w = 0; w < kNfWaves; ++w) r += sh.d[w];
    __syncthreads();
    return r;
}
__device__ int block_sum(int v, NfShared& sh)
{
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) sh.i[threadIdx.x >> 6] = v;
    __syncthreads();
    int r = 0;
    for (int w = 0; w < kNfWaves; ++w) r += sh.i[w];
    __syncthreads();
    return r;
}
__device__ float block_min(float v, NfShared& sh)
{
    v = wave_min(v);
    if ((threadIdx.x & 63) == 0) sh.f[threadIdx.x >> 6] = v;
    __syncthreads();
    float r = sh.f[0];
    for (int w = 1; w < kNfWaves; ++w) r = fminf(r, sh.f[w]);
    __syncthreads();
    return r;
}
__device__ float block_max(float v, NfShared& sh)
{
    v = wave_max(v);
    if ((threadIdx.x & 63) == 0) sh.f[threadIdx.x >> 6] = v;
    __syncthreads();
    float r = sh.f[0];
    for (int w = 1; w < kNfWaves; ++w) r = fmaxf(r, sh.f[w]);
    __syncthreads();
    return r;
}

// Waveform::mean_rms (util/src/Waveform.cxx:19-37): double sums.
__device__ void mean_rms(const float* s, int n, double& mean, double& rms, NfShared& sh)
{
    double a = 0, b = 0;
    for (int i = threadIdx.x; i < n; i += kNfBlock) {
        const double x = s[i];
        a += x;
        b += x * x;
    }
    a = block_sum(a, sh);
    b = block_sum(b, sh);
    mean = a / n;
    rms = sqrt((b - a * a / n) / n);
}

// Waveform::percentile_binned (Waveform.cxx:108-131) over the samples
// of s that pass valid (the reference's vector of those samples), for np
// probabilities from one histogram: m = number of such samples bins of
// (max - min) / m between their min and max (float), sample bin
// round((v - min) / width) clamped to [0, m), the lower edge of the bin
// where the running count first exceeds int(m p).  hist holds m ints.
// Returns the count m (0: no samples, out untouched).
template <class Valid>
__device__ int pct_binned(const float* s, int n, Valid valid, const float* p, int np, float* out, int* hist,
                          NfShared& sh)
{
    int cnt = 0;
    float mn = INFINITY, mx = -INFINITY;
    for (int i = threadIdx.x; i < n; i += kNfBlock) {
        const float v = s[i];
        if (valid(v)) {
            ++cnt;
            mn = fminf(mn, v);
            mx = fmaxf(mx, v);
        }
    }
    const int m = block_sum(cnt, sh);
    if (m == 0) return 0;
    mn = block_min(mn, sh);
    mx = block_max(mx, sh);
    const float width = __fdiv_rn(mx - mn, (float)m);
    for (int i = threadIdx.x; i < m; i += kNfBlock) hist[i] = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += kNfBlock) {
        const float v = s[i];
        if (valid(v)) {
            const float r = __fdiv_rn(v - mn, width);
            int b = (r != r) ? 0 : (int)roundf(r);   // 0/0 (all equal): bin 0, as x86's INT_MIN clamps
            b = max(0, min(m - 1, b));
            atomicAdd(&hist[b], 1);
        }
    }
    __syncthreads();
    const int C = (m + kNfBlock - 1) / kNfBlock;
    const int lo = min(m, threadIdx.x * C), hi = min(m, lo + C);
    int csum = 0;
    for (int i = lo; i < hi; ++i) csum += hist[i];
    // exclusive prefix of the chunk sums: scans within waves, then the
    // totals of the waves before this one
    int incl = csum;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(incl, o);
        if ((threadIdx.x & 63) >= o) incl += y;
    }
    if ((threadIdx.x & 63) == 63) sh.i[threadIdx.x >> 6] = incl;
    __syncthreads();
    int before = incl - csum;
    for (int w = 0; w < (int)(threadIdx.x >> 6); ++w) before += sh.i[w];
    __syncthreads();
    for (int k = 0; k < np; ++k) {
        const int imed = (int)((float)m * p[k]);
        if (threadIdx.x == 0) sh.res[k] = mn + (mx - mn) * p[k];   // not reached for p < 1
        __syncthreads();
        if (before <= imed && before + csum > imed) {
            int c = before;
            for (int i = lo; i < hi; ++i) {
                c += hist[i];
                if (c > imed) {
                    sh.res[k] = mn + (float)i * width;
                    break;
                }
            }
        }
        __syncthreads();
        out[k] = sh.res[k];
        __syncthreads();
    }
    return m;
}

struct BelowAdcMax {
    __device__ bool operator()(float v) const { return v < 4096.0f; }
};
struct AnySample {
    __device__ bool operator()(float) const { return true; }
};

// Microboone::CalcRMSWithFlags (Microboone.cxx:545-566).
__device__ float rms_with_flags(const float* s, int n, int* hist, NfShared& sh)
{
    const float p[3] = {(float)(0.5 - 0.34), 0.5f, (float)(0.5 + 0.34)};
    float q[3];
    if (!pct_binned(s, n, BelowAdcMax(), p, 3, q, hist, sh)) return 0.f;
    const double a = q[2] - q[1], b = q[1] - q[0];
    return (float)sqrt((a * a + b * b) / 2.);
}

// Microboone::SignalFilter (:568-610): samples within [i - 8, i + 8) of a
// sample beyond 4 RMS (and below 4096) get 20000 added, once.
__device__ void signal_filter(float* s, int n, int* hist, unsigned char* fl, NfShared& sh)
{
    const float thr = (float)(4.0 * rms_with_flags(s, n, hist, sh));
    for (int i = threadIdx.x; i < n; i += kNfBlock) {
        const float v = s[i];
        fl[i] = ((v > thr) || ((double)v < -1.0 * (double)thr)) && v < 4096.0f;
    }
    __syncthreads();
    for (int j = threadIdx.x; j < n; j += kNfBlock) {
        const float v = s[j];
        if (v < 4096.0f) {
            bool near = false;
#pragma unroll
            for (int d = -7; d <= 8; ++d) {
                const int i = j + d;
                near |= i >= 0 && i < n && fl[i] != 0;
            }
            if (near) s[j] = (float)((double)v + 20000.0);
        }
    }
    __syncthreads();
}

// Microboone::RawAdapativeBaselineAlg (:612-737): subtract the mean of the
// samples below 4096 in [j - 10, j + 10]; where that window holds fewer
// than 10 of them, interpolate (with the reference's weights) between the
// nearest bins d <= j <= u that are filled, at an edge or flagged
// (10000), or take the one that is filled.  base: n doubles; fl: n bytes.
__device__ void raw_adaptive_baseline(float* s, int n, double* base, unsigned char* fl, NfShared& sh)
{
    int nflag = 0;
    for (int i = threadIdx.x; i < n; i += kNfBlock) nflag += s[i] == kNfFlag;
    if (block_sum(nflag, sh) == n) return;
    for (int j = threadIdx.x; j < n; j += kNfBlock) {
        double a = 0;
        int c = 0;
#pragma unroll
        for (int d = -10; d <= 10; ++d) {
            const int k = j + d;
            const float v = (k >= 0 && k < n) ? s[k] : 4096.0f;
            if (v < 4096.0f) {
                a += v;
                ++c;
            }
        }
        base[j] = c ? a / c : 0.0;
        fl[j] = (c >= 10 ? 1 : 0) | (s[j] == kNfFlag ? 2 : 0);
    }
    __syncthreads();
    for (int j = threadIdx.x; j < n; j += kNfBlock) {
        if (fl[j] & 2) continue;
        double b = base[j];
        if (!(fl[j] & 1)) {
            int d = j;
            while (!(fl[d] & 1) && d > 0 && !(fl[d] & 2)) --d;
            int u = j;
            while (!(fl[u] & 1) && u < n - 1 && !(fl[u] & 2)) ++u;
            const bool dn = !(fl[d] & 1), up = !(fl[u] & 1);
            if (!dn && !up)
                b = ((j - d) * base[d] + (u - j) * base[u]) / ((double)u - d);
            else if (dn && !up)
                b = base[u];
            else if (!dn && up)
                b = base[d];
            else
                b = 0.0;
        }
        s[j] = (float)((double)s[j] - b);
    }
    __syncthreads();
}

// Microboone::RemoveFilterFlags (:739-757).
__device__ void remove_filter_flags(float* s, int n)
{
    for (int i = threadIdx.x; i < n; i += kNfBlock) {
        const float v = s[i];
        if (v > 4096.0f) s[i] = v > kNfFlag ? (float)((double)v - 20000.0) : 0.f;
    }
    __syncthreads();
}

// LDS of the per-channel passes: n floats (the samples), 8 n bytes of
// scratch (doubles of the adaptive baseline, or a float copy plus an int
// histogram), n flag bytes.
__host__ __device__ __forceinline__ size_t nf_lds_bytes(int n) { return (size_t)13 * n + 32; }
__device__ __forceinline__ void nf_lds(unsigned char* base, int n, float*& s, unsigned char*& scratch,
                                       unsigned char*& fl)
{
    s = reinterpret_cast<float*>(base);
    scratch = base + ((size_t)4 * n + 15) / 16 * 16;
    fl = scratch + (size_t)8 * n;
}

// A channel's sample as its filter starts from it: MicroBooNE's filter
// first subtracts the nominal baseline (Microboone.cxx:914-918, float).
__device__ __forceinline__ float nf_input(const float* in, int ld_in, const wct_nf_channel& ch, int c, int t)
{
    const float x = G(in)[(long)c * ld_in + t];
    return ch.mode == WCT_NF_MICROBOONE ? x + ch.baseline * (-1) : x;
}

__global__ __launch_bounds__(kNfBlock) void k_nf_load(const float* __restrict__ in, int ld_in,
                                                      const wct_nf_channel* __restrict__ chans,
                                                      float2* __restrict__ rows, int nch, int n)
{
    const int npair = (nch + 1) / 2;
    const long i = (long)blockIdx.x * kNfBlock + threadIdx.x;
    if (i >= (long)npair * n) return;
    const int p = (int)(i / n), t = (int)(i % n);
    const float a = nf_input(in, ld_in, chans[2 * p], 2 * p, t);
    const float b = 2 * p + 1 < nch ? nf_input(in, ld_in, chans[2 * p + 1], 2 * p + 1, t) : 0.f;
    G(rows)[i] = make_float2(a, b);
}

// Diagnostics::Chirp (sigproc/src/Diagnostics.cxx:33-137) on the
// gain-corrected samples of MicroBooNE channels: the RMS of each complete
// 20-sample window (double) in parallel, the run bookkeeping over the
// windows on one lane.  Writes every channel's three flag ints.
constexpr int kChirpWindow = 20;
__global__ __launch_bounds__(kNfBlock) void k_nf_chirp(const float* __restrict__ in, int ld_in, int n,
                                                       const wct_nf_channel* __restrict__ chans,
                                                       int* __restrict__ flags)
{
    extern __shared__ __align__(16) unsigned char nf_smem[];
    double* wrms = reinterpret_cast<double*>(nf_smem);
    const int c = blockIdx.x;
    const wct_nf_channel ch = chans[c];
    if (ch.mode != WCT_NF_MICROBOONE) {
        if (threadIdx.x == 0) {
            flags[3 * c] = 0;
            flags[3 * c + 1] = -1;
            flags[3 * c + 2] = -1;
        }
        return;
    }
    const int nwin = n / kChirpWindow;
    for (int w = threadIdx.x; w < nwin; w += kNfBlock) {
        double m = 0, m2 = 0;
        for (int k = 0; k < kChirpWindow; ++k) {
            const double x = nf_input(in, ld_in, ch, c, w * kChirpWindow + k) * ch.gain;
            m += x;
            m2 += x * x;
        }
        m /= kChirpWindow;
        m2 /= kChirpWindow;
        wrms[w] = sqrt(m2 - m * m);
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    const double min_rms = 0.9, max_frac = 0.20;
    int num_low = 0, first = -1, last = -1, normal = 0;
    bool low = false;
    double r1 = 0, r2 = 0, r3 = 0;
    for (int w = 0; w < nwin; ++w) {
        const int ibin = w * kChirpWindow + kChirpWindow - 1;
        r1 = r2;
        r2 = r3;
        r3 = wrms[w];
        if (r3 < min_rms) ++num_low;
        if (ibin >= 3 * kChirpWindow - 1) {
            if (r2 < min_rms && (r1 > min_rms || r3 > min_rms)) ++normal;
            if (!low) {
                if (r2 < min_rms && r3 < min_rms) {
                    low = true;
                    first = ibin - 2 * kChirpWindow + 1;
                    last = ibin - kChirpWindow + 1;
                }
                if (ibin == 3 * kChirpWindow - 1 && r1 < min_rms && r2 < min_rms) {
                    low = true;
                    first = ibin - 3 * kChirpWindow + 1;
                    last = ibin - 2 * kChirpWindow + 1;
                }
            } else if (r2 < min_rms && r3 < min_rms) {
                last = ibin - kChirpWindow + 1;
            }
        }
    }
    const double chirp_frac = (double)num_low / ((double)n / (double)kChirpWindow);
    const double normal_frac = (double)normal / (double)num_low;
    int bits = 0, beg = -1, end = -1;
    if (num_low > 4 && (normal_frac < max_frac ||
                        (num_low < 2.0 / max_frac && last - first == num_low * kChirpWindow))) {
        first = max(1, first - kChirpWindow);
        last = min(n, last + 2 * kChirpWindow);
        if (n - last < kChirpWindow) last = n;
        if (chirp_frac > 0.99) {
            first = 1;
            last = n;
        }
        bits = WCT_NF_CHIRP;
        beg = first - 1;
        end = last;
    }
    flags[3 * c] = bits;
    flags[3 * c + 1] = beg;
    flags[3 * c + 2] = end;
}

// One channel's bin k of the pair row Z: X_a(k) = (Z(k) + Z*(n-k)) / 2
// (which = 0) or X_b(k) = (Z(k) - Z*(n-k)) / 2i (which = 1).
template <typename P>
__device__ __forceinline__ float2 unpack_bin(P row, int n, int k, int which)
{
    const float2 z = row[k], w = row[k ? n - k : 0];
    return which == 0 ? make_float2(0.5f * (z.x + w.x), 0.5f * (z.y - w.y))
                      : make_float2(0.5f * (z.y + w.y), -0.5f * (z.x - w.x));
}

// a / b, rounded from double.
__device__ __forceinline__ float2 cdiv(float2 a, float2 b)
{
    const double den = (double)b.x * b.x + (double)b.y * b.y;
    return make_float2((float)(((double)a.x * b.x + (double)a.y * b.y) / den),
                       (float)(((double)a.y * b.x - (double)a.x * b.y) / den));
}

__device__ __forceinline__ float2 cmulf(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// LDS of the spectrum stage when some channel has flagged runs: the
// magnitudes, the adaptive-baseline scratch, and per channel of the pair
// a flag per lower-half bin that the harmonic removal zeroed.
__host__ __device__ __forceinline__ size_t nf_spec_lds_bytes(int n) { return (size_t)13 * n + 2 * (n / 2 + 1) + 64; }

// Spectrum stage of one channel pair (Protodune.cxx:767-876).
__global__ __launch_bounds__(kNfBlock) void k_nf_spec(float2* __restrict__ rows, int n, int nch,
                                                      const wct_nf_channel* __restrict__ chans,
                                                      const float2* __restrict__ rcrc, const int* __restrict__ bins,
                                                      int* __restrict__ flags)
{
    extern __shared__ __align__(16) unsigned char nf_smem[];
    __shared__ NfShared sh;
    __shared__ int part[2], nharm[2];
    const int p = blockIdx.x;
    const int nc = 2 * p + 1 < nch ? 2 : 1;
    wct_nf_channel ch[2];
    ch[0] = chans[2 * p];
    ch[1] = nc == 2 ? chans[2 * p + 1] : ch[0];
    const auto row = G(rows + (long)p * n);
    // Diagnostics::Partial (sigproc/src/Diagnostics.cxx:15-31), nfreqs 4,
    // maxpower 6000, one lane per channel
    if (threadIdx.x < 2) {
        const int c = threadIdx.x;
        const float2 z1 = unpack_bin(row, n, 1, c);
        const double mag0 = hypotf(z1.x, z1.y);
        double sum = mag0;
        int partial = 1;
        for (int ind = 1; ind <= 4 && ind < n; ++ind) {
            const float2 z = unpack_bin(row, n, ind + 1, c);
            const double magi = hypotf(z.x, z.y);
            if (mag0 <= magi) {
                partial = 0;
                break;
            }
            sum += magi;
        }
        part[c] = partial && sum / 5 > 6000.0;
        nharm[c] = 0;
    }
    __syncthreads();
    const bool divide[2] = {!part[0] && ch[0].rcrc >= 0, !part[1] && ch[1].rcrc >= 0};
    const int half = n / 2;
    unsigned char* zf = nullptr;
    bool harm[2] = {false, false};
    for (int c = 0; c < nc; ++c)
        harm[c] = bins && ch[c].mode == WCT_NF_PROTODUNE && ch[c].plane == 2 && ch[c].bins_end > ch[c].bins_begin;
    if (harm[0] || harm[1]) {
        float* mag;
        unsigned char *scratch, *fl;
        nf_lds(nf_smem, n, mag, scratch, fl);
        zf = fl + n;
        for (int k = threadIdx.x; k < 2 * (half + 1); k += kNfBlock) zf[k] = 0;
        for (int c = 0; c < nc; ++c) {
            if (!harm[c]) continue;
            __syncthreads();
            const auto R = G(rcrc + (long)max(ch[c].rcrc, 0) * n);
            for (int k = threadIdx.x; k < n; k += kNfBlock) {
                float2 x = unpack_bin(row, n, k, c);
                if (divide[c]) x = cdiv(x, R[k]);
                mag[k] = k == 0 ? 0.f : hypotf(x.x, x.y);
            }
            __syncthreads();
            raw_adaptive_baseline(mag, n, reinterpret_cast<double*>(scratch), fl, sh);
            // the flagged runs, five rounds (:808-846), on one lane: the
            // zeroed mirror bins feed later runs' statistics.  Only bins
            // 0..n/2 reach the output (the mirror replaces the rest), so a
            // zeroing is recorded there: bin j itself, or its partner
            // n + 1 - j when that one is in the lower half.
            if (threadIdx.x == 0) {
                unsigned char* z = zf + c * (half + 1);
                for (int iter = 0; iter < 5; ++iter) {
                    for (int r = ch[c].bins_begin; r < ch[c].bins_end; ++r) {
                        const int lo = G(bins)[2 * r], iend = G(bins)[2 * r + 1] + 1;
                        double a = 0, b = 0;
                        for (int j = lo; j < iend; ++j) {
                            a += mag[j];
                            b += (double)mag[j] * mag[j];
                        }
                        const int m = iend - lo;
                        const double mean = m == 1 ? (double)mag[lo] : (m ? a / m : 0.0);
                        const double rms = m > 1 ? sqrt((b - a * a / m) / m) : 0.0;
                        const double cut = lo > 1050 ? mean + 3 * rms : mean + 2.7 * rms;
                        for (int j = lo; j < iend; ++j) {
                            if (mag[j] > cut) {
                                const int jm = n + 1 - j;
                                mag[j] = 0;
                                if (j <= half) z[j] = 1;
                                if (jm >= 0 && jm < n) {
                                    mag[jm] = 0;
                                    if (jm <= half) z[jm] = 1;
                                }
                                ++nharm[c];
                            }
                        }
                    }
                }
            }
        }
    }
    __syncthreads();
    // per bin pair (k, n - k) on one thread: the channels' corrected bins,
    // DC removed (:880), mirrored as inv_c2r mirrors them
    const auto Ra = G(rcrc + (long)max(ch[0].rcrc, 0) * n);
    const auto Rb = G(rcrc + (long)max(ch[1].rcrc, 0) * n);
    // MicroBooNE: the config and noise spectra multiply the spectrum
    // (Microboone.cxx:958-974); only bins 0..n/2 reach the output
    const bool mb[2] = {ch[0].mode == WCT_NF_MICROBOONE, ch[1].mode == WCT_NF_MICROBOONE};
    const auto Ca = G(rcrc + (long)max(ch[0].config, 0) * n), Cb = G(rcrc + (long)max(ch[1].config, 0) * n);
    const auto Na = G(rcrc + (long)max(ch[0].noise, 0) * n), Nb = G(rcrc + (long)max(ch[1].noise, 0) * n);
    for (int k = threadIdx.x; k <= half; k += kNfBlock) {
        float2 xa = unpack_bin(row, n, k, 0), xb = unpack_bin(row, n, k, 1);
        if (divide[0]) xa = cdiv(xa, Ra[k]);
        if (divide[1]) xb = cdiv(xb, Rb[k]);
        if (mb[0] && ch[0].config >= 0) xa = cmulf(xa, Ca[k]);
        if (mb[0] && ch[0].noise >= 0) xa = cmulf(xa, Na[k]);
        if (mb[1] && ch[1].config >= 0) xb = cmulf(xb, Cb[k]);
        if (mb[1] && ch[1].noise >= 0) xb = cmulf(xb, Nb[k]);
        if (zf && harm[0] && zf[k]) xa = make_float2(0.f, 0.f);
        if (zf && harm[1] && zf[half + 1 + k]) xb = make_float2(0.f, 0.f);
        if (nc == 1) xb = make_float2(0.f, 0.f);
        if (k == 0) {
            row[0] = make_float2(0.f, 0.f);
        } else if (2 * k == n) {
            row[k] = make_float2(xa.x, xb.x);
        } else {
            row[k] = make_float2(xa.x - xb.y, xa.y + xb.x);
            row[n - k] = make_float2(xa.x + xb.y, xb.x - xa.y);
        }
    }
    if (threadIdx.x < nc)
        flags[3 * (2 * p + threadIdx.x)] |=
            (part[threadIdx.x] ? WCT_NF_PARTIAL : 0) | (nharm[threadIdx.x] > 5 ? WCT_NF_HARMONIC : 0);
}

// Time stage of one channel (Protodune.cxx:877-915).  rows hold the
// inverse transforms (real parts used).
__global__ __launch_bounds__(kNfBlock) void k_nf_time(const float2* __restrict__ rows, int n, float* __restrict__ out,
                                                      int ld_out, const wct_nf_channel* __restrict__ chans,
                                                      int* __restrict__ flags)
{
    extern __shared__ __align__(16) unsigned char nf_smem[];
    __shared__ NfShared sh;
    const int c = blockIdx.x;
    const wct_nf_channel ch = chans[c];
    float* s;
    unsigned char *scratch, *fl;
    nf_lds(nf_smem, n, s, scratch, fl);
    const auto row = G(reinterpret_cast<const float*>(rows + (long)(c / 2) * n) + (c & 1));
    for (int i = threadIdx.x; i < n; i += kNfBlock) s[i] = row[2 * i];
    __syncthreads();
    int flag = flags[3 * c];
    // baseline: binned median of the samples with those beyond 6 RMS set
    // to the mean (:880-890)
    double mean, rms;
    mean_rms(s, n, mean, rms, sh);
    float* temp = reinterpret_cast<float*>(scratch);
    int* hist = reinterpret_cast<int*>(scratch + (size_t)4 * n);
    for (int i = threadIdx.x; i < n; i += kNfBlock) {
        const float v = s[i];
        temp[i] = fabs((double)v - mean) > 6 * rms ? (float)mean : v;
    }
    __syncthreads();
    const float half = 0.5f;
    float baseline = 0.f;
    pct_binned(temp, n, AnySample(), &half, 1, &baseline, hist, sh);
    const float shift = baseline * (-1);
    for (int i = threadIdx.x; i < n; i += kNfBlock) s[i] = s[i] + shift;
    __syncthreads();
    int* hist2 = reinterpret_cast<int*>(scratch);
    if (ch.mode == WCT_NF_MICROBOONE) {
        // chirped samples raised to the flag, then the adaptive baseline
        // (Microboone.cxx:1000-1024; no flag removal until the end)
        if (flag & WCT_NF_CHIRP) {
            const int b0 = max(flags[3 * c + 1], 0), b1 = min(flags[3 * c + 2], n);
            for (int i = b0 + (int)threadIdx.x; i < b1; i += kNfBlock) s[i] = kNfFlag;
            __syncthreads();
            signal_filter(s, n, hist2, fl, sh);
            raw_adaptive_baseline(s, n, reinterpret_cast<double*>(scratch), fl, sh);
        }
        if (flag & WCT_NF_PARTIAL) {
            signal_filter(s, n, hist2, fl, sh);
            raw_adaptive_baseline(s, n, reinterpret_cast<double*>(scratch), fl, sh);
        }
    } else if (flag & WCT_NF_PARTIAL) {
        signal_filter(s, n, hist2, fl, sh);
        raw_adaptive_baseline(s, n, reinterpret_cast<double*>(scratch), fl, sh);
        remove_filter_flags(s, n);
    }
    signal_filter(s, n, hist2, fl, sh);
    // Microboone::NoisyFilterAlg (:464-533)
    const double r = rms_with_flags(s, n, hist2, sh);
    const bool noisy = r > (double)ch.max_rms || r < (double)ch.min_rms;
    if (noisy)
        for (int i = threadIdx.x; i < n; i += kNfBlock) s[i] = kNfFlag;
    __syncthreads();
    remove_filter_flags(s, n);
    const auto dst = G(out + (long)c * ld_out);
    for (int i = threadIdx.x; i < n; i += kNfBlock) dst[i] = s[i];
    if (threadIdx.x == 0) flags[3 * c] = flag | (noisy ? WCT_NF_NOISY : 0);
}

// SimpleRC samples at t = i tick, rounded to float as the reference's
// realseq_t holds them (util/src/Response.cxx:517-536, :267-276), as a
// double complex row.
__global__ __launch_bounds__(kNfBlock) void k_nf_rc_samples(double2* __restrict__ row, int n, double width,
                                                            double tick)
{
    const int i = blockIdx.x * kNfBlock + threadIdx.x;
    if (i >= n) return;
    const double t = i * ((n * tick) / n);
    double v = 0;
    if (width > 0) v += -tick / width * exp(-t / width);
    if (t < tick) v += 1.0;
    G(row)[i] = make_double2((double)(float)v, 0.0);
}

// The float spectrum raised to the power layers (OmniChannelNoiseDB.cxx:187-199).
__global__ __launch_bounds__(kNfBlock) void k_nf_rc_power(const double2* __restrict__ spec, float2* __restrict__ out,
                                                          int n, int layers)
{
    const int i = blockIdx.x * kNfBlock + threadIdx.x;
    if (i >= n) return;
    const double2 z = G(spec)[i];
    const float2 s1 = make_float2((float)z.x, (float)z.y);
    float2 acc = s1;
    for (int l = 1; l < layers; ++l) acc = make_float2(acc.x * s1.x - acc.y * s1.y, acc.x * s1.y + acc.y * s1.x);
    G(out)[i] = acc;
}

// Protodune::RelGainCalib::apply (Protodune.cxx:975-1009).
__global__ __launch_bounds__(kNfBlock) void k_nf_rel_gain(float* __restrict__ rows, int ld, int nch, int n,
                                                          const float* __restrict__ gain, float gmin, float gmax)
{
    const long i = (long)blockIdx.x * kNfBlock + threadIdx.x;
    if (i >= (long)nch * n) return;
    const int c = (int)(i / n), t = (int)(i % n);
    const float g = G(gain)[c];
    if (g > gmax || g < gmin) return;
    float* p = rows + (long)c * ld + t;
    const float v = g * (*p - 0.f);
    *p = v + 0.f;
}

// ---- MicroBooNE's coherent noise subtraction (Microboone.cxx:53-872) ----

// Derivations::CalcRMS (sigproc/src/Derivations.cxx:7-21): mean_rms, then
// mean_rms of the samples within 4.5 RMS of that mean.
__device__ void calc_rms(const float* s, int n, double& mean, double& rms, NfShared& sh)
{
    double m0, r0;
    mean_rms(s, n, m0, r0, sh);
    double a = 0, b = 0;
    int c = 0;
    for (int i = threadIdx.x; i < n; i += kNfBlock) {
        const double x = s[i];
        if (fabs(x - m0) < 4.5 * r0) {
            a += x;
            b += x * x;
            ++c;
        }
    }
    a = block_sum(a, sh);
    b = block_sum(b, sh);
    c = block_sum(c, sh);
    if (c == 0) {
        mean = rms = 0;
    } else if (c == 1) {
        mean = a;
        rms = 0;
    } else {
        mean = a / c;
        rms = sqrt((b - a * a / c) / c);
    }
}

// filter_time, filter_low, filter_low_loose (Microboone.cxx:31-51) at the
// frequency of bin i of n (2 MHz digitisation, :164-171), as the float
// factor the reference multiplies by.
__device__ __forceinline__ double cn_freq(int i, int n)
{
    return i < n / 2. ? i / (1. * n) * 2. : (n - i) / (1. * n) * 2.;
}
__device__ __forceinline__ double cn_filter_time(double f)
{
    return (f > 0) * exp(-0.5 * pow(f / 0.143555, 4.95096));
}
__device__ __forceinline__ double cn_filter_low(double f, double cut)
{
    if ((f > 0.177 && f < 0.18) || (f > 0.2143 && f < 0.215) || (f >= 0.106 && f <= 0.109) || (f > 0.25 && f < 0.251))
        return 0;
    return 1 - exp(-pow(f / cut, 8));
}
__device__ __forceinline__ double cn_filter_low_loose(double f) { return 1 - exp(-pow(f / 0.005, 2)); }

constexpr int kCnTicks = 128;        // ticks per median workgroup
constexpr int kMaxGroupRows = 128;   // channels per coherent group

// Per-channel mean_rms and CalcRMS (for the group's mean RMS and the
// coefficient's 4 RMS window).
__global__ __launch_bounds__(kNfBlock) void k_cn_rms(const float* __restrict__ rows, int ld, int n,
                                                     double* __restrict__ rms0, double* __restrict__ crms)
{
    extern __shared__ __align__(16) unsigned char nf_smem[];
    __shared__ NfShared sh;
    float* s = reinterpret_cast<float*>(nf_smem);
    const int c = blockIdx.x;
    const auto src = G(rows + (long)c * ld);
    for (int i = threadIdx.x; i < n; i += kNfBlock) s[i] = src[i];
    __syncthreads();
    double m, r;
    mean_rms(s, n, m, r, sh);
    double cm, cr;
    calc_rms(s, n, cm, cr, sh);
    if (threadIdx.x == 0) {
        rms0[c] = r;
        crms[c] = cr;
    }
}

// Derivations::CalcMedian (:23-69) per group and tick: the group's mean
// RMS accumulated in float in channel order, then per tick the binned
// median of the samples within 5 of it and above 0.001 in size, found as
// the int(m / 2)-th smallest bin index (the bin where the running count
// first passes it).
__global__ __launch_bounds__(kCnTicks) void k_cn_median(const float* __restrict__ rows, int ld, int n,
                                                        const wct_nf_group* __restrict__ groups,
                                                        const int* __restrict__ grows,
                                                        const double* __restrict__ rms0, float* __restrict__ med)
{
    const int g = blockIdx.x;
    const wct_nf_group grp = groups[g];
    const int t = blockIdx.y * kCnTicks + threadIdx.x;
    float maxrms = 0.f, count = 0.f;
    for (int k = 0; k < grp.count; ++k) {
        const double r = rms0[grows[grp.first + k]];
        if (r > 0) {
            maxrms = (float)(maxrms + r);
            count = count + 1.f;
        }
    }
    if (count > 0) maxrms = maxrms / count;
    const float lim = 5 * maxrms;
    if (t >= n) return;
    const int* gr = grows + grp.first;
    const int cnt = min(grp.count, kMaxGroupRows);   // the documented limit
    int m = 0;
    float vmin = INFINITY, vmax = -INFINITY;
    for (int k = 0; k < cnt; ++k) {
        const float v = G(rows + (long)gr[k] * ld)[t];
        if (fabsf(v) < lim && fabs((double)v) > 0.001) {
            ++m;
            vmin = fminf(vmin, v);
            vmax = fmaxf(vmax, v);
        }
    }
    float out = 0.f;
    if (m > 0) {
        // bin indices as bytes, four to a register; 127 marks a sample
        // outside the selection (and pads the last word): no search
        // threshold below exceeds m - 1 <= 127, so it is never counted
        constexpr int kWords = kMaxGroupRows / 4;
        const int nw = (cnt + 3) / 4;
        const float width = __fdiv_rn(vmax - vmin, (float)m);
        unsigned packed[kWords];
#pragma unroll
        for (int w = 0; w < kWords; ++w) {
            unsigned word = 0x7f7f7f7fu;
            if (w < nw) {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int k = 4 * w + j;
                    if (k < cnt) {
                        const float v = G(rows + (long)gr[k] * ld)[t];
                        unsigned b = 127;
                        if (fabsf(v) < lim && fabs((double)v) > 0.001) {
                            const float r = __fdiv_rn(v - vmin, width);
                            const int bi = (r != r) ? 0 : (int)roundf(r);
                            b = (unsigned)max(0, min(m - 1, bi));
                        }
                        word = (word & ~(0xffu << (8 * j))) | (b << (8 * j));
                    }
                }
            }
            packed[w] = word;
        }
        // the int(m / 2)-th smallest bin: the least b with more than imed
        // bins <= b, by bisection over [0, m - 1]; a byte x in [0, 127]
        // is below q in [1, 127] exactly when (x | 0x80) - q clears bit 7,
        // and no byte borrows from its neighbour
        const int imed = (int)((float)m * 0.5f);
        int lo = 0, hi = m - 1;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            const unsigned q = (unsigned)(mid + 1) * 0x01010101u;
            int below = 0;
#pragma unroll
            for (int w = 0; w < kWords; ++w)
                if (w < nw) below += 4 - __popc(((packed[w] | 0x80808080u) - q) & 0x80808080u);
            if (below > imed)
                hi = mid;
            else
                lo = mid + 1;
        }
        out = vmin + (float)lo * width;
    }
    G(med)[(long)g * n + t] = out;
}

// Run bounds of bin b in the marks (marked bins only).
__device__ __forceinline__ void run_of(const unsigned char* mk, int n, int b, int& s, int& e)
{
    s = b;
    while (s > 0 && mk[s - 1]) --s;
    e = b;
    while (e < n - 1 && mk[e + 1]) ++e;
}

// Marks of the samples beyond the limit, padded (:324-346, :388-410):
// bin b is marked when a source lies in [b - pad_b, b + pad_f] (the
// reference's padding clamps at both ends, which this covers).
__device__ void cn_pad_marks(const unsigned char* src, unsigned char* mk, int n, int pad_f, int pad_b)
{
    for (int b = threadIdx.x; b < n; b += kNfBlock) {
        bool any = mk[b] != 0;
        for (int j = max(b - pad_b, 0); j <= min(b + pad_f, n - 1) && !any; ++j) any = src[j] != 0;
        mk[b] = any;
    }
    __syncthreads();
}

// SignalProtection, first part (:276-346), one workgroup per group: the
// limit from the medians' CalcRMS, the marks; groups without the response
// path get their marked runs replaced by the line between the bins just
// outside (:439-459).  Groups with it keep the medians and hand them to
// the deconvolution (complex rows).
__global__ __launch_bounds__(kNfBlock) void k_cn_protect(float* __restrict__ med, int n,
                                                         const wct_nf_group* __restrict__ groups,
                                                         unsigned char* __restrict__ marks,
                                                         float2* __restrict__ grow_c)
{
    extern __shared__ __align__(16) unsigned char nf_smem[];
    __shared__ NfShared sh;
    const int g = blockIdx.x;
    const wct_nf_group grp = groups[g];
    float* m = reinterpret_cast<float*>(nf_smem);
    unsigned char* src = nf_smem + ((size_t)4 * n + 15) / 16 * 16;
    unsigned char* mk = src + n;
    const auto row = G(med + (long)g * n);
    for (int i = threadIdx.x; i < n; i += kNfBlock) m[i] = row[i];
    __syncthreads();
    double mean, rms;
    calc_rms(m, n, mean, rms, sh);
    const double pf = grp.protection_factor;
    float limit = pf * rms > (double)grp.adc_limit ? (float)(pf * rms) : grp.adc_limit;
    if (grp.min_adc_limit < limit) limit = grp.min_adc_limit;
    for (int j = threadIdx.x; j < n; j += kNfBlock) {
        src[j] = fabs((double)m[j] - mean) > (double)limit;
        mk[j] = 0;
    }
    __syncthreads();
    cn_pad_marks(src, mk, n, grp.pad_f, grp.pad_b);
    for (int j = threadIdx.x; j < n; j += kNfBlock) G(marks + (long)g * n)[j] = mk[j];
    if (grp.respec >= 0) {
        for (int j = threadIdx.x; j < n; j += kNfBlock) G(grow_c + (long)g * n)[j] = make_float2(m[j], 0.f);
        return;
    }
    for (int b = threadIdx.x; b < n; b += kNfBlock) {
        if (!mk[b]) continue;
        int s0, e0;
        run_of(mk, n, b, s0, e0);
        const int b0 = max(s0 - 1, 0), bf = min(e0 + 1, n - 1);
        const double m0 = m[b0], mf = m[bf];
        row[b] = (float)(m0 + (b - b0) / (double)(bf - b0) * (mf - m0));
    }
}

// A spectrum row divided by the response, times a float filter factor,
// then DftTools::inv_c2r's mirror (bins above n/2 from the lower half).
__device__ void cn_decon_spectrum(float2* rowp, int n, const float2* respec, bool loose, double cut)
{
    const auto row = G(rowp);
    const auto R = G(respec);
    for (int k = threadIdx.x; k <= n / 2; k += kNfBlock) {
        const double f = cn_freq(k, n);
        const float fac = (float)(cn_filter_time(f) * (loose ? cn_filter_low_loose(f) : cn_filter_low(f, cut)));
        float2 x = cdiv(row[k], R[k]);
        x = make_float2(x.x * fac, x.y * fac);
        if (k == 0) x.y = 0.f;
        if (2 * k == n) x.y = 0.f;
        row[k] = x;
        if (k > 0 && 2 * k != n) row[n - k] = make_float2(x.x, -x.y);
    }
}

__global__ __launch_bounds__(kNfBlock) void k_cn_median_decon(float2* __restrict__ grow_c, int n,
                                                              const wct_nf_group* __restrict__ groups,
                                                              const float2* __restrict__ spectra)
{
    const wct_nf_group grp = groups[blockIdx.x];
    if (grp.respec < 0) return;
    cn_decon_spectrum(grow_c + (long)blockIdx.x * n, n, spectra + (long)grp.respec * n, false, grp.decon_lf_cutoff);
}

// SignalProtection, the deconvolution part (:348-411): marks where the
// deconvolved medians exceed their limit, shifted by the response offset.
__global__ __launch_bounds__(kNfBlock) void k_cn_protect_decon(const float2* __restrict__ grow_c, int n,
                                                               const wct_nf_group* __restrict__ groups,
                                                               unsigned char* __restrict__ marks)
{
    extern __shared__ __align__(16) unsigned char nf_smem[];
    __shared__ NfShared sh;
    const int g = blockIdx.x;
    const wct_nf_group grp = groups[g];
    if (grp.respec < 0) return;
    float* d = reinterpret_cast<float*>(nf_smem);
    unsigned char* src = nf_smem + ((size_t)4 * n + 15) / 16 * 16;
    unsigned char* mk = src + n;
    for (int i = threadIdx.x; i < n; i += kNfBlock) {
        d[i] = G(grow_c + (long)g * n)[i].x;
        mk[i] = G(marks + (long)g * n)[i];
        src[i] = 0;
    }
    __syncthreads();
    double mean, rms;
    calc_rms(d, n, mean, rms, sh);
    const double pf = grp.protection_factor;
    const float limit = pf * rms > (double)grp.decon_limit ? (float)(pf * rms) : grp.decon_limit;
    for (int j = threadIdx.x; j < n; j += kNfBlock) {
        if ((double)d[j] - mean > (double)limit) {
            int tb = j + grp.res_offset;
            if (tb >= n) tb -= n;
            if (tb >= 0 && tb < n) src[tb] = 1;
        }
    }
    __syncthreads();
    cn_pad_marks(src, mk, n, grp.pad_f, grp.pad_b);
    for (int j = threadIdx.x; j < n; j += kNfBlock) G(marks + (long)g * n)[j] = mk[j];
}

// Subtract_WScaling's coefficient per channel (:67-96): sums over the
// samples within 4 CalcRMS of zero, products in float.
__global__ __launch_bounds__(kNfBlock) void k_cn_coef(const float* __restrict__ rows, int ld, int n,
                                                      const int* __restrict__ row_group,
                                                      const float* __restrict__ med,
                                                      const double* __restrict__ crms, double* __restrict__ coef)
{
    __shared__ NfShared sh;
    const int c = blockIdx.x;
    const int g = row_group[c];
    if (g < 0) return;
    const auto s = G(rows + (long)c * ld);
    const auto m = G(med + (long)g * n);
    const double r4 = 4 * crms[c];
    double a = 0, b = 0;
    for (int j = threadIdx.x; j < n; j += kNfBlock) {
        const float x = s[j], y = m[j];
        if (fabsf(x) < r4) {
            a += x * y;
            b += y * y;
        }
    }
    a = block_sum(a, sh);
    b = block_sum(b, sh);
    if (threadIdx.x == 0) coef[c] = b > 0 ? a / b : 0.0;
}

// The group's mean coefficient and each channel's scaling (:97-123).
__global__ void k_cn_scaling(const wct_nf_group* __restrict__ groups, int ngroups, const int* __restrict__ grows,
                             const double* __restrict__ coef, float* __restrict__ scaling)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngroups) return;
    const wct_nf_group grp = groups[g];
    double ave = 0, cnt = 0;
    for (int k = 0; k < grp.count; ++k) {
        const double c = coef[grows[grp.first + k]];
        if (c != 0) {
            ave += c;
            cnt += 1;
        }
    }
    if (cnt > 0) ave /= cnt;
    for (int k = 0; k < grp.count; ++k) {
        const int r = grows[grp.first + k];
        float sc = 0.f;
        if (ave != 0) {
            sc = (float)(coef[r] / ave);
            if (sc < 0) sc = 0;
            if (sc > 1.5f) sc = 1.5f;
        }
        scaling[r] = sc;
    }
}

// The channel's samples in the group's marked runs with the line between
// the bins just outside subtracted, zero elsewhere (:130-145), as a
// complex row.
// A channel row on the response path with its median-protected runs kept
// and everything else zero: in a marked run, the sample less the line
// between the samples just outside the run.
__device__ __forceinline__ float cn_roi_value(const float* s, const unsigned char* mk, int n, int b)
{
    if (!mk[b]) return 0.f;
    int s0 = b, e0 = b;
    while (s0 > 0 && mk[s0 - 1]) --s0;
    while (e0 < n - 1 && mk[e0 + 1]) ++e0;
    const int b0 = max(s0 - 1, 0), bf = min(e0 + 1, n - 1);
    const double m0 = s[b0], mf = s[bf];
    return (float)(s[b] - (m0 + (b - b0) / (double)(bf - b0) * (mf - m0)));
}

// Rows 2p and 2p + 1 as the real and imaginary parts of complex row p;
// rows off the response path (and the missing partner of an odd last row)
// are zero.
__global__ __launch_bounds__(kNfBlock) void k_cn_roi_rows(const float* __restrict__ rows, int ld, int n,
                                                          int nrows, const int* __restrict__ row_group,
                                                          const wct_nf_group* __restrict__ groups,
                                                          const unsigned char* __restrict__ marks,
                                                          float2* __restrict__ crow)
{
    const int p = blockIdx.x;
    const int ca = 2 * p, cb = 2 * p + 1;
    const int ga = row_group[ca], gb = cb < nrows ? row_group[cb] : -1;
    const bool ra = ga >= 0 && groups[ga].respec >= 0, rb = gb >= 0 && groups[gb].respec >= 0;
    const float* sa = rows + (long)ca * ld;
    const float* sb = rows + (long)min(cb, nrows - 1) * ld;
    const unsigned char* ma = marks + (long)(ra ? ga : 0) * n;
    const unsigned char* mb = marks + (long)(rb ? gb : 0) * n;
    const auto dst = G(crow + (long)p * n);
    for (int b = threadIdx.x; b < n; b += kNfBlock)
        dst[b] = make_float2(ra ? cn_roi_value(sa, ma, n, b) : 0.f, rb ? cn_roi_value(sb, mb, n, b) : 0.f);
}

// The two rows of a pair deconvolved for the ROI judgement
// (cn_decon_spectrum's steps on each): their spectra separated from the
// pair's, X_a(k) = (Z(k) + Z*(n - k)) / 2 and X_b(k) = (Z(k) - Z*(n - k))
// / 2i, each divided by its group's response and filtered, then repacked
// as Y_a + i Y_b (both Hermitian), so the inverse gives the two rows as
// its real and imaginary parts.  Bins k and n - k belong to one thread.
__global__ __launch_bounds__(kNfBlock) void k_cn_roi_decon(float2* __restrict__ crow, int n, int nrows,
                                                           const int* __restrict__ row_group,
                                                           const wct_nf_group* __restrict__ groups,
                                                           const float2* __restrict__ spectra)
{
    const int p = blockIdx.x;
    const int ca = 2 * p, cb = 2 * p + 1;
    const int ga = row_group[ca], gb = cb < nrows ? row_group[cb] : -1;
    const int ra = ga >= 0 ? groups[ga].respec : -1, rb = gb >= 0 ? groups[gb].respec : -1;
    if (ra < 0 && rb < 0) return;
    const auto row = G(crow + (long)p * n);
    const auto Ra = G(spectra + (long)max(ra, 0) * n), Rb = G(spectra + (long)max(rb, 0) * n);
    for (int k = threadIdx.x; k <= n / 2; k += kNfBlock) {
        const int kn = k == 0 ? 0 : n - k;
        const float2 z = row[k], zn = row[kn];
        const double f = cn_freq(k, n);
        const float fac = (float)(cn_filter_time(f) * cn_filter_low_loose(f));
        float2 ya = make_float2(0.f, 0.f), yb = make_float2(0.f, 0.f);
        if (ra >= 0) {
            const float2 xa = make_float2(0.5f * (z.x + zn.x), 0.5f * (z.y - zn.y));
            ya = cdiv(xa, Ra[k]);
            ya = make_float2(ya.x * fac, ya.y * fac);
            if (k == 0 || 2 * k == n) ya.y = 0.f;
        }
        if (rb >= 0) {
            const float2 xb = make_float2(0.5f * (z.y + zn.y), -0.5f * (z.x - zn.x));
            yb = cdiv(xb, Rb[k]);
            yb = make_float2(yb.x * fac, yb.y * fac);
            if (k == 0 || 2 * k == n) yb.y = 0.f;
        }
        row[k] = make_float2(ya.x - yb.y, ya.y + yb.x);
        if (k > 0 && 2 * k != n) row[kn] = make_float2(ya.x + yb.y, -ya.y + yb.x);
    }
}

// The subtraction (:124-263): with the response path, the median's runs
// whose deconvolved channel (shifted back by the offset) peaks above
// decon_limit1 without a comparable dip are replaced by the line between
// the bins just outside; samples above 0.001 in size lose the scaled
// median.
__global__ __launch_bounds__(kNfBlock) void k_cn_subtract(float* __restrict__ rows, int ld, int n,
                                                          const int* __restrict__ row_group,
                                                          const wct_nf_group* __restrict__ groups,
                                                          const unsigned char* __restrict__ marks,
                                                          const float* __restrict__ med,
                                                          const float2* __restrict__ crow,
                                                          const float* __restrict__ scaling)
{
    const int c = blockIdx.x;
    const int g = row_group[c];
    if (g < 0) return;
    const wct_nf_group grp = groups[g];
    const auto s = G(rows + (long)c * ld);
    const auto m = G(med + (long)g * n);
    const auto mk = G(marks + (long)g * n);
    const float sc = scaling[c];
    const bool resp = grp.respec >= 0;
    const auto cr = G(reinterpret_cast<const float*>(crow) + (long)(c >> 1) * n * 2 + (c & 1));
    for (int b = threadIdx.x; b < n; b += kNfBlock) {
        float mv = m[b];
        if (resp && mk[b]) {
            int s0 = b, e0 = b;
            while (s0 > 0 && mk[s0 - 1]) --s0;
            while (e0 < n - 1 && mk[e0 + 1]) ++e0;
            const int b0 = max(s0 - 1, 0), bf = min(e0 + 1, n - 1);
            double mx = 0, mn = 0;
            for (int i = b0; i <= bf; ++i) {
                int tb = i - grp.res_offset;
                if (tb < 0) tb += n;
                if (tb >= n) tb -= n;
                const double v = cr[2 * tb];
                if (i == b0) {
                    mx = mn = v;
                } else {
                    mx = fmax(mx, v);
                    mn = fmin(mn, v);
                }
            }
            if (mx > (double)grp.decon_limit1 && fabs(mn) < mx * (double)grp.roi_min_max_ratio) {
                const double m0 = m[b0], mf = m[bf];
                mv = (float)(m0 + (b - b0) / (double)(bf - b0) * (mf - m0));
            }
        }
        const float x = s[b];
        if (fabs((double)x) > 0.001) s[b] = x - mv * sc;
    }
}

// ---- Microboone::OneChannelStatus (Microboone.cxx:1263-1334) ----------

// ID_lf_noisy's samples (:1287-1318) for channel c: the binned median and
// 16/84 % points of the row from one histogram, the rms from them in
// double, the count of non-zero samples, and the row with every sample
// within [i - window, i + window) of a sample i beyond threshold x rms
// from the median set to the median (marks from the original samples);
// written as the real (even c) or imaginary (odd c) part of complex row
// c / 2 (the partner of an odd last row zero).
__global__ __launch_bounds__(kNfBlock) void k_st_load(const float* __restrict__ in, int ld, int nch, int n,
                                                      double threshold, int window, float2* __restrict__ rows,
                                                      int* __restrict__ valid)
{
    extern __shared__ __align__(16) unsigned char nf_smem[];
    __shared__ NfShared sh;
    const int c = blockIdx.x;
    const size_t a4 = ((size_t)4 * n + 15) / 16 * 16;
    float* s = reinterpret_cast<float*>(nf_smem);
    int* hist = reinterpret_cast<int*>(nf_smem + a4);
    unsigned char* mk = nf_smem + a4;   // over the histogram, which pct_binned is done with
    const auto src = G(in + (long)c * ld);
    int nz = 0;
    for (int i = threadIdx.x; i < n; i += kNfBlock) {
        const float v = src[i];
        s[i] = v;
        nz += v != 0.f;
    }
    __syncthreads();
    const float p[3] = {0.5f, (float)(0.5 - 0.34), (float)(0.5 + 0.34)};
    float q[3];
    pct_binned(s, n, AnySample(), p, 3, q, hist, sh);
    nz = block_sum(nz, sh);
    const double mean = q[0], da = (double)q[1] - mean, db = (double)q[2] - mean;
    const double lim = threshold * sqrt((da * da + db * db) / 2.);
    for (int i = threadIdx.x; i < n; i += kNfBlock) mk[i] = fabs((double)s[i] - mean) > lim;
    __syncthreads();
    const auto dst = G(reinterpret_cast<float*>(rows) + (long)(c >> 1) * n * 2 + (c & 1));
    const bool lone = (c & 1) == 0 && c + 1 == nch;
    for (int j = threadIdx.x; j < n; j += kNfBlock) {
        bool hit = false;   // a marked i with j in [i - window, i + window)
        const int i1 = min(n - 1, j + window);
        for (int i = max(0, j - window + 1); i <= i1; ++i) hit |= mk[i] != 0;
        dst[2 * j] = hit ? (float)mean : s[j];
        if (lone) dst[2 * j + 1] = 0.f;
    }
    if (threadIdx.x == 0) valid[c] = nz;
}

// :1320-1333 for the two channels of complex row p: each spectrum
// separated from the pair's (X_a(k) = (Z(k) + Z*(n - k)) / 2, X_b(k) =
// (Z(k) - Z*(n - k)) / 2i), the float magnitudes of bins 1..nbins summed
// in double, over the channel's non-zero count, against cut.
__global__ __launch_bounds__(kNfBlock) void k_st_ratio(const float2* __restrict__ rows, int nch, int n, int nbins,
                                                       double cut, const int* __restrict__ valid,
                                                       int* __restrict__ flag, double* __restrict__ ratio)
{
    __shared__ NfShared sh;
    const int p = blockIdx.x;
    const auto row = G(rows + (long)p * n);
    double ca = 0, cb = 0;
    for (int k = 1 + (int)threadIdx.x; k <= nbins; k += kNfBlock) {
        const float2 z = row[k], zn = row[n - k];
        ca += (double)hypotf(0.5f * (z.x + zn.x), 0.5f * (z.y - zn.y));
        cb += (double)hypotf(0.5f * (z.y + zn.y), -0.5f * (z.x - zn.x));
    }
    ca = block_sum(ca, sh);
    cb = block_sum(cb, sh);
    if (threadIdx.x < 2) {
        const int c = 2 * p + (int)threadIdx.x;
        if (c < nch) {
            const int v = G(valid)[c];
            const double r = v ? (threadIdx.x ? cb : ca) / v : 0.0;
            G(flag)[c] = v && r > cut;
            if (ratio) G(ratio)[c] = r;
        }
    }
}

// ---- Microboone::ADCBitShift (Microboone.cxx:1106-1216) ---------------

// One channel per workgroup, in place.  Detection: the low nbits bits of
// each of the first exam samples (as int, truncated) against the previous
// sample's (zero before the first), flips counted per bit by wave ballots;
// nshift = the low bits, from bit 0 up, flipping fewer than nbin / 2 -
// nbin / 2 sqrt(1 / nbin) threshold_sigma times.  For 0 < nshift < 11:
// every sample rotated back with its own low bits as filling
// (Bits::shift_right(x, nshift, lowest_bits(x, nshift), 12)), the set of
// fillings and the integer mean formed in parallel, then the examination
// of :1164-1208 on one lane (each sample's previous two are the corrected
// ones, so it is sequential): a sample far from the mean and from a
// neighbour takes, fillings in increasing order, each candidate closer to
// the neighbours' extrapolation.
__global__ __launch_bounds__(kNfBlock) void k_bs(float* __restrict__ rows, int ld, int n, int nbits, int exam,
                                                 double threshold_sigma, double threshold_fix,
                                                 int* __restrict__ nshift_out)
{
    extern __shared__ __align__(16) unsigned char nf_smem[];
    __shared__ NfShared sh;
    __shared__ int cnt[32];
    __shared__ unsigned fill[32];   // up to 2^10 fillings
    const int c = blockIdx.x;
    const auto row = G(rows + (long)c * ld);
    int* x = reinterpret_cast<int*>(nf_smem);
    if (threadIdx.x < 32) {
        cnt[threadIdx.x] = 0;
        fill[threadIdx.x] = 0;
    }
    __syncthreads();
    for (int i0 = 0; i0 < exam; i0 += kNfBlock) {
        const int i = i0 + (int)threadIdx.x;
        unsigned d = 0;
        if (i < exam) d = (unsigned)((int)row[i] ^ (i ? (int)row[i - 1] : 0));
        for (int j = 0; j < nbits; ++j) {
            const unsigned long long m = __ballot((d >> j) & 1u);
            if ((threadIdx.x & 63) == 0 && m) atomicAdd(&cnt[j], __popcll(m));
        }
    }
    __syncthreads();
    const double lim = exam / 2. - exam / 2. * sqrt(1. / exam) * threshold_sigma;
    int ns = 0;
    while (ns < nbits && cnt[ns] < lim) ++ns;
    if (ns == 0 || ns >= 11) {   // uniform: every thread reads the same counts
        if (threadIdx.x == 0) G(nshift_out)[c] = 0;
        return;
    }
    int part = 0;
    for (int i = threadIdx.x; i < n; i += kNfBlock) {
        const int xo = (int)row[i];
        const int f = xo & ((1 << ns) - 1);
        const int y = (xo >> ns) | (f << (12 - ns));
        x[i] = y;
        part += y;
        atomicOr(&fill[f >> 5], 1u << (f & 31));
    }
    const int total = block_sum(part, sh);   // int arithmetic, as the reference sums
    const int mean = total / n;
    const int exp_diff = (int)(ldexp(1.0, nbits - ns) * threshold_fix);
    __syncthreads();
    if (threadIdx.x == 0) {
        int prev1 = mean, prev = mean, next = mean, next1 = mean;
        const int nf = 1 << ns;
        for (int i = 0; i < n; ++i) {
            const int cur = x[i];
            if (abs(cur - mean) > exp_diff && (abs(cur - prev) > exp_diff || abs(cur - next) > exp_diff)) {
                const int ev =
                    (int)(((2 * prev - prev1) + (prev + next) / 2. + (prev * 2. / 3. + next1 / 3.)) / 3.);
                const int xo = (int)row[i];
                for (int f = 0; f < nf; ++f) {
                    if (!((fill[f >> 5] >> (f & 31)) & 1u)) continue;
                    const int y = (xo >> ns) | (f << (12 - ns));
                    if (abs(y - ev) < abs(x[i] - ev)) x[i] = y;
                }
            }
            prev1 = prev;
            prev = x[i];
            next = i + 2 < n ? x[i + 2] : mean;
            next1 = i + 3 < n ? x[i + 3] : mean;
        }
        G(nshift_out)[c] = ns;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += kNfBlock) row[i] = (float)x[i];
}

}  // namespace

}  // namespace wct

using namespace wct;

extern "C" {

size_t wct_nf_workspace_size(int nch, int nt)
{
    if (nch <= 0 || nt <= 0) return 0;
    return ((size_t)((nch + 1) / 2) * nt * sizeof(float2) + 255) / 256 * 256;
}

int wct_nf_rcrc_spectrum(double width, double tick, int nt, int layers, float* out, void* stream)
{
    WCT_ENTRY();
    WCT_REQUIRE(out && nt >= 2 && tick > 0 && layers >= 1, WCT_ERR_ARG, "nf rcrc spectrum: bad arguments");
    hipStream_t s = as_stream(stream);
    DevTemp<double2> row(s);
    WCT_TRY(row.alloc(nt));
    const unsigned nb = (nt + kNfBlock - 1) / kNfBlock;
    hipLaunchKernelGGL(k_nf_rc_samples, dim3(nb), dim3(kNfBlock), 0, s, row.p, nt, width, tick);
    WCT_HIP_CHECK(hipGetLastError());
    WCT_TRY(launch_rows<double>(row.p, row.p, nt, 1, nt, -1, 1.0, s));
    hipLaunchKernelGGL(k_nf_rc_power, dim3(nb), dim3(kNfBlock), 0, s, row.p, reinterpret_cast<float2*>(out), nt,
                       layers);
    WCT_HIP_CHECK(hipGetLastError());
    return WCT_OK;
}

int wct_nf_one_channel(const float* in, int ld_in, float* out, int ld_out, int nch, int nt,
                       const wct_nf_channel* chans, const float* rcrc, const int* bins, int* flags,
                       void* workspace, size_t ws_bytes, void* stream)
{
    WCT_ENTRY();
    WCT_REQUIRE(in && out && chans && flags && nch >= 1 && nt >= 32 && ld_in >= nt && ld_out >= nt, WCT_ERR_ARG,
                "nf: bad arguments (nch %d, nt %d, ld_in %d, ld_out %d)", nch, nt, ld_in, ld_out);
    WCT_REQUIRE(workspace && ws_bytes >= wct_nf_workspace_size(nch, nt), WCT_ERR_ARG,
                "nf: workspace of %zu bytes, %zu needed", ws_bytes, wct_nf_workspace_size(nch, nt));
    const size_t lds = nf_lds_bytes(nt);
    WCT_REQUIRE((long)lds + (long)sizeof(NfShared) <= lds_limit_bytes(), WCT_ERR_UNSUPPORTED,
                "nf: %d ticks need %zu bytes of LDS per channel", nt, lds);
    hipStream_t s = as_stream(stream);
    float2* rows = reinterpret_cast<float2*>(workspace);
    const int npair = (nch + 1) / 2;
    const long total = (long)npair * nt;
    const size_t chirp_lds = sizeof(double) * (nt / kChirpWindow + 1);
    WCT_TRY(allow_lds(k_nf_chirp, chirp_lds));
    hipLaunchKernelGGL(k_nf_chirp, dim3(nch), dim3(kNfBlock), chirp_lds, s, in, ld_in, nt, chans, flags);
    WCT_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_nf_load, dim3((unsigned)((total + kNfBlock - 1) / kNfBlock)), dim3(kNfBlock), 0, s, in,
                       ld_in, chans, rows, nch, nt);
    WCT_HIP_CHECK(hipGetLastError());
    WCT_TRY(launch_rows<float>(rows, rows, nt, npair, nt, -1, 1.0f, s));
    // LDS only when some channel has flagged runs (bins given)
    const size_t spec_lds = bins ? nf_spec_lds_bytes(nt) : 0;
    WCT_REQUIRE((long)spec_lds + (long)sizeof(NfShared) <= lds_limit_bytes(), WCT_ERR_UNSUPPORTED,
                "nf: %d ticks need %zu bytes of LDS per channel pair", nt, spec_lds);
    WCT_TRY(allow_lds(k_nf_spec, spec_lds));
    hipLaunchKernelGGL(k_nf_spec, dim3(npair), dim3(kNfBlock), spec_lds, s, rows, nt, nch, chans,
                       reinterpret_cast<const float2*>(rcrc), bins, flags);
    WCT_HIP_CHECK(hipGetLastError());
    WCT_TRY(launch_rows<float>(rows, rows, nt, npair, nt, +1, 1.0f / nt, s));
    WCT_TRY(allow_lds(k_nf_time, lds));
    hipLaunchKernelGGL(k_nf_time, dim3(nch), dim3(kNfBlock), lds, s, rows, nt, out, ld_out, chans, flags);
    WCT_HIP_CHECK(hipGetLastError());
    return WCT_OK;
}

}  // extern "C"

extern "C" int wct_nf_rel_gain(float* rows, int ld, int nch, int nt, const float* gain, float gain_min,
                               float gain_max, void* stream)
{
    WCT_ENTRY();
    WCT_REQUIRE(rows && gain && nch >= 0 && nt >= 0 && ld >= nt, WCT_ERR_ARG, "nf rel gain: bad arguments");
    if (nch == 0 || nt == 0) return WCT_OK;
    const long total = (long)nch * nt;
    hipLaunchKernelGGL(k_nf_rel_gain, dim3((unsigned)((total + kNfBlock - 1) / kNfBlock)), dim3(kNfBlock), 0,
                       as_stream(stream), rows, ld, nch, nt, gain, gain_min, gain_max);
    WCT_HIP_CHECK(hipGetLastError());
    return WCT_OK;
}

extern "C" size_t wct_nf_coherent_workspace_size(int ngroups, int nrows, int nt)
{
    if (ngroups <= 0 || nrows <= 0 || nt <= 0) return 0;
    auto al = [](size_t b) { return (b + 255) / 256 * 256; };
    return al(sizeof(float) * (size_t)ngroups * nt)          // medians
           + al((size_t)ngroups * nt)                         // marks
           + al(sizeof(float2) * (size_t)ngroups * nt)        // group rows (complex)
           + al(sizeof(float2) * (size_t)nrows * nt)          // channel rows (complex)
           + al(sizeof(double) * 3 * (size_t)nrows)           // rms, CalcRMS, coefficients
           + al(sizeof(float) * (size_t)nrows)                // scaling
           + al(sizeof(int) * (size_t)nrows);                 // group of each row
}

extern "C" int wct_nf_coherent(float* rows, int ld, int nrows, int nt, const wct_nf_group* groups, int ngroups,
                               const int* group_rows, const int* row_group, const float* spectra, int any_response,
                               void* workspace, size_t ws_bytes, void* stream)
{
    WCT_ENTRY();
    WCT_REQUIRE(rows && groups && group_rows && row_group && nrows >= 1 && ngroups >= 1 && nt >= 32 && ld >= nt,
                WCT_ERR_ARG, "nf coherent: bad arguments");
    WCT_REQUIRE(!any_response || spectra, WCT_ERR_ARG, "nf coherent: response spectra missing");
    WCT_REQUIRE(workspace && ws_bytes >= wct_nf_coherent_workspace_size(ngroups, nrows, nt), WCT_ERR_ARG,
                "nf coherent: workspace of %zu bytes, %zu needed", ws_bytes,
                wct_nf_coherent_workspace_size(ngroups, nrows, nt));
    hipStream_t s = as_stream(stream);
    auto al = [](size_t b) { return (b + 255) / 256 * 256; };
    unsigned char* p = reinterpret_cast<unsigned char*>(workspace);
    float* med = reinterpret_cast<float*>(p);
    p += al(sizeof(float) * (size_t)ngroups * nt);
    unsigned char* marks = p;
    p += al((size_t)ngroups * nt);
    float2* grow_c = reinterpret_cast<float2*>(p);
    p += al(sizeof(float2) * (size_t)ngroups * nt);
    float2* crow = reinterpret_cast<float2*>(p);
    p += al(sizeof(float2) * (size_t)nrows * nt);
    double* rms0 = reinterpret_cast<double*>(p);
    double* crms = rms0 + nrows;
    double* coef = crms + nrows;
    p += al(sizeof(double) * 3 * (size_t)nrows);
    float* scaling = reinterpret_cast<float*>(p);
    // every row's group (row_group: -1 for rows in no group)
    const size_t lds_row = ((size_t)4 * nt + 15) / 16 * 16 + 2 * (size_t)nt + 16;
    WCT_REQUIRE((long)lds_row + (long)sizeof(NfShared) <= lds_limit_bytes(), WCT_ERR_UNSUPPORTED,
                "nf coherent: %d ticks need %zu bytes of LDS", nt, lds_row);
    WCT_TRY(allow_lds(k_cn_rms, lds_row));
    hipLaunchKernelGGL(k_cn_rms, dim3(nrows), dim3(kNfBlock), lds_row, s, rows, ld, nt, rms0, crms);
    WCT_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_cn_median, dim3(ngroups, (nt + kCnTicks - 1) / kCnTicks), dim3(kCnTicks), 0, s, rows,
                       ld, nt, groups, group_rows, rms0, med);
    WCT_HIP_CHECK(hipGetLastError());
    WCT_TRY(allow_lds(k_cn_protect, lds_row));
    hipLaunchKernelGGL(k_cn_protect, dim3(ngroups), dim3(kNfBlock), lds_row, s, med, nt, groups, marks, grow_c);
    WCT_HIP_CHECK(hipGetLastError());
    if (any_response) {
        WCT_TRY(launch_rows<float>(grow_c, grow_c, nt, ngroups, nt, -1, 1.0f, s));
        hipLaunchKernelGGL(k_cn_median_decon, dim3(ngroups), dim3(kNfBlock), 0, s, grow_c, nt, groups,
                           reinterpret_cast<const float2*>(spectra));
        WCT_HIP_CHECK(hipGetLastError());
        WCT_TRY(launch_rows<float>(grow_c, grow_c, nt, ngroups, nt, +1, 1.0f / nt, s));
        WCT_TRY(allow_lds(k_cn_protect_decon, lds_row));
        hipLaunchKernelGGL(k_cn_protect_decon, dim3(ngroups), dim3(kNfBlock), lds_row, s, grow_c, nt, groups, marks);
        WCT_HIP_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_cn_coef, dim3(nrows), dim3(kNfBlock), 0, s, rows, ld, nt, row_group, med, crms, coef);
    WCT_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_cn_scaling, dim3((ngroups + 63) / 64), dim3(64), 0, s, groups, ngroups, group_rows, coef,
                       scaling);
    WCT_HIP_CHECK(hipGetLastError());
    if (any_response) {
        // two channel rows per complex row
        const int npair = (nrows + 1) / 2;
        hipLaunchKernelGGL(k_cn_roi_rows, dim3(npair), dim3(kNfBlock), 0, s, rows, ld, nt, nrows, row_group, groups,
                           marks, crow);
        WCT_HIP_CHECK(hipGetLastError());
        WCT_TRY(launch_rows<float>(crow, crow, nt, npair, nt, -1, 1.0f, s));
        hipLaunchKernelGGL(k_cn_roi_decon, dim3(npair), dim3(kNfBlock), 0, s, crow, nt, nrows, row_group, groups,
                           reinterpret_cast<const float2*>(spectra));
        WCT_HIP_CHECK(hipGetLastError());
        WCT_TRY(launch_rows<float>(crow, crow, nt, npair, nt, +1, 1.0f / nt, s));
    }
    hipLaunchKernelGGL(k_cn_subtract, dim3(nrows), dim3(kNfBlock), 0, s, rows, ld, nt, row_group, groups, marks, med,
                       crow, scaling);
    WCT_HIP_CHECK(hipGetLastError());
    return WCT_OK;
}

extern "C" size_t wct_nf_status_workspace_size(int nch, int nt)
{
    if (nch <= 0 || nt <= 0) return 0;
    auto al = [](size_t b) { return (b + 255) / 256 * 256; };
    return al(sizeof(float2) * (size_t)((nch + 1) / 2) * nt) + al(sizeof(int) * (size_t)nch);
}

extern "C" int wct_nf_lf_noisy(const float* rows, int ld, int nch, int nt, double threshold, int window, int nbins,
                               double cut, int* lf_noisy, double* ratio, void* workspace, size_t ws_bytes,
                               void* stream)
{
    WCT_ENTRY();
    WCT_REQUIRE(rows && lf_noisy && nch >= 1 && nt >= 2 && ld >= nt && window >= 0 && nbins >= 1 && nbins < nt,
                WCT_ERR_ARG, "nf lf_noisy: bad arguments (nch %d, nt %d, ld %d, window %d, nbins %d)", nch, nt, ld,
                window, nbins);
    WCT_REQUIRE(workspace && ws_bytes >= wct_nf_status_workspace_size(nch, nt), WCT_ERR_ARG,
                "nf lf_noisy: workspace of %zu bytes, %zu needed", ws_bytes, wct_nf_status_workspace_size(nch, nt));
    const size_t lds = 2 * (((size_t)4 * nt + 15) / 16 * 16);
    WCT_REQUIRE((long)lds + (long)sizeof(NfShared) <= lds_limit_bytes(), WCT_ERR_UNSUPPORTED,
                "nf lf_noisy: %d ticks need %zu bytes of LDS per channel", nt, lds);
    hipStream_t s = as_stream(stream);
    auto al = [](size_t b) { return (b + 255) / 256 * 256; };
    const int npair = (nch + 1) / 2;
    float2* crow = reinterpret_cast<float2*>(workspace);
    int* valid = reinterpret_cast<int*>(reinterpret_cast<unsigned char*>(workspace) +
                                        al(sizeof(float2) * (size_t)npair * nt));
    WCT_TRY(allow_lds(k_st_load, lds));
    hipLaunchKernelGGL(k_st_load, dim3(nch), dim3(kNfBlock), lds, s, rows, ld, nch, nt, threshold, window, crow,
                       valid);
    WCT_HIP_CHECK(hipGetLastError());
    WCT_TRY(launch_rows<float>(crow, crow, nt, npair, nt, -1, 1.0f, s));
    hipLaunchKernelGGL(k_st_ratio, dim3(npair), dim3(kNfBlock), 0, s, crow, nch, nt, nbins, cut, valid, lf_noisy,
                       ratio);
    WCT_HIP_CHECK(hipGetLastError());
    return WCT_OK;
}

extern "C" int wct_nf_adc_bit_shift(float* rows, int ld, int nch, int nt, int nbits, int exam_nticks,
                                    double threshold_sigma, double threshold_fix, int* nshift, void* stream)
{
    WCT_ENTRY();
    WCT_REQUIRE(rows && nshift && nch >= 1 && nt >= 1 && ld >= nt && nbits >= 1 && nbits <= 31 &&
                    exam_nticks >= 1 && exam_nticks <= nt,
                WCT_ERR_ARG, "nf adc bit shift: bad arguments (nch %d, nt %d, ld %d, nbits %d, exam %d)", nch, nt,
                ld, nbits, exam_nticks);
    const size_t lds = sizeof(int) * (size_t)nt;
    WCT_REQUIRE((long)lds + (long)sizeof(NfShared) + 256 <= lds_limit_bytes(), WCT_ERR_UNSUPPORTED,
                "nf adc bit shift: %d ticks need %zu bytes of LDS per channel", nt, lds);
    hipStream_t s = as_stream(stream);
    WCT_TRY(allow_lds(k_bs, lds));
    hipLaunchKernelGGL(k_bs, dim3(nch), dim3(kNfBlock), lds, s, rows, ld, nt, nbits, exam_nticks, threshold_sigma,
                       threshold_fix, nshift);
    WCT_HIP_CHECK(hipGetLastError());
    return WCT_OK;
}
