/* CPU oracle for the 2D signal-processing deconvolution of Wire-Cell's
 * OmnibusSigProc (sigproc/src/OmnibusSigProc.cxx:707-1380) and its tight
 * ROI pass (sigproc/src/ROI_formation.cxx:364-531).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
 * the smoke check and the CPU-baseline leg of bench.py load it, and
 * never as the thing measured or shipped.  The product path lives in
 * the HIP library and must fail loudly when that library is missing.
 *
 * Contents
 *   - own mixed-radix Stockham FFT, float and double (osp_real.inc)
 *   - IDFT methods with FftwDFT semantics (aux/src/FftwDFT.cxx:88-217)
 *   - Hermitian mirror (util/inc/WireCellUtil/Spectrum.h:56-89)
 *   - Waveform::percentile / median (util/src/Waveform.cxx:85-106)
 *   - restore_baseline (OmnibusSigProc.cxx:867-906)
 *   - init_overall_response (OmnibusSigProc.cxx:707-865)
 *   - decon_2D_init / decon_2D_*ROI step by step (:944-1117)
 *   - a fused FP64 decon (mathematically identical; the FP64 truth)
 *   - cal_RMS + tight-ROI threshold (ROI_formation.cxx:364-531)
 *   - ROI run list, pad/merge and connect (ROI_formation.cxx:167-360)
 *   - loose ROIs and their extension by the tight ones
 *     (ROI_formation.cxx:538-937), ROI_refinement::apply_roi
 *     (ROI_refinement.cxx:118-218)
 *   - Response::coldelec / ColdElec sampling (util/src/Response.cxx:277-335,
 *     gen/src/ColdElecResponse.cxx:44-49), HfFilter/LfFilter
 *     (sigproc/src/HfFilter.cxx:38-52, LfFilter.cxx:31-45,
 *     util/src/Response.cxx:435-444), wire_region_average
 *     (util/src/Response.cxx:79-196), fft_best_length
 *     (util/src/FFTBestLength.cxx:12-58)
 *   - a seeded synthetic frame generator (tracks + noise, FP64)
 *
 * Pinning: every function here is checked in tests/ against a golden
 * vector, a known answer of the reference's tests, a closed form or a
 * brute-force computation (see tests/test_oracle_*.py).
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "wct_oracle.h"

#define REAL float
#define SUF(x) x##_f
#include "osp_real.inc"
#undef REAL
#undef SUF

#define REAL double
#define SUF(x) x##_d
#include "osp_real.inc"
#undef REAL
#undef SUF

/* ------------------------------------------------------------------ */
/* Brute-force DFT in long double (pins the FFTs).                     */
/* ------------------------------------------------------------------ */
void wo_dft_naive(int n, const double* in, double* out, int sign)
{
    const long double twopi = 6.283185307179586476925286766559L;
    for (int k = 0; k < n; ++k) {
        long double sr = 0, si = 0;
        for (int j = 0; j < n; ++j) {
            long q = ((long)j * k) % n;
            long double ang = sign * twopi * (long double)q / (long double)n;
            long double c = cosl(ang), s = sinl(ang);
            sr += in[2 * j] * c - in[2 * j + 1] * s;
            si += in[2 * j] * s + in[2 * j + 1] * c;
        }
        out[2 * k] = (double)sr;
        out[2 * k + 1] = (double)si;
    }
}

/* ------------------------------------------------------------------ */
/* Filters and electronics                                             */
/* ------------------------------------------------------------------ */

/* Response::hf_filter (util/src/Response.cxx:435-442) on the frequency
 * grid of SigProc::HfFilter::filter_waveform (HfFilter.cxx:38-52). */
void wo_hf_filter(int n, double max_freq, double sigma, double power, int flag,
                  double* out)
{
    for (int i = 0; i < n; ++i) {
        double freq = i * 1.0 / n * 2 * max_freq;
        if (freq > max_freq) freq = freq - 2 * max_freq;
        double f = fabs(freq);
        if (flag && f == 0) out[i] = 0;
        else out[i] = exp(-0.5 * pow(f / sigma, power));
    }
}

/* Response::lf_filter (Response.cxx:444) on LfFilter's grid (LfFilter.cxx:31-45). */
void wo_lf_filter(int n, double max_freq, double tau, double* out)
{
    for (int i = 0; i < n; ++i) {
        double freq = i * 1.0 / n * 2 * max_freq;
        if (freq > max_freq) freq = freq - 2 * max_freq;
        double f = fabs(freq);
        out[i] = 1 - exp(-pow(f / tau, 2));
    }
}

/* Response::coldelec (Response.cxx:277-308): closed-form cold
 * electronics impulse response.  The coefficients are the reference's
 * own (an inverse Laplace transform of the shaper); they define the
 * response and so must be the same numbers.  Time unit: ns (WCT units). */
double wo_coldelec(double time, double gain, double shaping)
{
    if (time <= 0 || time >= 10 * 1000.0) return 0.0;
    const double x = time / shaping;
    gain *= 10 * 1.012;
    const double e1 = exp(-2.94809 * x), e2 = exp(-2.82833 * x), e3 = exp(-2.40318 * x);
    const double c1 = cos(1.19361 * x), s1 = sin(1.19361 * x);
    const double c2 = cos(2.38722 * x), s2 = sin(2.38722 * x);
    const double c3 = cos(2.5928 * x), s3 = sin(2.5928 * x);
    const double c4 = cos(5.18561 * x), s4 = sin(5.18561 * x);
    double v = 4.31054 * e1;
    v += -2.6202 * e2 * c1 - 2.6202 * e2 * c1 * c2;
    v += 0.464924 * e3 * c3 + 0.464924 * e3 * c3 * c4;
    v += 0.762456 * e2 * s1 - 0.762456 * e2 * c2 * s1 + 0.762456 * e2 * c1 * s2;
    v += -2.620200 * e2 * s1 * s2;
    v += -0.327684 * e3 * s3 + 0.327684 * e3 * c4 * s3 - 0.327684 * e3 * c3 * s4;
    v += 0.464924 * e3 * s3 * s4;
    return v * gain;
}

/* ColdElecResponse::waveform_samples(Binning) with Binning(n, tmin, tmax):
 * values at bin centres, stored as float and scaled by the component's
 * postgain in float (Waveform::scale(Sequence<float>&, double)). */
void wo_elec_samples(int n, double tmin, double tmax, double gain, double shaping,
                     double postgain, float* out_f, double* out_d)
{
    const double bs = (tmax - tmin) / n;
    for (int i = 0; i < n; ++i) {
        double t = tmin + (i + 0.5) * bs;
        double v = wo_coldelec(t, gain, shaping);
        if (out_f) out_f[i] = (float)v * (float)postgain;
        if (out_d) out_d[i] = v * postgain;
    }
}

/* util/src/FFTBestLength.cxx:12-58 (table copied as data: the lengths
 * are part of the reference's behaviour). */
static const int fbl_any[] = {33, 35, 40, 42, 44, 56, 64, 66, 70, 77, 80, 96, 98, 100, 112, 121, 128, 140, 143, 160, 168, 169, 200,
    224, 256, 280, 320, 400, 448, 512, 560, 640, 686, 768, 800, 847, 896, 1024, 1120, 1280, 1331, 1372, 1400,
    1536, 1600, 1792, 2048, 2560, 2662, 2744, 3200, 3584, 4096, 4116, 4480, 4802, 5120, 5324, 5376, 5544, 5600,
    5632, 5929, 6174, 6272, 6292, 6655, 7168, 7840, 8192, 8232, 8960, 9317, 10240, 10648, 11979, 14336, 15972,
    16384};
static const int fbl_even[] = {40, 42, 44, 56, 64, 66, 70, 72, 80, 96, 98, 100, 112, 120, 128, 140, 160, 168, 200, 224, 256, 280,
    320, 400, 448, 512, 560, 640, 686, 768, 800, 896, 1024, 1120, 1280, 1372, 1400, 1536, 1600, 1792, 2048,
    2560, 2662, 2744, 3200, 3584, 4096, 4116, 4480, 4802, 5120, 5324, 5376, 5544, 5600, 5632, 6174, 6272,
    6292, 6468, 7168, 7840, 8192, 8232, 8960, 10240, 10648, 11200, 11858, 11880, 14336, 15972, 16384};
static const int fbl_odd[] = {33, 35, 39, 49, 55, 63, 65, 77, 81, 91, 99, 121, 143, 169, 175, 187, 225, 231, 245, 275, 297, 343,
    363, 375, 385, 405, 441, 455, 495, 539, 605, 637, 693, 735, 847, 875, 891, 945, 1001, 1089, 1331, 1573,
    1859, 1875, 2197, 2299, 2431, 2541, 2625, 2695, 3025, 3267, 3993, 4095, 4125, 4719, 4725, 5005, 5145,
    5445, 5929, 6655, 7007, 7623, 7865, 9317, 9801, 11979, 14157, 14175, 15379, 16335};

long wo_fft_best_length(long n, int keep_odd_even)
{
    const int* e; size_t ne;
    if (!keep_odd_even) { e = fbl_any; ne = sizeof(fbl_any) / sizeof(int); }
    else if (n % 2 == 0) { e = fbl_even; ne = sizeof(fbl_even) / sizeof(int); }
    else { e = fbl_odd; ne = sizeof(fbl_odd) / sizeof(int); }
    if (n <= e[ne - 1])
        for (size_t i = 0; i < ne; ++i) if (e[i] >= n) return e[i];
    return n;
}

/* ------------------------------------------------------------------ */
/* Response::wire_region_average (util/src/Response.cxx:79-196), float */
/* arithmetic as the reference (realseq_t = float).                    */
/* Output rows are ordered by increasing region number.                */
/* Returns the number of regions (<= max_regions), -1 on error.         */
/* ------------------------------------------------------------------ */
typedef struct { int key; float* resp; } wo_fr_entry;

static int wo_find_key(wo_fr_entry* m, int nm, int key)
{
    for (int i = 0; i < nm; ++i) if (m[i].key == key) return i;
    return -1;
}

static int wo_cmp_entry(const void* a, const void* b)
{
    int x = ((const wo_fr_entry*)a)->key, y = ((const wo_fr_entry*)b)->key;
    return (x > y) - (x < y);
}

int wo_wire_region_average(int npaths, const double* pitchpos, const float* current,
                           int nsamp, double pitch, int max_regions,
                           int* regions_out, float* avg_out)
{
    wo_fr_entry* m = (wo_fr_entry*)calloc(2 * npaths + 2, sizeof(wo_fr_entry));
    int nm = 0, nsamples = 0;
    for (int p = 0; p < npaths; ++p) {
        const float* cur = current + (size_t)p * nsamp;
        int eff = (int)(pitchpos[p] / (0.01 * pitch));
        int keys[2] = {eff, -eff};
        for (int s = 0; s < 2; ++s) {
            int i = wo_find_key(m, nm, keys[s]);
            if (i < 0) {
                m[nm].key = keys[s];
                m[nm].resp = (float*)malloc(sizeof(float) * nsamp);
                memcpy(m[nm].resp, cur, sizeof(float) * nsamp);
                ++nm;
            } else {
                if (s == 0) nsamples = nsamp;  /* only the first branch sets it */
                for (int k = 0; k < nsamp; ++k)
                    m[i].resp[k] = (float)(((double)m[i].resp[k] + (double)cur[k]) / 2.);
            }
        }
    }
    qsort(m, nm, sizeof(wo_fr_entry), wo_cmp_entry);
    double* lo = (double*)malloc(sizeof(double) * nm);
    double* hi = (double*)malloc(sizeof(double) * nm);
    for (int i = 0; i < nm; ++i) {
        if (i == 0) { lo[i] = -1e9; hi[i] = (m[i].key + m[i + 1].key) / 2. * 0.01 * pitch; }
        else if (i == nm - 1) { lo[i] = (m[i].key + m[i - 1].key) / 2. * 0.01 * pitch; hi[i] = 1e9; }
        else {
            lo[i] = (m[i].key + m[i - 1].key) / 2. * 0.01 * pitch;
            hi[i] = (m[i].key + m[i + 1].key) / 2. * 0.01 * pitch;
        }
    }
    /* std::set of wire regions, in increasing order */
    int* reg = (int*)malloc(sizeof(int) * nm);
    int nreg = 0;
    for (int i = 0; i < nm; ++i) {
        double v = m[i].key > 0 ? round((m[i].key * 0.01 * pitch - 0.001 * pitch) / pitch)
                                : round((m[i].key * 0.01 * pitch + 0.001 * pitch) / pitch);
        int r = (int)v, dup = 0;
        for (int q = 0; q < nreg; ++q) if (reg[q] == r) { dup = 1; break; }
        if (!dup) reg[nreg++] = r;
    }
    for (int a = 0; a < nreg; ++a)
        for (int b = a + 1; b < nreg; ++b)
            if (reg[b] < reg[a]) { int t = reg[a]; reg[a] = reg[b]; reg[b] = t; }
    if (nreg > max_regions) nreg = -1;
    for (int r = 0; nreg > 0 && r < nreg; ++r) {
        const int wire_no = reg[r];
        float* avg = avg_out + (size_t)r * nsamp;
        for (int k = 0; k < nsamp; ++k) avg[k] = 0;
        regions_out[r] = wire_no;
        for (int i = 0; i < nm; ++i) {
            double l = lo[i], h = hi[i];
            if (l < (wire_no - 0.5) * pitch) l = (wire_no - 0.5) * pitch;
            if (h > (wire_no + 0.5) * pitch) h = (wire_no + 0.5) * pitch;
            if (h > l) {
                /* realseq_t(nsamples): entries past nsamples stay 0 (quirk) */
                for (int k = 0; k < nsamples && k < nsamp; ++k)
                    avg[k] = (float)((double)avg[k] + (double)m[i].resp[k] * (h - l) / pitch);
            }
        }
    }
    for (int i = 0; i < nm; ++i) free(m[i].resp);
    free(m); free(lo); free(hi); free(reg);
    return nreg;
}

/* ------------------------------------------------------------------ */
/* Fused FP64 decon: the same linear map as decon_2D_init followed by  */
/* one decon_2D_*ROI variant, without the intermediate round trips.    */
/* Y = roll_{w,t}( c2r_t( IFFT_w( FFT_w(FFT_t(D)) * Fw / R ) * Ft ) )   */
/* where R = FFT_w(FFT_t(resp padded to nw rows)), entries with R == 0  */
/* contribute 0 (the NaN/Inf fix-up of :1010-1017), and the t-mirror    */
/* keeps bins 0..nt/2 with real DC/Nyquist (Spectrum.h:56-89).         */
/* ------------------------------------------------------------------ */
int wo_decon_fused_d(const double* data, int nw, int nt, const double* resp, int nr,
                     const double* wire_filter, const double* time_filter,
                     int wire_shift, int time_shift, double* out)
{
    const size_t N = (size_t)nw * nt;
    double* s = (double*)malloc(sizeof(double) * 2 * N);
    double* r = (double*)calloc(2 * N, sizeof(double));
    if (!s || !r) { free(s); free(r); return -1; }
    for (size_t i = 0; i < N; ++i) { s[2 * i] = data[i]; s[2 * i + 1] = 0; }
    for (int w = 0; w < nr && w < nw; ++w)
        for (int t = 0; t < nt; ++t) r[2 * ((size_t)w * nt + t)] = resp[(size_t)w * nt + t];
    fft_many_d(s, s, nt, nw, 1, nt, -1);
    fft_many_d(s, s, nw, nt, nt, 1, -1);
    fft_many_d(r, r, nt, nw, 1, nt, -1);
    fft_many_d(r, r, nw, nt, nt, 1, -1);
    const int half = nt / 2;
    for (int w = 0; w < nw; ++w) {
        for (int t = 0; t <= half; ++t) {
            size_t i = (size_t)w * nt + t;
            double br = r[2 * i], bi = r[2 * i + 1];
            double den = br * br + bi * bi;
            double qr = 0, qi = 0;
            if (den != 0) {
                double g = wire_filter[w] / den;
                qr = (s[2 * i] * br + s[2 * i + 1] * bi) * g;
                qi = (s[2 * i + 1] * br - s[2 * i] * bi) * g;
            }
            s[2 * i] = qr; s[2 * i + 1] = qi;
        }
    }
    /* inverse along wires for the half spectrum columns */
    for (int t = 0; t <= half; ++t)
        fft_many_d(s + 2 * t, s + 2 * t, nw, 1, nt, 0, +1);
    for (int w = 0; w < nw; ++w) {
        double* row = s + 2 * (size_t)w * nt;
        for (int t = 0; t <= half; ++t) { row[2 * t] *= time_filter[t]; row[2 * t + 1] *= time_filter[t]; }
        hermitian_mirror_row_d(row, nt);
    }
    fft_many_d(s, s, nt, nw, 1, nt, +1);
    const double norm = 1.0 / ((double)nw * (double)nt);
    for (int w = 0; w < nw; ++w) {
        int wd = ((w + wire_shift) % nw + nw) % nw;
        for (int t = 0; t < nt; ++t) {
            int td = time_shift > 0 ? (t + time_shift) % nt : t;
            out[(size_t)wd * nt + td] = s[2 * ((size_t)w * nt + t)] * norm;
        }
    }
    free(s); free(r);
    return 0;
}

/* Forward model for the synthetic frames: the circular 2D convolution
 * whose deconvolution (above, with unit filters) returns q:
 * D = IFFT2( FFT2(q) * R * exp(+2 pi i (m s/nw + k ts/nt)) ), i.e.
 * D(w,t) = sum_{i,tau} resp(i,tau) q(w - i + s, t - tau + ts). */
int wo_forward_conv_d(const double* q, int nw, int nt, const double* resp, int nr,
                      int wire_shift, int time_shift, double* out)
{
    const size_t N = (size_t)nw * nt;
    double* s = (double*)malloc(sizeof(double) * 2 * N);
    double* r = (double*)calloc(2 * N, sizeof(double));
    if (!s || !r) { free(s); free(r); return -1; }
    const int ts = time_shift > 0 ? time_shift : 0;
    /* roll q by (-s, -ts) first, then convolve */
    for (int w = 0; w < nw; ++w) {
        int ws = ((w - wire_shift) % nw + nw) % nw;
        for (int t = 0; t < nt; ++t) {
            int tt = ((t - ts) % nt + nt) % nt;
            s[2 * ((size_t)ws * nt + tt)] = q[(size_t)w * nt + t];
            s[2 * ((size_t)ws * nt + tt) + 1] = 0;
        }
    }
    for (int w = 0; w < nr && w < nw; ++w)
        for (int t = 0; t < nt; ++t) r[2 * ((size_t)w * nt + t)] = resp[(size_t)w * nt + t];
    fft_many_d(s, s, nt, nw, 1, nt, -1);
    fft_many_d(s, s, nw, nt, nt, 1, -1);
    fft_many_d(r, r, nt, nw, 1, nt, -1);
    fft_many_d(r, r, nw, nt, nt, 1, -1);
    for (size_t i = 0; i < N; ++i) {
        double ar = s[2 * i], ai = s[2 * i + 1], br = r[2 * i], bi = r[2 * i + 1];
        s[2 * i] = ar * br - ai * bi;
        s[2 * i + 1] = ar * bi + ai * br;
    }
    fft_many_d(s, s, nw, nt, nt, 1, +1);
    fft_many_d(s, s, nt, nw, 1, nt, +1);
    const double norm = 1.0 / ((double)nw * (double)nt);
    for (size_t i = 0; i < N; ++i) out[i] = s[2 * i] * norm;
    free(s); free(r);
    return 0;
}

/* ------------------------------------------------------------------ */
/* ROI run list, pad/merge and connect                                  */
/* (find_ROI_by_decon_itself run scan :458-505, extend_ROI_self         */
/* :167-250, create_ROI_connect_info :252-360).                         */
/* mask: raw threshold mask (nw x nt).  Output ROIs flattened:          */
/* roi_count[nw], roi_pairs[2*cap] in row order.  Returns total count   */
/* or -1 if cap is exceeded.                                            */
/* ------------------------------------------------------------------ */
typedef struct { int n, cap; int* v; } wo_pairs;

static void wo_push(wo_pairs* p, int a, int b)
{
    if (p->n == p->cap) {
        p->cap = p->cap ? 2 * p->cap : 8;
        p->v = (int*)realloc(p->v, sizeof(int) * 2 * p->cap);
    }
    p->v[2 * p->n] = a; p->v[2 * p->n + 1] = b; p->n++;
}

int wo_roi_lists(const unsigned char* mask, int nw, int nt, int pad, float asy,
                 int* roi_count, int* roi_pairs, int cap)
{
    wo_pairs* rows = (wo_pairs*)calloc(nw, sizeof(wo_pairs));
    for (int w = 0; w < nw; ++w) {
        const unsigned char* m = mask + (size_t)w * nt;
        wo_pairs tmp = {0, 0, NULL};
        for (int j = 0; j < nt - 1; j++) {
            if (m[j]) {
                int roi_begin = j, roi_end = j;
                for (int k = j + 1; k < nt; k++) {
                    if (m[k]) roi_end = k; else break;
                }
                int b = roi_begin < 0 ? 0 : roi_begin;
                int e = roi_end > nt - 1 ? nt - 1 : roi_end;
                if (tmp.n == 0) wo_push(&tmp, b, e);
                else if (b <= tmp.v[2 * (tmp.n - 1) + 1]) tmp.v[2 * (tmp.n - 1) + 1] = e;
                else wo_push(&tmp, b, e);
                j = roi_end + 1;
            }
        }
        /* extend_ROI_self: pad and merge with a strict '<' */
        wo_pairs ext = {0, 0, NULL};
        for (int i = 0; i < tmp.n; ++i) {
            int b = tmp.v[2 * i] - pad; if (b < 0) b = 0;
            int e = tmp.v[2 * i + 1] + pad; if (e >= nt) e = nt - 1;
            if (ext.n == 0) wo_push(&ext, b, e);
            else if (b < ext.v[2 * (ext.n - 1) + 1]) {
                if (e > ext.v[2 * (ext.n - 1) + 1]) ext.v[2 * (ext.n - 1) + 1] = e;
            } else wo_push(&ext, b, e);
        }
        free(tmp.v);
        rows[w] = ext;
    }
    /* create_ROI_connect_info: sequential over i, lists of i+1 grow */
    for (int i = 0; i < nw - 2; i++) {
        for (int j = 0; j < rows[i].n; j++) {
            int start1 = rows[i].v[2 * j], end1 = rows[i].v[2 * j + 1];
            int length1 = end1 - start1 + 1;
            for (int k = 0; k < rows[i + 2].n; k++) {
                int start2 = rows[i + 2].v[2 * k], end2 = rows[i + 2].v[2 * k + 1];
                int length2 = end2 - start2 + 1;
                if (abs(length2 - length1) < (length2 + length1) * asy) {
                    int start3 = (int)((start1 + start2) / 2.);
                    int end3 = (int)((end1 + end2) / 2.);
                    if (start3 < end3 && start3 <= end1 && start3 <= end2 && end3 >= start1 && end3 >= start2) {
                        int flag = 0;
                        for (int i1 = 0; i1 < rows[i + 1].n; i1++) {
                            int max_start = start3;
                            if (rows[i + 1].v[2 * i1] > max_start) max_start = rows[i + 1].v[2 * i1];
                            int min_end = end3;
                            if (rows[i + 1].v[2 * i1 + 1] < min_end) min_end = rows[i + 1].v[2 * i1 + 1];
                            if (max_start < min_end) { flag = 1; break; }
                        }
                        if (flag == 0) wo_push(&rows[i + 1], start3, end3);
                    }
                }
            }
        }
    }
    int tot = 0, over = 0;
    for (int w = 0; w < nw; ++w) {
        roi_count[w] = rows[w].n;
        for (int i = 0; i < rows[w].n; ++i) {
            if (tot < cap) { roi_pairs[2 * tot] = rows[w].v[2 * i]; roi_pairs[2 * tot + 1] = rows[w].v[2 * i + 1]; }
            else over = 1;
            ++tot;
        }
        free(rows[w].v);
    }
    free(rows);
    return over ? -1 : tot;
}

/* ------------------------------------------------------------------ */
/* Loose ROIs (ROI_formation::find_ROI_loose :687-937 with local_ave   */
/* :598-618, find_ROI_end :620-653, find_ROI_begin :654-685 and        */
/* extend_ROI_loose :538-596) and ROI_refinement::apply_roi            */
/* (ROI_refinement.cxx:118-218).                                       */
/* ------------------------------------------------------------------ */
static double wo_local_ave(const float* s, int n, int bin, int width)
{
    double sum1 = 0, sum2 = 0;
    for (int i = -width; i < width + 1; i++) {
        int c = bin + i;
        while (c < 0) c += n;
        while (c >= n) c -= n;
        sum1 += s[c];
        sum2++;
    }
    return sum2 > 0 ? sum1 / sum2 : 0;
}

static int wo_roi_end(const float* s, int n, int bin, double th, int jump)
{
    int end = bin;
    double content = s[end];
    while (content > th) {
        end++;
        content = end >= n ? s[end - n] : s[end];
        if (end >= n - 1) { end = n - 1; break; }
    }
    while (wo_local_ave(s, n, end + 1, 1) < wo_local_ave(s, n, end, 1) + 25 ||
           (jump && wo_local_ave(s, n, end + 2, 1) < wo_local_ave(s, n, end, 1) + 25)) {
        end++;
        if (end >= n - 1) { end = n - 1; break; }
    }
    return end;
}

static int wo_roi_begin(const float* s, int n, int bin, double th, int jump)
{
    int begin = bin;
    double content = s[begin];
    while (content > th) {
        begin--;
        content = begin < 0 ? s[begin + n] : s[begin];
        if (begin <= 0) { begin = 0; break; }
    }
    while (wo_local_ave(s, n, begin - 1, 1) < wo_local_ave(s, n, begin, 1) + 25 ||
           (jump && wo_local_ave(s, n, begin - 2, 1) < wo_local_ave(s, n, begin, 1) + 25)) {
        begin--;
        if (begin <= 0) { begin = 0; break; }
    }
    return begin;
}

static int wo_in_ranges(const int* lo, const int* hi, int q0, int q1, int t)
{
    for (int q = q0; q < q1; ++q)
        if (t >= lo[q] && t <= hi[q]) return 1;
    return 0;
}

/* r_data: nw x nt (the loose decon).  bad_ptr[nw+1], bad_lo, bad_hi:
 * per-row inclusive tick ranges, or NULL.  rms: per-row RMS to use in
 * place of cal_RMS of the row without its bad ticks, or NULL.
 * tight_ptr[nw+1], tight_pairs: the plane's tight ROIs (inclusive) for
 * extend_ROI_loose, or NULL for none (the collection plane).  Output
 * flattened as wo_roi_lists.  The merge loop of :861-900 is not run (its
 * flag starts at 0). */
int wo_loose_rois(const float* r_data, int nw, int nt, const int* bad_ptr, const int* bad_lo,
                  const int* bad_hi, const double* rms, const int* tight_ptr, const int* tight_pairs,
                  int rebin, double l_factor, double l_max_th, double l_factor1, int l_short_length,
                  int l_jump_one_bin, int* roi_count, int* roi_pairs, int cap)
{
    float* sig = (float*)malloc(sizeof(float) * (nt > 0 ? nt : 1));
    float* sig1 = (float*)malloc(sizeof(float) * (nt > 0 ? nt : 1));
    const int n2 = rebin > 0 ? nt / rebin : 0;
    float* s2 = (float*)malloc(sizeof(float) * (n2 > 0 ? n2 : 1));
    int tot = 0, over = 0;
    for (int w = 0; w < nw; ++w) {
        const float* row = r_data + (size_t)w * nt;
        const int q0 = bad_ptr ? bad_ptr[w] : 0, q1 = bad_ptr ? bad_ptr[w + 1] : 0;
        int nc = 0;
        for (int t = 0; t < nt; ++t) {
            if (q1 > q0 && wo_in_ranges(bad_lo, bad_hi, q0, q1, t)) {
                sig1[t] = 0;
            } else {
                sig[nc++] = row[t];
                sig1[t] = row[t];
            }
        }
        for (int i = 0; i < n2; ++i) {
            double temp = 0;
            for (int j = 0; j < rebin; ++j) temp += sig1[rebin * i + j];
            s2[i] = (float)temp;
        }
        const double r = rms ? rms[w] : wo_cal_rms_f(sig, (size_t)nc);
        float th = (float)(r * rebin * l_factor);
        if (th > l_max_th) th = (float)l_max_th;
        wo_pairs rois = {0, 0, NULL};
        for (int j = 1; j < n2 - 1; j++) {
            const double content = s2[j], prev = s2[j - 1], next = s2[j + 1];
            int flag = 0, begin = 0, end = 0, max_bin = 0;
            if (content > th) {
                begin = wo_roi_begin(s2, n2, j, th * l_factor1, l_jump_one_bin);
                end = wo_roi_end(s2, n2, j, th * l_factor1, l_jump_one_bin);
                max_bin = begin;
                for (int k = begin; k <= end; k++)
                    if (s2[k] > s2[max_bin]) max_bin = k;
                flag = 1;
            } else if (content > prev && content > next) {
                begin = wo_roi_begin(s2, n2, j, prev, l_jump_one_bin);
                end = wo_roi_end(s2, n2, j, next, l_jump_one_bin);
                max_bin = begin;
                for (int k = begin; k <= end; k++)
                    if (s2[k] > s2[max_bin]) max_bin = k;
                const float rise = s2[max_bin] - s2[begin] + s2[max_bin] - s2[end];
                if (rise > th * 2) flag = 1;
                int tb = max_bin - l_short_length;
                if (tb < begin) tb = begin;
                int te = max_bin + l_short_length;
                if (te > end) te = end;
                const float d0 = s2[max_bin] - s2[tb], d1 = s2[max_bin] - s2[te];
                if (d0 > th * l_factor1 && d1 > th * l_factor1) flag = 1;
            }
            if (flag) {
                if (rois.n > 0 && begin <= rois.v[2 * (rois.n - 1) + 1]) rois.v[2 * (rois.n - 1) + 1] = end;
                else wo_push(&rois, begin, end);
                j = end < n2 ? end : n2;
            }
        }
        for (int i = 0; i < rois.n; ++i) {
            rois.v[2 * i] *= rebin;
            rois.v[2 * i + 1] = rois.v[2 * i + 1] * rebin + (rebin - 1);
        }
        if (tight_ptr) {
            wo_pairs ext = {0, 0, NULL};
            for (int i = 0; i < rois.n; ++i) {
                int start = rois.v[2 * i], stop = rois.v[2 * i + 1];
                for (int k = tight_ptr[w]; k < tight_ptr[w + 1]; ++k) {
                    const int ts = tight_pairs[2 * k], te = tight_pairs[2 * k + 1];
                    if (start > ts && start < te) start = ts;
                    if (stop > ts && stop < te) stop = te;
                }
                if (ext.n == 0) wo_push(&ext, start, stop);
                else if (start < ext.v[2 * (ext.n - 1) + 1]) ext.v[2 * (ext.n - 1) + 1] = stop;
                else wo_push(&ext, start, stop);
            }
            free(rois.v);
            rois = ext;
        }
        roi_count[w] = rois.n;
        for (int i = 0; i < rois.n; ++i) {
            if (tot < cap) { roi_pairs[2 * tot] = rois.v[2 * i]; roi_pairs[2 * tot + 1] = rois.v[2 * i + 1]; }
            else over = 1;
            ++tot;
        }
        free(rois.v);
    }
    free(sig);
    free(sig1);
    free(s2);
    return over ? -1 : tot;
}

/* The int conversion of a float as x86-64 performs it (cvttss2si):
 * truncation toward zero, NaN and values outside the int range give
 * INT_MIN. */
static int wo_x86_trunc(float v)
{
    if (v >= -2147483648.0f && v < 2147483648.0f) return (int)v;
    return (int)0x80000000u;
}

/* apply_roi on one plane (r_data nw x nt, in place): samples outside
 * every ROI become 0; ROIs (CSR roi_ptr[nw+1], inclusive pairs) are
 * applied in order, so a later one overwrites an overlap.  linear = 1
 * (induction, :120-166): sample i of ROI [s, e] becomes the int of
 * r(i) - ((r(e) - r(s)) * (i - s) / (e - s) + r(s)) in float arithmetic
 * (a one-tick ROI computes 0 * 0 / 0); linear = 0 (collection,
 * :167-202): the int of r(i). */
void wo_apply_roi(float* r_data, int nw, int nt, const int* roi_ptr, const int* roi_pairs, int linear)
{
    float* out = (float*)malloc(sizeof(float) * (nt > 0 ? nt : 1));
    for (int w = 0; w < nw; ++w) {
        float* row = r_data + (size_t)w * nt;
        for (int t = 0; t < nt; ++t) out[t] = 0;
        for (int k = roi_ptr[w]; k < roi_ptr[w + 1]; ++k) {
            const int sb = roi_pairs[2 * k], eb = roi_pairs[2 * k + 1];
            const float sc = row[sb], ec = row[eb];
            for (int i = sb; i < eb + 1; i++) {
                float v;
                if (linear) {
                    const float slope_part = (ec - sc) * (float)(i - sb) / (float)(eb - sb);
                    v = row[i] - (slope_part + sc);
                } else {
                    v = row[i];
                }
                out[i] = (float)wo_x86_trunc(v);
            }
        }
        for (int t = 0; t < nt; ++t) row[t] = out[t];
    }
    free(out);
}

/* ------------------------------------------------------------------ */
/* Synthetic frames: counter-based RNG (splitmix64 of seed, stream and */
/* index), Box-Muller normals, straight tracks rasterised at           */
/* q_per_tick electrons per tick.                                      */
/* ------------------------------------------------------------------ */
static uint64_t wo_mix(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double wo_uniform(uint64_t seed, uint64_t stream, uint64_t idx)
{
    uint64_t z = wo_mix(seed ^ wo_mix(stream ^ wo_mix(idx)));
    return ((z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

void wo_normal(uint64_t seed, uint64_t stream, size_t n, double sigma, double* out)
{
    for (size_t i = 0; i < n; ++i) {
        double u1 = wo_uniform(seed, stream, 2 * i), u2 = wo_uniform(seed, stream, 2 * i + 1);
        out[i] = sigma * sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
    }
}

/* Adds ntracks straight tracks with uniform endpoints in
 * [0,nw) x [0,nt) to q (nw x nt), q_per_tick electrons per tick. */
void wo_tracks(uint64_t seed, uint64_t stream, int ntracks, int nw, int nt,
               double q_per_tick, double* q)
{
    for (int k = 0; k < ntracks; ++k) {
        double w0 = wo_uniform(seed, stream, 4 * k) * nw;
        double t0 = wo_uniform(seed, stream, 4 * k + 1) * nt;
        double w1 = wo_uniform(seed, stream, 4 * k + 2) * nw;
        double t1 = wo_uniform(seed, stream, 4 * k + 3) * nt;
        double dw = fabs(w1 - w0), dt = fabs(t1 - t0);
        int nstep = (int)ceil(dw > dt ? dw : dt) + 1;
        double per = q_per_tick * (dt + 1.0) / nstep;
        for (int s = 0; s < nstep; ++s) {
            double f = nstep > 1 ? (double)s / (nstep - 1) : 0;
            int w = (int)floor(w0 + f * (w1 - w0));
            int t = (int)floor(t0 + f * (t1 - t0));
            if (w < 0) w = 0;
            if (w >= nw) w = nw - 1;
            if (t < 0) t = 0;
            if (t >= nt) t = nt - 1;
            q[(size_t)w * nt + t] += per;
        }
    }
}
