// Fused 2D deconvolution of OmnibusSigProc (sigproc/src/OmnibusSigProc.cxx:
// 944-1160) plus the tight-ROI RMS/threshold of ROI_formation
// (sigproc/src/ROI_formation.cxx:364-531), batched over planes.
//
// The reference runs, per plane, 7 full complex transforms in
// decon_2D_init and 3 more per time-filter variant, with element-wise
// passes in between.  All of them are linear and their composition is
//   Y = roll_w,t( c2r_t( IFFT_w( FFT_w( r2c_t(D) ) * Fw/R ) * Ft ) )
// so the same output comes from three stages with one spectrum buffer:
//   K1  row pass:    real FFT of every wire row (r2c, half spectrum),
//                    keeping the kc bins the time filter passes;
//   K2  wire pass:   for tiles of TK bins, FFT along wires in LDS,
//                    multiply by K = Fw Ft phase / (R nw nt), inverse FFT;
//   K3  row pass:    inverse real FFT, restore_baseline, cal_RMS,
//                    threshold mask, outputs.
// K is built once per plane type in FP64 (wct_decon_kernel_create).
#include <algorithm>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "decon_kernels.hpp"

namespace wct {

// ---- spectrum mode: the time spectrum of the deconvolved frame -------
// X = c2r(Y) unnormalised (K carries 1/(nw nt)), so FFT_t(X)[k] is nt
// times the Hermitian extension of the kept half spectrum Y: bins
// 0..nt/2 from Y (zero past kc) with real DC and Nyquist, bin nt - k the
// conjugate of bin k.  One workgroup per row, grid-stride over rows.
__global__ __launch_bounds__(256) void k_row_spectrum(const DJob* __restrict__ jobs, int njobs, long rbeg,
                                                      long rend)
{
    for (long r = rbeg + blockIdx.x; r < rend; r += gridDim.x) {
        const DJob& J = jobs[find_job_rows(jobs, njobs, r)];
        const int w = (int)(r - J.row0);
        const auto Y = G(J.S + (long)w * J.ldk);
        const auto o = G(J.spec + (long)w * J.ld_spec);
        const float sc = (float)J.nt;
        for (int k = threadIdx.x; k < J.nt; k += blockDim.x) {
            const int kk = k <= J.M ? k : J.nt - k;
            float2 v;
            v.x = v.y = 0.f;
            if (kk < J.kc) v = Y[kk];
            if (kk == 0 || kk == J.M) v.y = 0.f;
            if (k > J.M) v.y = -v.y;
            v.x *= sc;
            v.y *= sc;
            o[k] = v;
        }
    }
}

// ---- FP64 kernel build ------------------------------------------------

// K(m,k) for k <= nt/2, stored k-major in double with |K| per entry.
// Ft NULL: no time filter in K (the outputs apply theirs); Fenv (or
// NULL) then weights |K| for the spectral cut.  |R|^2 <= den_floor gives
// K = 0 (den_floor 0: only R == 0).
__global__ void k_build_K(const double2* __restrict__ R, int nw, int nt, int kh,
                          const double* __restrict__ Fw, const double* __restrict__ Ft,
                          const double* __restrict__ Fenv, double den_floor,
                          int wshift, int tshift, double2* __restrict__ Kd,
                          double* __restrict__ Kabs)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)kh * nw) return;
    const int k = (int)(i / nw), m = (int)(i % nw);
    const double2 r = R[(long)m * nt + k];
    const double den = r.x * r.x + r.y * r.y;
    double2 out;
    out.x = out.y = 0.0;
    if (den > den_floor) {
        const long qm = ((long)m * wshift) % nw, qk = ((long)k * tshift) % nt;
        const double ang = -2.0 * M_PI * ((double)qm / nw + (double)qk / nt);
        double sn, cs;
        sincos(ang, &sn, &cs);
        const double g = Fw[m] * (Ft ? Ft[k] : 1.0) / (den * (double)nw * (double)nt);
        // phase / R = phase * conj(R) / |R|^2
        const double pr = cs * r.x + sn * r.y;   // (cs + i sn)(r.x - i r.y)
        const double pi = sn * r.x - cs * r.y;
        out.x = pr * g;
        out.y = pi * g;
    }
    Kd[i] = out;
    Kabs[i] = sqrt(out.x * out.x + out.y * out.y) * (Fenv ? fabs(Fenv[k]) : 1.0);
}

// Per-block maxima of |R|^2 over n entries (grid-stride).
__global__ void k_den_max(const double2* __restrict__ R, long n, double* __restrict__ bmax)
{
    __shared__ double red[kNT];
    double m = 0;
    for (long i = (long)blockIdx.x * kNT + threadIdx.x; i < n; i += (long)gridDim.x * kNT)
        m = fmax(m, R[i].x * R[i].x + R[i].y * R[i].y);
    red[threadIdx.x] = m;
    lds_barrier();
    for (int s = kNT / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        lds_barrier();
    }
    if (threadIdx.x == 0) bmax[blockIdx.x] = red[0];
}

// Forward response spectrum for the synthetic-frame generator and the
// adjoint's forward model: K = R exp(+2 pi i (m ws/nw + k ts/nt)) / (nw nt).
__global__ void k_build_fwd(const double2* __restrict__ R, int nw, int nt, int kh, int wshift,
                            int tshift, double2* __restrict__ Kd, double* __restrict__ Kabs)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)kh * nw) return;
    const int k = (int)(i / nw), m = (int)(i % nw);
    const double2 r = R[(long)m * nt + k];
    const long qm = ((long)m * wshift) % nw, qk = ((long)k * tshift) % nt;
    const double ang = 2.0 * M_PI * ((double)qm / nw + (double)qk / nt);
    double sn, cs;
    sincos(ang, &sn, &cs);
    const double g = 1.0 / ((double)nw * (double)nt);
    double2 out;
    out.x = (r.x * cs - r.y * sn) * g;
    out.y = (r.x * sn + r.y * cs) * g;
    Kd[i] = out;
    Kabs[i] = sqrt(out.x * out.x + out.y * out.y);
}

__global__ void k_colmax(const double* __restrict__ Kabs, int nw, int kh, double* __restrict__ cmax)
{
    __shared__ double red[kNT];
    const int k = blockIdx.x;
    double m = 0;
    for (int i = threadIdx.x; i < nw; i += kNT) m = fmax(m, Kabs[(long)k * nw + i]);
    red[threadIdx.x] = m;
    lds_barrier();
    for (int s = kNT / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        lds_barrier();
    }
    if (threadIdx.x == 0) cmax[k] = red[0];
}

__global__ void k_to_f32(const double2* __restrict__ in, float2* __restrict__ out, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        float2 v;
        v.x = (float)in[i].x;
        v.y = (float)in[i].y;
        out[i] = v;
    }
}

__global__ void k_real_to_cplx(const double* __restrict__ in, double2* __restrict__ out, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        double2 v;
        v.x = in[i];
        v.y = 0.0;
        out[i] = v;
    }
}

// Response convolution of init_overall_response (:793-813): each row's
// spectrum times the scaled electronics spectrum times fine_period, then
// the Hermitian mirror (bins above n/2 from the conjugate of those below,
// DC and Nyquist real) ahead of the inverse.
__global__ void k_resp_mul_mirror(double2* __restrict__ rows, const double2* __restrict__ elec,
                                  int nr, int n, double fine_period)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)nr * n) return;
    const int r = (int)(i / n), k = (int)(i % n);
    double2* row = rows + (long)r * n;
    // only the lower half is computed here; the mirror fills the rest
    if (k > n / 2) return;
    double2 c = row[k], e = elec[k], p;
    p.x = (c.x * e.x - c.y * e.y) * fine_period;
    p.y = (c.x * e.y + c.y * e.x) * fine_period;
    if (k == 0 || (n % 2 == 0 && k == n / 2)) p.y = 0.0;
    row[k] = p;
    if (k > 0 && n - k != k) {
        double2 q;
        q.x = p.x;
        q.y = -p.y;
        row[n - k] = q;
    }
}

// Default bins per wire-pass tile: 4 (the autotuner's pick on the
// 150-APA workload, 21.9 against 22.3 ms per step with 8; tiles of 2, 4
// and 8 give the same bits).  Planes above 2047 wires are refused.
static int pick_tk(int) { return 4; }

// Complex slots per thread of the row kernels for rows of up to Mmax
// complex samples (M = nt/2), at kNT threads per row.
template <typename F>
static int dispatch_vh(int Mmax, F&& f)
{
    const int vh = (Mmax + kNT - 1) / kNT;
    if (vh <= 4) return f(std::integral_constant<int, 4>());
    if (vh <= 8) return f(std::integral_constant<int, 8>());
    if (vh <= 12) return f(std::integral_constant<int, 12>());
    if (vh <= 16) return f(std::integral_constant<int, 16>());
    if (vh <= 20) return f(std::integral_constant<int, 20>());
    if (vh <= 24) return f(std::integral_constant<int, 24>());
    if (vh <= 32) return f(std::integral_constant<int, 32>());
    set_error("time extent %d above the supported 16384 ticks", 2 * Mmax);
    return WCT_ERR_UNSUPPORTED;
}

static int run_rows(wct_decon_plan p, hipStream_t s, bool c2r)
{
    if (c2r && p->spectrum) {
        const long rows = p->rows;
        const int grid = (int)std::min<long>(rows, (long)p->ncu * 8);
        hipLaunchKernelGGL(k_row_spectrum, dim3(grid), dim3(256), 0, s, p->jobs, p->njobs, 0L, rows);
        WCT_HIP_CHECK(hipGetLastError());
        return WCT_OK;
    }
    for (const auto& g : p->rgroups) {
        const int rc = dispatch_vh(g.len, [&](auto V) -> int {
            constexpr int VH = decltype(V)::value;
            return row_pass_vh<VH>(p, g, s, c2r);
        });
        if (rc) return rc;
    }
    return WCT_OK;
}

static int run_wire(wct_decon_plan p, hipStream_t s)
{
    for (const auto& g : p->tgroups) {
        int rc = WCT_OK;
        switch (p->tk) {
        case 16: rc = wire_pass_tk<16>(p, g, s); break;
        case 8: rc = wire_pass_tk<8>(p, g, s); break;
        case 4: rc = wire_pass_tk<4>(p, g, s); break;
        case 2: rc = wire_pass_tk<2>(p, g, s); break;
        default:
            set_error("decon_plan_run: wire tile %d not supported", p->tk);
            rc = WCT_ERR_UNSUPPORTED;
        }
        if (rc) return rc;
    }
    return WCT_OK;
}

int plan_run(wct_decon_plan p, hipStream_t s)
{
    int rc = run_rows(p, s, false);
    if (rc) return rc;
    rc = run_wire(p, s);
    if (rc) return rc;
    return run_rows(p, s, true);
}

// The job of a one-off launch, passed by value and stored in stream-
// ordered scratch (no host buffer has to outlive the call).
__global__ void k_put_job(const DJob job, DJob* dst)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *dst = job;
}

// g[k] = f[k] * sc: inv_c2r's 1/nt folded into the per-bin filter that
// the c2r stage applies.
__global__ void k_scale_filter(const float* __restrict__ f, float* __restrict__ g, int n, float sc)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) g[k] = f[k] * sc;
}

// Job table (one DJob) then the scaled filter (M + 1 floats).
size_t spectrum_filter_scratch_bytes(int M)
{
    const size_t jb = (sizeof(DJob) + 255) & ~(size_t)255;
    return jb + sizeof(float) * (size_t)(M + 1);
}

// Decon2DFilter's pass through K3's c2r stage over the caller's spectrum
// rows, with the one-job table and the scaled filter in scratch (device
// memory of spectrum_filter_scratch_bytes(nt / 2)).  Enqueues only.
int spectrum_filter_run(const float* spectrum, int ld_spectrum, int nw, int nt, const float* filter,
                        int restore_baseline, float* out, int ld_out, const DevPlan* pM, const float2* line,
                        void* scratch, hipStream_t s)
{
    const int M = nt / 2;
    WCT_REQUIRE(ld_spectrum >= M + 1 && ld_out >= nt && ld_out % 2 == 0 &&
                    (reinterpret_cast<uintptr_t>(spectrum) & 7) == 0 && (reinterpret_cast<uintptr_t>(out) & 7) == 0,
                WCT_ERR_ARG, "spectrum_filter: bad strides or misaligned buffers");
    const size_t jb = (sizeof(DJob) + 255) & ~(size_t)255;
    float* g = reinterpret_cast<float*>(static_cast<char*>(scratch) + jb);
    DJob d;
    memset(&d, 0, sizeof(d));
    d.out = out;
    d.ld_out = ld_out;
    d.S = const_cast<float2*>(reinterpret_cast<const float2*>(spectrum));
    d.ldk = ld_spectrum;
    d.lf1 = g;
    d.nw = nw;
    d.nt = nt;
    d.M = M;
    d.kc = M + 1;
    d.restore = restore_baseline ? 1 : 0;
    d.planM = pM->plan;
    d.twM = (const float2*)pM->tw;
    d.wN = line;
    // the job by value into the table, inv_c2r's 1/nt into the filter
    hipLaunchKernelGGL(k_put_job, dim3(1), dim3(1), 0, s, d, static_cast<DJob*>(scratch));
    hipLaunchKernelGGL(k_scale_filter, dim3((M + 1 + 255) / 256), dim3(256), 0, s, filter, g, M + 1,
                       1.0f / (float)nt);
    WCT_HIP_CHECK(hipGetLastError());
    // a transient one-job plan on the stack, on the runtime row plan (the
    // job carries no twiddles of the compile-time schedule)
    wct_decon_plan_s tmp;
    tmp.row_plan = 0;
    tmp.jobs = static_cast<DJob*>(scratch);
    tmp.njobs = 1;
    tmp.rows = nw;
    tmp.Mmax = M;
    tmp.twM = d.planM.twcount;
    tmp.rgroups.push_back({M, 0, (long)nw, d.planM.twcount, is_smooth5(M), false, false});
    int dev = 0;
    WCT_HIP_CHECK(hipGetDevice(&dev));
    WCT_HIP_CHECK(hipDeviceGetAttribute(&tmp.ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int rc = run_rows(&tmp, s, true);
    tmp.jobs = nullptr;
    return rc;
}

}  // namespace wct

using namespace wct;

extern "C" {

int wct_overall_response(const double* fr_avg, int nr, int nin, const double* elec, int nfine,
                         double fine_period, double coarse_period, double inter_gain,
                         double adc_mv, double fine_time_offset, int nt, double* resp_out,
                         void* stream)
{
    WCT_ENTRY();
    WCT_REQUIRE(fr_avg && elec && resp_out && nr > 0 && nin > 0 && nfine >= nin && nt > 0 &&
                    fine_period > 0 && coarse_period > 0,
                WCT_ERR_ARG, "overall_response: bad arguments");
    WCT_REQUIRE(fft_supported_length(nfine), WCT_ERR_UNSUPPORTED,
                "fine length %d has a prime factor above 13", nfine);
    hipStream_t s = as_stream(stream);
    // host staging: rows zero-padded to nfine, electronics scaled by
    // -inter_gain*adc_mv (Waveform::scale at :768)
    std::vector<double> h((size_t)(nr + 1) * nfine * 2, 0.0);
    for (int r = 0; r < nr; ++r)
        for (int i = 0; i < nin; ++i) h[2 * ((size_t)r * nfine + i)] = fr_avg[(size_t)r * nin + i];
    const double esc = inter_gain * adc_mv * (-1);
    for (int i = 0; i < nfine; ++i) h[2 * ((size_t)nr * nfine + i)] = elec[i] * esc;
    DevTemp<double2> d(s);
    const size_t bytes = h.size() * sizeof(double);
    WCT_TRY(d.alloc(h.size() / 2));
    WCT_HIP_CHECK(hipMemcpyAsync(d, h.data(), bytes, hipMemcpyHostToDevice, s));
    int rc = launch_rows<double>(d, d, nfine, nr + 1, nfine, -1, 1.0, s);
    if (rc) return rc;
    const long tot = (long)nr * nfine;
    hipLaunchKernelGGL(k_resp_mul_mirror, dim3((tot + 255) / 256), dim3(256), 0, s, d,
                       d + (size_t)nr * nfine, nr, nfine, fine_period);
    WCT_HIP_CHECK(hipGetLastError());
    rc = launch_rows<double>(d, d, nfine, nr, nfine, +1, 1.0 / nfine, s);
    if (rc) return rc;
    WCT_HIP_CHECK(hipMemcpyAsync(h.data(), d, (size_t)nr * nfine * 2 * sizeof(double),
                                 hipMemcpyDeviceToHost, s));
    WCT_HIP_CHECK(hipStreamSynchronize(s));
    // fine shift (:816-831) and redigitise (:834-859) on the host: a
    // short serial interpolation per row, kept in the reference's form
    // (float tick grids, swapped interpolation weights).
    std::vector<double> arr((size_t)nfine);
    const int fts = (int)(fine_time_offset / fine_period);
    for (int r = 0; r < nr; ++r) {
        for (int i = 0; i < nfine; ++i) {
            const int dst = fts > 0 ? (i + fts) % nfine : i;
            arr[dst] = h[2 * ((size_t)r * nfine + i)];
        }
        size_t fcount = 1;
        for (int i = 0; i < nt; ++i) {
            const double ctime = (float)(i * coarse_period);
            if (fcount < (size_t)nfine) {
                while (ctime > (double)(float)(fcount * fine_period)) {
                    fcount++;
                    if (fcount >= (size_t)nfine) break;
                }
            }
            double v = 0;
            if (fcount < (size_t)nfine) {
                const double f0 = (float)((fcount - 1) * fine_period);
                const double f1 = (float)(fcount * fine_period);
                v = (ctime - f0) / fine_period * arr[fcount - 1] + (f1 - ctime) / fine_period * arr[fcount];
            }
            resp_out[(size_t)r * nt + i] = v;
        }
    }
    return WCT_OK;
}

// Correction factors of one chunk of channel responses: spectra C (nc
// rows of nt, row 0 the electronics spectrum E), out row r = E / C_{r+1}
// over bins [0, nb), 0 where C is exactly 0 (OmnibusSigProc.cxx:972-978).
__global__ void k_chan_factor(const double2* __restrict__ C, int nc, int nt, int nb, float2* __restrict__ out,
                              int ld)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)nc * nb) return;
    const int r = (int)(i / nb), k = (int)(i % nb);
    const double2 e = C[k], c = C[(long)(r + 1) * nt + k];
    float2 o;
    o.x = o.y = 0.f;
    if (c.x != 0.0 || c.y != 0.0) {
        // e / c with Smith's scaling
        double qr, qi;
        if (fabs(c.x) >= fabs(c.y)) {
            const double rat = c.y / c.x, den = c.x + c.y * rat;
            qr = (e.x + e.y * rat) / den;
            qi = (e.y - e.x * rat) / den;
        } else {
            const double rat = c.x / c.y, den = c.x * rat + c.y;
            qr = (e.x * rat + e.y) / den;
            qi = (e.y * rat - e.x) / den;
        }
        o.x = (float)qr;
        o.y = (float)qi;
    }
    out[(long)r * ld + k] = o;
}

int wct_channel_factors(float* out, int ld_out, const float* responses, int nresp, int nsamp,
                        const double* elec, int nt, void* stream)
{
    WCT_ENTRY();
    WCT_REQUIRE(out && responses && elec && nresp > 0 && nsamp > 0 && nt >= 2, WCT_ERR_ARG,
                "channel_factors: bad arguments");
    WCT_REQUIRE(nt <= kMaxDftLength, WCT_ERR_UNSUPPORTED, "channel_factors: %d ticks above %d", nt,
                kMaxDftLength);
    const int nb = nt / 2 + 1;
    WCT_REQUIRE(ld_out >= nb && (reinterpret_cast<uintptr_t>(out) & 7) == 0, WCT_ERR_ARG,
                "channel_factors: row stride %d below %d bins or misaligned table", ld_out, nb);
    hipStream_t s = as_stream(stream);
    // chunks of responses, each transformed with the electronics row
    const int chunk = std::min(nresp, 256);
    std::vector<double> h((size_t)(chunk + 1) * nt * 2);
    DevTemp<double2> d(s);
    WCT_TRY(d.alloc(h.size() / 2));
    int rc = WCT_OK;
    for (int r0 = 0; r0 < nresp && rc == WCT_OK; r0 += chunk) {
        const int nc = std::min(chunk, nresp - r0);
        std::fill(h.begin(), h.end(), 0.0);
        for (int t = 0; t < nt; ++t) h[2 * (size_t)t] = elec[t];
        const int ns = std::min(nsamp, nt);   // cut or zero-padded to nt (:966)
        for (int r = 0; r < nc; ++r)
            for (int t = 0; t < ns; ++t)
                h[2 * ((size_t)(r + 1) * nt + t)] = responses[(size_t)(r0 + r) * nsamp + t];
        const size_t used = (size_t)(nc + 1) * nt * 2 * sizeof(double);
        rc = hipMemcpyAsync(d, h.data(), used, hipMemcpyHostToDevice, s) == hipSuccess ? WCT_OK : WCT_ERR_HIP;
        if (rc) {
            set_error("channel_factors: upload failed");
            break;
        }
        rc = launch_rows<double>(d, d, nt, nc + 1, nt, -1, 1.0, s);
        if (rc) break;
        const long tot = (long)nc * nb;
        hipLaunchKernelGGL(k_chan_factor, dim3((tot + 255) / 256), dim3(256), 0, s, d, nc, nt, nb,
                           reinterpret_cast<float2*>(out) + (size_t)r0 * ld_out, ld_out);
        if (hipGetLastError() != hipSuccess) {
            set_error("channel_factors: launch failed");
            rc = WCT_ERR_HIP;
            break;
        }
        // the host staging is refilled by the next chunk
        if (hipStreamSynchronize(s) != hipSuccess) {
            set_error("channel_factors: stream failed");
            rc = WCT_ERR_HIP;
        }
    }
    return rc;
}

// forward: the response convolution.  Otherwise the decon kernel, with
// time_filter baked in, or (time_filter NULL) without a time filter and
// envelope (or NULL) weighting the spectral cut.
static int kernel_create(wct_decon_kernel* out, const double* resp, int nr, int nw, int nt,
                         const double* wire_filter, const double* time_filter, const double* envelope,
                         int wire_shift, int time_shift, double cut_tol, double r_floor, bool forward,
                         void* stream)
{
    WCT_REQUIRE(out && resp && nr > 0 && nw > 0 && nt > 1 && (forward || wire_filter) && r_floor >= 0 &&
                    r_floor < 1,
                WCT_ERR_ARG, "kernel_create: bad arguments");
    WCT_REQUIRE(nt % 2 == 0, WCT_ERR_UNSUPPORTED, "odd tick count %d not supported", nt);
    WCT_REQUIRE(fft_supported_length(nw) && fft_supported_length(nt / 2), WCT_ERR_UNSUPPORTED,
                "plane %d x %d: extents need prime factors <= 13", nw, nt);
    WCT_REQUIRE(nw <= 2047, WCT_ERR_UNSUPPORTED, "more than 2047 wires in a plane");
    *out = nullptr;
    hipStream_t s = as_stream(stream);
    const int kh = nt / 2 + 1;
    const size_t N = (size_t)nw * nt;
    DevTemp<double2> R(s);
    DevTemp<double> tmp(s);
    WCT_TRY(R.alloc(N));
    WCT_HIP_CHECK(hipMemsetAsync(R, 0, N * sizeof(double2), s));
    const int nrr = std::min(nr, nw);
    WCT_TRY(tmp.alloc((size_t)nrr * nt));
    WCT_HIP_CHECK(hipMemcpyAsync(tmp, resp, (size_t)nrr * nt * sizeof(double), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_real_to_cplx, dim3(((long)nrr * nt + 255) / 256), dim3(256), 0, s, tmp, R,
                       (long)nrr * nt);
    WCT_HIP_CHECK(hipGetLastError());
    // R = FFT_w(FFT_t(resp)) in FP64 (:988-998)
    int rc = launch_rows<double>(R, R, nt, nrr, nt, -1, 1.0, s);
    if (rc) return rc;
    rc = launch_cols_any<double>(R, R, nw, nt, -1, 1.0, s);
    if (rc) return rc;
    // |R|^2 floor of the regularised inverse: (r_floor max|R|)^2
    double den_floor = 0.0;
    if (r_floor > 0 && !forward) {
        constexpr int nb = 512;
        DevTemp<double> bmax(s);
        WCT_TRY(bmax.alloc(nb));
        hipLaunchKernelGGL(k_den_max, dim3(nb), dim3(kNT), 0, s, R, (long)N, bmax);
        WCT_HIP_CHECK(hipGetLastError());
        std::vector<double> hb(nb);
        WCT_HIP_CHECK(hipMemcpyAsync(hb.data(), bmax, nb * sizeof(double), hipMemcpyDeviceToHost, s));
        WCT_HIP_CHECK(hipStreamSynchronize(s));
        double dmax = 0;
        for (double v : hb) dmax = std::max(dmax, v);
        den_floor = r_floor * r_floor * dmax;
    }
    DevTemp<double> Fw(s), Ft(s), Fe(s), Kabs(s), cmax(s);
    DevTemp<double2> Kd(s);
    WCT_TRY(Fw.alloc(nw));
    WCT_TRY(Ft.alloc(nt));
    WCT_TRY(Fe.alloc(nt));
    WCT_TRY(Kd.alloc((size_t)kh * nw));
    WCT_TRY(Kabs.alloc((size_t)kh * nw));
    WCT_TRY(cmax.alloc(kh));
    const long nk = (long)kh * nw;
    if (forward) {
        hipLaunchKernelGGL(k_build_fwd, dim3((nk + 255) / 256), dim3(256), 0, s, R, nw, nt, kh,
                           wire_shift, time_shift > 0 ? time_shift : 0, Kd, Kabs);
    } else {
        WCT_HIP_CHECK(hipMemcpyAsync(Fw, wire_filter, nw * sizeof(double), hipMemcpyHostToDevice, s));
        if (time_filter)
            WCT_HIP_CHECK(hipMemcpyAsync(Ft, time_filter, nt * sizeof(double), hipMemcpyHostToDevice, s));
        if (!time_filter && envelope)
            WCT_HIP_CHECK(hipMemcpyAsync(Fe, envelope, nt * sizeof(double), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_build_K, dim3((nk + 255) / 256), dim3(256), 0, s, R, nw, nt, kh, Fw,
                           time_filter ? Ft : nullptr, (!time_filter && envelope) ? Fe : nullptr, den_floor,
                           wire_shift, time_shift > 0 ? time_shift : 0, Kd, Kabs);
    }
    WCT_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_colmax, dim3(kh), dim3(kNT), 0, s, Kabs, nw, kh, cmax);
    WCT_HIP_CHECK(hipGetLastError());
    std::vector<double> hc(kh);
    WCT_HIP_CHECK(hipMemcpyAsync(hc.data(), cmax, kh * sizeof(double), hipMemcpyDeviceToHost, s));
    WCT_HIP_CHECK(hipStreamSynchronize(s));
    double gmax = 0;
    for (double v : hc) gmax = std::max(gmax, v);
    int kc = kh;
    if (cut_tol > 0) {
        kc = 1;
        for (int k = kh - 1; k >= 0; --k)
            if (hc[k] > cut_tol * gmax) { kc = k + 1; break; }
    }
    auto* kk = new wct_decon_kernel_s();
    kk->nw = nw;
    kk->nt = nt;
    kk->kc = kc;
    if (hipGetDevice(&kk->device) != hipSuccess) {
        delete kk;
        set_error("kernel_create: no current device");
        return WCT_ERR_HIP;
    }
    if (const int rc_a = dev_alloc((void**)&kk->K, (size_t)kc * nw * sizeof(float2))) {
        delete kk;
        return rc_a;
    }
    const long nkc = (long)kc * nw;
    hipLaunchKernelGGL(k_to_f32, dim3((nkc + 255) / 256), dim3(256), 0, s, Kd, kk->K, nkc);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
        (void)hipFree(kk->K);
        delete kk;
        set_error("kernel_create: conversion kernel failed");
        return WCT_ERR_HIP;
    }
    *out = kk;
    return WCT_OK;
}

int wct_decon_kernel_create(wct_decon_kernel* out, const double* resp, int nr, int nw, int nt,
                            const double* wire_filter, const double* time_filter, int wire_shift,
                            int time_shift, double cut_tol, void* stream)
{
    WCT_ENTRY();
    WCT_REQUIRE(time_filter, WCT_ERR_ARG, "kernel_create: no time filter");
    return kernel_create(out, resp, nr, nw, nt, wire_filter, time_filter, nullptr, wire_shift, time_shift,
                         cut_tol, 0.0, false, stream);
}

int wct_decon_kernel_create_unfiltered(wct_decon_kernel* out, const double* resp, int nr, int nw, int nt,
                                       const double* wire_filter, const double* envelope, int wire_shift,
                                       int time_shift, double cut_tol, double r_floor, void* stream)
{
    WCT_ENTRY();
    return kernel_create(out, resp, nr, nw, nt, wire_filter, nullptr, envelope, wire_shift, time_shift,
                         cut_tol, r_floor, false, stream);
}

int wct_forward_kernel_create(wct_decon_kernel* out, const double* resp, int nr, int nw, int nt,
                              int wire_shift, int time_shift, void* stream)
{
    WCT_ENTRY();
    return kernel_create(out, resp, nr, nw, nt, nullptr, nullptr, nullptr, wire_shift, time_shift, 0.0, 0.0,
                         true, stream);
}

int wct_decon_kernel_info(wct_decon_kernel k, int* nw, int* nt, int* kc)
{
    WCT_REQUIRE(k, WCT_ERR_ARG, "null kernel");
    if (nw) *nw = k->nw;
    if (nt) *nt = k->nt;
    if (kc) *kc = k->kc;
    return WCT_OK;
}

int wct_decon_kernel_download(wct_decon_kernel k, float* host_out)
{
    WCT_REQUIRE(k && host_out, WCT_ERR_ARG, "kernel_download: bad arguments");
    WCT_HIP_CHECK(hipMemcpy(host_out, k->K, (size_t)k->kc * k->nw * sizeof(float2), hipMemcpyDeviceToHost));
    return WCT_OK;
}

int wct_decon_kernel_destroy(wct_decon_kernel k)
{
    if (!k) return WCT_OK;
    if (k->K) WCT_HIP_CHECK(hipFree(k->K));
    delete k;
    return WCT_OK;
}

// Spectrum row length: kc rounded up to 16 bins, a multiple of every
// wire-pass tile width.
static int job_ldk(const wct_decon_job& j) { return (j.kernel->kc + 15) / 16 * 16; }

static int wire_tile_max_nw(int tk) { return tk <= 4 ? 2047 : 1023; }

size_t wct_decon_workspace_size(const wct_decon_job* jobs, int njobs)
{
    if (!jobs || njobs <= 0) return 0;
    size_t tot = 0;
    for (int i = 0; i < njobs; ++i) {
        if (!jobs[i].kernel) return 0;
        tot += (size_t)jobs[i].kernel->nw * job_ldk(jobs[i]) * sizeof(float2);
        tot = (tot + 255) & ~(size_t)255;
    }
    return tot + 256;
}

int wct_decon_plan_create(wct_decon_plan* out, const wct_decon_job* jobs, int njobs,
                          void* workspace, size_t ws_bytes)
{
    WCT_ENTRY();
    return wct_decon_plan_create_tuned(out, jobs, njobs, workspace, ws_bytes, nullptr);
}

int wct_decon_plan_create_tuned(wct_decon_plan* out, const wct_decon_job* jobs, int njobs,
                                void* workspace, size_t ws_bytes, const wct_decon_tuning* tuning)
{
    WCT_ENTRY();
    WCT_REQUIRE(out && jobs && njobs > 0, WCT_ERR_ARG, "decon_plan: no jobs");
    *out = nullptr;
    const size_t need = wct_decon_workspace_size(jobs, njobs);
    WCT_REQUIRE(need > 0, WCT_ERR_ARG, "decon_plan: a job has no kernel");
    WCT_REQUIRE(workspace && ws_bytes >= need, WCT_ERR_ARG,
                "decon_plan: workspace of %zu bytes, %zu needed", ws_bytes, need);
    std::vector<DJob> h(njobs);
    size_t off = 0;
    long rows = 0, tiles = 0;
    int nw_all = 0;
    for (int i = 0; i < njobs; ++i) {
        WCT_REQUIRE(jobs[i].kernel, WCT_ERR_ARG, "decon_plan: job %d has no kernel", i);
        nw_all = std::max(nw_all, jobs[i].kernel->nw);
    }
    int Mmax = 0, tk = pick_tk(nw_all), nwmax = 0, twM = 0, twW = 0;
    int row_pct = 100, row_plan = 1, r2c_blocks = 0, c2r_blocks = 0;
    if (tuning) {
        if (tuning->wire_tile) {
            WCT_REQUIRE(tuning->wire_tile == 2 || tuning->wire_tile == 4 || tuning->wire_tile == 8 ||
                            tuning->wire_tile == 16,
                        WCT_ERR_ARG, "decon_plan: wire tile %d not in {2, 4, 8, 16}", tuning->wire_tile);
            tk = tuning->wire_tile;
        }
        if (tuning->row_grid_pct) {
            WCT_REQUIRE(tuning->row_grid_pct >= 10 && tuning->row_grid_pct <= 100, WCT_ERR_ARG,
                        "decon_plan: row grid percent %d not in [10, 100]", tuning->row_grid_pct);
            row_pct = tuning->row_grid_pct;
        }
        WCT_REQUIRE(tuning->row_plan >= -1 && tuning->row_plan <= 1, WCT_ERR_ARG,
                    "decon_plan: row plan %d not in {-1, 0, 1}", tuning->row_plan);
        WCT_REQUIRE(tuning->r2c_blocks == 0 || (tuning->row_plan >= 0 && tuning->r2c_blocks >= 3 &&
                                                tuning->r2c_blocks <= 5),
                    WCT_ERR_ARG, "decon_plan: r2c blocks %d not in {3, 4, 5} with the compile-time row plan",
                    tuning->r2c_blocks);
        WCT_REQUIRE(tuning->c2r_blocks == 0 || (tuning->row_plan >= 0 && tuning->c2r_blocks >= 2 &&
                                                tuning->c2r_blocks <= 4),
                    WCT_ERR_ARG, "decon_plan: c2r blocks %d not in {2, 3, 4} with the compile-time row plan",
                    tuning->c2r_blocks);
        row_plan = tuning->row_plan < 0 ? 0 : 1;
        r2c_blocks = tuning->r2c_blocks;
        c2r_blocks = tuning->c2r_blocks;
    }
    for (int i = 0; i < njobs; ++i) {
        const wct_decon_job& j = jobs[i];
        const wct_decon_kernel K = j.kernel;
        WCT_REQUIRE(K->nw <= wire_tile_max_nw(tk), WCT_ERR_UNSUPPORTED,
                    "decon_plan: %d wires exceed the %d-bin wire tile's limit of %d", K->nw, tk,
                    wire_tile_max_nw(tk));
        const bool spec = j.spectrum != nullptr;
        WCT_REQUIRE(spec == (jobs[0].spectrum != nullptr), WCT_ERR_ARG,
                    "decon_plan: jobs 0 and %d differ in spectrum mode (a plan uses it for all jobs or none)", i);
        if (spec) {
            WCT_REQUIRE(j.adc && j.ld_adc >= K->nt && j.ld_adc % 2 == 0 && j.ld_spectrum >= K->nt &&
                            (reinterpret_cast<uintptr_t>(j.adc) & 7) == 0 &&
                            (reinterpret_cast<uintptr_t>(j.spectrum) & 7) == 0,
                        WCT_ERR_ARG, "decon_plan: spectrum job %d has bad pointers or strides", i);
            WCT_REQUIRE(!j.out && !j.mask && !j.rms && !j.bad_ptr && !j.lf_tight && !j.lf_tighter &&
                            !j.restore_baseline && !j.save_mode,
                        WCT_ERR_ARG, "decon_plan: spectrum job %d sets outputs of the decon mode", i);
        } else {
            WCT_REQUIRE(j.adc && j.out && j.ld_adc >= K->nt && j.ld_out >= K->nt && j.ld_adc % 2 == 0 &&
                            j.ld_out % 2 == 0 && (!j.mask || (j.ld_mask >= K->nt && j.ld_mask % 2 == 0)),
                        WCT_ERR_ARG, "decon_plan: job %d has bad pointers or strides", i);
            WCT_REQUIRE((reinterpret_cast<uintptr_t>(j.adc) & 7) == 0 &&
                            (reinterpret_cast<uintptr_t>(j.out) & 7) == 0 &&
                            (!j.mask || (reinterpret_cast<uintptr_t>(j.mask) & 1) == 0),
                        WCT_ERR_ARG, "decon_plan: job %d buffers are misaligned", i);
        }
        WCT_REQUIRE(!j.adjoint || (!j.bad_ptr && !j.restore_baseline && !j.mask && !j.save_mode &&
                                   j.noutputs == 0),
                    WCT_ERR_ARG, "decon_plan: adjoint job %d must be linear (no bad ranges, baseline, "
                    "mask, save_mode or extra outputs)", i);
        WCT_REQUIRE(j.noutputs >= 0 && (j.noutputs == 0 || j.outputs) && j.noutputs <= WCT_DECON_MAX_OUTPUTS,
                    WCT_ERR_ARG, "decon_plan: job %d asks for %d extra outputs (at most %d)", i, j.noutputs,
                    WCT_DECON_MAX_OUTPUTS);
        WCT_REQUIRE(!spec || j.noutputs == 0, WCT_ERR_ARG,
                    "decon_plan: spectrum job %d sets outputs of the decon mode", i);
        WCT_REQUIRE(!j.chan_factors == !j.chan_factor_row, WCT_ERR_ARG,
                    "decon_plan: job %d sets only one of chan_factors and chan_factor_row", i);
        WCT_REQUIRE(!j.chan_factors || (!j.adjoint && j.ld_chan_factors >= K->kc &&
                                        (reinterpret_cast<uintptr_t>(j.chan_factors) & 7) == 0),
                    WCT_ERR_ARG, "decon_plan: job %d has a bad per-channel correction (adjoint, row stride %d "
                    "below %d kept bins, or a misaligned table)", i, j.ld_chan_factors, K->kc);
        for (int q = 0; q < j.noutputs; ++q) {
            const wct_decon_output& o = j.outputs[q];
            WCT_REQUIRE(o.filter && (!o.alt_rows || o.filter_alt) &&
                            (!o.out || (o.ld_out >= K->nt && o.ld_out % 2 == 0 &&
                                        (reinterpret_cast<uintptr_t>(o.out) & 7) == 0)),
                        WCT_ERR_ARG, "decon_plan: job %d output %d has a bad filter or frame", i, q);
        }
        DJob& d = h[i];
        memset(&d, 0, sizeof(d));
        d.adc = j.adc;
        d.out = j.out;
        d.mask = j.mask;
        d.rms = j.rms;
        d.bad_ptr = j.bad_ptr;
        d.bad_lo = j.bad_lo;
        d.bad_hi = j.bad_hi;
        d.lf1 = j.lf_tight;
        d.lf2 = j.lf_tighter;
        for (int q = 0; q < j.noutputs; ++q) {
            const wct_decon_output& o = j.outputs[q];
            DVar& v = d.var[d.nvar++];
            v.g = o.filter;
            v.g_alt = o.filter_alt;
            v.alt_rows = o.alt_rows;
            v.out = o.out;
            v.ld_out = o.ld_out;
            v.rms = o.rms;
            v.restore = o.restore_baseline;
            v.save_mode = o.save_mode;
        }
        d.cf = reinterpret_cast<const float2*>(j.chan_factors);
        d.cf_row = j.chan_factor_row;
        d.ld_cf = j.ld_chan_factors;
        d.K = K->K;
        d.ld_adc = j.ld_adc;
        d.ld_out = j.ld_out;
        d.ld_mask = j.ld_mask;
        d.nw = K->nw;
        d.nt = K->nt;
        d.M = K->nt / 2;
        d.kc = K->kc;
        d.ldk = job_ldk(j);
        d.th_factor = j.th_factor;
        d.restore = j.restore_baseline;
        d.save_mode = j.save_mode;
        d.adjoint = j.adjoint ? 1 : 0;
        d.spec = reinterpret_cast<float2*>(j.spectrum);
        d.ld_spec = j.ld_spectrum;
        d.S = reinterpret_cast<float2*>(static_cast<char*>(workspace) + off);
        off += (size_t)d.nw * d.ldk * sizeof(float2);
        off = (off + 255) & ~(size_t)255;
        rows += d.nw;
        tiles += d.ldk / tk;
        const DevPlan* pM = nullptr;
        const DevPlan* pW = nullptr;
        int rc = get_plan(d.M, false, &pM);
        if (rc) return rc;
        rc = get_plan(d.nw, false, &pW);
        if (rc) return rc;
        d.planM = pM->plan;
        d.twM = (const float2*)pM->tw;
        if (d.M == kStaticRowLen) {   // the compile-time row plan's schedule
            const DevPlan* pL = nullptr;
            rc = get_plan(d.M, false, &pL, true);
            if (rc) return rc;
            d.twL = (const float2*)pL->tw;
        }
        d.planW = pW->plan;
        d.twW = (const float2*)pW->tw;
        const void* line = nullptr;
        rc = get_twiddle_line(d.nt, d.M + 1, false, &line);
        if (rc) return rc;
        d.wN = (const float2*)line;
        Mmax = std::max(Mmax, d.M);
        nwmax = std::max(nwmax, d.nw);
        twM = std::max(twM, d.planM.twcount);
        twW = std::max(twW, d.planW.twcount);
    }
    WCT_REQUIRE(Mmax <= 8192, WCT_ERR_UNSUPPORTED, "decon_plan: %d ticks above 16384", 2 * Mmax);
    auto* p = new wct_decon_plan_s();
    p->smooth5 = true;
    p->spectrum = jobs[0].spectrum != nullptr;
    for (int i = 0; i < njobs; ++i)
        if (!is_smooth5(h[i].M) || !is_smooth5(h[i].nw)) p->smooth5 = false;
    p->twM = twM;
    p->twW = twW;
    {
        int dev = 0;
        WCT_HIP_CHECK(hipGetDevice(&dev));
        WCT_HIP_CHECK(hipDeviceGetAttribute(&p->ncu, hipDeviceAttributeMultiprocessorCount, dev));
    }
    p->njobs = njobs;
    p->rows = rows;
    p->tiles = tiles;
    p->Mmax = Mmax;
    p->nwmax = nwmax;
    p->tk = tk;
    p->row_pct = row_pct;
    p->row_plan = row_plan;
    p->r2c_blocks = r2c_blocks;
    p->c2r_blocks = c2r_blocks;
    // two orderings of the same jobs: by row length for the row passes,
    // by wire count for the wire pass, each with its own cumulative row
    // and tile starts and its groups of equal length
    std::vector<int> ord(njobs);
    for (int i = 0; i < njobs; ++i) ord[i] = i;
    std::vector<DJob> hr(njobs), ht(njobs);
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return h[a].M < h[b].M; });
    long acc = 0;
    for (int i = 0; i < njobs; ++i) {
        hr[i] = h[ord[i]];
        hr[i].row0 = acc;
        if (p->rgroups.empty() || p->rgroups.back().len != hr[i].M)
            p->rgroups.push_back({hr[i].M, acc, acc, hr[i].planM.twcount, true, false, false});
        acc += hr[i].nw;
        p->rgroups.back().end = acc;
        if (!is_smooth5(hr[i].M)) p->rgroups.back().smooth5 = false;
        if (hr[i].bad_ptr) p->rgroups.back().bad = true;
        if (hr[i].nvar) p->rgroups.back().var = true;
    }
    for (int i = 0; i < njobs; ++i) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return h[a].nw < h[b].nw; });
    acc = 0;
    for (int i = 0; i < njobs; ++i) {
        ht[i] = h[ord[i]];
        ht[i].tile0 = acc;
        if (p->tgroups.empty() || p->tgroups.back().len != ht[i].nw)
            p->tgroups.push_back({ht[i].nw, acc, acc, ht[i].planW.twcount, true, false, false});
        acc += ht[i].ldk / tk;
        p->tgroups.back().end = acc;
        if (!is_smooth5(ht[i].nw)) p->tgroups.back().smooth5 = false;
    }
    int ra = dev_alloc((void**)&p->jobs, njobs * sizeof(DJob));
    if (!ra) ra = dev_alloc((void**)&p->jobs_t, njobs * sizeof(DJob));
    if (ra) {
        if (p->jobs) (void)hipFree(p->jobs);
        delete p;
        return ra;
    }
    hipError_t e = hipMemcpy(p->jobs, hr.data(), njobs * sizeof(DJob), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->jobs_t, ht.data(), njobs * sizeof(DJob), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(p->jobs);
        (void)hipFree(p->jobs_t);
        delete p;
        set_error("decon_plan: job table upload failed: %s", hipGetErrorString(e));
        return WCT_ERR_HIP;
    }
    *out = p;
    return WCT_OK;
}

int wct_decon_plan_run(wct_decon_plan p, void* stream)
{
    WCT_ENTRY();
    WCT_REQUIRE(p && p->jobs, WCT_ERR_ARG, "decon_plan_run: null plan");
    return plan_run(p, as_stream(stream));
}

int wct_decon_plan_run_stage(wct_decon_plan p, int stage, void* stream)
{
    WCT_ENTRY();
    WCT_REQUIRE(p && p->jobs, WCT_ERR_ARG, "decon_plan_run_stage: null plan");
    WCT_REQUIRE(stage >= 0 && stage <= 2, WCT_ERR_ARG, "decon_plan_run_stage: stage %d not in [0, 2]", stage);
    hipStream_t s = as_stream(stream);
    if (stage == 1) return run_wire(p, s);
    return run_rows(p, s, stage == 2);
}

int wct_decon_plan_destroy(wct_decon_plan p)
{
    if (!p) return WCT_OK;
    if (p->jobs) WCT_HIP_CHECK(hipFree(p->jobs));
    if (p->jobs_t) WCT_HIP_CHECK(hipFree(p->jobs_t));
    delete p;
    return WCT_OK;
}

int wct_decon_graph_create(wct_decon_graph* out, const wct_decon_plan* plans, int nplans, void* stream)
{
    WCT_ENTRY();
    WCT_REQUIRE(out && plans && nplans > 0, WCT_ERR_ARG, "decon_graph: no plans");
    *out = nullptr;
    for (int i = 0; i < nplans; ++i) WCT_REQUIRE(plans[i] && plans[i]->jobs, WCT_ERR_ARG, "decon_graph: null plan");
    hipStream_t s = as_stream(stream);
    hipStream_t own = nullptr;
    if (!s) {   // the legacy default stream cannot be captured
        WCT_HIP_CHECK(hipStreamCreateWithFlags(&own, hipStreamNonBlocking));
        s = own;
    }
    auto* g = new wct_decon_graph_s();
    int rc = WCT_OK;
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        for (int i = 0; i < nplans && rc == WCT_OK; ++i) rc = wct_decon_plan_run(plans[i], s);
        hipGraph_t graph = nullptr;
        const hipError_t e2 = hipStreamEndCapture(s, &graph);
        g->graph = graph;
        if (rc == WCT_OK && e2 != hipSuccess) {
            set_error("decon_graph: capture failed: %s", hipGetErrorString(e2));
            rc = WCT_ERR_HIP;
        }
        if (rc == WCT_OK) {
            e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                set_error("decon_graph: instantiate failed: %s", hipGetErrorString(e));
                rc = WCT_ERR_HIP;
            }
        }
    } else {
        set_error("decon_graph: begin capture failed: %s", hipGetErrorString(e));
        rc = WCT_ERR_HIP;
    }
    if (own) (void)hipStreamDestroy(own);
    if (rc != WCT_OK) {
        if (g->exec) (void)hipGraphExecDestroy(g->exec);
        if (g->graph) (void)hipGraphDestroy(g->graph);
        delete g;
        return rc;
    }
    *out = g;
    return WCT_OK;
}

int wct_decon_graph_launch(wct_decon_graph g, void* stream)
{
    WCT_ENTRY();
    WCT_REQUIRE(g && g->exec, WCT_ERR_ARG, "decon_graph_launch: null graph");
    WCT_HIP_CHECK(hipGraphLaunch(g->exec, as_stream(stream)));
    return WCT_OK;
}

int wct_decon_graph_destroy(wct_decon_graph g)
{
    if (!g) return WCT_OK;
    if (g->exec) WCT_HIP_CHECK(hipGraphExecDestroy(g->exec));
    if (g->graph) WCT_HIP_CHECK(hipGraphDestroy(g->graph));
    delete g;
    return WCT_OK;
}

int wct_decon_run(const wct_decon_job* jobs, int njobs, void* workspace, size_t ws_bytes,
                  void* stream)
{
    WCT_ENTRY();
    wct_decon_plan p = nullptr;
    int rc = wct_decon_plan_create(&p, jobs, njobs, workspace, ws_bytes);
    if (rc) return rc;
    rc = wct_decon_plan_run(p, stream);
    // the job table must outlive the launches: wait before freeing it
    hipError_t e = hipStreamSynchronize(as_stream(stream));
    int rc2 = wct_decon_plan_destroy(p);
    if (rc) return rc;
    if (e != hipSuccess) {
        set_error("decon_run: %s", hipGetErrorString(e));
        return WCT_ERR_HIP;
    }
    return rc2;
}

int wct_spectrum_filter(const float* spectrum, int ld_spectrum, int nw, int nt, const float* filter,
                        int restore_baseline, float* out, int ld_out, void* stream)
{
    WCT_ENTRY();
    WCT_REQUIRE(spectrum && filter && out && nw > 0 && nt > 0 && nt % 2 == 0, WCT_ERR_ARG,
                "spectrum_filter: bad arguments (nw %d, nt %d)", nw, nt);
    const int M = nt / 2;
    WCT_REQUIRE(M <= 8192, WCT_ERR_UNSUPPORTED, "spectrum_filter: %d ticks above 16384", nt);
    hipStream_t s = as_stream(stream);
    const DevPlan* pM = nullptr;
    int rc = get_plan(M, false, &pM);
    if (rc) return rc;
    const void* line = nullptr;
    rc = get_twiddle_line(nt, M + 1, false, &line);
    if (rc) return rc;
    DevTemp<unsigned char> scratch(s);
    WCT_TRY(scratch.alloc(spectrum_filter_scratch_bytes(M)));
    return spectrum_filter_run(spectrum, ld_spectrum, nw, nt, filter, restore_baseline, out, ld_out, pM,
                               (const float2*)line, scratch, s);
}

// ---- the same without allocation ----------------------------------------
// A fixed device scratch for the job table and the scaled filter of
// wct_spectrum_filter_noalloc.  Host code never passes the array's name
// to a kernel (that would be a host-side stand-in, not a device
// address): its address on each device is looked up once with
// hipGetSymbolAddress and used from then on.
static __device__ __align__(256) unsigned char scratch[WCT_NOALLOC_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mutex;
constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};

// The scratch's device address on the current device (looked up on the
// first use per device).
int scratch_address(void** out)
{
    int dev = 0;
    WCT_HIP_CHECK(hipGetDevice(&dev));
    WCT_REQUIRE(dev >= 0 && dev < kMaxDevices, WCT_ERR_UNSUPPORTED, "device %d beyond %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        WCT_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        WCT_REQUIRE(p, WCT_ERR_HIP, "no device address for the static scratch");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return WCT_OK;
}
}  // namespace

int wct_noalloc_prepare(int nt)
{
    WCT_ENTRY();
    WCT_REQUIRE(nt > 0 && nt % 2 == 0 && nt / 2 <= 8192, WCT_ERR_ARG, "noalloc_prepare: bad length %d", nt);
    const int M = nt / 2;
    WCT_REQUIRE(spectrum_filter_scratch_bytes(M) <= WCT_NOALLOC_SCRATCH_BYTES, WCT_ERR_UNSUPPORTED,
                "noalloc_prepare: %d ticks need %zu bytes of scratch, %d available", nt,
                spectrum_filter_scratch_bytes(M), WCT_NOALLOC_SCRATCH_BYTES);
    void* p = nullptr;
    int rc = scratch_address(&p);
    if (rc) return rc;
    const DevPlan* pM = nullptr;
    rc = get_plan(M, false, &pM);
    if (rc) return rc;
    const void* line = nullptr;
    return get_twiddle_line(nt, M + 1, false, &line);
}

int wct_spectrum_filter_noalloc(const float* spectrum, int ld_spectrum, int nw, int nt, const float* filter,
                                int restore_baseline, float* out, int ld_out, void* stream)
{
    WCT_ENTRY();
    WCT_REQUIRE(spectrum && filter && out && nw > 0 && nt > 0 && nt % 2 == 0, WCT_ERR_ARG,
                "spectrum_filter_noalloc: bad arguments (nw %d, nt %d)", nw, nt);
    const int M = nt / 2;
    WCT_REQUIRE(M <= 8192 && spectrum_filter_scratch_bytes(M) <= WCT_NOALLOC_SCRATCH_BYTES, WCT_ERR_UNSUPPORTED,
                "spectrum_filter_noalloc: %d ticks need %zu bytes of scratch, %d available", nt,
                spectrum_filter_scratch_bytes(M), WCT_NOALLOC_SCRATCH_BYTES);
    const DevPlan* pM = nullptr;
    int rc = find_plan(M, false, &pM);
    if (rc) return rc;
    const void* line = nullptr;
    rc = find_twiddle_line(nt, M + 1, false, &line);
    if (rc) return rc;
    void* p = nullptr;
    rc = scratch_address(&p);
    if (rc) return rc;
    return spectrum_filter_run(spectrum, ld_spectrum, nw, nt, filter, restore_baseline, out, ld_out, pM,
                               (const float2*)line, p, as_stream(stream));
}

}  // extern "C"
