/* C ABI of the MI355X (gfx950) implementation of Wire-Cell's 2D
 * signal-processing deconvolution (the hot path of OmnibusSigProc).
 *
 * Everything here is plain C: device pointers, sizes, doubles, an int
 * status and a thread-local error string.  No C++ or runtime types
 * cross this boundary; "stream" is an opaque hipStream_t (NULL = the
 * default stream).  Status: 0 on success, a negative WCT_ERR_* code
 * otherwise, with wct_last_error() describing it.
 *
 * Reference interfaces each group replaces are cited per declaration
 * (paths relative to the Wire-Cell Toolkit source tree).
 */
#ifndef WCT_HIP_H
#define WCT_HIP_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WCT_OK 0
#define WCT_ERR_ARG (-1)
#define WCT_ERR_HIP (-2)
#define WCT_ERR_UNSUPPORTED (-3)
#define WCT_ERR_NOMEM (-4)
#define WCT_ERR_STATE (-5)

/* Message of the last failing call on this thread ("" if none).
 * Replaces the reference's exceptions (util/inc/WireCellUtil/Exceptions.h:43),
 * which the C++ host classes re-raise from it. */
const char* wct_last_error(void);
int wct_abi_version(void);

/* ---- device and memory plumbing (no reference counterpart: the
 *      reference is host-only; these let a C/C++ host drive the
 *      device without runtime headers) ---- */
int wct_device_count(int* n);
int wct_set_device(int device);
int wct_malloc(void** dptr, size_t bytes);
int wct_free(void* dptr);
int wct_memcpy_htod(void* dst, const void* src, size_t bytes, void* stream);
int wct_memcpy_dtoh(void* dst, const void* src, size_t bytes, void* stream);
int wct_memcpy_dtod(void* dst, const void* src, size_t bytes, void* stream);
int wct_memset(void* dptr, int value, size_t bytes, void* stream);
/* Page-locked host memory for the staging buffers of the C++ host
 * components (hipHostMalloc; freed with wct_host_free). */
int wct_host_malloc(void** hptr, size_t bytes);
int wct_host_free(void* hptr);
/* Asynchronous copies on a stream (the wct_memcpy_htod/dtoh pair above
 * waits for the copy). */
int wct_memcpy_htod_async(void* dst, const void* src, size_t bytes, void* stream);
int wct_memcpy_dtoh_async(void* dst, const void* src, size_t bytes, void* stream);
int wct_stream_create(void** stream);
int wct_stream_destroy(void* stream);
int wct_stream_sync(void* stream);

/* ---- IDFT drop-in: iface/inc/WireCellIface/IDFT.h:80-127 ----
 * Device pointers to interleaved single-precision complex arrays
 * (layout of std::complex<float>), row-major.  in == out is allowed.
 * Forward transforms are unnormalised, inverse ones scale by 1/n of the
 * transformed extent (IDFT.h:40-66, aux/src/FftwDFT.cxx:113-116,177-182,
 * 212-216). */
/* IDFT::fwd1d / inv1d (IDFT.h:93-97; FftwDFT.cxx:87-117) */
int wct_dft_fwd1d(const float* in, float* out, int size, void* stream);
int wct_dft_inv1d(const float* in, float* out, int size, void* stream);
/* IDFT::fwd1b / inv1b (IDFT.h:101-107; FftwDFT.cxx:149-183):
 * axis = 1 transforms each row, axis = 0 each column. */
int wct_dft_fwd1b(const float* in, float* out, int nrows, int ncols, int axis, void* stream);
int wct_dft_inv1b(const float* in, float* out, int nrows, int ncols, int axis, void* stream);
/* IDFT::fwd2d / inv2d (IDFT.h:111-116; FftwDFT.cxx:186-217).  FftwDFT
 * plans fftwf_plan_dft_2d(ncols, nrows, ...): the buffer is transformed
 * as a row-major (ncols x nrows) array; these follow that. */
int wct_dft_fwd2d(const float* in, float* out, int nrows, int ncols, void* stream);
int wct_dft_inv2d(const float* in, float* out, int nrows, int ncols, void* stream);
/* IDFT::transpose (IDFT.h:120-125; iface/src/IDFT.cxx:53-95):
 * out (ncols x nrows) = in (nrows x ncols)^T, in == out allowed. */
int wct_dft_transpose_real(const float* in, float* out, int nrows, int ncols, void* stream);
int wct_dft_transpose_complex(const float* in, float* out, int nrows, int ncols, void* stream);
/* Double-precision batched transforms (same conventions as 1b), used
 * to build response spectra in FP64. */
int wct_dft_fwd1b_f64(const double* in, double* out, int nrows, int ncols, int axis, void* stream);
int wct_dft_inv1b_f64(const double* in, double* out, int nrows, int ncols, int axis, void* stream);
/* ---- failure-path hooks (tests) ----
 * The reference has none (its errors are exceptions); these force the
 * error paths of a caller from outside.
 * wct_debug_inject: the entry point named entry (e.g. "wct_decon_plan_run")
 * fails with status (< 0) on its (after + 1)-th call from now, once, with
 * the error text "<entry>: injected failure (status N)"; a NULL or empty
 * entry, or status 0, disarms it.  Entry points doing device work check
 * it first, before their arguments.
 * wct_debug_alloc_cap: every later device allocation the library makes
 * above bytes (0: no cap) fails with WCT_ERR_NOMEM, as when the device is
 * out of memory.  Both are process-wide. */
int wct_debug_inject(const char* entry, int after, int status);
int wct_debug_alloc_cap(size_t bytes);

/* Whether the IDFT functions above take length n (1 <= n <= 2^20; any
 * length: those with a prime factor above 13 run Bluestein's chirp-z
 * algorithm on 5-smooth transforms, FftwDFT's any-length plans,
 * aux/src/FftwDFT.cxx:120-146), and whether n is native to the LDS
 * kernels and the fused deconvolution (all prime factors in
 * {2,3,5,7,11,13}). */
int wct_dft_supported(int n);
int wct_dft_native_length(int n);

/* ---- response spectra, FP64 on the device ----
 * OmnibusSigProc::init_overall_response
 * (sigproc/src/OmnibusSigProc.cxx:707-865) after
 * Response::wire_region_average: each region-averaged field-response row
 * (nin samples at fine_period, zero-padded to nfine) is convolved with
 * the electronics samples elec[nfine] scaled by -inter_gain*adc_mv and
 * by fine_period, shifted by fine_time_offset and redigitised onto nt
 * coarse ticks (including the reference's interpolation weights).
 * Host arrays in and out. */
int wct_overall_response(const double* fr_avg, int nr, int nin, const double* elec, int nfine,
                         double fine_period, double coarse_period, double inter_gain,
                         double adc_mv, double fine_time_offset, int nt, double* resp_out,
                         void* stream);

/* Per-plane deconvolution kernel, the fused form of decon_2D_init and a
 * decon_2D_*ROI time filter (OmnibusSigProc.cxx:944-1117):
 *   K(m,k) = Fw[m] Ft[k] exp(-2 pi i (m ws/nw + k ts/nt)) / (R(m,k) nw nt)
 * with R = FFT_w(FFT_t(resp rows padded to nw)) built in FP64 on the
 * device, K = 0 where R == 0 (the NaN/Inf fix-up of :1010-1017), ts = 0
 * unless time_shift > 0 (:1040-1048), for k <= nt/2 (the c2r mirror of
 * aux/src/DftTools.cxx:133-139 only reads those bins).  Bins k >= kc are
 * dropped when max_m |K(m,k)| <= cut_tol * max |K| (cut_tol = 0 keeps
 * all).  The kernel is stored in FP32 on the device.
 * resp, wire_filter (nw) and time_filter (nt) are host arrays. */
typedef struct wct_decon_kernel_s* wct_decon_kernel;
int wct_decon_kernel_create(wct_decon_kernel* out, const double* resp, int nr, int nw, int nt,
                            const double* wire_filter, const double* time_filter,
                            int wire_shift, int time_shift, double cut_tol, void* stream);
/* The forward response convolution with the same layout, for synthetic
 * frames and the adjoint's forward model: the job output is
 *   D = IFFT2( FFT2(q) R exp(+2 pi i (m ws/nw + k ts/nt)) )
 * i.e. the frame whose decon with unit filters returns q (the field x
 * electronics convolution of gen/src/ImpactTransform.cxx:100-308 on the
 * region-averaged response). */
int wct_forward_kernel_create(wct_decon_kernel* out, const double* resp, int nr, int nw, int nt,
                              int wire_shift, int time_shift, void* stream);
int wct_decon_kernel_info(wct_decon_kernel k, int* nw, int* nt, int* kc);
/* Copies K (kc x nw complex, k-major) to a host buffer of 2*kc*nw floats. */
int wct_decon_kernel_download(wct_decon_kernel k, float* host_out);
int wct_decon_kernel_destroy(wct_decon_kernel k);

/* The kernel of a job whose outputs each apply their whole time filter
 * (the job's lf_tight for the tight output, wct_decon_output.filter for
 * the others), so one spectrum serves filters whose ratios are not
 * bounded (Wiener_tight, Wiener_wide, Gaus_wide; OmnibusSigProc.cxx:
 * 1053-1380):
 *   K(m,k) = Fw[m] exp(-2 pi i (m ws/nw + k ts/nt)) / (R(m,k) nw nt)
 * built in FP64 as above.  envelope (nt host doubles, or NULL) bounds
 * the outputs' filters: bins k >= kc with
 * max_m |K(m,k)| envelope[k] <= cut_tol * max are dropped.  r_floor
 * regularises the inverse: K = 0 where |R| <= r_floor * max|R|
 * (r_floor = 0: only where R == 0, the reference's NaN/Inf fix-up).  A
 * response that falls below FP64 rounding at high frequency (the
 * synthetic Gaussian responses) needs r_floor > 0 for a filter that
 * does not fall faster than the response (Gaus_wide). */
int wct_decon_kernel_create_unfiltered(wct_decon_kernel* out, const double* resp, int nr, int nw, int nt,
                                       const double* wire_filter, const double* envelope, int wire_shift,
                                       int time_shift, double cut_tol, double r_floor, void* stream);

/* An extra output of a decon job: the same deconvolved spectrum through
 * another per-bin time filter, as OmnibusSigProc's other decon variants
 * of a plane (decon_2D_tighterROI :1120-1160, decon_2D_looseROI
 * :1162-1238, decon_2D_ROI_refine :1053-1075, decon_2D_hits :1316-1347,
 * decon_2D_charge :1349-1380), then optionally restore_baseline, the
 * per-wire cal_RMS and save_data's zeroing.
 *   filter       device, kc floats: per-bin time factor (times the
 *                kernel's own)
 *   filter_alt   device, kc floats, or NULL: the factor of the rows
 *                flagged in alt_rows (decon_2D_looseROI's tight-LF filter
 *                for channels next to bad or lf_noisy ones, :1216-1231)
 *   alt_rows     device, nw bytes, or NULL
 *   out          device, nw x ld_out floats, or NULL
 *   rms          device, nw floats, or NULL: cal_RMS of this output with
 *                the job's bad ranges excluded (ROI_formation.cxx:364-389)
 *   restore_baseline, save_mode: as for the job's own output
 * (The tighter ROI of the mask is the job's lf_tighter, not an output.) */
typedef struct {
    const float* filter;
    const float* filter_alt;
    const unsigned char* alt_rows;
    float* out;
    int ld_out;
    float* rms;
    int restore_baseline;
    int save_mode;
} wct_decon_output;
/* Extra outputs of one job. */
#define WCT_DECON_MAX_OUTPUTS 6

/* One plane of one frame.  Replaces, per plane, load_data + decon_2D_init
 * + decon_2D_tightROI (+ decon_2D_tighterROI) + restore_baseline +
 * ROI_formation::find_ROI_by_decon_itself's RMS and threshold
 * (OmnibusSigProc.cxx:304-352,944-1160; ROI_formation.cxx:364-531).
 *   adc   : device, nw rows of nt floats, row stride ld_adc
 *   out   : device, nw x ld_out floats: the tight decon after
 *           restore_baseline (save_mode 1 also zeroes negatives and
 *           [lo,hi) bad ranges as save_data does, :383-418)
 *   mask  : device, nw x ld_mask bytes or NULL; bit 0 = tight > th,
 *           bit 1 = tighter > th (th = th_factor * rms + 1)
 *   rms   : device, nw floats or NULL (ROI_formation's per-wire rms)
 *   bad_* : device CSR of per-wire bad tick ranges or NULL:
 *           bad_ptr[nw+1], ranges [bad_lo, bad_hi) are zeroed on input
 *           (load_data) and [bad_lo, bad_hi] are excluded from the RMS and
 *           the mask (find_ROI_by_decon_itself)
 *   lf_tight / lf_tighter : device, kc floats or NULL: extra per-bin
 *           time factors for the tight and the tighter output. */
typedef struct {
    wct_decon_kernel kernel;
    const float* adc;
    int ld_adc;
    float* out;
    int ld_out;
    unsigned char* mask;
    int ld_mask;
    float* rms;
    const int* bad_ptr;
    const int* bad_lo;
    const int* bad_hi;
    const float* lf_tight;
    const float* lf_tighter;
    float th_factor;
    int restore_baseline;
    int save_mode;
    int adjoint;        /* 1: apply the adjoint operator (conj(K)); the
                           job must then have no bad ranges, baseline,
                           mask or save_mode (linear part only) */
    /* Spectrum mode, the tensor-set Decon2DResponse
     * (sig/src/Decon2DResponse.cxx:203-390): when set, the third stage
     * writes, instead of out/mask/rms, the full time spectrum of the
     * deconvolved and rolled frame, FFT_t(X) with X the reference's
     * r_data before its final fwd_r2c: nw rows of nt complex64 values
     * (interleaved floats), row stride ld_spectrum complex elements.  The
     * kernel then carries no time filter (all ones) and no spectral cut
     * (cut_tol 0); out, mask, rms, bad ranges, lf filters, baseline and
     * save_mode must be unset.  A plan's jobs all use the mode or none. */
    float* spectrum;
    int ld_spectrum;
    /* Extra outputs (host array of noutputs entries, copied when the
     * plan is created; at most WCT_DECON_MAX_OUTPUTS). */
    const wct_decon_output* outputs;
    int noutputs;
    /* Channel-by-channel electronics correction (decon_2D_init
     * :949-982; Decon2DResponse.cxx:280-308): each wire's time spectrum
     * is multiplied, bin by bin, by row chan_factor_row[w] of the
     * chan_factors table before the wire transform.
     *   chan_factors    : device table from wct_channel_factors, rows of
     *                     ld_chan_factors complex64 values (>= kc)
     *   chan_factor_row : device, nw ints; -1 leaves a wire uncorrected
     * Both NULL: no correction.  Not with adjoint. */
    const float* chan_factors;
    int ld_chan_factors;
    const int* chan_factor_row;
} wct_decon_job;

/* The correction factors of channel responses (OmnibusSigProc.cxx:
 * 949-982): row r of out holds elec(k) / C_r(k) for k in [0, nt/2], or 0
 * where C_r(k) == 0, with elec and C_r the length-nt forward DFTs of the
 * nominal electronics response and of channel response r (cut or
 * zero-padded to nt, :966).  Built in FP64 on the device, stored as
 * complex64 rows of ld_out values (ld_out >= nt/2 + 1).
 *   responses : host, nresp x nsamp floats (IChannelResponse samples)
 *   elec      : host, nt doubles (the electronics response sampled on
 *               the channel responses' binning over nt ticks) */
int wct_channel_factors(float* out, int ld_out, const float* responses, int nresp, int nsamp,
                        const double* elec, int nt, void* stream);

/* Device scratch needed to run the jobs (spectrum buffers). */
size_t wct_decon_workspace_size(const wct_decon_job* jobs, int njobs);
/* A batch of jobs bound to a workspace: the job table is validated and
 * uploaded once (synchronously); wct_decon_plan_run then only enqueues
 * one launch per stage over every plane of the batch (no allocation, no
 * copy), so it can be replayed and captured in a hipGraph.  The plan
 * keeps the device pointers of the jobs; runs that share a workspace
 * must be ordered on one stream. */
typedef struct wct_decon_plan_s* wct_decon_plan;
int wct_decon_plan_create(wct_decon_plan* out, const wct_decon_job* jobs, int njobs,
                          void* workspace, size_t ws_bytes);
int wct_decon_plan_run(wct_decon_plan plan, void* stream);
/* One stage of wct_decon_plan_run (0: the forward row transforms, 1: the
 * wire pass, 2: the inverse row pass with the statistics and outputs),
 * for timing a stage on its own.  A run is the three stages in order on
 * one stream; other orders leave the outputs undefined. */
int wct_decon_plan_run_stage(wct_decon_plan plan, int stage, void* stream);
int wct_decon_plan_destroy(wct_decon_plan plan);
/* Launch-shape choices of a plan (0 = the default for that field):
 *   wire_tile     bins per wire-pass workgroup: 2, 4, 8 or 16 (16 and 8
 *                 need <= 1023 wires, 2 and 4 take up to 2047; default 4)
 *   row_grid_pct  persistent grid of the row passes as a percentage of
 *                 the most workgroups the device keeps resident
 *   row_plan      0 (default) or 1: rows of 3000 complex samples (6000
 *                 ticks) use transforms compiled for that length (radix
 *                 15, 20, 10 on an unpadded row: every LDS address a base
 *                 plus an immediate); -1: the runtime radix plan for
 *                 every length
 *   r2c_blocks    with the compiled plan, workgroups per CU of the
 *                 forward row pass: 3, 4 (default) or 5
 *   c2r_blocks    with the compiled plan, workgroups per CU of the
 *                 inverse row pass: 2, 3 (default) or 4
 * The autotuner (wct_amd.tune) times these per workload. */
typedef struct {
    int wire_tile;
    int row_grid_pct;
    int row_plan;
    int r2c_blocks;
    int c2r_blocks;
} wct_decon_tuning;
int wct_decon_plan_create_tuned(wct_decon_plan* out, const wct_decon_job* jobs, int njobs,
                                void* workspace, size_t ws_bytes, const wct_decon_tuning* tuning);
/* hipGraph of a sequence of plan runs (captured on stream, or on a
 * private stream when stream is NULL); a launch replays every stage of
 * every plan.  The plans and their buffers must outlive the graph. */
typedef struct wct_decon_graph_s* wct_decon_graph;
int wct_decon_graph_create(wct_decon_graph* out, const wct_decon_plan* plans, int nplans, void* stream);
int wct_decon_graph_launch(wct_decon_graph graph, void* stream);
int wct_decon_graph_destroy(wct_decon_graph graph);
/* One-shot convenience: create a plan, run it, wait, destroy it.
 * jobs is a host array; workspace is device memory of at least
 * wct_decon_workspace_size bytes. */
int wct_decon_run(const wct_decon_job* jobs, int njobs, void* workspace, size_t ws_bytes,
                  void* stream);

/* Decon2DFilter's device pass (sig/src/Decon2DFilter.cxx:90-125):
 * per row of a complex64 spectrum (nw rows of nt complex values,
 * interleaved floats, row stride ld_spectrum complex elements), multiply
 * bin k by filter[k], inverse real transform as DftTools::inv_c2r does
 * (Hermitian mirror of bins 0..nt/2 with real DC and Nyquist, 1/nt
 * normalisation; aux/src/DftTools.cxx:133-139, Spectrum.h:56-89), then
 * Sig::restore_baseline when restore_baseline is set
 * (sig/src/Util.cxx:5-35).  filter holds nt/2 + 1 floats (bins above
 * nt/2 do not reach the output).  All pointers are device memory; out
 * has nw rows of nt floats, row stride ld_out.  nt even, up to 16384. */
int wct_spectrum_filter(const float* spectrum, int ld_spectrum, int nw, int nt, const float* filter,
                        int restore_baseline, float* out, int ld_out, void* stream);

/* wct_spectrum_filter without any device allocation.  No counterpart in
 * the reference: this is for callers that cannot allocate device memory
 * (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The one-job table and the scaled filter that
 * wct_spectrum_filter puts in stream-ordered scratch go into a fixed
 * static device array of WCT_NOALLOC_SCRATCH_BYTES per device instead
 * (held by each device the library runs on until the process ends);
 * tick counts whose job and filter need more are refused.  The length's
 * transform tables must exist first: wct_noalloc_prepare(nt), a set-up
 * call that may allocate, creates them and resolves the scratch address
 * on the current device; without it the entry point returns
 * WCT_ERR_STATE.  Every call on a device shares that device's scratch:
 * calls that overlap on one GPU (concurrent streams or threads) race. */
#define WCT_NOALLOC_SCRATCH_BYTES 65536
int wct_noalloc_prepare(int nt);
int wct_spectrum_filter_noalloc(const float* spectrum, int ld_spectrum, int nw, int nt, const float* filter,
                                int restore_baseline, float* out, int ld_out, void* stream);

/* ---- the plain per-plane path ----
 * OmnibusSigProc's per-plane control flow (load_data, decon_2D_init,
 * decon_2D_tighterROI/tightROI, restore_baseline, cal_RMS + threshold,
 * save_data; sigproc/src/OmnibusSigProc.cxx:304-472,944-1160,
 * ROI_formation.cxx:364-456) with one device launch per reference step,
 * full c2c transforms (aux/src/DftTools.cxx:127-139), FP32 throughout
 * and the response spectrum rebuilt on every call (:988-998).  All
 * pointers are device memory:
 *   resp           nr x nt overall response rows (nr <= nw)
 *   wire_filter    nw values; tight_filter / tighter_filter nt values
 *                  (tighter NULL for a collection plane)
 *   out, mask, rms as for wct_decon_job (mask needs rms); save_mode as
 *                  there.
 * The workspace holds wct_plain_workspace_size(nw, nt) bytes. */
size_t wct_plain_workspace_size(int nw, int nt);
int wct_plain_decon_plane(const float* adc, int ld_adc, int nw, int nt, const float* resp, int nr,
                          const float* wire_filter, const float* tight_filter, const float* tighter_filter,
                          int wire_shift, int time_shift, float th_factor, const int* bad_ptr,
                          const int* bad_lo, const int* bad_hi, float* out, int ld_out,
                          unsigned char* mask, int ld_mask, float* rms, int save_mode, void* workspace,
                          size_t ws_bytes, void* stream);

/* ---- forward simulation convolution ----
 * Gen::ImpactTransform (gen/src/ImpactTransform.cxx:19-347), the forward
 * model of the deconvolution: the charges BinnedDiffusion_transform::
 * get_charge_vec sorts into impact groups (gen/src/
 * BinnedDiffusion_transform.cxx:227-443) convolved with the plane's impact
 * responses over wires and ticks.
 *
 * wct_fft_best_length: WireCell::fft_best_length
 * (util/src/FFTBestLength.cxx:12-62). */
long wct_fft_best_length(long n, int keep_odd_even);

/* The frame of one transform (ImpactTransform.cxx:84-115): channels
 * [start_ch, end_ch) padded by npad_wire each side to the nw wires
 * [frame_start_ch, frame_end_ch) (m_start_ch, m_end_ch), ticks
 * [frame_start_tick, frame_end_tick) (m_start_tick, m_end_tick), nt of
 * them. */
typedef struct wct_impact_frame {
    int start_ch, end_ch, npad_wire;
    int frame_start_ch, frame_end_ch;
    int frame_start_tick, frame_end_tick;
    int nw, nt;
} wct_impact_frame;
/* From BinnedDiffusion_transform's impact_bin_range and time_bin_range
 * (impact_lo/hi, time_lo/hi), the response's waveform_pad and the
 * response geometry: num_group = round(pitch / impact) + 1 impact groups,
 * nwires_resp the response's wire count (:28-29). */
int wct_impact_geometry(int num_group, int nwires_resp, int impact_lo, int impact_hi, int time_lo, int time_hi,
                        int waveform_pad, wct_impact_frame* out);

/* The response side (ImpactTransform.cxx:34-74): resp holds
 * num_group x nwires_resp rows of nresp floats (device or host memory,
 * copied), row (i, j) the waveform() of the impact response that
 * m_vec_map_resp[i][j - P] holds, P = (nwires_resp - 1) / 2.  The
 * reference re-derives that waveform from spectrum() with inv_c2r
 * (:158-161); both are the same samples.  nwires_resp must be odd (an
 * even count leaves the reference's map without the rows it reads). */
typedef struct wct_impact_s* wct_impact;
int wct_impact_create(wct_impact* out, int num_group, int nwires_resp, const float* resp, int nresp,
                      void* stream);
int wct_impact_destroy(wct_impact h);
/* The constructor's convolution (:117-306) into out (nw x nt floats,
 * device, m_decon_data row-major by wire).  The charges of group g are
 * entries [group_offsets[g], group_offsets[g + 1]) of ch, tick and q
 * (device arrays; group_offsets is a host array of num_group + 1), the
 * tuples of m_vec_vec_charge[g].  Entries outside the frame are skipped
 * (the reference's Eigen access would be out of bounds).  Charges add up
 * with float atomics, so sums over one cell may round differently from
 * run to run.  The handle keeps frame-sized work buffers (grown on
 * demand); calls on one handle must not overlap. */
int wct_impact_run(wct_impact h, const wct_impact_frame* frame, const long* group_offsets, const int* ch,
                   const int* tick, const double* q, float* out, void* stream);
/* ImpactTransform::waveform(iwire) (:312-347) for every wire of the
 * frame: out has nw rows of nsamples floats (row r: channel
 * frame_start_ch + r); decon is wct_impact_run's output.  With nlong > 0
 * each row is convolved with long_aux (device, nlong floats; the
 * response's long_aux_waveform) over fft_best_length(nsamples +
 * long_pad) samples. */
int wct_impact_waveforms(wct_impact h, const wct_impact_frame* frame, const float* decon, int nsamples,
                         const float* long_aux, int nlong, int long_pad, float* out, void* stream);

/* ---- noise filter ----
 * The one-channel filters that OmnibusNoiseFilter::operator() runs per
 * channel (sigproc/src/OmnibusNoiseFilter.cxx:170-188), over every
 * channel of a frame at once: Protodune::OneChannelNoise::apply
 * (sigproc/src/Protodune.cxx:753-918, mode WCT_NF_PROTODUNE) and
 * Microboone::OneChannelNoise::apply (sigproc/src/Microboone.cxx:902-1064,
 * mode WCT_NF_MICROBOONE).  Per channel, as the noise DB and the anode
 * give them:
 *   plane      wire-plane index (2: collection)
 *   rcrc       row of the spectra table dividing the spectrum
 *              (IChannelNoiseDatabase::rcrc; not for partial waveforms),
 *              -1 for none
 *   bins_begin, bins_end   ProtoDUNE, collection plane: the channel's runs
 *              of flagged bins (noise spectrum real part < 0.5) cleaned of
 *              harmonic lines, [begin, end) of the bins table's (first,
 *              last) inclusive pairs
 *   min_rms, max_rms       min_rms_cut / max_rms_cut
 *   mode       WCT_NF_PROTODUNE or WCT_NF_MICROBOONE
 *   baseline, gain         MicroBooNE: nominal_baseline subtracted first,
 *              gain_correction applied for the chirp test
 *   config, noise          MicroBooNE: rows of the spectra table the
 *              spectrum is multiplied by (IChannelNoiseDatabase::config,
 *              ::noise), -1 for none */
typedef struct wct_nf_channel {
    int plane;
    int rcrc;
    int bins_begin, bins_end;
    float min_rms, max_rms;
    int mode;
    float baseline, gain;
    int config, noise;
} wct_nf_channel;
#define WCT_NF_PROTODUNE 0
#define WCT_NF_MICROBOONE 1
/* Flags, three ints per channel: bits, then the chirped samples [begin,
 * end) (-1, -1 without).  Bits: the spectrum looked like a partial
 * waveform (Diagnostics::Partial; masked "lf_noisy" off the collection
 * plane by the caller), the RMS test failed ("noisy"), more than 5
 * harmonic bins were removed ("harmonic"), the waveform has a chirp
 * (Diagnostics::Chirp, MicroBooNE; "chirp"). */
#define WCT_NF_PARTIAL 1
#define WCT_NF_NOISY 2
#define WCT_NF_HARMONIC 4
#define WCT_NF_CHIRP 8
/* The rcrc spectrum of OmniChannelNoiseDB::parse_rcrc
 * (sigproc/src/OmniChannelNoiseDB.cxx:171-208): Response::SimpleRC of
 * width (time units) sampled at tick over nt samples, its spectrum (FP64
 * on the device, rounded to float) raised to the power layers; out is nt
 * complex floats (device). */
int wct_nf_rcrc_spectrum(double width, double tick, int nt, int layers, float* out, void* stream);
/* in: nch rows of nt samples (row stride ld_in, channels whose signal the
 * filter sees, zeros for bad ones); out: the filtered rows (stride
 * ld_out); chans: nch entries (device); spectra: rows of nt complex floats
 * that the channels' rcrc, config and noise index (device, may be NULL
 * when no channel uses one); bins: (first, last) pairs (device, may be
 * NULL when no channel has any); flags: 3 nch ints (device).  The
 * workspace holds wct_nf_workspace_size(nch, nt) bytes.  Up to 12000
 * ticks (one channel's samples and scratch in LDS). */
size_t wct_nf_workspace_size(int nch, int nt);
int wct_nf_one_channel(const float* in, int ld_in, float* out, int ld_out, int nch, int nt,
                       const wct_nf_channel* chans, const float* spectra, const int* bins, int* flags,
                       void* workspace, size_t ws_bytes, void* stream);
/* Microboone::CoherentNoiseSub::apply (sigproc/src/Microboone.cxx:798-872)
 * on every channel group of a frame at once: per group the binned median
 * over its channels (Derivations::CalcMedian), the signal protection
 * (SignalProtection, :276-462), the scaled subtraction
 * (Subtract_WScaling, :53-274).  A group lists rows [first, first +
 * count) of group_rows (row indices into rows, in channel order, up to
 * 128 per group; groups disjoint); row_group gives each row's group (-1:
 * in none).  Per group, from the DB values of its first channel: respec
 * (row of the spectra table, or -1 when the reference skips the response
 * path: no response, a unit first bin or a zero offset), res_offset,
 * pad_f / pad_b (pad_window_front / back) and the coherent_nf_* limits.
 * rows are filtered in place.  rms_threshold is taken as 0 (the
 * reference's default; its per-channel limit chain is not built). */
typedef struct wct_nf_group {
    int first, count;
    int respec, res_offset, pad_f, pad_b;
    float decon_limit, decon_lf_cutoff, adc_limit, decon_limit1, roi_min_max_ratio, protection_factor,
        min_adc_limit;
} wct_nf_group;
size_t wct_nf_coherent_workspace_size(int ngroups, int nrows, int nt);
int wct_nf_coherent(float* rows, int ld, int nrows, int nt, const wct_nf_group* groups, int ngroups,
                    const int* group_rows, const int* row_group, const float* spectra, int any_response,
                    void* workspace, size_t ws_bytes, void* stream);
/* Protodune::RelGainCalib::apply (Protodune.cxx:975-1009) on nch rows in
 * place: row c scaled by gain[c] (float) when gain[c] lies within
 * [gain_min, gain_max]; gain is a device array. */
int wct_nf_rel_gain(float* rows, int ld, int nch, int nt, const float* gain, float gain_min, float gain_max,
                    void* stream);
/* Microboone::OneChannelStatus::ID_lf_noisy (sigproc/src/Microboone.cxx:
 * 1281-1334) on nch rows of nt samples (device, row stride ld): per row
 * the binned median and 16/84 % points (Waveform::percentile_binned), the
 * samples within [i - window, i + window) of a sample i beyond threshold x
 * rms from the median set to the median, the float magnitudes of bins
 * 1..nbins of their spectrum summed, over the row's count of non-zero
 * samples; lf_noisy[c] (device, nch ints) = 1 when that ratio exceeds cut
 * (0 for a row with no non-zero sample), ratio (device, nch doubles, may
 * be NULL) receives it.  window >= 0, 1 <= nbins < nt (the reference's
 * Threshold, Window, Nbins and Cut).  The caller picks the channels
 * (OneChannelStatus::apply, :1263-1279, tests channels off the collection
 * plane only). */
/* Microboone::ADCBitShift::apply (sigproc/src/Microboone.cxx:1106-1216)
 * on nch rows of nt samples in place (device, row stride ld): the low
 * bits that rarely flip over the first exam_nticks samples give nshift;
 * for 0 < nshift < 11 every sample is rotated back (Bits::shift_right with
 * its own low nshift bits as filling, util/src/Bits.cxx:4-9) and samples
 * far from the mean and a neighbour take the filling nearest the
 * neighbours' extrapolation.  nshift (device, nch ints) receives each
 * row's count (0: unchanged).  1 <= nbits <= 31, 1 <= exam_nticks <= nt
 * (the reference's Number_of_ADC_bits and Exam_number_of_ticks_test). */
int wct_nf_adc_bit_shift(float* rows, int ld, int nch, int nt, int nbits, int exam_nticks, double threshold_sigma,
                         double threshold_fix, int* nshift, void* stream);
size_t wct_nf_status_workspace_size(int nch, int nt);
int wct_nf_lf_noisy(const float* rows, int ld, int nch, int nt, double threshold, int window, int nbins, double cut,
                    int* lf_noisy, double* ratio, void* workspace, size_t ws_bytes, void* stream);

#ifdef __cplusplus
}
#endif
#endif
