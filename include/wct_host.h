/* C ABI of the host-side helpers of libWireCellHip.so (the C++ plugin
 * that sits above the HIP library).  They restate the reference's
 * response and filter functions so a C/Python caller can build the
 * inputs of wct_overall_response / wct_decon_kernel_create without the
 * plugin classes.  Plain C types only; status 0 = success. */
#ifndef WCT_HOST_H
#define WCT_HOST_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Response::coldelec (util/src/Response.cxx:277-308) sampled at bin
 * centres of Binning(n, tmin, tmax) times postgain
 * (gen/src/ColdElecResponse.cxx:44-49). */
int wcth_coldelec_samples(int n, double tmin, double tmax, double gain, double shaping,
                          double postgain, double* out);
/* SigProc::HfFilter::filter_waveform (sigproc/src/HfFilter.cxx:38-52). */
int wcth_hf_filter(int n, double max_freq, double sigma, double power, int flag, double* out);
/* SigProc::LfFilter::filter_waveform (sigproc/src/LfFilter.cxx:31-45). */
int wcth_lf_filter(int n, double max_freq, double tau, double* out);
/* WireCell::fft_best_length (util/src/FFTBestLength.cxx:12-58). */
long wcth_fft_best_length(long n, int keep_odd_even);
/* Response::wire_region_average (util/src/Response.cxx:79-196) for one
 * plane: paths given by pitch position and current (npaths x nsamp).
 * Writes at most max_regions region numbers and averaged rows
 * (max_regions x nsamp); returns the region count or a negative status. */
int wcth_wire_region_average(int npaths, const double* pitchpos, const double* current, int nsamp,
                             double pitch, int max_regions, int* regions_out, double* avg_out);

/* ---- the component layer (C++ classes registered by type name) ---- */
const char* wcth_last_error(void);
/* Configure components from a JSON list of {"type", "name", "data"}
 * objects, in order, each with default_configuration() updated by
 * "data" (apps/src/Main.cxx:330-350). */
int wcth_configure(const char* json);
/* Drop all component instances. */
int wcth_reset(void);
/* Newline-separated registered type names; returns the string length. */
int wcth_known_types(char* buf, size_t cap);

/* Calls an IDFT component ("type:name") through its virtual interface
 * on host arrays (interleaved complex float, row-major).  1d ops use
 * nrows as the size. */
enum {
    WCTH_FWD1D = 0, WCTH_INV1D, WCTH_FWD1B, WCTH_INV1B, WCTH_FWD2D, WCTH_INV2D,
    WCTH_TRANSPOSE_REAL, WCTH_TRANSPOSE_COMPLEX
};
int wcth_dft(const char* tn, int op, const float* in, float* out, int nrows, int ncols, int axis);

/* IFilterWaveform::filter_waveform(n) of component tn into out[n]. */
int wcth_filter_waveform(const char* tn, int n, float* out);
/* IChannelResponse "type:name" (iface/inc/WireCellIface/IChannelResponse.h:
 * 17-27): the response of a channel (its first min(n, cap) samples into
 * out, n its length) and the binning's start and bin size; a channel
 * without a response fails with the component's KeyError text. */
int wcth_channel_response(const char* tn, int channel, float* out, int cap, int* n, double* tmin,
                          double* binsize);

/* Frames: build an input frame, run an IFrameFilter ("type:name") on it
 * (in = NULL is end-of-stream; *out = NULL then), read the result. */
void* wcth_frame_new(int ident, double time, double tick);
void wcth_frame_free(void* frame);
int wcth_frame_add_trace(void* frame, int channel, int tbin, int n, const float* charge);
int wcth_frame_add_mask(void* frame, const char* name, int channel, int first, int second);
/* Tags of an input frame, applied when it is first used
 * (Aux::SimpleFrame::tag_frame / tag_traces, aux/inc/WireCellAux/
 * SimpleFrame.h:42-70): a frame tag, or a trace tag over trace indices
 * with an optional per-trace summary. */
int wcth_frame_tag_frame(void* frame, const char* tag);
int wcth_frame_tag_traces(void* frame, const char* tag, const size_t* idx, int n, const double* summary,
                          int nsummary);
int wcth_frame_filter(const char* tn, void* in, void** out);
int wcth_frame_info(void* frame, int* ident, double* time, double* tick, int* ntraces);
int wcth_frame_trace(void* frame, int i, int* channel, int* tbin, int* n, const float** charge);
/* which = 0: frame tags, 1: trace tags.  i < 0 returns the count. */
int wcth_frame_tags(void* frame, int which, int i, const char** tag);
int wcth_frame_tagged(void* frame, const char* tag, const size_t** idx, int* n);
int wcth_frame_summary(void* frame, const char* tag, const double** val, int* n);
/* Ranges of one channel-mask-map entry; returns the total count (only
 * the first cap are written). */
int wcth_frame_masks(void* frame, const char* name, int* channels, int* firsts, int* seconds, int cap);

/* Frame sinks and sources (IFrameSink / IFrameSource, iface/inc/
 * WireCellIface/IFrameSink.h:11-21, IFrameSource.h:12-20), e.g. the
 * frame files FrameFileSink / FrameFileSource (sio/src/FrameFileSink.cxx,
 * FrameFileSource.cxx).  wcth_frame_sink passes frame (NULL = end of
 * stream) to the sink "type:name".  wcth_frame_source sets *out to the
 * next frame, or NULL at end of stream; it returns 1 once the source has
 * reported end of stream and is called again. */
int wcth_frame_sink(const char* tn, void* frame);
int wcth_frame_source(const char* tn, void** out);

/* Tensor sets (ITensorSet / ITensor, iface/inc/WireCellIface/
 * ITensorSet.h:9-20, ITensor.h:14-65): build an input set, run an
 * ITensorSetFilter ("type:name", e.g. Decon2DResponse or Decon2DFilter,
 * sig/src/Decon2DResponse.cxx, Decon2DFilter.cxx) on it (in = NULL is
 * end-of-stream; *out = NULL then), read the result.  Metadata is JSON;
 * dtype strings follow util/src/Dtype.cxx ("f4", "f8", "c8", "i4", ...);
 * tensor data is copied in, and read back in place (valid while the set
 * lives).  The metadata calls return the JSON length (only cap - 1
 * characters are written). */
void* wcth_tensorset_new(int ident, const char* metadata_json);
void wcth_tensorset_free(void* set);
int wcth_tensorset_add(void* set, const char* dtype, int ndim, const size_t* shape, const void* data,
                       const char* metadata_json);
int wcth_tensorset_filter(const char* tn, void* in, void** out);
int wcth_tensorset_info(void* set, int* ident, int* ntensors, char* metadata_json, size_t cap);
/* Tensor i: dtype (up to 7 characters), ndim, shape (up to 8 extents),
 * data pointer and size in bytes. */
int wcth_tensorset_tensor(void* set, int i, char* dtype, int* ndim, size_t* shape, const void** data,
                          size_t* nbytes);
int wcth_tensorset_tensor_md(void* set, int i, char* metadata_json, size_t cap);

/* ROI_formation::find_ROI_by_decon_itself's run scan, extend_ROI_self
 * and create_ROI_connect_info (sigproc/src/ROI_formation.cxx:167-360,
 * 458-531) over a threshold mask (nonzero = above threshold), nw rows
 * of nt bytes at row stride ld.  roi_count[nw] and up to cap inclusive
 * (first, last) pairs; returns the pair count, -2 if cap was exceeded. */
int wcth_roi_lists(const unsigned char* mask, int nw, int nt, long ld, int pad, float asy,
                   int* roi_count, int* roi_pairs, int cap);
/* ROI_formation::find_ROI_loose + extend_ROI_loose
 * (sigproc/src/ROI_formation.cxx:538-937) over a plane's loose decon r
 * (nw rows of nt floats at row stride ld).  bad_ptr[nw+1], bad_lo,
 * bad_hi: per-row inclusive bad tick ranges (CSR) or NULL.  rms: per-row
 * RMS (the row's cal_RMS without its bad ticks, as the decon's extra
 * output computes it) or NULL to compute it here.  tight_count[nw] +
 * tight_pairs: the plane's tight ROIs (inclusive) to extend by, or NULL
 * for none (the collection plane).  rebin, l_factor, l_max_th,
 * l_factor1, l_short_length, l_jump_one_bin: OmnibusSigProc's keys of
 * those names.  Output as wcth_roi_lists. */
int wcth_loose_rois(const float* r, int nw, int nt, long ld, const int* bad_ptr, const int* bad_lo,
                    const int* bad_hi, const float* rms, const int* tight_count, const int* tight_pairs,
                    int rebin, double l_factor, double l_max_th, double l_factor1, int l_short_length,
                    int l_jump_one_bin, int* roi_count, int* roi_pairs, int cap);
/* ROI_refinement::apply_roi (sigproc/src/ROI_refinement.cxx:118-218) in
 * place on r (nw rows of nt at row stride ld) with roi_count[nw] +
 * roi_pairs inclusive ROIs per row; linear = 1 for the induction planes
 * (a straight baseline between the ROI's end samples is subtracted),
 * 0 for the collection plane.  Results are the reference's int
 * truncations, stored as floats. */
int wcth_apply_roi(float* r, int nw, int nt, long ld, const int* roi_count, const int* roi_pairs, int linear);

/* PeakFinding::find_peak (sigproc/src/PeakFinding.cxx:34-474) on n
 * samples: up to cap positions x and heights y (the source at the
 * rounded position); returns the peak count, -2 if cap was exceeded. */
int wcth_find_peaks(const float* signal, int n, int max_peaks, double sigma, double threshold, double* x,
                    double* y, int cap);

/* ROI_refinement (sigproc/src/ROI_refinement.cxx) of one plane, as a
 * handle.  plane 0/1 (induction: the loose ROIs are refined) or 2
 * (collection: the tight ones).  params: the 13 values r_th_factor,
 * r_fake_signal_low_th, r_fake_signal_high_th,
 * r_fake_signal_low_th_ind_factor, r_fake_signal_high_th_ind_factor,
 * r_pad, r_break_roi_loop, r_th_peak, r_sep_peak,
 * r_low_peak_sep_threshold_pre, r_max_npeaks, r_sigma, r_th_percent
 * (OmnibusSigProc's keys), or NULL for their defaults; wrapped: the
 * isWrapped key.  NULL on a bad argument (wcth_last_error). */
void* wcth_refine_new(int plane, int nw, int nt, const double* params, int wrapped);
void wcth_refine_free(void* h);
/* ROI_refinement::load_data (:269-699): the plane's decon r (nw rows of
 * nt at row stride ld), per-row inclusive bad tick ranges (CSR or NULL),
 * the tight stage's per-wire RMS, and the tight and loose ROI lists
 * (count per wire + inclusive pairs; loose NULL on the collection
 * plane). */
int wcth_refine_load(void* h, const float* r, long ld, const int* bad_ptr, const int* bad_lo, const int* bad_hi,
                     const float* rms, const int* tight_count, const int* tight_pairs, const int* loose_count,
                     const int* loose_pairs);
/* One step by name: "cleanup" (CleanUpROIs), "merge"
 * (generate_merge_ROIs), "check" (CheckROIs), "break" (BreakROIs),
 * "shrink" (ShrinkROIs), "cleanup_collection", "cleanup_induction",
 * "extend" (ExtendROIs), or "all" for OmnibusSigProc's sequence. */
int wcth_refine_step(void* h, const char* step);
/* The plane's ROIs per wire in list order: extended ranges (after
 * "extend") or the ROI ranges; output as wcth_roi_lists. */
int wcth_refine_rois(void* h, int extended, int* roi_count, int* roi_pairs, int cap);
/* The contents of wire w's ROIs, concatenated in list order; returns
 * the value count, -2 if cap was exceeded. */
int wcth_refine_contents(void* h, int w, float* out, int cap);

#ifdef __cplusplus
}
#endif
#endif
