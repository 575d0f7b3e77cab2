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

#ifdef __cplusplus
}
#endif
#endif
