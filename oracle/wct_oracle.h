/* CPU oracle for the OmnibusSigProc 2D deconvolution hot path.
 * TEST INFRASTRUCTURE ONLY: the checker, never the product. */
#ifndef WCT_ORACLE_H
#define WCT_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

void wo_dft_naive(int n, const double* in, double* out, int sign);

/* IDFT methods with FftwDFT semantics; _f float, _d double.
 * Complex arrays are interleaved (re, im). */
int wo_fwd1d_f(const float* in, float* out, int n);
int wo_inv1d_f(const float* in, float* out, int n);
int wo_fwd1b_f(const float* in, float* out, int nrows, int ncols, int axis);
int wo_inv1b_f(const float* in, float* out, int nrows, int ncols, int axis);
int wo_fwd2d_f(const float* in, float* out, int nrows, int ncols);
int wo_inv2d_f(const float* in, float* out, int nrows, int ncols);
int wo_fwd1d_d(const double* in, double* out, int n);
int wo_inv1d_d(const double* in, double* out, int n);
int wo_fwd1b_d(const double* in, double* out, int nrows, int ncols, int axis);
int wo_inv1b_d(const double* in, double* out, int nrows, int ncols, int axis);
int wo_fwd2d_d(const double* in, double* out, int nrows, int ncols);
int wo_inv2d_d(const double* in, double* out, int nrows, int ncols);
void wo_hermitian_mirror_f(float* v, int n);
void wo_hermitian_mirror_d(double* v, int n);

float wo_percentile_f(const float* wave, size_t n, float p);
double wo_percentile_d(const double* wave, size_t n, float p);
void wo_restore_baseline_f(float* arr, int nrows, int ncols, float* base_out);
void wo_restore_baseline_d(double* arr, int nrows, int ncols, double* base_out);

int wo_overall_response_f(const float* avg, int nrows, int nin, const float* elec, int nfine,
                          double fine_period, double coarse_period, double inter_gain,
                          double adc_mv, double fine_time_offset, int nticks, float* resp);
int wo_overall_response_d(const double* avg, int nrows, int nin, const double* elec, int nfine,
                          double fine_period, double coarse_period, double inter_gain,
                          double adc_mv, double fine_time_offset, int nticks, double* resp);

int wo_decon_init_f(const float* data, int nw, int nt, const float* resp, int nr,
                    const float* wire_filter, int wire_shift, int time_shift, float* c_data);
int wo_decon_init_d(const double* data, int nw, int nt, const double* resp, int nr,
                    const double* wire_filter, int wire_shift, int time_shift, double* c_data);
int wo_decon_init_cr_f(const float* data, int nw, int nt, const float* resp, int nr,
                       const float* wire_filter, int wire_shift, int time_shift,
                       const float* ch_resp, int nsamp, const float* ewave, float* c_data);
int wo_decon_init_cr_d(const double* data, int nw, int nt, const double* resp, int nr,
                       const double* wire_filter, int wire_shift, int time_shift,
                       const double* ch_resp, int nsamp, const double* ewave, double* c_data);
int wo_decon_apply_filter_f(const float* c_data, int nw, int nt, const float* tf, float* out);
int wo_decon_apply_filter_d(const double* c_data, int nw, int nt, const double* tf, double* out);
int wo_decon_fused_d(const double* data, int nw, int nt, const double* resp, int nr,
                     const double* wire_filter, const double* time_filter,
                     int wire_shift, int time_shift, double* out);
int wo_forward_conv_d(const double* q, int nw, int nt, const double* resp, int nr,
                      int wire_shift, int time_shift, double* out);

double wo_cal_rms_f(const float* signal, size_t n);
double wo_cal_rms_d(const double* signal, size_t n);
void wo_roi_threshold_f(const float* r_data, const float* r_tight, int nw, int nt,
                        float th_factor, const int* bad_row, const int* bad_lo,
                        const int* bad_hi, int nbad, double* rms_out, double* th_out,
                        unsigned char* mask);
void wo_roi_threshold_d(const double* r_data, const double* r_tight, int nw, int nt,
                        float th_factor, const int* bad_row, const int* bad_lo,
                        const int* bad_hi, int nbad, double* rms_out, double* th_out,
                        unsigned char* mask);
int wo_roi_lists(const unsigned char* mask, int nw, int nt, int pad, float asy,
                 int* roi_count, int* roi_pairs, int cap);

int wo_loose_rois(const float* r_data, int nw, int nt, const int* bad_ptr, const int* bad_lo,
                  const int* bad_hi, const double* rms, const int* tight_ptr, const int* tight_pairs,
                  int rebin, double l_factor, double l_max_th, double l_factor1, int l_short_length,
                  int l_jump_one_bin, int* roi_count, int* roi_pairs, int cap);
void wo_apply_roi(float* r_data, int nw, int nt, const int* roi_ptr, const int* roi_pairs, int linear);

void wo_hf_filter(int n, double max_freq, double sigma, double power, int flag, double* out);
void wo_lf_filter(int n, double max_freq, double tau, double* out);
double wo_coldelec(double time, double gain, double shaping);
void wo_elec_samples(int n, double tmin, double tmax, double gain, double shaping,
                     double postgain, float* out_f, double* out_d);
long wo_fft_best_length(long n, int keep_odd_even);
int wo_wire_region_average(int npaths, const double* pitchpos, const float* current,
                           int nsamp, double pitch, int max_regions, int* regions_out,
                           float* avg_out);

double wo_uniform(uint64_t seed, uint64_t stream, uint64_t idx);
void wo_normal(uint64_t seed, uint64_t stream, size_t n, double sigma, double* out);
void wo_tracks(uint64_t seed, uint64_t stream, int ntracks, int nw, int nt,
               double q_per_tick, double* q);

#ifdef __cplusplus
}
#endif
#endif
