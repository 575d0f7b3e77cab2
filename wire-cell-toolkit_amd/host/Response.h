// Host-side response and filter functions the deconvolution needs,
// restated from the reference (util/src/Response.cxx, sigproc/src/
// HfFilter.cxx, sigproc/src/LfFilter.cxx, util/src/FFTBestLength.cxx,
// gen/src/ColdElecResponse.cxx).  All double precision; the FP64
// response spectra are then built on the device.
#pragma once
#include <map>
#include <string>
#include <vector>

namespace WireCell {
namespace Hip {

// WCT system of units (util/inc/WireCellUtil/Units.h): ns, mm, e+.
namespace units {
constexpr double nanosecond = 1.0;
constexpr double second = 1e9;
constexpr double microsecond = 1e-6 * second;
constexpr double hertz = 1.0 / second;
constexpr double megahertz = 1e6 * hertz;
constexpr double millimeter = 1.0;
constexpr double eplus = 1.0;
constexpr double e_SI = 1.602176487e-19;
constexpr double coulomb = eplus / e_SI;
constexpr double fC = 1.0e-15 * coulomb;
constexpr double volt = 1e-6;
constexpr double mV = 1e-3 * volt;
}  // namespace units

// One Garfield-style path (impact position) of a plane's field response.
struct PathResponse {
    std::vector<double> current;
    double pitchpos;
};
struct PlaneResponse {
    std::vector<PathResponse> paths;
    int planeid;
    double location;
    double pitch;
};
struct FieldResponse {
    std::vector<PlaneResponse> planes;
    double origin, tstart, period, speed;
};

// Response::coldelec (util/src/Response.cxx:277-308).
double coldelec(double time, double gain, double shaping);

// Gen::ColdElecResponse::waveform_samples(Binning(n, tmin, tmax)):
// bin centres times postgain (gen/src/ColdElecResponse.cxx:44-49).
std::vector<double> coldelec_samples(int n, double tmin, double tmax, double gain, double shaping,
                                     double postgain);

// SigProc::HfFilter / LfFilter ::filter_waveform (HfFilter.cxx:38-52,
// LfFilter.cxx:31-45) with Response::hf_filter / lf_filter
// (util/src/Response.cxx:435-444).
std::vector<double> hf_filter_waveform(int n, double max_freq, double sigma, double power, bool flag);
std::vector<double> lf_filter_waveform(int n, double max_freq, double tau);

// util/src/FFTBestLength.cxx:12-58.
size_t fft_best_length(size_t n, bool keep_odd_even = false);

// Response::wire_region_average (util/src/Response.cxx:79-196): the
// region-averaged paths, ordered by increasing wire region.  The averages
// are accumulated in float as the reference's realseq_t are.
struct RegionAverage {
    std::vector<int> regions;
    std::vector<std::vector<double>> rows;
};
RegionAverage wire_region_average(const PlaneResponse& plane);

}  // namespace Hip
}  // namespace WireCell
