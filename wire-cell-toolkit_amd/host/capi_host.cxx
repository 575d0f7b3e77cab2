// extern "C" front of the host-side response helpers (include/wct_host.h).
#include "../../include/wct_host.h"

#include <algorithm>
#include <cstring>

#include "Response.h"

using namespace WireCell::Hip;

extern "C" {

int wcth_coldelec_samples(int n, double tmin, double tmax, double gain, double shaping,
                          double postgain, double* out)
{
    if (n <= 0 || !out || !(tmax > tmin)) return -1;
    auto v = coldelec_samples(n, tmin, tmax, gain, shaping, postgain);
    std::copy(v.begin(), v.end(), out);
    return 0;
}

int wcth_hf_filter(int n, double max_freq, double sigma, double power, int flag, double* out)
{
    if (n <= 0 || !out) return -1;
    auto v = hf_filter_waveform(n, max_freq, sigma, power, flag != 0);
    std::copy(v.begin(), v.end(), out);
    return 0;
}

int wcth_lf_filter(int n, double max_freq, double tau, double* out)
{
    if (n <= 0 || !out) return -1;
    auto v = lf_filter_waveform(n, max_freq, tau);
    std::copy(v.begin(), v.end(), out);
    return 0;
}

long wcth_fft_best_length(long n, int keep_odd_even)
{
    return (long)fft_best_length((size_t)n, keep_odd_even != 0);
}

int wcth_wire_region_average(int npaths, const double* pitchpos, const double* current, int nsamp,
                             double pitch, int max_regions, int* regions_out, double* avg_out)
{
    if (npaths <= 0 || !pitchpos || !current || nsamp <= 0 || pitch <= 0) return -1;
    PlaneResponse pr;
    pr.pitch = pitch;
    pr.planeid = 0;
    pr.location = 0;
    for (int p = 0; p < npaths; ++p) {
        PathResponse path;
        path.pitchpos = pitchpos[p];
        path.current.assign(current + (size_t)p * nsamp, current + (size_t)(p + 1) * nsamp);
        pr.paths.push_back(path);
    }
    RegionAverage ra = wire_region_average(pr);
    const int n = (int)ra.regions.size();
    if (n > max_regions) return -2;
    for (int r = 0; r < n; ++r) {
        regions_out[r] = ra.regions[r];
        for (int k = 0; k < nsamp; ++k)
            avg_out[(size_t)r * nsamp + k] = k < (int)ra.rows[r].size() ? ra.rows[r][k] : 0.0;
    }
    return n;
}

}  // extern "C"
