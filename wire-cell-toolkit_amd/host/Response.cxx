#include "Response.h"

#include <algorithm>
#include <cmath>
#include <set>

namespace WireCell {
namespace Hip {

double coldelec(double time, double gain, double shaping)
{
    // Closed form of the cold-electronics shaper (util/src/Response.cxx:
    // 277-308): zero outside (0, 10 us), gain scaled by 10*1.012, and the
    // reference's inverse-Laplace coefficients (they are the definition of
    // the response, so they are the same numbers).
    if (time <= 0 || time >= 10 * units::microsecond) return 0.0;
    const double x = time / shaping;
    gain *= 10 * 1.012;
    const double e1 = std::exp(-2.94809 * x);
    const double e2 = std::exp(-2.82833 * x);
    const double e3 = std::exp(-2.40318 * x);
    const double c1 = std::cos(1.19361 * x), s1 = std::sin(1.19361 * x);
    const double c2 = std::cos(2.38722 * x), s2 = std::sin(2.38722 * x);
    const double c3 = std::cos(2.5928 * x), s3 = std::sin(2.5928 * x);
    const double c4 = std::cos(5.18561 * x), s4 = std::sin(5.18561 * x);
    double a = 4.31054 * e1;
    a -= 2.6202 * e2 * c1 * (1.0 + c2);
    a += 0.464924 * e3 * c3 * (1.0 + c4);
    a += 0.762456 * e2 * (s1 - c2 * s1 + c1 * s2);
    a -= 2.620200 * e2 * s1 * s2;
    a += 0.327684 * e3 * (c4 * s3 - s3 - c3 * s4);
    a += 0.464924 * e3 * s3 * s4;
    return a * gain;
}

std::vector<double> coldelec_samples(int n, double tmin, double tmax, double gain, double shaping,
                                     double postgain)
{
    std::vector<double> out(n);
    const double bin = (tmax - tmin) / n;
    for (int i = 0; i < n; ++i) out[i] = coldelec(tmin + (i + 0.5) * bin, gain, shaping) * postgain;
    return out;
}

static double folded_freq(int i, int n, double max_freq)
{
    double f = i * 1.0 / n * 2 * max_freq;
    if (f > max_freq) f -= 2 * max_freq;
    return std::fabs(f);
}

std::vector<double> hf_filter_waveform(int n, double max_freq, double sigma, double power, bool flag)
{
    std::vector<double> out(n);
    for (int i = 0; i < n; ++i) {
        const double f = folded_freq(i, n, max_freq);
        out[i] = (flag && f == 0) ? 0.0 : std::exp(-0.5 * std::pow(f / sigma, power));
    }
    return out;
}

std::vector<double> lf_filter_waveform(int n, double max_freq, double tau)
{
    std::vector<double> out(n);
    for (int i = 0; i < n; ++i) {
        const double f = folded_freq(i, n, max_freq);
        out[i] = 1 - std::exp(-std::pow(f / tau, 2));
    }
    return out;
}

size_t fft_best_length(size_t n, bool keep_odd_even)
{
    // The reference's table of preferred lengths (data, not code).
    static const size_t any[] = {33, 35, 40, 42, 44, 56, 64, 66, 70, 77, 80, 96, 98, 100, 112, 121, 128, 140,
        143, 160, 168, 169, 200, 224, 256, 280, 320, 400, 448, 512, 560, 640, 686, 768, 800, 847, 896, 1024,
        1120, 1280, 1331, 1372, 1400, 1536, 1600, 1792, 2048, 2560, 2662, 2744, 3200, 3584, 4096, 4116, 4480,
        4802, 5120, 5324, 5376, 5544, 5600, 5632, 5929, 6174, 6272, 6292, 6655, 7168, 7840, 8192, 8232, 8960,
        9317, 10240, 10648, 11979, 14336, 15972, 16384};
    static const size_t even[] = {40, 42, 44, 56, 64, 66, 70, 72, 80, 96, 98, 100, 112, 120, 128, 140, 160,
        168, 200, 224, 256, 280, 320, 400, 448, 512, 560, 640, 686, 768, 800, 896, 1024, 1120, 1280, 1372,
        1400, 1536, 1600, 1792, 2048, 2560, 2662, 2744, 3200, 3584, 4096, 4116, 4480, 4802, 5120, 5324, 5376,
        5544, 5600, 5632, 6174, 6272, 6292, 6468, 7168, 7840, 8192, 8232, 8960, 10240, 10648, 11200, 11858,
        11880, 14336, 15972, 16384};
    static const size_t odd[] = {33, 35, 39, 49, 55, 63, 65, 77, 81, 91, 99, 121, 143, 169, 175, 187, 225,
        231, 245, 275, 297, 343, 363, 375, 385, 405, 441, 455, 495, 539, 605, 637, 693, 735, 847, 875, 891,
        945, 1001, 1089, 1331, 1573, 1859, 1875, 2197, 2299, 2431, 2541, 2625, 2695, 3025, 3267, 3993, 4095,
        4125, 4719, 4725, 5005, 5145, 5445, 5929, 6655, 7007, 7623, 7865, 9317, 9801, 11979, 14157, 14175,
        15379, 16335};
    const size_t* b;
    const size_t* e;
    if (!keep_odd_even) { b = any; e = any + sizeof(any) / sizeof(any[0]); }
    else if (n % 2 == 0) { b = even; e = even + sizeof(even) / sizeof(even[0]); }
    else { b = odd; e = odd + sizeof(odd) / sizeof(odd[0]); }
    const size_t* it = std::lower_bound(b, e, n);
    return it == e ? n : *it;
}

RegionAverage wire_region_average(const PlaneResponse& plane)
{
    const double pitch = plane.pitch;
    // Paths are folded onto +-pitchpos in units of 1% of the pitch; a
    // second path landing on an occupied slot is averaged in.  Only a
    // repeat met through the +pitchpos slot fixes the sample count that
    // sizes the averages (Response.cxx:98-119,161).
    std::map<int, std::vector<float>> slots;
    size_t nsamples = 0;
    for (const auto& path : plane.paths) {
        const int eff = (int)(path.pitchpos / (0.01 * pitch));
        for (int side = 0; side < 2; ++side) {
            const int key = side == 0 ? eff : -eff;
            auto it = slots.find(key);
            if (it == slots.end()) {
                slots[key] = std::vector<float>(path.current.begin(), path.current.end());
                continue;
            }
            if (side == 0) nsamples = path.current.size();
            for (size_t k = 0; k < path.current.size(); ++k)
                it->second[k] = (float)(((double)it->second[k] + (double)(float)path.current[k]) / 2.);
        }
    }
    std::vector<int> keys;
    for (auto& kv : slots) keys.push_back(kv.first);
    std::map<int, std::pair<double, double>> range;
    for (size_t i = 0; i < keys.size(); ++i) {
        const double lo = i == 0 ? -1e9 : (keys[i] + keys[i - 1]) / 2. * 0.01 * pitch;
        const double hi = i + 1 == keys.size() ? 1e9 : (keys[i] + keys[i + 1]) / 2. * 0.01 * pitch;
        range[keys[i]] = {lo, hi};
    }
    std::set<int> regions;
    for (int k : keys) {
        const double off = k > 0 ? -0.001 * pitch : 0.001 * pitch;
        regions.insert((int)std::round((k * 0.01 * pitch + off) / pitch));
    }
    RegionAverage ra;
    for (int wire : regions) {
        std::vector<float> avg(nsamples, 0.0f);
        for (auto& kv : slots) {
            double lo = range[kv.first].first, hi = range[kv.first].second;
            lo = std::max(lo, (wire - 0.5) * pitch);
            hi = std::min(hi, (wire + 0.5) * pitch);
            if (hi > lo)
                for (size_t k = 0; k < nsamples && k < kv.second.size(); ++k)
                    avg[k] = (float)((double)avg[k] + (double)kv.second[k] * (hi - lo) / pitch);
        }
        ra.regions.push_back(wire);
        ra.rows.emplace_back(avg.begin(), avg.end());
    }
    return ra;
}

}  // namespace Hip
}  // namespace WireCell
