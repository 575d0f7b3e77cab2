// High-resolution peak search of ROI_refinement::BreakROI: the
// sigproc PeakFinding class (sigproc/src/PeakFinding.cxx:34-474), a
// restatement of the TSpectrum "SearchHighRes" method: the signal is
// extended at both ends, smoothed by the Markov-chain method, deconvolved
// by the Gold algorithm with a Gaussian response of width sigma, and the
// local maxima of the deconvolved spectrum above threshold percent of its
// maximum (and of the smoothed source) are the peaks, ordered by the
// height of the smoothed source at them.  Double-precision arithmetic in
// the reference's order of operations.
#pragma once
#include <vector>

namespace WireCell {
namespace Hip {

struct PeakSearch {
    int max_peaks{200};
    double sigma{1};
    double threshold{0.05};    // percent of the deconvolved maximum
    bool background_remove{false};
    int decon_iterations{3};
    bool markov{true};
    int aver_window{3};
};

struct Peaks {
    std::vector<double> x;     // positions (fractional bins, 0 .. n-1)
    std::vector<double> y;     // the source at the rounded position
};

// The peaks of signal (PeakFinding::find_peak).  Fewer than 2 sigma + 1
// samples give none.
Peaks find_peaks(const std::vector<float>& signal, const PeakSearch& ps);

}  // namespace Hip
}  // namespace WireCell
