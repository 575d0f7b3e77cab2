// Tight ROI lists from the device's threshold mask: the run scan of
// ROI_formation::find_ROI_by_decon_itself (sigproc/src/ROI_formation.cxx:
// 458-531), its extend_ROI_self padding (:167-250) and
// create_ROI_connect_info gap filling (:252-360).  The per-sample
// threshold test itself (th = factor * rms + 1, tight or tighter above
// it) runs on the device and arrives as a nonzero mask byte.
#pragma once
#include <cstdint>
#include <utility>
#include <vector>

namespace WireCell {
namespace Hip {

typedef std::pair<int, int> Roi;   // inclusive tick range [first, second]
typedef std::vector<Roi> RoiList;

// Runs of nonzero mask bytes in one row of nt ticks.  As in the
// reference a run can only start before the last tick and the scan
// resumes two ticks after a run ends.
RoiList row_runs(const uint8_t* mask, int nt);

// Pads every ROI by pad ticks (clamped to [0, nt-1]) and merges an ROI
// into the previous one when it starts strictly before that one's end.
void extend_rois(RoiList& rois, int pad, int nt);

// For wires i and i+2 holding ROIs of similar length
// (|l2 - l1| < (l1 + l2) * asy) whose midpoint range is consistent, adds
// that range to wire i+1 unless it overlaps an ROI already there.
// Sequential in i: ranges added to i+1 take part when i+1 is the lower
// wire of the next step.
void connect_rois(std::vector<RoiList>& rows, float asy);

// All three steps over a plane's mask (nw rows, row stride ld bytes).
std::vector<RoiList> tight_rois(const uint8_t* mask, int nw, int nt, long ld, int pad, float asy);

// ROI_formation::cal_RMS (sigproc/src/ROI_formation.cxx:364-389) of a
// sample set: the 16/50/84 % points by Waveform::percentile
// (util/src/Waveform.cxx:89-106), then the float-accumulated RMS of the
// samples within 5 times their spread.  0 for an empty set.
double cal_rms(std::vector<float> samples);

// Loose-ROI parameters (OmnibusSigProc's m_rebin, m_l_factor,
// m_l_max_th, m_l_factor1, m_l_short_length, m_l_jump_one_bin,
// sigproc/inc/WireCellSigProc/OmnibusSigProc.h:134-139).
struct LooseParams {
    int rebin{6};
    double factor{3.5};
    double max_th{10000};
    double factor1{0.7};
    int short_length{3};
    int jump_one_bin{0};
};

// ROI_formation::find_ROI_loose on one row of nt ticks
// (ROI_formation.cxx:687-937): the row is summed in groups of rebin ticks
// (bad ticks as 0); a group above th = rms * rebin * factor (at most
// max_th), or a local maximum that stands out of its neighbourhood by
// more than 2 th or by th * factor1 within short_length groups on both
// sides, opens an ROI that is walked outwards past th * factor1 (or the
// neighbour) and then downhill while the next 3-group average does not
// rise by 25 or more (find_ROI_begin / find_ROI_end, :620-685);
// overlapping ROIs merge, and the ROI is scaled back to ticks.  rms is
// the row's cal_RMS with its bad ticks left out.  (The reference's
// pairwise merge loop, :861-900, never runs: its flag starts at 0.)
RoiList loose_row(const float* row, int nt, const RoiList& bad, double rms, const LooseParams& lp);

// extend_ROI_loose (:538-596): each loose ROI whose start or end falls
// strictly inside a tight ROI is widened to it, then ROIs that start
// before the previous one's end merge.
void extend_loose(RoiList& loose, const RoiList& tight);

// A plane: rows of r (row stride ld) with per-row bad tick ranges
// (inclusive; CSR bad_ptr[nw+1] into bad_lo/bad_hi, or null), rms per
// row (or null: cal_RMS of the row without its bad ticks) and the tight
// ROIs to extend by (induction planes; null for none).
std::vector<RoiList> loose_rois(const float* r, int nw, int nt, long ld, const int* bad_ptr, const int* bad_lo,
                                const int* bad_hi, const float* rms, const std::vector<RoiList>* tight,
                                const LooseParams& lp);

// ROI_refinement::apply_roi (sigproc/src/ROI_refinement.cxx:118-218) on a
// plane in place: ticks outside the row's ROIs become 0; inside ROI
// [s, e], applied in list order, tick i becomes the int of r(i) minus the
// straight line through r(s) and r(e) (linear, the induction planes) or
// of r(i) (the collection plane), in float arithmetic, with the
// conversion x86-64 makes (truncation; NaN, as a one-tick ROI's 0/0
// gives, and values outside the int range become INT_MIN).
void apply_roi(float* r, int nw, int nt, long ld, const std::vector<RoiList>& rois, bool linear);

}  // namespace Hip
}  // namespace WireCell
