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

}  // namespace Hip
}  // namespace WireCell
