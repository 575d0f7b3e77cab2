// ROI refinement of OmnibusSigProc (sigproc/src/ROI_refinement.cxx):
// the signal ROIs of one plane, their links to overlapping ROIs on the
// neighbouring wires and to the tight ROIs they contain, and the steps
// OmnibusSigProc::operator() runs on them (OmnibusSigProc.cxx:1506-1590)
// before ROI_refinement::apply_roi cuts the hits and charge decons to
// them.
//
// An ROI keeps the decon of its wire on its tick range minus the straight
// line through its end samples (float arithmetic, SignalROI.cxx:9-22).
// ROIs are held by index in creation order; every wire keeps its list in
// the reference's list order (new ROIs appended, ExtendROIs sorting by
// start tick, stably), and every link list keeps its insertion order, so
// each step visits ROIs in the order the reference does.  A link list
// also records whether its key exists (the reference's map lookups test
// the key, not the list), which matters where a key is left with an
// empty list.
//
// Not restated: the multi-plane protection (MP2ROI/MP3ROI), which needs
// the anode's wire geometry; BreakROI then has no protected zones.
#pragma once
#include <utility>
#include <vector>

#include "PeakFinding.h"
#include "RoiFormation.h"

namespace WireCell {
namespace Hip {

// OmnibusSigProc's r_* configuration (sigproc/inc/WireCellSigProc/
// OmnibusSigProc.h:142-154) and isWrapped (:229).
struct RefineParams {
    float th_factor{3.0f};
    float fake_signal_low_th{500};
    float fake_signal_high_th{1000};
    float fake_signal_low_th_ind_factor{1.0f};
    float fake_signal_high_th_ind_factor{1.0f};
    int pad{5};
    int break_roi_loop{2};
    float th_peak{3.0f};
    float sep_peak{6.0f};
    float low_peak_sep_threshold_pre{1200};
    int max_npeaks{200};
    float sigma{2};
    float th_percent{0.1f};
    bool wrapped{false};
};

class RoiRefinement {
public:
    // plane 0/1 (induction: loose ROIs are refined) or 2 (collection:
    // tight ROIs); nw wires, nt ticks.
    RoiRefinement(int plane, int nw, int nt, const RefineParams& p);

    // ROI_refinement::load_data (:269-699).  r: the plane's decon (nw rows
    // of nt, row stride ld); bad: per-wire inclusive tick ranges zeroed in
    // the ROI contents; rms: the per-wire RMS of the tight stage; tight,
    // loose: the ROI lists of ROI_formation (loose ignored on the
    // collection plane).
    void load(const float* r, long ld, const std::vector<RoiList>& bad, const std::vector<float>& rms,
              const std::vector<RoiList>& tight, const std::vector<RoiList>& loose);

    void cleanup_rois();              // CleanUpROIs (:701-873)
    void generate_merge_rois();       // generate_merge_ROIs (:875-1159)
    void check_rois();                // CheckROIs (:1161-1284)
    void break_rois();                // BreakROIs (:2448-2481)
    void shrink_rois();               // ShrinkROIs (:1844-1865)
    void cleanup_collection_rois();   // CleanUpCollectionROIs (:1286-1366)
    void cleanup_induction_rois();    // CleanUpInductionROIs (:1368-1601)
    void extend_rois();               // ExtendROIs (:2946-3136)

    // The sequence of OmnibusSigProc::operator() for this plane (without
    // the multi-plane protection): cleanup and merge, break_roi_loop
    // rounds of break / check / cleanup, shrink / check / cleanup, the
    // plane's fake-signal cleanup, extend.
    void refine();

    // The refined ROIs of each wire in list order: (ext_start, ext_end)
    // after extend_rois, else (start, end).
    std::vector<RoiList> rois(bool extended = true) const;

    // The contents of the ROIs of a wire (list order), for inspection.
    std::vector<std::vector<float>> contents(int wire) const;

    int plane() const { return m_plane; }
    int nwires() const { return m_nw; }
    int nticks() const { return m_nt; }

private:
    struct Sroi {
        int wire;
        int start, end;
        int ext_start, ext_end;
        std::vector<float> contents;
    };
    struct Links {
        bool key = false;
        std::vector<int> ids;
    };

    int make_roi(int wire, int start, int end, const std::vector<float>& signal);
    int copy_roi(int id);
    bool overlap(int a, int b) const;
    bool overlap_th(int a, int b, float th, float th1) const;
    std::vector<std::pair<int, int>> above(int id, float th) const;
    double average(int id) const;
    void add(std::vector<Links>& m, int key, int val);   // push_back, creating the key
    void link(int prev, int next);
    void unlink(int prev, int next);
    void drop_links(int id);   // unlink from every neighbour, erase both keys
    void replace(int id, const std::vector<int>& repl);   // list, links and containment move to repl
    void break_roi(int id, float rms);
    void break_roi1(int id);
    void shrink_roi(int id);
    void remove_unsupported(bool use_mean, float high_th, float mean_th);
    std::vector<int>& list(int w) { return m_plane == 2 ? m_tight[w] : m_loose[w]; }
    const std::vector<int>& list(int w) const { return m_plane == 2 ? m_tight[w] : m_loose[w]; }
    Links& grow(std::vector<Links>& m, int id);

    int m_plane, m_nw, m_nt;
    RefineParams m_p;
    std::vector<float> m_rms;
    std::vector<Sroi> m_pool;
    std::vector<std::vector<int>> m_tight, m_loose;
    std::vector<Links> m_front, m_back, m_contained;
};

}  // namespace Hip
}  // namespace WireCell
