"""ROI refinement (sigproc/src/ROI_refinement.cxx, SignalROI.cxx) and its
peak search (sigproc/src/PeakFinding.cxx).

The oracle (oracle/roi_refine.py, a statement-by-statement restatement)
is pinned by hand-worked cases of the reference's rules: the ROI
baseline, the overlap test (more than one common tick), CleanUpROIs
keeping only loose ROIs linked to one that contains a tight ROI,
generate_merge_ROIs turning uncovered tight ROIs into loose ones,
ShrinkROI cutting to the contained tight span plus pad, BreakROI1
splitting at zero samples, the fake-signal cleanups (above the high
threshold, or above the low one on average, or next to such an ROI),
ExtendROIs' neighbour span and midpoint rule, and the peak search on
single and double Gaussians (position, order by height, too-short
signals).  The product (host/RoiRefinement.cxx, PeakFinding.cxx, through
the C front) must then equal the oracle bit for bit, after every step,
on seeded decon-like planes of every plane type, with bad channels and a
wrapped plane, and on random peak-search inputs.
"""
import math

import numpy as np
import pytest

import oracle as O
import roi_refine as RR

F = np.float32


def _hostlib():
    from wct_amd import components
    components.lib()
    return components


# ---- hand-worked pins of the oracle --------------------------------------

def test_signalroi_baseline_and_overlap():
    sig = [F(v) for v in (0, 1, 5, 9, 3, 1, 0)]
    r = RR.SignalROI(0, 0, 1, 5, sig)
    # line through sig[1] = 1 and sig[5] = 1: contents = sig - 1
    assert [float(v) for v in r.contents] == [0.0, 4.0, 8.0, 2.0, 0.0]
    r2 = RR.SignalROI(0, 0, 0, 6, [F(v) for v in range(7)])   # a straight line leaves zeros
    assert all(v == 0 for v in r2.contents)
    one = RR.SignalROI(0, 0, 3, 3, sig)                      # one tick: 0/0
    assert math.isnan(one.contents[0])
    a = RR.SignalROI(0, 0, 10, 20, [F(0)] * 40)
    assert not a.overlap(RR.SignalROI(0, 1, 20, 30, [F(0)] * 40))   # one common tick
    assert a.overlap(RR.SignalROI(0, 1, 19, 30, [F(0)] * 40))
    assert a.get_above_threshold(F(0)) == []
    assert r.get_above_threshold(F(3)) == [(1, 2)]


def _plane_from(rows, nt):
    r = np.zeros((len(rows), nt), np.float32)
    for w, pulses in enumerate(rows):
        for (c, a, s) in pulses:
            r[w] += (a * np.exp(-0.5 * ((np.arange(nt) - c) / s) ** 2)).astype(np.float32)
    return r


def test_cleanup_and_merge_rules():
    nt = 200
    # wire 0: a loose ROI with a tight ROI inside; wire 1: a loose ROI
    # overlapping it (kept through the link); wire 3: an isolated loose ROI
    # with no tight ROI (removed); wire 2: a tight ROI no loose ROI covers
    # (becomes a loose ROI of its own)
    r = _plane_from([[(60, 3000, 4)], [(62, 3000, 4)], [(150, 3000, 4)], [(100, 3000, 4)]], nt)
    rms = np.full(4, 10.0, np.float32)
    tight = [[(55, 65)], [], [(140, 160)], []]
    loose = [[(40, 80)], [(45, 85)], [], [(90, 110)]]
    Q = RR.RoiRefinement(0, 4)
    Q.load_data(r, rms, tight, loose)
    assert Q.rois(False) == [[(40, 80)], [(45, 85)], [], [(90, 110)]]
    Q.CleanUpROIs()
    assert Q.rois(False) == [[(40, 80)], [(45, 85)], [], []]
    Q.generate_merge_ROIs()
    assert Q.rois(False) == [[(40, 80)], [(45, 85)], [(140, 160)], []]
    roi2 = Q.rois_loose[2][0]
    assert Q.contained_rois[roi2] == [Q.rois_tight[2][0]]


def test_shrink_break1_and_extend_rules():
    nt = 300
    r = _plane_from([[(100, 4000, 3)], [(205, 500, 2)]], nt)
    rms = np.full(2, 10.0, np.float32)
    Q = RR.RoiRefinement(0, 2, pad=5)
    Q.load_data(r, rms, [[(95, 105)], []], [[(60, 140)], [(200, 210)]])
    Q.ShrinkROIs()
    # wire 0 shrinks to its tight ROI +- pad; wire 1 has nothing marked and
    # keeps its range (rebuilt with a new baseline)
    assert Q.rois(False) == [[(90, 110)], [(200, 210)]]
    # BreakROI1: the ROI splits at its zero samples
    roi = Q.rois_loose[0][0]
    roi.contents = [F(0), F(5), F(0), F(7), F(8), F(0)] + [F(3)] * (len(roi.contents) - 6)
    Q.BreakROI1(roi)
    assert Q.rois(False)[0] == [(90, 92), (92, 95)]
    # ExtendROIs (collection plane): wire 1's ROI C = (55, 80) overlaps both
    # of wire 0's, A = (40, 60) and B = (75, 90); A extends to (40, 80), B
    # to (55, 90), and as they now overlap both end at the gap's midpoint
    # int(60 * 0.5 + 75 * 0.5) = 67; C extends to (40, 90)
    Q2 = RR.RoiRefinement(2, 2)
    rr = _plane_from([[(50, 3000, 3), (80, 3000, 3)], [(67, 3000, 3)]], 200)
    Q2.load_data(rr, np.full(2, 10.0, np.float32), [[(40, 60), (75, 90)], [(55, 80)]])
    Q2.ExtendROIs()
    assert Q2.rois(True) == [[(40, 67), (67, 90)], [(40, 90)]]


def _box(nt, lo, hi, h):
    row = np.zeros(nt, np.float32)
    row[lo:hi + 1] = h
    return row


def test_fake_signal_cleanup_rules():
    nt = 200
    # collection plane: a strong ROI (above 1000), a weak one linked to it
    # (kept), a weak isolated one (removed), one averaging above 500 with
    # no sample above 1000 (kept): a box of 800 on ticks 85-115 in the
    # ROI 80-120 averages 800 * 31 / 41 = 605
    r = _plane_from([[(50, 3000, 3)], [(52, 300, 3)], [(150, 300, 3)], []], nt)
    r[3] = _box(nt, 85, 115, 800)
    rms = np.full(4, 10.0, np.float32)
    tight = [[(45, 55)], [(47, 57)], [(145, 155)], [(80, 120)]]
    Q = RR.RoiRefinement(2, 4)
    Q.load_data(r, rms, tight)
    Q.CleanUpCollectionROIs()
    assert Q.rois(False) == [[(45, 55)], [(47, 57)], [], [(80, 120)]]
    # V plane: the average criterion is not used (only above the high one)
    Q1 = RR.RoiRefinement(1, 1)
    Q1.load_data(_box(nt, 85, 115, 800)[None], np.full(1, 10.0, np.float32), [[]], [[(80, 120)]])
    Q1.CleanUpInductionROIs()
    assert Q1.rois(False) == [[]]
    Q0 = RR.RoiRefinement(0, 1)
    Q0.load_data(_box(nt, 85, 115, 800)[None], np.full(1, 10.0, np.float32), [[]], [[(80, 120)]])
    Q0.CleanUpInductionROIs()
    assert Q0.rois(False) == [[(80, 120)]]


def test_peak_search_known_answers():
    t = np.arange(200)
    one = (1000 * np.exp(-0.5 * ((t - 80.0) / 4) ** 2)).astype(np.float32)
    x, y = RR.find_peak(one, 200, 2.0, 0.1)
    assert len(x) == 1 and abs(x[0] - 80) < 0.5 and y[0] == float(one[int(round(x[0]))])
    two = one + (400 * np.exp(-0.5 * ((t - 140.0) / 4) ** 2)).astype(np.float32)
    x, y = RR.find_peak(two, 200, 2.0, 0.1)
    assert len(x) == 2
    assert abs(x[0] - 80) < 0.5 and abs(x[1] - 140) < 0.5     # the higher first
    assert RR.find_peak(np.ones(4, np.float32), 200, 2.0, 0.1) == ([], [])   # below 2 sigma + 1
    assert RR.find_peak(np.zeros(50, np.float32), 200, 2.0, 0.1) == ([], [])


def test_break_roi_splits_two_peaks():
    nt = 400
    # two strong peaks with a deep valley: BreakROI then BreakROI1 split
    # the loose ROI in two at the valley
    r = _plane_from([[(150, 5000, 4), (230, 5000, 4)]], nt)
    Q = RR.RoiRefinement(0, 1)
    Q.load_data(r, np.full(1, 20.0, np.float32), [[(140, 160), (220, 240)]], [[(120, 260)]])
    Q.BreakROIs()
    got = Q.rois(False)[0]
    assert len(got) >= 2
    assert min(s for s, e in got) >= 120 and max(e for s, e in got) <= 260
    assert any(s <= 150 <= e for s, e in got) and any(s <= 230 <= e for s, e in got)
    assert not any(s <= 150 and 230 <= e for s, e in got)


# ---- product against the oracle ------------------------------------------

def _decon_plane(seed, nw, nt, ntr, sigma=80.0):
    rng = np.random.default_rng(seed)
    r = rng.normal(0, sigma, (nw, nt))
    t = np.arange(nt)
    for _ in range(ntr):
        w0, w1 = sorted(rng.integers(0, nw, 2))
        t0 = rng.uniform(100, nt - 100)
        slope = rng.uniform(-20, 20)
        amp, wid = rng.uniform(1500, 6000), rng.uniform(2, 8)
        for w in range(w0, w1 + 1):
            c = t0 + slope * (w - w0)
            if 0 < c < nt:
                r[w] += amp * np.exp(-0.5 * ((t - c) / wid) ** 2) * rng.uniform(0.5, 1.5)
                if rng.random() < 0.3:
                    r[w] += 0.7 * amp * np.exp(-0.5 * ((t - c - rng.uniform(8, 30)) / wid) ** 2)
    return r.astype(np.float32)


STEPS = ([("cleanup", "CleanUpROIs"), ("merge", "generate_merge_ROIs")]
         + [("break", "BreakROIs"), ("check", "CheckROIs"), ("cleanup", "CleanUpROIs")] * 2
         + [("shrink", "ShrinkROIs"), ("check", "CheckROIs"), ("cleanup", "CleanUpROIs")])


def _compare(P, Q, nw, where):
    assert P.rois(extended=False) == Q.rois(extended=False), where
    for w in range(nw):
        a = P.contents(w)
        b = np.array([v for c in Q.contents(w) for v in c], np.float32)
        np.testing.assert_array_equal(a, b, err_msg=f"{where}, wire {w}")


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("plane", [0, 1, 2])
def test_refinement_product_matches_oracle(seed, plane):
    K = _hostlib()
    nw, nt = 40, 1200
    r = _decon_plane(seed, nw, nt, 6)
    bad = None
    if seed % 2:   # bad channels: zeroed in the ROI contents, inclusive ranges
        bad = [[] for _ in range(nw)]
        bad[7] = [(0, nt - 1)]
        bad[12] = [(300, 420), (800, 805)]
    rms = np.array([O.cal_rms(r[w]) for w in range(nw)], np.float32)
    thf = 3.0 if plane != 2 else 5.0
    tight = O.roi_lists((r > (thf * rms[:, None] + 1)).astype(np.uint8))
    loose = O.loose_rois(r, rms=rms, tight=tight, bad=bad) if plane != 2 else None
    wrapped = seed == 4
    P = K.RoiRefinement(plane, nw, nt, wrapped=wrapped)
    P.load(r, rms, tight, loose, bad=bad)
    Q = RR.RoiRefinement(plane, nw, isWrapped=wrapped)
    Q.load_data(r, rms, tight, loose, bad=bad)
    _compare(P, Q, nw, "load")
    fake = ("cleanup_collection", "CleanUpCollectionROIs") if plane == 2 else \
           ("cleanup_induction", "CleanUpInductionROIs")
    for ps, qs in STEPS + [fake, ("extend", "ExtendROIs")]:
        P.step(ps)
        getattr(Q, qs)()
        _compare(P, Q, nw, ps)
    assert P.rois(True) == Q.rois(True)
    assert sum(len(x) for x in Q.rois(True)) > 0


def test_refinement_all_step_equals_sequence():
    K = _hostlib()
    nw, nt = 32, 1000
    r = _decon_plane(11, nw, nt, 5)
    rms = np.array([O.cal_rms(r[w]) for w in range(nw)], np.float32)
    tight = O.roi_lists((r > (3 * rms[:, None] + 1)).astype(np.uint8))
    loose = O.loose_rois(r, rms=rms, tight=tight)
    P = K.RoiRefinement(0, nw, nt, params={"r_break_roi_loop": 3, "r_pad": 4})
    P.load(r, rms, tight, loose)
    P.step("all")
    Q = RR.RoiRefinement(0, nw, break_roi_loop=3, pad=4)
    Q.load_data(r, rms, tight, loose)
    Q.refine()
    assert P.rois(True) == Q.rois(True)


def test_peak_search_product_matches_oracle():
    K = _hostlib()
    rng = np.random.default_rng(1)
    for trial in range(120):
        n = int(rng.integers(3, 400))
        t = np.arange(n)
        sig = rng.normal(0, 30, n)
        for _ in range(int(rng.integers(0, 5))):
            sig += rng.uniform(100, 3000) * np.exp(-0.5 * ((t - rng.uniform(0, n)) / rng.uniform(1, 8)) ** 2)
        sig = sig.astype(np.float32)
        if trial % 7 == 0:
            sig = np.abs(sig)
        th = float(F(0.1))
        x1, y1 = K.find_peaks(sig, 200, 2.0, th)
        x2, y2 = RR.find_peak(sig, 200, 2.0, th)
        np.testing.assert_array_equal(x1, np.array(x2))
        np.testing.assert_array_equal(y1, np.array(y2))


def test_refinement_argument_errors():
    K = _hostlib()
    from wct_amd.capi import WctError
    with pytest.raises(WctError):
        K.RoiRefinement(3, 4, 10)
    P = K.RoiRefinement(0, 2, 10)
    with pytest.raises(WctError, match="outside the row"):
        P.load(np.zeros((2, 10), np.float32), np.ones(2), [[(5, 12)], []], [[], []])
    with pytest.raises(WctError, match="unknown step"):
        P.step("nope")
