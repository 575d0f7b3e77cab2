"""Pins the oracle's loose ROIs (ROI_formation::find_ROI_loose with
find_ROI_begin / find_ROI_end / local_ave and extend_ROI_loose,
sigproc/src/ROI_formation.cxx:538-937) and ROI_refinement::apply_roi
(sigproc/src/ROI_refinement.cxx:118-218) with cases worked by hand from
the reference's rules, then compares the product's host functions with
the oracle on seeded planes (exact equality).

Hand cases (rebin 1 unless noted; rms 40 so th = 40 * 1 * 3.5 = 140 and
the walk threshold th * 0.7 = 98; walls of 100 stop the downhill walks
because a step of 100 raises the 3-sample average by 33 >= 25):

  0..2 = 100, 5/6/7 = 200/400/200, 10..14 = 100, 17 = 300, 20..29 = 100
    group 5 opens an ROI: back past 98 to 4, downhill to 3 (the average
    at 2 is 66.7 >= 33.3 + 25), forward to 8 then 9 -> (3, 9); group 17:
    (15, 19), not merged (15 > 9).
  the same with group 12 = 130 (below th, a local maximum): walked back
    from 12 past its left neighbour's 100 to 9 and forward to 15; 130 -
    0 + 130 - 0 = 260 is not above 2 th = 280, but 130 - s(9) and 130 -
    s(15) exceed 98 within 3 groups -> (9, 15), which merges with (3, 9)
    and then with group 17's (15, 19) -> (3, 19).
"""
import numpy as np
import pytest

import oracle as O


def _row():
    s = np.zeros(30, np.float32)
    s[0:3] = 100
    s[5], s[6], s[7] = 200, 400, 200
    s[10:15] = 100
    s[17] = 300
    s[20:30] = 100
    return s


def test_hand_cases_rebin_1():
    s = _row()
    assert O.loose_rois(s[None, :], rms=[40.0], rebin=1) == [[(3, 9), (15, 19)]]
    s2 = s.copy()
    s2[12] = 130
    assert O.loose_rois(s2[None, :], rms=[40.0], rebin=1) == [[(3, 19)]]


def test_small_bump_is_not_an_roi():
    # a 30 bump on zeros, rms 40: 60 < 280 and 30 < 98
    s = np.zeros(30, np.float32)
    s[12] = 30
    assert O.loose_rois(s[None, :], rms=[40.0], rebin=1) == [[]]


def test_threshold_cap():
    s = _row()
    # rms 1e6 puts th far above every group, unless l_max_th caps it at
    # 150 (walk threshold 105: the same walls, the same ROIs)
    assert O.loose_rois(s[None, :], rms=[1e6], rebin=1) == [[]]
    assert O.loose_rois(s[None, :], rms=[1e6], rebin=1, l_max_th=150.0) == [[(3, 9), (15, 19)]]


def _rebinned_row(rebin=6):
    # each group's value on its first tick: the group sums are exact
    g = _row()
    s = np.zeros(30 * rebin, np.float32)
    s[::rebin] = g
    return s


def test_hand_case_rebin_6_and_extension():
    s = _rebinned_row()
    # rms 7: th = 7 * 6 * 3.5 = 147, walk threshold 102.9: as above in
    # groups, scaled back to ticks as (6 b, 6 e + 5)
    loose = O.loose_rois(s[None, :], rms=[7.0], rebin=6)
    assert loose == [[(18, 59), (90, 119)]]
    # extend_ROI_loose: starts/ends strictly inside a tight ROI move to
    # its edges; (10, 95) and (55, 119) then merge (55 < 95)
    assert O.loose_rois(s[None, :], rms=[7.0], rebin=6, tight=[[(10, 20), (55, 95)]]) == [[(10, 119)]]
    # an edge equal to a tight ROI's edge is not inside it
    assert O.loose_rois(s[None, :], rms=[7.0], rebin=6, tight=[[(18, 30)]]) == [[(18, 59), (90, 119)]]


def test_rms_of_row_without_bad_ticks():
    rng = np.random.default_rng(2)
    s = rng.normal(0, 100, 600).astype(np.float32)
    s[100:130] += 1500
    # a noisy bad range: left out of the RMS (th = 21 rms opens the
    # pulse's ROI only with the quiet part's RMS), 0 in the groups
    s[300:] = rng.normal(0, 1000, 300).astype(np.float32)
    bad = [[(300, 599)]]
    r = O.cal_rms(s[:300])
    a = O.loose_rois(s[None, :], bad=bad)
    assert a == O.loose_rois(s[None, :], bad=bad, rms=[r])
    assert a[0] and any(x <= 110 <= y for x, y in a[0])
    assert O.loose_rois(s[None, :], bad=bad, rms=[O.cal_rms(s)]) == [[]]


def test_apply_roi_hand_cases():
    r = np.array([[1.5, 2.5, 10.0, 4.0, 7.9, -3.7, 0.2, 5.0]], np.float32)
    # induction: r(i) - line through r(1) = 2.5 and r(4) = 7.9, truncated
    np.testing.assert_array_equal(O.apply_roi(r, [[(1, 4)]], 1), [[0, 0, 5, -2, 0, 0, 0, 0]])
    # collection: truncation only
    np.testing.assert_array_equal(O.apply_roi(r, [[(1, 4), (5, 5)]], 0), [[0, 2, 10, 4, 7, -3, 0, 0]])
    # a one-tick induction ROI: 0 * 0 / 0 is NaN, converted as x86-64 does
    np.testing.assert_array_equal(O.apply_roi(r, [[(6, 6)]], 1), [[0, 0, 0, 0, 0, 0, -2147483648.0, 0]])
    # overlapping ROIs: the later one wins on the overlap
    np.testing.assert_array_equal(O.apply_roi(r, [[(0, 3), (2, 5)]], 1), [[0, 0, 0, -1, 7, 0, 0, 0]])


# ---- the product's host functions against the oracle -------------------

def _plane(seed, nw=24, nt=1200):
    rng = np.random.default_rng(seed)
    # decon-like units (electrons): noise 300, pulses of 2e3-2e4
    r = rng.normal(0, 300, (nw, nt)).astype(np.float32)
    for _ in range(3 * nw):
        w, t = rng.integers(0, nw), rng.integers(0, nt - 40)
        r[w, t:t + rng.integers(5, 40)] += rng.uniform(2e3, 2e4)
    bad = [[] for _ in range(nw)]
    for w in rng.choice(nw, nw // 4, replace=False):
        a = int(rng.integers(0, nt - 100))
        bad[w] = [(a, a + int(rng.integers(0, 100)))]
    bad[0] = [(0, nt - 1)]    # a whole dead row
    return r, bad


@pytest.mark.parametrize("seed,rebin,jump", [(1, 6, 0), (2, 6, 1), (3, 4, 0), (4, 1, 1)])
def test_host_loose_rois_match_oracle(seed, rebin, jump):
    from wct_amd import components as K
    r, bad = _plane(seed)
    # tight ROIs: the tight rules on the plane plus random ranges, so
    # that loose edges fall inside them
    rng = np.random.default_rng(seed + 100)
    tight = O.roi_lists((r > 1500).astype(np.uint8))
    for w in range(len(tight)):
        extra = sorted((int(a), int(a) + int(rng.integers(5, 80))) for a in rng.integers(0, 1100, 4))
        tight[w] = sorted(tight[w] + extra)
    rms = np.abs(np.random.default_rng(seed).normal(300, 30, r.shape[0])).astype(np.float32)
    results = []
    for kw in (dict(), dict(rms=rms), dict(tight=tight), dict(rms=rms, tight=tight, bad=bad), dict(bad=bad)):
        kw2 = dict(kw)
        if "rms" in kw2:
            kw2["rms"] = kw2["rms"].astype(np.float64)
        want = O.loose_rois(r, rebin=rebin, l_jump_one_bin=jump, **kw2)
        got = K.loose_rois(r, rebin=rebin, l_jump_one_bin=jump, **kw)
        assert got == want, kw.keys()
        results.append(want)
    assert any(results[0])
    assert results[2] != results[0] and results[4] != results[0]   # the extension and the bad ranges act


def test_host_apply_roi_matches_oracle():
    from wct_amd import components as K
    rng = np.random.default_rng(7)
    r, _ = _plane(7)
    nw, nt = r.shape
    rois = []
    for w in range(nw):
        rows = []
        for _ in range(rng.integers(0, 5)):
            a = int(rng.integers(0, nt - 1))
            rows.append((a, min(nt - 1, a + int(rng.integers(0, 60)))))
        rois.append(rows)
    rois[3] = [(10, 10), (20, 20)]
    for linear in (0, 1):
        want = O.apply_roi(r, rois, linear)
        got = K.apply_roi(r, rois, linear)
        np.testing.assert_array_equal(got, want)
    from wct_amd.capi import WctError
    with pytest.raises(WctError, match="outside"):
        K.apply_roi(r, [[(5, nt)]], 1)
