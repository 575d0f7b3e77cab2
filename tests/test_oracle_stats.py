"""Pins the oracle's order statistics, restore_baseline, cal_RMS and the
tight-ROI threshold / run-list logic.

Order statistics are checked against sorting (the value nth_element
returns is the rank-th element of the sorted set), including the float
rank arithmetic of Waveform::percentile (util/src/Waveform.cxx:85-106).
restore_baseline and cal_RMS are checked on rows whose answers follow
from their definitions (constant rows, symmetric noise, rows with large
outliers, all-zero rows).  ROI lists are checked on hand-built masks.
"""
import numpy as np
import pytest

import oracle as O


def _rank(p, n):
    # (size_t)(p * (float)n) in float arithmetic, clamped to n-1
    mid = int(np.float32(p) * np.float32(n))
    return min(mid, n - 1)


@pytest.mark.parametrize("n", [1, 2, 3, 10, 101, 6000, 6001])
def test_percentile_is_sorted_rank(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n).astype(np.float32)
    for p in (np.float32(0.5 - 0.34), 0.5, np.float32(0.5 + 0.34), 0.0, 1.0):
        got = O.percentile(x, float(p))
        assert got == np.sort(x)[_rank(p, n)]


def test_percentile_rank_uses_float_arithmetic():
    # 0.16f * 6000 rounds to 960.0 in float; 0.84f * 6000 to 5040.0
    assert _rank(np.float32(0.5 - 0.34), 6000) == 960
    assert _rank(np.float32(0.5 + 0.34), 6000) == 5040
    x = np.arange(6000, dtype=np.float32)
    assert O.percentile(x, float(np.float32(0.5 - 0.34))) == 960.0


def test_percentile_with_ties():
    x = np.array([3, 1, 3, 3, 2, 3, 0], np.float32)
    assert O.percentile(x, 0.5) == 3.0
    assert O.percentile(x, 0.0) == 0.0


def test_restore_baseline_constant_offset():
    # a row = c + small symmetric noise: the baseline returned is the
    # median, and zeros (dead samples) stay zero and are ignored
    rng = np.random.default_rng(1)
    row = (7.25 + rng.standard_normal(1001)).astype(np.float32)
    row[::10] = 0.0
    arr = row[None, :].copy()
    out, base = O.restore_baseline(arr)
    nz = row[row != 0]
    assert base[0] == np.sort(nz)[len(nz) // 2]
    assert np.all(out[0][row == 0] == 0)
    assert np.allclose(out[0][row != 0], row[row != 0] - base[0])


def test_restore_baseline_second_median_excludes_far_samples():
    # 60% of the samples near 0, 40% near +2000: the first median lands in
    # the low group, the second one only sees samples within 500 of it
    rng = np.random.default_rng(2)
    lo = (rng.standard_normal(600) * 3).astype(np.float32)
    hi = (2000 + rng.standard_normal(400)).astype(np.float32)
    row = np.concatenate([lo, hi])
    rng.shuffle(row)
    m1 = np.sort(row)[len(row) // 2]
    near = row[np.abs(row - m1) < 500]
    m2 = np.sort(near)[len(near) // 2]
    _, base = O.restore_baseline(row[None, :])
    assert base[0] == m2
    assert m2 != m1


def test_restore_baseline_zero_row_untouched():
    arr = np.zeros((2, 50), np.float32)
    arr[1, 3] = 5
    out, base = O.restore_baseline(arr)
    assert np.all(out[0] == 0) and base[0] == 0
    assert out[1, 3] == 0 and base[1] == 5


def test_cal_rms_gaussian_and_clip():
    rng = np.random.default_rng(3)
    x = rng.standard_normal(200000).astype(np.float32) * 2.0
    r = O.cal_rms(x)
    assert abs(r - 2.0) < 0.02
    # outliers beyond 5 sigma of the percentile estimate are excluded
    y = x.copy()
    y[:100] = 1e4
    assert abs(O.cal_rms(y) - r) < 0.01
    assert O.cal_rms(np.zeros(0, np.float32)) == 0.0
    assert O.cal_rms(np.zeros(10, np.float32)) == 0.0   # r0 = 0: nothing below 0


def test_roi_threshold_definition():
    rng = np.random.default_rng(4)
    a = rng.standard_normal((3, 500)).astype(np.float32)
    a[1, 100:110] = 50
    rms, th, mask = O.roi_threshold(a, 3.0)
    for w in range(3):
        assert rms[w] == O.cal_rms(a[w])
        assert th[w] == 3.0 * rms[w] + 1
        assert np.array_equal(mask[w], (a[w].astype(np.float64) > th[w]).astype(np.uint8))
    # bad ranges are excluded from the RMS and the mask (inclusive ends)
    rms2, th2, mask2 = O.roi_threshold(a, 3.0, bad=[(1, 100, 109)])
    keep = np.ones(500, bool)
    keep[100:110] = False
    assert rms2[1] == O.cal_rms(a[1][keep])
    assert np.all(mask2[1, 100:110] == 0)
    # a second ("tighter") frame ORs into the mask
    b = np.zeros_like(a)
    b[0, 7] = 100
    _, _, mask3 = O.roi_threshold(a, 3.0, r_tight=b)
    assert mask3[0, 7] == 1


def _mask(nt, runs_per_row):
    m = np.zeros((len(runs_per_row), nt), np.uint8)
    for w, runs in enumerate(runs_per_row):
        for a, b in runs:
            m[w, a:b + 1] = 1
    return m


def test_roi_lists_pad_and_merge():
    nt = 100
    m = _mask(nt, [[(10, 12), (20, 21)], [(0, 0), (50, 50)], [(99, 99)], [(95, 99)]])
    rois = O.roi_lists(m, pad=5, asy=0.1)
    # row 0: [5,17] and [15,26] overlap (15 < 17): merged to [5,26]
    assert rois[0] == [(5, 26)]
    # row 1: [0,5] and [45,55]
    assert (0, 5) in rois[1] and (45, 55) in rois[1]
    # row 2: a lone last sample never starts a run (the scan stops at nt-2)
    assert rois[2] == []
    # row 3: a run reaching the end is kept, clamped to nt-1
    assert rois[3] == [(90, 99)]


def test_roi_lists_touching_not_merged():
    # merge test is strict: a begin equal to the previous end is a new ROI
    nt = 100
    m = _mask(nt, [[(10, 10), (21, 21)]])
    rois = O.roi_lists(m, pad=5, asy=0.1)
    assert rois[0] == [(5, 15), (16, 26)]
    m = _mask(nt, [[(10, 10), (20, 20)]])
    assert O.roi_lists(m, pad=5, asy=0.1)[0] == [(5, 15), (15, 25)]


def test_roi_connect_fills_gap_wire():
    nt = 200
    # rows 0 and 2 carry similar ROIs, row 1 is empty: row 1 gets the mean
    m = _mask(nt, [[(50, 60)], [], [(52, 62)]])
    rois = O.roi_lists(m, pad=5, asy=0.1)
    assert rois[0] == [(45, 65)] and rois[2] == [(47, 67)]
    assert rois[1] == [(46, 66)]
    # lengths too different: no connection
    m = _mask(nt, [[(50, 60)], [], [(50, 90)]])
    assert O.roi_lists(m, pad=5, asy=0.1)[1] == []
    # row 1 already overlaps: nothing added
    m = _mask(nt, [[(50, 60)], [(55, 56)], [(52, 62)]])
    assert O.roi_lists(m, pad=5, asy=0.1)[1] == [(50, 61)]
