"""Pins the oracle's electronics sampling and length table, and compares
the product's host-side response helpers (libWireCellHip: coldelec
samples, wire_region_average, fft_best_length) with the oracle.

  - electronics: ColdElecResponse::waveform_samples(Binning) samples
    Response::coldelec at bin centres tmin + (i + 1/2)(tmax - tmin)/n
    (util/inc/WireCellUtil/Binning.h:79), stores float and scales by the
    postgain in float (util/inc/WireCellUtil/Waveform.h:153-157,
    gen/src/ColdElecResponse.cxx:44-49).  The closed form is evaluated
    here term by term as util/src/Response.cxx:295-308 writes it (the
    oracle and the product factor it differently).
  - fft_best_length: every boundary of the three tables parsed from the
    reference's source (tests/golden/fft_best_length.json), both
    keep_odd_even branches, and product == oracle for every n <= 17000.
  - wire_region_average: hand-derived averages of small asymmetric path
    sets (util/src/Response.cxx:79-196), the half set of quirk A4, and
    product == oracle on random path sets with repeated positions.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle as O

US = 1000.0
HERE = os.path.dirname(os.path.abspath(__file__))


def coldelec_expanded(t, gain, shaping):
    """Response::coldelec written out as the reference's sum of terms."""
    if t <= 0 or t >= 10 * US:
        return 0.0
    x = t / shaping
    g = gain * 10 * 1.012
    e, c, s = math.exp, math.cos, math.sin
    return (4.31054 * e(-2.94809 * x) * g - 2.6202 * e(-2.82833 * x) * c(1.19361 * x) * g
            - 2.6202 * e(-2.82833 * x) * c(1.19361 * x) * c(2.38722 * x) * g
            + 0.464924 * e(-2.40318 * x) * c(2.5928 * x) * g
            + 0.464924 * e(-2.40318 * x) * c(2.5928 * x) * c(5.18561 * x) * g
            + 0.762456 * e(-2.82833 * x) * s(1.19361 * x) * g
            - 0.762456 * e(-2.82833 * x) * c(2.38722 * x) * s(1.19361 * x) * g
            + 0.762456 * e(-2.82833 * x) * c(1.19361 * x) * s(2.38722 * x) * g
            - 2.620200 * e(-2.82833 * x) * s(1.19361 * x) * s(2.38722 * x) * g
            - 0.327684 * e(-2.40318 * x) * s(2.5928 * x) * g
            + 0.327684 * e(-2.40318 * x) * c(5.18561 * x) * s(2.5928 * x) * g
            - 0.327684 * e(-2.40318 * x) * c(2.5928 * x) * s(5.18561 * x) * g
            + 0.464924 * e(-2.40318 * x) * s(2.5928 * x) * s(5.18561 * x) * g)


@pytest.mark.parametrize("n,tmin,tmax,postgain", [(1024, 0.0, 102.4 * US, 1.0),
                                                  (200, -3 * US, 17 * US, 1.2),
                                                  (37, 0.25 * US, 11 * US, 0.7)])
def test_elec_samples_bin_centres_float_postgain(n, tmin, tmax, postgain):
    gain, shaping = 14.0 * 1e-9 / (1e-15 / 1.602176487e-19), 2.0 * US
    of, od = O.elec_samples(n, tmin, tmax, gain, shaping, postgain)
    bs = (tmax - tmin) / n
    centres = [coldelec_expanded(tmin + (i + 0.5) * bs, gain, shaping) for i in range(n)]
    want_f = np.array([np.float32(np.float32(v) * np.float32(postgain)) for v in centres], np.float32)
    peak = max(abs(v) for v in centres)
    # float storage then float scaling (one rounding each): equal up to
    # the ulp the factored closed form may move the double value across
    assert np.max(np.abs(of - want_f)) <= 2 * np.spacing(np.float32(peak * postgain))
    np.testing.assert_allclose(od, np.array(centres) * postgain, rtol=0, atol=1e-12 * peak)
    # sampling at the bin edges instead would be visibly different
    edges = np.array([coldelec_expanded(tmin + i * bs, gain, shaping) for i in range(n)]) * postgain
    assert np.max(np.abs(od - edges)) > 1e-3 * peak


def test_elec_samples_support_and_peak():
    gain, shaping = 1.0, 2.0 * US
    n = 1000
    of, od = O.elec_samples(n, -10 * US, 20 * US, gain, shaping, 1.0)
    t = -10 * US + (np.arange(n) + 0.5) * 30.0
    assert np.all(od[(t <= 0) | (t >= 10 * US)] == 0)
    assert abs(od.max() - 1.0) < 0.01         # the 10*1.012 scaling puts the peak at the gain
    assert 1.5 * US < t[od.argmax()] < 2.5 * US


def test_product_coldelec_samples_match_oracle():
    from wct_amd import host
    gain, shaping = 14.0 * 1e-9 / (1e-15 / 1.602176487e-19), 2.0 * US
    for n, tmin, tmax, pg in [(1024, 0.0, 102.4 * US, 1.0), (300, -1 * US, 14 * US, 1.3)]:
        of, od = O.elec_samples(n, tmin, tmax, gain, shaping, pg)
        prod = host.coldelec_samples(n, tmin, tmax, gain, shaping, pg)
        peak = np.abs(od).max()
        np.testing.assert_allclose(prod, od, rtol=0, atol=1e-13 * peak)
        # the reference's float samples are the product's rounded to float
        # (to within the double rounding of the two factorisations)
        pf = np.array([np.float32(np.float32(v / pg) * np.float32(pg)) for v in prod])
        assert np.max(np.abs(pf - of)) <= 2 * np.spacing(np.float32(peak))


def _golden_fbl():
    with open(os.path.join(HERE, "golden", "fft_best_length.json")) as f:
        return json.load(f)


def test_fft_best_length_reference_tables():
    g = _golden_fbl()
    for n, keep, want in g["cases"]:
        assert O.fft_best_length(n, bool(keep)) == want, (n, keep)
    # named boundaries
    assert O.fft_best_length(1000) == 1024
    assert O.fft_best_length(16384) == 16384 and O.fft_best_length(16385) == 16385
    assert O.fft_best_length(1000, True) == 1024 and O.fft_best_length(1001, True) == 1001
    assert O.fft_best_length(16336, True) == 16384 and O.fft_best_length(16337, True) == 16337
    # every table entry is 2-3-5-7-11-13-smooth except three of the odd
    # table's: 187 = 11 17, 2299 = 11^2 19, 2431 = 11 13 17 (lengths the
    # device FFTs refuse)
    rough = []
    for tab in g["tables"].values():
        assert tab == sorted(tab)
        for e in tab:
            m = e
            for q in (2, 3, 5, 7, 11, 13):
                while m % q == 0:
                    m //= q
            if m != 1:
                rough.append(e)
    assert rough == [187, 2299, 2431]


def test_product_fft_best_length_matches_oracle():
    from wct_amd import host
    for keep in (False, True):
        for n in range(1, 17001):
            assert host.fft_best_length(n, keep) == O.fft_best_length(n, keep), (n, keep)


def _cur(rng, npaths, nsamp):
    return rng.standard_normal((npaths, nsamp)).astype(np.float32)


def test_wire_region_average_hand_derived():
    pitch = 5.0
    rng = np.random.default_rng(5)
    # paths at -p, 0, +p with different currents: the +p path repeats the
    # slot of -p's mirror (sizing the averages), both outer slots hold
    # the mean of the two, so the result is symmetric
    cur = _cur(rng, 3, 16)
    regs, avg = O.wire_region_average(np.array([-pitch, 0.0, pitch]), cur, pitch)
    assert list(regs) == [-1, 0, 1]
    m = ((cur[0].astype(np.float64) + cur[2]) / 2).astype(np.float32)
    np.testing.assert_array_equal(avg[0], m)
    np.testing.assert_array_equal(avg[1], cur[1])
    np.testing.assert_array_equal(avg[2], m)
    # Garfield-style sub-pitch positions 0, p/4, p/2 plus -p/4 (the repeat):
    # slots at 0, +-25, +-50 (1% of the pitch) split the one region 0
    # into widths 1/8, 1/4, 1/4, 1/4, 1/8
    cur = _cur(rng, 4, 16)
    pos = np.array([0.0, pitch / 4, pitch / 2, -pitch / 4])
    regs, avg = O.wire_region_average(pos, cur, pitch)
    assert list(regs) == [0]
    c = cur.astype(np.float64)
    q = ((c[1] + c[3]) / 2).astype(np.float32).astype(np.float64)
    want = 0.25 * c[2] + 0.5 * q + 0.25 * c[0]
    np.testing.assert_allclose(avg[0], want, rtol=2e-6, atol=2e-6)


def test_wire_region_average_half_set_is_empty():
    # quirk A4: unique half-set positions never size the averages
    pitch = 5.0
    rng = np.random.default_rng(6)
    pos = np.array([0.0, 1.0, 2.5, 5.0, 10.0])
    regs, avg = O.wire_region_average(pos, _cur(rng, 5, 8), pitch)
    assert len(regs) > 0 and avg.shape[1] == 8 and np.all(avg == 0)


def test_product_wire_region_average_matches_oracle():
    from wct_amd import host
    pitch = 4.71
    rng = np.random.default_rng(11)
    sets = [
        np.array([d * pitch for d in range(-10, 11)]),                    # full symmetric set
        np.array([d * pitch for d in range(-3, 6)]),                      # asymmetric
        np.array([0.0, 0.3, 0.6, 0.9, 1.2, 1.5, 1.8, 2.1, 2.4, -0.6]),    # sub-pitch, one repeat
        np.array([0.0, 1.0, 2.5, 4.71, 9.42]),                            # half set (A4)
        np.array([1.0, 1.0, 2.0, -2.0, 7.0, 7.0, -9.0]),                  # duplicated positions
    ]
    for k in range(6):                                                    # random sets with repeats
        p = np.round(rng.uniform(-3 * pitch, 3 * pitch, 9), 2)
        sets.append(np.concatenate([p, -p[:3]]))
    for pos in sets:
        cur = _cur(rng, len(pos), 24)
        r_o, a_o = O.wire_region_average(pos, cur, pitch)
        r_p, a_p = host.wire_region_average(pos, cur.astype(np.float64), pitch)
        assert list(r_o) == list(r_p), pos
        np.testing.assert_array_equal(a_p.astype(np.float32), a_o)
