"""The ProtoDUNE noise filter (OmnibusNoiseFilter + Protodune::OneChannelNoise,
sigproc/src/OmnibusNoiseFilter.cxx, Protodune.cxx:753-918, Microboone.cxx
helpers, OmniChannelNoiseDB spectra): the oracle (oracle/noise_filter.py)
pinned on the CPU by hand-worked cases and closed forms.  The reference
ships no test vectors for these filters.
"""
import math

import numpy as np
import pytest

import noise_filter as NF

F = np.float32


def test_percentile_binned_hand_worked():
    # 0..9: bin width 0.9, 5 -> round(5.56) = 6, 9 -> 10 clamped to 9;
    # counts per bin 1 1 1 1 1 0 1 1 1 2; the running count first passes
    # 10 * 0.5 = 5 at bin 6: 6 * 0.9
    w = np.arange(10, dtype=F)
    assert NF.percentile_binned(w, 0.5) == F(F(6) * F(0.9))
    # p = 0: passes 0 at the first occupied bin
    assert NF.percentile_binned(w, 0.0) == 0
    # constant wave: zero bin width, every sample in bin 0
    assert NF.percentile_binned(np.full(7, 3.25, F), 0.5) == F(3.25)
    # n p rounds in float: 6000 * 0.16f is 960 in float (959.99998 in double)
    w = np.arange(6000, dtype=F)
    got = NF.percentile_binned(w, F(0.5 - 0.34))
    binsize = F(F(5999) / F(6000))
    # bins of k: round(k / 0.99983) = k except where the ratio crosses .5;
    # the 961st sample (960 counted before it) sits in bin 960
    assert got == F(F(960) * binsize)


def test_mean_rms_and_calc_rms():
    assert NF.mean_rms([1, 2, 3, 4]) == (2.5, math.sqrt(1.25))
    rng = np.random.default_rng(1)
    x = rng.normal(0, 5, 20000).astype(F)
    r = float(NF.calc_rms_with_flags(x))
    assert abs(r - 5) < 0.25
    # flagged samples (>= 4096) do not count
    y = np.concatenate([x, np.full(500, 10000, F), np.full(300, 25000, F)])
    assert float(NF.calc_rms_with_flags(y)) == pytest.approx(r, rel=0.02)
    assert NF.calc_rms_with_flags(np.full(10, 5000, F)) == 0


def test_signal_filter_pads_eight_each_side():
    rng = np.random.default_rng(2)
    x = rng.normal(0, 1, 200).astype(F)
    x[50] = 100
    y = NF.signal_filter(x.copy())
    assert (y[42:58] > 10000).all()
    assert (y[:42] < 4096).all() and (y[58:] < 4096).all()
    np.testing.assert_array_equal(y[42:58], (x[42:58].astype(np.float64) + 20000.0).astype(F))
    # already-flagged samples are left alone
    x2 = x.copy(); x2[45] = 10000
    assert NF.signal_filter(x2.copy())[45] == 10000


def test_raw_adaptive_baseline_cases():
    n = 300
    # constant: zero everywhere
    np.testing.assert_array_equal(NF.raw_adaptive_baseline(np.full(n, 7.0, F)), 0)
    # linear ramp: interior ~0, first sample minus the mean of samples 0..10
    a, b = F(0.5), F(3.0)
    x = (a * np.arange(n, dtype=F) + b).astype(F)
    y = NF.raw_adaptive_baseline(x.copy())
    assert np.abs(y[10:n - 10]).max() < 1e-4
    assert y[0] == pytest.approx(-5 * 0.5, abs=1e-5)
    # flagged samples stay; all-flagged input is returned unchanged
    x = np.full(n, 2.0, F); x[100] = 10000
    y = NF.raw_adaptive_baseline(x.copy())
    assert y[100] == 10000 and np.abs(np.delete(y, 100)).max() == 0
    np.testing.assert_array_equal(NF.raw_adaptive_baseline(np.full(30, 10000, F)), 10000)
    # a block of 5000s (>= 4096, not flags) between backgrounds 0 and 100:
    # windows inside the block see fewer than 10 valid samples and take
    # ((j - d) base[d] + (u - j) base[u]) / (u - d) from the nearest
    # filled windows d < j < u (the reference's weights, the nearer
    # background weighted by the far distance)
    x = np.zeros(n, F); x[180:] = 100; x[120:180] = 5000
    y = NF.raw_adaptive_baseline(x.copy())
    # window [j-10, j+10] holds 130 - j valid samples left of the block
    # (filled up to j = 120, mean 0) and j - 169 right of it (filled from
    # j = 179, mean of x[180..189] = 100); j = 140 walks down to 120 and
    # up to 179: base = ((140 - 120) * 0 + (179 - 140) * 100) / 59
    assert y[140] == F(5000 - 3900 / 59)
    assert y[120] == F(5000 - 0) and y[179] == F(5000 - 100)
    assert y[121] == F(5000 - (1 * 0 + 58 * 100) / 59)


def test_remove_filter_flags_and_noisy():
    x = np.array([5000, 20005, 100, 10000, 4096, -3], F)
    np.testing.assert_array_equal(NF.remove_filter_flags(x.copy()), [0, 5, 100, 0, 4096, -3])
    rng = np.random.default_rng(3)
    quiet = rng.normal(0, 0.1, 1000).astype(F)
    assert NF.noisy_filter(quiet, 0.5, 10.0) and (quiet == 10000).all()
    ok = rng.normal(0, 3, 1000).astype(F)
    assert not NF.noisy_filter(ok.copy(), 0.5, 10.0)
    loud = rng.normal(0, 40, 1000).astype(F)
    assert NF.noisy_filter(loud, 0.5, 10.0)


def test_partial_detection():
    spec = np.zeros(100, np.complex64)
    spec[1:7] = [9e4, 8e4, 7e4, 6e4, 5e4, 1]
    assert NF.is_partial(spec)
    spec[3] = 9.5e4                          # not falling
    assert not NF.is_partial(spec)
    spec[1:7] = [9e3, 8e3, 7e3, 6e3, 5e3, 1]   # falling but mean 7000 > 6000
    assert NF.is_partial(spec)
    spec[1:7] = [7e3, 6e3, 5e3, 4e3, 3e3, 1]   # mean 5000
    assert not NF.is_partial(spec)


def test_rcrc_spectrum_closed_form():
    # x[0] = 1 - t/tau, x[k] = -(t/tau) e^{-k t/tau}: the DFT is
    # 1 - (t/tau) (1 - e^{-N a}) / (1 - e^{-a} e^{-2 pi i f/N}), a = t/tau
    tick, tau, n = 500.0, 1.1e6, 1000
    got = NF.rcrc_spectrum(tau, tick, n).astype(np.complex128)
    a = tick / tau
    f = np.arange(n)
    want = 1 - a * (1 - np.exp(-n * a)) / (1 - np.exp(-a) * np.exp(-2j * np.pi * f / n))
    assert np.abs(got - want).max() < 1e-5
    sq = NF.rcrc_spectrum(tau, tick, n, layers=2).astype(np.complex128)
    assert np.abs(sq - want * want).max() < 1e-5


def test_freqmasks_and_flagged_bins():
    s = NF.freqmask_spectrum([(1.0, 0, 99), (0.0, 10, 12), (0.0, 40, 40), (0.0, 98, 200)], 100)
    assert NF.flagged_bins(s) == [(10, 12), (40, 40), (98, 99)]
    assert NF.flagged_bins(np.ones(10)) == []


def test_inv_c2r_mirrors_the_lower_half():
    rng = np.random.default_rng(4)
    for n in (16, 17):
        x = rng.normal(size=n).astype(F)
        np.testing.assert_allclose(NF.inv_c2r(NF.fwd_r2c(x)), x, atol=1e-5)
        s = NF.fwd_r2c(x)
        s2 = s.copy(); s2[n // 2 + 1:] = 123 + 45j   # upper half is replaced by the mirror
        np.testing.assert_allclose(NF.inv_c2r(s2), x, atol=1e-5)


def test_fft_scaling_resamples():
    # a cosine on bin k of 5996 samples comes back on bin k of 6000, scaled
    # by 5996/6000 (forward unnormalised, inverse by 1/6000, integer factor 1)
    k, n0, n1 = 37, 5996, 6000
    x = np.cos(2 * np.pi * k * np.arange(n0) / n0).astype(F)
    y = NF.fft_scaling(x, n1)
    want = np.cos(2 * np.pi * k * np.arange(n1) / n1) * n0 / n1
    assert np.abs(y - want).max() < 1e-4


def test_one_channel_baseline_and_masks():
    rng = np.random.default_rng(5)
    n = 2000
    x = (900 + rng.normal(0, 3, n)).astype(F)
    y, m = NF.one_channel_noise(x, 0, min_rms=1.0, max_rms=30.0)
    assert m == {} and abs(float(np.median(y))) < 0.5 and float(np.std(y)) == pytest.approx(3, rel=0.1)
    quiet = (900 + rng.normal(0, 0.2, n)).astype(F)
    y, m = NF.one_channel_noise(quiet, 1, min_rms=1.0, max_rms=30.0)
    assert m == {"noisy": True} and not y.any()


def test_one_channel_rc_undershoot_removed():
    rng = np.random.default_rng(6)
    n, tick, tau = 2000, 500.0, 1.1e6
    rc = NF.rcrc_spectrum(tau, tick, n)
    q = np.zeros(n); q[[300, 900, 1500]] = [800.0, 500.0, 1000.0]   # peaks up to 1000 ADC
    q = np.convolve(q, np.exp(-0.5 * (np.arange(-20, 21) / 4.0) ** 2), "same")
    x = np.fft.ifft(np.fft.fft(q) * rc.astype(np.complex128)).real + 700 + rng.normal(0, 1, n)
    y, m = NF.one_channel_noise(x.astype(F), 2, rcrc=rc, min_rms=0.5, max_rms=30.0)
    assert m == {}
    # the undershoot after the last pulse is gone and the baseline is at
    # zero: what is left is the noise
    rms = lambda v: float(np.sqrt(np.mean(np.square(v))))
    assert rms(y - q) < 1.1
    assert abs(float(np.mean((y - q)[1530:1990]))) < 0.3
    assert float(np.mean((x - 700 - q)[1530:1990])) < -1.0


def test_one_channel_harmonics_removed_on_collection():
    rng = np.random.default_rng(7)
    n = 2000
    t = np.arange(n)
    bins = [300, 340, 380, 420, 460, 500]
    x = 500 + rng.normal(0, 3, n)
    for b in bins:
        x = x + 6 * np.cos(2 * np.pi * b * t / n)
    noise = NF.freqmask_spectrum([(1.0, 0, n - 1)] + [(0.0, b - 5, b + 5) for b in bins], n)
    y, m = NF.one_channel_noise(x.astype(F), 2, noise=noise, min_rms=1.0, max_rms=30.0)
    assert m == {"harmonic": True}
    spec = np.abs(np.fft.fft(y))
    assert spec[bins].max() < 1e-2 * np.abs(np.fft.fft(x - 500))[bins].min()
    # the same on an induction plane leaves the lines alone
    y2, m2 = NF.one_channel_noise(x.astype(F), 0, noise=noise, min_rms=1.0, max_rms=30.0)
    assert m2 == {} and np.abs(np.fft.fft(y2))[bins].min() > 1000


def test_one_channel_partial_waveform():
    # a large slow decay: its spectrum falls over bins 1..5 with a mean
    # above 6000, so the channel is partial: no RC correction, lf_noisy on
    # induction planes, the adaptive baseline instead
    rng = np.random.default_rng(8)
    n = 2000
    x = (3000 * np.exp(-np.arange(n) / 300.0) + rng.normal(0, 3, n)).astype(F)
    assert NF.is_partial(NF.fwd_r2c(x))
    y, m = NF.one_channel_noise(x, 0, rcrc=NF.rcrc_spectrum(1.1e6, 500.0, n), min_rms=1.0, max_rms=30.0)
    assert m == {"lf_noisy": True}
    assert abs(float(np.median(y[200:]))) < 3
    y, m = NF.one_channel_noise(x, 2, min_rms=1.0, max_rms=30.0)
    assert "lf_noisy" not in m


def _chirp_signal(rng, n, lo, hi, sigma=3.0):
    x = rng.normal(0, sigma, n).astype(F)
    x[lo:hi] = rng.normal(0, 0.1, hi - lo)      # a dead (chirped) stretch
    return x


def test_chirp_hand_worked():
    rng = np.random.default_rng(21)
    n = 2000
    # low-RMS windows 20..49 (samples 400..999): the run is first seen at
    # window 21 (first = 400, last = 420) and last extended at window 49
    # (last = 980); 30 low windows, 2 next to normal ones (0.067 < 0.2);
    # widened to first - 20 = 380 and last + 40 = 1020, and the 1-based
    # first bin turned into beg = 379
    ok, beg, end = NF.chirp(_chirp_signal(rng, n, 400, 1000))
    assert ok and (beg, end) == (379, 1020)
    # a dead channel (every window low): the whole readout
    ok, beg, end = NF.chirp(rng.normal(0, 0.1, n).astype(F))
    assert ok and (beg, end) == (0, n)
    # too few low windows (4): no chirp
    ok, _, _ = NF.chirp(_chirp_signal(rng, n, 400, 480))
    assert not ok
    # ordinary noise
    assert not NF.chirp(rng.normal(0, 3, n).astype(F))[0]


def test_mb_one_channel_chirp_and_masks():
    rng = np.random.default_rng(22)
    n = 2000
    x = (2048 + _chirp_signal(rng, n, 400, 1000)).astype(F)
    y, m = NF.mb_one_channel_noise(x, 0, baseline=2048, min_rms=1.0, max_rms=30.0)
    assert m["chirp"] == [(379, 1020)] and m["lf_noisy"] == [(379, 1020)] and "noisy" not in m
    assert not y[379:1020].any()                                 # raised to the flag, then removed
    assert abs(float(np.median(y[:300]))) < 0.5 and abs(float(np.median(y[1100:]))) < 0.5
    # collection: chirp but no lf_noisy
    y, m = NF.mb_one_channel_noise(x, 2, baseline=2048, min_rms=1.0, max_rms=30.0)
    assert "chirp" in m and "lf_noisy" not in m
    # a notch from the noise spectrum removes a line; the config spectrum scales
    t = np.arange(n)
    x = (400 + rng.normal(0, 3, n) + 20 * np.cos(2 * np.pi * 300 * t / n)).astype(F)
    notch = NF.freqmask_spectrum([(1.0, 0, n - 1), (0.0, 295, 305), (0.0, n - 305, n - 295)], n)
    y, m = NF.mb_one_channel_noise(x, 2, baseline=400, noise=notch, min_rms=1.0, max_rms=30.0)
    assert m == {} and np.abs(np.fft.fft(y))[300] < 1.0
    half = np.full(n, 0.5, np.complex64)
    y2, _ = NF.mb_one_channel_noise(x, 2, baseline=400, config=half, noise=notch, min_rms=0.5, max_rms=30.0)
    assert float(np.std(y2)) == pytest.approx(float(np.std(y)) / 2, rel=0.05)
    # noisy clears lf_noisy
    q = (900 + rng.normal(0, 0.05, n)).astype(F)
    y, m = NF.mb_one_channel_noise(q, 0, baseline=900, min_rms=1.0, max_rms=30.0)
    assert "noisy" in m and "lf_noisy" not in m and not y.any()


def test_rel_gain_calib():
    x = np.array([1.0, -2.0, 3.5], F)
    np.testing.assert_array_equal(NF.rel_gain_calib(x, 1.1), (F(1.1) * x).astype(F))
    np.testing.assert_array_equal(NF.rel_gain_calib(x, 1.3), x)      # above the 1.25 cut: unchanged
    np.testing.assert_array_equal(NF.rel_gain_calib(x, 0.7), x)      # below 0.8


def test_microboone_spectral_filters():
    assert NF.filter_time(0.0) == 0.0
    assert NF.filter_time(0.143555) == pytest.approx(math.exp(-0.5))
    assert NF.filter_low(0.178) == 0.0 and NF.filter_low(0.107) == 0.0 and NF.filter_low(0.2145) == 0.0
    assert NF.filter_low(0.08) == pytest.approx(1 - math.exp(-1))
    assert NF.filter_low(0.05, 0.05) == pytest.approx(1 - math.exp(-1))
    assert NF.filter_low_loose(0.005) == pytest.approx(1 - math.exp(-1))


def test_calc_rms_and_median_hand_worked():
    # CalcRMS drops samples beyond 4.5 RMS: 20 zeros and 100 give mean 4.76,
    # RMS 21.3; 100 is 4.47 RMS out, so kept; with 40 zeros (RMS 15.4) it goes
    x = np.zeros(41, F); x[0] = 100
    assert NF.calc_rms(x) == (0.0, 0.0)
    y = np.zeros(21, F); y[0] = 100
    assert NF.calc_rms(y) == NF.mean_rms(y)
    # per tick: [1, 2, 3] -> bins of 2/3 at 0, 2 (1.5 rounds up), 2
    # (clamped); counts 1 0 2 pass int(3 * 0.5) = 1 at bin 2: 1 + 2 * 2/3
    sig = [np.array([1.0, 5], F), np.array([2.0, 5], F), np.array([3.0, 0.0005], F)]
    med = NF.calc_median(sig)
    assert med[0] == F(F(1) + F(2) * F(F(2) / F(3)))
    assert med[1] == 5     # 0.0005 is below 0.001 and left out; 5 is within 5 RMS
    # samples beyond 5 mean-RMS are left out: mean RMS of the channels
    # below is 0.5 (two of 0..1 alternating), so 3 is out
    a = np.tile(np.array([0.0, 1.0], F), 50)
    c = a.copy(); c[10] = 3
    assert NF.calc_median([a, a, c])[10] == F(0)


def test_signal_protection_without_response():
    n = 200
    med = np.zeros(n, F); med[100] = 80; med[[99, 101]] = 2
    rois = NF.signal_protection(med, None, 0, pad_f=3, pad_b=4, upper_decon_limit=0.02, decon_lf_cutoff=0.08,
                                upper_adc_limit=15, protection_factor=5, min_adc_limit=50)
    # |80 - mean| > 15: ticks 97..104 protected and replaced by the line
    # between ticks 96 and 105 (both 0)
    assert rois == [list(range(97, 105))]
    assert not med.any()


def test_subtract_wscaling_removes_scaled_common_mode():
    rng = np.random.default_rng(41)
    n = 1500
    common = (10 * np.sin(np.arange(n) / 30.0)).astype(F)
    scales = [0.8, 1.0, 1.2, 1.0]
    sig = [(s * common + rng.normal(0, 0.5, n)).astype(F) for s in scales]
    med = NF.calc_median(sig)
    out = NF.subtract_wscaling([x.copy() for x in sig], med, None, 0, [], 0.08, 0.8)
    for s, o in zip(scales, out):
        assert float(np.std(o)) < 1.0      # the common mode is gone, scaled per channel
    # the scaling follows each channel's share of the common mode
    x = NF.subtract_wscaling([sig[0].copy()] + [x.copy() for x in sig[1:]], med, None, 0, [], 0.08, 0.8)
    assert abs(float(np.mean(x[0] * common)) / float(np.mean(common * common))) < 0.05


def test_coherent_noise_sub_protects_signal():
    rng = np.random.default_rng(42)
    n, nch = 1500, 16
    common = (8 * np.sin(np.arange(n) / 25.0)).astype(F)
    sig = [(common + rng.normal(0, 1.0, n)).astype(F) for _ in range(nch)]
    pulse = (300 * np.exp(-0.5 * ((np.arange(n) - 700) / 5.0) ** 2)).astype(F)
    for c in range(nch):     # a track crossing every channel at the same time
        sig[c] = (sig[c] + pulse).astype(F)
    out = NF.coherent_noise_sub(sig, adc_limit=15, protection_factor=5, min_adc_limit=50, pad_f=10, pad_b=10)
    for o in out:
        assert float(np.std(o[:600])) < 1.5               # common mode removed away from the pulse
        assert float(o[700]) > 250                         # the pulse itself survives (protected)


# ---- Microboone::OneChannelStatus (Microboone.cxx:1263-1334) ------------

def _sine(n, k0, amp, phase=0.3):
    return (amp * np.sin(2 * np.pi * k0 * np.arange(n) / n + phase)).astype(F)


def test_lf_noisy_sine_closed_form():
    """A sine at bin k0 of amplitude A has |X(k0)| = A n / 2 and no other
    bin in 1..n/2: the ratio is A / 2 when 1 <= k0 <= nbins (every sample
    non-zero at this phase, none beyond 3.5 rms of a sine), 0 when k0 is
    past nbins."""
    n = 1000
    for k0, inside in ((1, True), (250, True), (251, False), (400, False)):
        s = _sine(n, k0, 40.0)
        assert np.count_nonzero(s) == n
        flag, ratio = NF.lf_noisy(s)
        if inside:
            assert abs(ratio - 20.0) < 2e-3, (k0, ratio)
            assert flag
        else:
            assert ratio < 1e-3, (k0, ratio)
            assert not flag
    flag, ratio = NF.lf_noisy(_sine(n, 5, 20.0))        # ratio 10 below the cut of 14
    assert not flag and abs(ratio - 10.0) < 1e-3
    flag, ratio = NF.lf_noisy(_sine(n, 5, 20.0), cut=9.5)
    assert flag


def test_lf_noisy_dc_and_empty():
    # only bin 0 (never summed): a constant row, whatever its level
    flag, ratio = NF.lf_noisy(np.full(1000, 500.0, F))
    assert not flag and ratio < 1e-6
    # no non-zero sample: not noisy (:1299-1301)
    assert NF.lf_noisy(np.zeros(1000, F)) == (False, 0.0)


def test_lf_noisy_mask_window():
    """A spike at i sets samples [i - 5, i + 5) to the median, clipped at
    the row's ends; marks come from the original samples."""
    n = 2000
    for at, lo, hi in ((100, 95, 105), (2, 0, 7), (n - 3, n - 8, n)):
        s = _sine(n, 3, 10.0)
        s[at] = 1000.0
        masked, med, rms, valid = NF.lf_noisy_masked(s, 3.5, 5)
        assert valid == n
        assert np.flatnonzero(np.abs(s.astype(np.float64) - med) > 3.5 * rms).tolist() == [at]
        assert np.flatnonzero(masked != s).tolist() == [i for i in range(lo, hi) if s[i] != F(med)]
        assert np.all(masked[lo:hi] == F(med))
    # window 0: nothing changes
    masked, *_ = NF.lf_noisy_masked(s, 3.5, 0)
    assert np.array_equal(masked, s)


def test_one_channel_status_planes():
    s = _sine(1000, 3, 40.0)
    assert NF.one_channel_status(s, plane=0) == [(0, 1000)]
    assert NF.one_channel_status(s, plane=1) == [(0, 1000)]
    assert NF.one_channel_status(s, plane=2) == []      # collection is not tested
    assert NF.one_channel_status(_sine(1000, 3, 10.0), plane=0) == []


# ---- Microboone::ADCBitShift (Microboone.cxx:1106-1216) -----------------

def _rotl12(x, n):
    """The readout fault the filter undoes: the 12 ADC bits rotated left
    by n (the top n bits land in the low n)."""
    x = np.asarray(x, dtype=np.int64)
    return (((x << n) | (x >> (12 - n))) & 0xFFF).astype(np.int64)


def test_bits_hand_worked():
    # Bits::shift_right(0b101101, 2, 0b11, 12) = 0b1011 | 0b11 << 10
    assert NF.bits_shift_right(0b101101, 2, 0b11, 12) == 11 + 3072
    assert NF.bits_lowest(0b101101, 3) == 0b101
    # shift_right with the value's own low bits undoes a left rotation
    for x in (0, 1, 1500, 2048, 4095):
        for n in (1, 2, 5, 10):
            v = int(_rotl12(x, n))
            assert NF.bits_shift_right(v, n, NF.bits_lowest(v, n), 12) == x


def test_adc_bit_shift_undoes_rotation():
    """A baseline of 1500 +- 20 rotated by n = 1..3: its low n bits (the
    top bits of 1500, all zero) never flip in the first 500 ticks, bit n
    (bit 0 of the sample) flips about half the time, so nshift = n and
    every sample is rotated back exactly (nothing is far from the mean)."""
    rng = np.random.default_rng(1106)
    x = np.rint(1500 + rng.normal(0, 20, 3000)).astype(np.int64)
    for n in (1, 2, 3):
        out, ns = NF.adc_bit_shift(_rotl12(x, n).astype(F))
        assert ns == n
        assert np.array_equal(out, x.astype(F))
    out, ns = NF.adc_bit_shift(x.astype(F))            # no fault: unchanged
    assert ns == 0 and np.array_equal(out, x.astype(F))


def test_adc_bit_shift_restores_a_glitch():
    """One sample with bit 11 set (1500 + 2048) rotated by 1 lands, rotated
    back with its own filling, 2048 away from the mean and its neighbours
    (beyond 2^11 x 0.8 = 1638): the other filling (0) is nearer the
    neighbours' extrapolation, so it becomes 1500 again."""
    x = np.full(2000, 1500, np.int64)
    x[1::2] += 1          # bit 0 of the samples flips every tick
    x[700] = 1500 + 2048
    out, ns = NF.adc_bit_shift(_rotl12(x, 1).astype(F))
    assert ns == 1
    want = x.copy()
    want[700] = 1500
    assert np.array_equal(out, want.astype(F))
