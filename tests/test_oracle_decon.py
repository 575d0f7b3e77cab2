"""Pins the oracle's response build and deconvolution.

  - coldelec: peak equals the gain at the shaping-time scale, zero
    outside (0, 10 us) (util/src/Response.cxx:277-308)
  - filters: closed-form values (util/src/Response.cxx:435-444)
  - wire_region_average: symmetric path sets return the paths; a half
    set with unique positions gives empty averages (Response.cxx:98-119)
  - overall response: an impulse field response returns the scaled
    electronics response; the redigitisation picks fine sample 1 for
    coarse bin 0 and 5i-1 for bin i (OmnibusSigProc.cxx:834-859)
  - decon: the reference's step order (FP64) equals the fused form on a
    well-conditioned response; a point charge deconvolves to the
    separable filter kernel; linearity and shift equivariance.
"""
import numpy as np
import pytest

import oracle as O

US = 1000.0
MHZ = 1e-3


def test_coldelec_peak_and_support():
    gain, shaping = 2.0, 2.0 * US
    t = np.arange(1, 10000) * 1.0
    v = np.array([O.coldelec(x, gain, shaping) for x in t])
    assert abs(v.max() / gain - 1) < 0.01
    tpk = t[v.argmax()]
    assert 1.5 * US < tpk < 2.5 * US
    assert O.coldelec(0.0, gain, shaping) == 0.0
    assert O.coldelec(-5.0, gain, shaping) == 0.0
    assert O.coldelec(10 * US, gain, shaping) == 0.0
    assert O.coldelec(9.99 * US, gain, shaping) != 0.0


def test_filters_closed_form():
    n = 1000
    hf = O.hf_filter(n, 1 * MHZ, 0.1 * MHZ, 2, True)
    assert hf[0] == 0.0                     # flag zeroes DC
    # bin 50: f = 0.1 MHz = sigma -> exp(-1/2)
    assert abs(hf[50] - np.exp(-0.5)) < 1e-15
    assert np.allclose(hf[1:], hf[1:][::-1])    # folded, symmetric
    hf2 = O.hf_filter(n, 1 * MHZ, 0.1 * MHZ, 2, False)
    assert hf2[0] == 1.0
    lf = O.lf_filter(n, 1 * MHZ, 0.05 * MHZ)
    assert lf[0] == 0.0 and abs(lf[25] - (1 - np.exp(-1))) < 1e-15


def test_wire_region_average_paths():
    nsamp = 20
    pitch = 4.0
    pos = np.array([d * pitch for d in range(-3, 4)])
    cur = np.array([np.full(nsamp, 1.0 / (1 + d * d), np.float32) for d in range(-3, 4)])
    regs, avg = O.wire_region_average(pos, cur, pitch)
    assert list(regs) == list(range(-3, 4))
    assert np.allclose(avg, cur, rtol=1e-6)
    # half set, unique positions: the reference never sizes its averages
    half = pos[3:]
    regs, avg = O.wire_region_average(half, cur[3:], pitch)
    assert np.all(avg == 0)


def test_overall_response_impulse_and_redigitise():
    nfine, nin, nt = 100, 100, 20
    fine, coarse = 100.0, 500.0
    avg = np.zeros((1, nin))
    avg[0, 0] = 1.0
    elec = np.sin(np.arange(nfine) * 0.37) + 0.1 * np.arange(nfine)
    r = O.overall_response(avg, elec, fine, coarse, 1.5, 2.0, 0.0, nt, dtype=np.float64)
    # impulse convolution: fine response = -1.5*2.0*fine*elec; coarse bin 0
    # takes fine sample 1, bin i >= 1 takes fine sample 5i - 1 (quirk A3)
    fr = -1.5 * 2.0 * fine * elec
    want = np.array([fr[1]] + [fr[5 * i - 1] for i in range(1, nt)])
    assert np.allclose(r[0], want, rtol=1e-12, atol=1e-12 * np.abs(fr).max())
    r32 = O.overall_response(avg, elec, fine, coarse, 1.5, 2.0, 0.0, nt, dtype=np.float32)
    assert np.abs(r32 - r).max() < 1e-5 * np.abs(r).max()


def _rand_resp(nr, nt, rng):
    # a short, well-conditioned response: dominant central tap
    r = np.zeros((nr, nt))
    r[nr // 2, 0] = 1.0
    r += 0.05 * rng.standard_normal((nr, nt)) * (np.arange(nt) < 8)
    return r


def test_stepwise_fp64_equals_fused():
    rng = np.random.default_rng(11)
    nw, nt, nr = 24, 60, 5
    resp = _rand_resp(nr, nt, rng)
    d = rng.standard_normal((nw, nt))
    wf = O.hf_filter(nw, 1.0, 0.75, 2, False)
    tf = O.hf_filter(nt, 1 * MHZ, 0.3 * MHZ, 2, True)
    for ts in (0, 7, -3):
        fused = O.decon_fused(d, resp, wf, tf, nr // 2, ts)
        step = O.decon_stepwise(d, resp, wf, tf, nr // 2, ts, dtype=np.float64)
        assert np.abs(fused - step).max() < 1e-11 * np.abs(fused).max()
        step32 = O.decon_stepwise(d, resp, wf, tf, nr // 2, ts, dtype=np.float32)
        assert np.abs(fused - step32).max() < 1e-4 * np.abs(fused).max()


def test_point_charge_gives_filter_kernel():
    rng = np.random.default_rng(12)
    nw, nt, nr = 32, 64, 5
    resp = _rand_resp(nr, nt, rng)
    s, ts = nr // 2, 9
    q = np.zeros((nw, nt))
    w0, t0 = 10, 20
    q[w0, t0] = 1.0
    adc = O.forward_conv(q, resp, s, ts)
    wf = O.hf_filter(nw, 1.0, 0.5, 2, False)
    tf = O.hf_filter(nt, 1 * MHZ, 0.25 * MHZ, 2, True)
    y = O.decon_fused(adc, resp, wf, tf, s, ts)
    kw = np.real(np.fft.ifft(wf))
    kt = np.fft.irfft(tf[:nt // 2 + 1], n=nt)
    want = np.outer(np.roll(kw, w0), np.roll(kt, t0))
    assert np.abs(y - want).max() < 1e-12
    # unit filters return the charge itself
    y1 = O.decon_fused(adc, resp, np.ones(nw), np.ones(nt), s, ts)
    assert np.abs(y1 - q).max() < 1e-12


def test_linearity_and_shift():
    rng = np.random.default_rng(13)
    nw, nt, nr = 16, 40, 3
    resp = _rand_resp(nr, nt, rng)
    wf = O.hf_filter(nw, 1.0, 0.5, 2, False)
    tf = O.hf_filter(nt, 1 * MHZ, 0.3 * MHZ, 3, True)
    a, b = rng.standard_normal((2, nw, nt))
    ya = O.decon_fused(a, resp, wf, tf, 1, 4)
    yb = O.decon_fused(b, resp, wf, tf, 1, 4)
    yab = O.decon_fused(2 * a - b, resp, wf, tf, 1, 4)
    assert np.abs(yab - (2 * ya - yb)).max() < 1e-12
    ys = O.decon_fused(np.roll(a, (3, 5), axis=(0, 1)), resp, wf, tf, 1, 4)
    assert np.abs(ys - np.roll(ya, (3, 5), axis=(0, 1))).max() < 1e-12


def test_tensorset_pair_composes_to_the_fused_decon():
    """Decon2DResponse then Decon2DFilter (sig/src/Decon2DResponse.cxx,
    Decon2DFilter.cxx) is the fused decon with the filter product as the
    time filter; the response stage's spectrum is Hermitian (the FFT of a
    real frame) and equals the FFT of the unit-time-filter decon; the
    cmm ranges zero [tmin, tmax) of row ch and skip rows past the end."""
    rng = np.random.default_rng(14)
    nw, nt, nr = 20, 48, 5
    resp = _rand_resp(nr, nt, rng)
    d = rng.standard_normal((nw, nt)) + 3.0
    wf = O.hf_filter(nw, 1.0, 0.75, 2, False)
    tf = O.hf_filter(nt, 1 * MHZ, 0.3 * MHZ, 2, True) * O.lf_filter(nt, 1 * MHZ, 0.02 * MHZ)
    for ts in (0, 6):
        spec = O.decon2d_response(d, resp, wf, nr // 2, ts)
        assert spec.shape == (nw, nt) and np.iscomplexobj(spec)
        herm = np.conj(spec[:, (-np.arange(nt)) % nt])
        assert np.abs(spec - herm).max() < 1e-10 * np.abs(spec).max()
        y1 = O.decon_fused(d, resp, wf, np.ones(nt), nr // 2, ts)
        assert np.abs(spec - np.fft.fft(y1, axis=1)).max() < 1e-10 * np.abs(spec).max()
        y = O.decon2d_filter(spec, tf, restore=False)
        want = O.decon_fused(d, resp, wf, tf, nr // 2, ts)
        assert np.abs(y - want).max() < 1e-10 * np.abs(want).max()
        yb = O.decon2d_filter(spec, tf)
        wb, _ = O.restore_baseline(want, np.float64)
        assert np.abs(yb - wb).max() < 1e-10 * np.abs(wb).max()
        yc = O.decon2d_filter(spec, tf, cmm=[(3, 5, 17), (nw + 2, 0, 4), (0, 40, 48)])
        assert np.all(yc[3, 5:17] == 0) and np.all(yc[0, 40:48] == 0)
        keep = np.ones((nw, nt), bool)
        keep[3, 5:17] = keep[0, 40:48] = False
        assert np.array_equal(yc[keep], yb[keep])
