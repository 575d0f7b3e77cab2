"""Pins the oracle's channel-by-channel electronics correction
(sigproc/src/OmnibusSigProc.cxx:949-982) by what the mathematics fixes,
so that a swapped ratio, a wrong sign of the bin phase, a wrong row or a
missed zero rule fails one of them:

  - a channel response equal to the nominal electronics response leaves
    the decon unchanged (factor 1);
  - a channel response that is the nominal one circularly delayed by d
    ticks advances that channel's data by d ticks (factor exp(+2 pi i d
    k / nt)): the corrected decon equals the plain decon of the rolled
    data;
  - twice the nominal response halves the channel's data;
  - an all-zero channel response zeroes the channel (|ch_elec| == 0);
  - samples past nt are cut (tch_resp.resize(nt), :966).
"""
import numpy as np

import oracle as O


def _setup(nw=12, nt=60, nr=5, seed=3):
    rng = np.random.default_rng(seed)
    data = rng.normal(size=(nw, nt))
    t = np.arange(nt)
    resp = np.zeros((nr, nt))
    for r in range(nr):
        resp[r] = np.exp(-0.5 * ((t - 8.0) / (1.5 + r)) ** 2) / (1 + r * r)
    wf = np.linspace(1.0, 0.5, nw)
    ewave = np.exp(-t / 6.0) * (t / 6.0) ** 2
    return data, resp, wf, ewave


def _init(data, resp, wf, ch=None, ewave=None):
    return O.decon_init(data, resp, wf, 2, 3, np.float64, ch, ewave)


def test_identity_response_is_no_correction():
    data, resp, wf, ewave = _setup()
    nw, nt = data.shape
    ch = np.tile(ewave, (nw, 1))
    a = _init(data, resp, wf, ch, ewave)
    b = _init(data, resp, wf)
    assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b))


def test_delayed_response_advances_the_channel():
    data, resp, wf, ewave = _setup()
    nw, nt = data.shape
    delays = (np.arange(nw) * 7) % nt
    ch = np.stack([np.roll(ewave, d) for d in delays])
    a = _init(data, resp, wf, ch, ewave)
    rolled = np.stack([np.roll(data[w], -d) for w, d in enumerate(delays)])
    b = _init(rolled, resp, wf)
    assert np.max(np.abs(a - b)) <= 1e-11 * np.max(np.abs(b))
    # the opposite direction differs (pins the sign of the phase)
    wrong = np.stack([np.roll(data[w], d) for w, d in enumerate(delays)])
    assert np.max(np.abs(a - _init(wrong, resp, wf))) > 1e-3 * np.max(np.abs(b))


def test_scaled_response_and_zero_response():
    data, resp, wf, ewave = _setup()
    nw, nt = data.shape
    ch = np.tile(2.0 * ewave, (nw, 1))
    ch[4] = 0.0
    a = _init(data, resp, wf, ch, ewave)
    d2 = data / 2.0
    d2[4] = 0.0
    b = _init(d2, resp, wf)
    assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b))


def test_samples_past_nt_are_cut():
    data, resp, wf, ewave = _setup()
    nw, nt = data.shape
    rng = np.random.default_rng(9)
    ch = np.tile(ewave, (nw, 1)) * (1.0 + 0.1 * rng.random((nw, 1)))
    longer = np.concatenate([ch, rng.normal(size=(nw, 17))], axis=1)
    a = _init(data, resp, wf, longer, ewave)
    b = _init(data, resp, wf, ch, ewave)
    assert np.array_equal(a, b)
    # and shorter responses are zero-padded
    short = ch[:, : nt // 2]
    c = _init(data, resp, wf, short, ewave)
    padded = np.concatenate([short, np.zeros((nw, nt - nt // 2))], axis=1)
    assert np.array_equal(c, _init(data, resp, wf, padded, ewave))
