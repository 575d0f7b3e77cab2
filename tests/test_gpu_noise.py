"""The per-channel noise filter on the device (wct_nf_one_channel) against
the oracle (oracle/noise_filter.py), which is pinned on the CPU
(tests/test_noise_filter.py).

The outputs are float32 and go through FP32 transforms on both sides (the
oracle's in FP64, rounded), so samples agree to about 1e-4 ADC.  The
binned median and percentiles return bin edges; a transform rounding
that moves one sample across a bin edge can move the crossing bin by
one, which shifts a channel's whole baseline by one bin width (the
sample range over the sample count, 0.01-0.5 ADC here); on noise-only
channels, whose bins are narrow, that happens to about a third of the
channels.  The test therefore requires, per channel, the difference minus
its median within 1e-2 ADC, the median within one bin width (estimated
from the output's range, with a margin of half a bin since the median's
own input has its outliers clipped), and identical flags.
"""
import numpy as np
import pytest

import noise_filter as NF

pytestmark = pytest.mark.gpu

TICK = 500.0        # ns
TAU = 1.1e6         # 1.1 ms


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def make_frame(rng, nt):
    """A mixed set of channels: plain noise on each plane, RC-shaped
    pulses, partial (slow-decay) waveforms, a quiet (noisy-tagged)
    channel, a bad (all-zero) channel, collection channels with harmonic
    lines in flagged runs."""
    rc = NF.rcrc_spectrum(TAU, TICK, nt).astype(np.complex128)
    t = np.arange(nt)
    rows, chans = [], []
    lines = [300, 340, 380, 420, 460, 500]
    noise_mask = NF.freqmask_spectrum([(1.0, 0, nt - 1)] + [(0.0, b - 5, b + 5) for b in lines], nt)
    runs = NF.flagged_bins(noise_mask)
    for i in range(40):
        plane = i % 3
        base = 900.0 if plane < 2 else 400.0
        x = base + rng.normal(0, 2.5 + 0.5 * plane, nt)
        kind = i % 8
        if kind in (1, 2):      # pulses through the RC
            q = np.zeros(nt)
            for p in rng.integers(100, nt - 100, size=3):
                q[p] = rng.uniform(200, 900)
            q = np.convolve(q, np.exp(-0.5 * (np.arange(-20, 21) / 4.0) ** 2), "same")
            x = x + np.fft.ifft(np.fft.fft(q) * rc).real
        if kind == 3:           # partial waveform
            x = x + 3000 * np.exp(-t / 300.0)
        if kind == 4 and i > 8:  # quiet: noisy-tagged
            x = base + rng.normal(0, 0.2, nt)
        if kind == 5:           # bad channel, zeroed by the caller
            x = np.zeros(nt)
        if plane == 2 and kind in (0, 6):
            for b in lines:
                x = x + 6 * np.cos(2 * np.pi * b * t / nt + rng.uniform(0, 6.28))
        rows.append(x.astype(np.float32))
        chans.append(dict(plane=plane, rcrc=0, bins=runs if plane == 2 else [], min_rms=1.0, max_rms=30.0,
                          noise=noise_mask if plane == 2 else None))
    return np.stack(rows), chans, rc.astype(np.complex64)


def test_rcrc_spectrum_matches_oracle():
    _torch()
    from wct_amd import noise as N
    for nt, layers in ((2000, 1), (6000, 1), (6000, 2), (1999, 1)):
        got = N.rcrc_spectrum(TAU, TICK, nt, layers).astype(np.complex128)
        want = NF.rcrc_spectrum(TAU, TICK, nt, layers).astype(np.complex128)
        assert np.abs(got - want).max() <= 2e-6 * np.abs(want).max(), (nt, layers)


@pytest.mark.parametrize("nt", [2000, 1998])
def test_one_channel_matches_oracle(nt):
    torch = _torch()
    from wct_amd import noise as N
    rng = np.random.default_rng(nt)
    frame, chans, rc = make_frame(rng, nt)
    print(f"noise filter {len(chans)} channels x {nt}", flush=True)
    out, flags = N.one_channel(frame, chans, rcrc=rc[None, :])
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    for i, c in enumerate(chans):
        want, masks = NF.one_channel_noise(frame[i], c["plane"], rcrc=rc, noise=c["noise"],
                                           min_rms=c["min_rms"], max_rms=c["max_rms"])
        assert sorted(N.mask_names(int(flags[i][0]), c["plane"])) == sorted(masks), (i, flags[i], masks)
        d = got[i].astype(np.float64) - want
        med = float(np.median(d))
        width = (float(want.max()) - float(want.min())) / nt + 1e-6
        assert np.abs(d - med).max() <= 1e-2 and abs(med) <= 1.5 * width, (i, np.abs(d - med).max(), med, width)
    # the designed cases came out as designed
    kinds = [i % 8 for i in range(len(chans))]
    assert all(flags[i][0] & 1 for i in range(len(chans)) if kinds[i] == 3)
    assert all(flags[i][0] & 2 for i in range(len(chans)) if kinds[i] == 5)
    assert any(flags[i][0] & 4 for i in range(len(chans)))


def test_one_channel_full_length_properties():
    """2560 channels x 6000 ticks (one APA): every output finite, bad
    channels zero and noisy, plain noise channels at zero baseline."""
    torch = _torch()
    from wct_amd import noise as N
    rng = np.random.default_rng(3)
    nt, nch = 6000, 2560
    frame = (900 + rng.normal(0, 3, (nch, nt))).astype(np.float32)
    frame[17] = 0
    chans = [dict(plane=0 if c < 800 else (1 if c < 1600 else 2), rcrc=0, min_rms=1.0, max_rms=30.0)
             for c in range(nch)]
    rc = N.rcrc_spectrum(TAU, TICK, nt)
    out, flags = N.one_channel(frame, chans, rcrc=rc[None, :])
    o = out.cpu().numpy()
    assert np.isfinite(o).all()
    assert not o[17].any() and flags[17][0] == 2
    med = np.median(np.delete(o, 17, axis=0), axis=1)
    assert np.abs(med).max() < 0.5
    assert (np.delete(flags[:, 0], 17) == 0).all()


def _nf_config(nt, lines):
    W_mask = [{"value": 1.0, "lobin": 0, "hibin": nt - 1}] + \
             [{"value": 0.0, "lobin": b - 5, "hibin": b + 5} for b in lines]
    return [
        {"type": "AnodeChannels", "name": "nfa", "data": {"ident": 0, "plane_ranges": [[0, 8], [8, 8], [16, 8]]}},
        {"type": "OmniChannelNoiseDB", "name": "db", "data": {
            "anode": "AnodeChannels:nfa", "tick": TICK, "nsamples": nt, "bad": [3], "groups": [],
            "channel_info": [
                {"channels": {"first": 0, "last": 23}, "rcrc": TAU, "rc_layers": 1,
                 "min_rms_cut": 1.0, "max_rms_cut": 30.0},
                {"channels": {"wpid": 4}, "freqmasks": W_mask},
            ]}},
        {"type": "pdOneChannelNoise", "name": "one", "data": {
            "anode": "AnodeChannels:nfa", "noisedb": "OmniChannelNoiseDB:db",
            "resmp": [{"channels": [5], "sample_from": nt - 4}]}},
        {"type": "OmnibusNoiseFilter", "name": "nf", "data": {
            "nticks": 0, "maskmap": {"sticky": "bad", "ledge": "bad", "noisy": "bad"},
            "channel_filters": ["pdOneChannelNoise:one"], "grouped_filters": [], "channel_status_filters": [],
            "noisedb": "OmniChannelNoiseDB:db", "intraces": "orig", "outtraces": "raw"}},
    ]


def test_omnibus_noise_filter_component():
    """OmnibusNoiseFilter with pdOneChannelNoise and OmniChannelNoiseDB
    through the component layer: output traces equal the oracle's
    one-channel filter per channel (the bad channel's samples zeroed, the
    resampled channel through FftScaling), masks merged and renamed as
    the reference merges them, the output tags."""
    _torch()
    from wct_amd import components as CO
    rng = np.random.default_rng(11)
    nt = 2000
    lines = [300, 340, 380, 420, 460, 500]
    frame_rows, _, _ = make_frame(rng, nt)
    rows = frame_rows[:24].copy()
    rows[8] = (900 + rng.normal(0, 0.2, nt)).astype(np.float32)        # quiet induction channel: noisy
    rows[10] = (900 + 3000 * np.exp(-np.arange(nt) / 300.0) + rng.normal(0, 3, nt)).astype(np.float32)  # partial
    t = np.arange(nt)
    for ch in (16, 17):     # collection channels with lines in the flagged runs
        rows[ch] = (400 + rng.normal(0, 3, nt) + sum(6 * np.cos(2 * np.pi * b * t / nt + ch) for b in lines)
                    ).astype(np.float32)
    CO.reset()
    CO.configure(_nf_config(nt, lines))
    fr = CO.Frame(ident=7, time=0.0, tick=TICK)
    for ch in range(24):
        fr.add_trace(ch, rows[ch], 0)
    fr.tag_traces("orig", list(range(24)))
    out = CO.frame_filter("OmnibusNoiseFilter:nf", fr)
    assert "noisefilter" in out.frame_tags()
    assert list(out.tagged_traces("raw")) == list(range(24))
    rc = NF.rcrc_spectrum(TAU, TICK, nt, 1)
    noise_w = NF.freqmask_spectrum([(1.0, 0, nt - 1)] + [(0.0, b - 5, b + 5) for b in lines], nt)
    want_masks = {"bad": {3}, "lf_noisy": set(), "harmonic": set()}
    got = {ch: q for ch, tb, q in out.traces()}
    for ch in range(24):
        plane = ch // 8
        sig = np.zeros(nt, np.float32) if ch == 3 else rows[ch]
        want, masks = NF.one_channel_noise(sig, plane, rcrc=rc, noise=noise_w if plane == 2 else None,
                                           min_rms=1.0, max_rms=30.0, resmp=nt - 4 if ch == 5 else None)
        for name in masks:
            want_masks["bad" if name == "noisy" else name].add(ch)
        d = got[ch].astype(np.float64) - want
        med = float(np.median(d))
        width = (float(want.max()) - float(want.min())) / nt + 1e-6
        assert np.abs(d - med).max() <= 1e-2 and abs(med) <= 1.5 * width, (ch, np.abs(d - med).max(), med)
    assert 8 in want_masks["bad"] and 10 in want_masks["lf_noisy"] and {16, 17} <= want_masks["harmonic"]
    for name, chans in want_masks.items():
        m = out.masks(name)
        assert set(m) == chans, (name, set(m), chans)
        assert all(v == [(0, nt)] for v in m.values())
    # end of stream passes through
    assert CO.frame_filter("OmnibusNoiseFilter:nf", None) is None
    CO.reset()


def _mb_frame(rng, nt):
    """MicroBooNE-style channels: baselines 2048 (induction) and 400
    (collection), gain corrections, dead (chirped) stretches, partial
    waveforms, quiet channels, a line removed by a notch spectrum, a
    config spectrum scaling the signal."""
    t = np.arange(nt)
    rows, chans = [], []
    notch = NF.freqmask_spectrum([(1.0, 0, nt - 1), (0.0, 295, 305), (0.0, nt - 305, nt - 295)], nt)
    ramp = (0.5 + 0.5 * np.cos(np.pi * np.minimum(np.arange(nt), nt - np.arange(nt)) / nt)).astype(np.complex64)
    for i in range(32):
        plane = i % 3
        base = 2048.0 if plane < 2 else 400.0
        x = base + rng.normal(0, 3.0, nt)
        kind = i % 8
        if kind == 1:                                     # chirped stretch
            lo = int(rng.integers(100, nt // 2))
            x[lo:lo + 600] = base + rng.normal(0, 0.1, 600)
        if kind == 2:
            x = x + 3000 * np.exp(-t / 300.0)             # partial
        if kind == 3 and i > 8:
            x = base + rng.normal(0, 0.2, nt)             # noisy
        if kind in (4, 5):
            x = x + 20 * np.cos(2 * np.pi * 300 * t / nt)
        rows.append(x.astype(np.float32))
        chans.append(dict(plane=plane, mode="microboone", rcrc=0, baseline=base,
                          gain=1.2 if kind == 6 else 1.0, config=2 if kind == 7 else -1,
                          noise_row=1 if kind in (4, 6) else -1, min_rms=1.0, max_rms=30.0))
    return np.stack(rows), chans, np.stack([NF.rcrc_spectrum(TAU, TICK, nt), notch, ramp]).astype(np.complex64)


@pytest.mark.parametrize("nt", [2000, 1999])
def test_mb_one_channel_matches_oracle(nt):
    torch = _torch()
    from wct_amd import noise as N
    rng = np.random.default_rng(100 + nt)
    frame, chans, spectra = _mb_frame(rng, nt)
    print(f"mb noise filter {len(chans)} channels x {nt}", flush=True)
    out, flags = N.one_channel(frame, chans, rcrc=spectra)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    kinds = {}
    for i, c in enumerate(chans):
        want, masks = NF.mb_one_channel_noise(
            frame[i], c["plane"], baseline=c["baseline"], gain=c["gain"], rcrc=spectra[0],
            config=spectra[c["config"]] if c["config"] >= 0 else None,
            noise=spectra[c["noise_row"]] if c["noise_row"] >= 0 else None, min_rms=1.0, max_rms=30.0)
        assert N.masks(flags[i], c["plane"], nt, "microboone") == masks, (i, flags[i], masks)
        for name in masks:
            kinds[name] = kinds.get(name, 0) + 1
        d = got[i].astype(np.float64) - want
        med = float(np.median(d))
        width = (float(want.max()) - float(want.min())) / nt + 1e-6
        assert np.abs(d - med).max() <= 1e-2 and abs(med) <= 1.5 * width, (i, np.abs(d - med).max(), med)
    assert kinds.get("chirp", 0) >= 3 and kinds.get("lf_noisy", 0) >= 2 and kinds.get("noisy", 0) >= 2, kinds


def test_rel_gain_matches_oracle():
    _torch()
    from wct_amd import noise as N
    rng = np.random.default_rng(5)
    x = rng.normal(0, 10, (6, 300)).astype(np.float32)
    g = [1.0, 1.1, 0.9, 1.3, 0.7, 1.25]
    got = N.rel_gain(x.copy(), g).cpu().numpy()
    for i in range(6):
        np.testing.assert_array_equal(got[i], NF.rel_gain_calib(x[i], g[i]))


def test_mb_component_chain():
    """OmnibusNoiseFilter with mbOneChannelNoise then pdRelGainCalib, the
    DB's nominal baselines, gain corrections, unit config spectra
    (reconfig {}) and frequency masks: traces equal the oracle's two
    filters in sequence, masks merged (chirp and noisy renamed bad)."""
    _torch()
    from wct_amd import components as CO
    rng = np.random.default_rng(31)
    nt = 2000
    frame, _, _ = _mb_frame(rng, nt)
    rows = frame[:24].copy()
    for ch in range(24):     # plane by channel block: 0-7 U, 8-15 V, 16-23 W
        base = 2048.0 if ch < 16 else 400.0
        rows[ch] = rows[ch] - (2048.0 if ch % 3 < 2 else 400.0) + base
    gains = [1.0 + 0.01 * c for c in range(24)]
    notch = [{"value": 1.0, "lobin": 0, "hibin": nt - 1}, {"value": 0.0, "lobin": 295, "hibin": 305},
             {"value": 0.0, "lobin": nt - 305, "hibin": nt - 295}]
    cfg = [
        {"type": "AnodeChannels", "name": "mba", "data": {"ident": 0, "plane_ranges": [[0, 8], [8, 8], [16, 8]]}},
        {"type": "OmniChannelNoiseDB", "name": "mbdb", "data": {
            "anode": "AnodeChannels:mba", "tick": TICK, "nsamples": nt, "bad": [], "groups": [],
            "channel_info": [
                {"channels": {"first": 0, "last": 15}, "nominal_baseline": 2048.0, "gain_correction": 1.0,
                 "rcrc": TAU, "rc_layers": 1, "min_rms_cut": 1.0, "max_rms_cut": 30.0, "reconfig": {},
                 "freqmasks": notch},
                {"channels": {"first": 16, "last": 23}, "nominal_baseline": 400.0, "gain_correction": 1.2,
                 "rcrc": TAU, "min_rms_cut": 1.0, "max_rms_cut": 30.0, "reconfig": {}, "freqmasks": notch},
            ]}},
        {"type": "mbOneChannelNoise", "name": "mb", "data": {"anode": "AnodeChannels:mba",
                                                            "noisedb": "OmniChannelNoiseDB:mbdb"}},
        {"type": "pdRelGainCalib", "name": "g", "data": {"anode": "AnodeChannels:mba",
                                                          "noisedb": "OmniChannelNoiseDB:mbdb", "rel_gain": gains}},
        {"type": "OmnibusNoiseFilter", "name": "mbnf", "data": {
            "nticks": 0, "maskmap": {"chirp": "bad", "noisy": "bad"},
            "channel_filters": ["mbOneChannelNoise:mb", "pdRelGainCalib:g"], "grouped_filters": [],
            "channel_status_filters": [], "noisedb": "OmniChannelNoiseDB:mbdb", "intraces": "orig",
            "outtraces": "quiet"}},
    ]
    CO.reset()
    CO.configure(cfg)
    fr = CO.Frame(ident=3, time=0.0, tick=TICK)
    for ch in range(24):
        fr.add_trace(ch, rows[ch], 0)
    fr.tag_traces("orig", list(range(24)))
    out = CO.frame_filter("OmnibusNoiseFilter:mbnf", fr)
    got = {ch: q for ch, tb, q in out.traces()}
    rc = NF.rcrc_spectrum(TAU, TICK, nt, 1)
    ones = np.ones(nt, np.complex64)
    nmask = NF.freqmask_spectrum([(1.0, 0, nt - 1), (0.0, 295, 305), (0.0, nt - 305, nt - 295)], nt)
    want_masks = {}
    for ch in range(24):
        plane = ch // 8
        want, masks = NF.mb_one_channel_noise(rows[ch], plane, baseline=2048.0 if ch < 16 else 400.0,
                                              gain=1.0 if ch < 16 else 1.2, rcrc=rc, config=ones, noise=nmask,
                                              min_rms=1.0, max_rms=30.0)
        want = NF.rel_gain_calib(want, gains[ch])
        for name, ranges in masks.items():
            want_masks.setdefault({"chirp": "bad", "noisy": "bad"}.get(name, name), {}).setdefault(ch, []).extend(ranges)
        d = got[ch].astype(np.float64) - want
        med = float(np.median(d))
        width = (float(want.max()) - float(want.min())) / nt + 1e-6
        assert np.abs(d - med).max() <= 1e-2 and abs(med) <= 1.5 * width * 1.25, (ch, np.abs(d - med).max(), med)
    for name, chans in want_masks.items():
        m = out.masks(name)
        assert set(m) == set(chans), (name, set(m), set(chans))
    assert "noisefilter" in out.frame_tags() and list(out.tagged_traces("quiet")) == list(range(24))
    CO.reset()


def _coherent_case(rng, nt, nch, with_response):
    """One group: a common waveform on every channel (scaled per channel),
    independent noise, and a pulse at the same tick on every channel (a
    track parallel to the wire plane), through an induction-like response
    when with_response."""
    t = np.arange(nt)
    common = 6 * np.sin(t / 23.0) + 3 * np.sin(t / 7.0)
    resp = np.zeros(nt)
    resp[100:140] = np.sin(np.linspace(0, 2 * np.pi, 40))          # bipolar, offset ~120
    pulse = np.zeros(nt)
    pulse[900] = 400.0
    pulse = np.convolve(pulse, np.exp(-0.5 * (np.arange(-15, 16) / 3.0) ** 2), "same")
    if with_response:
        pulse = np.fft.ifft(np.fft.fft(pulse) * np.fft.fft(resp)).real
    rows = []
    for c in range(nch):
        x = rng.uniform(0.8, 1.2) * common + rng.normal(0, 1.0, nt)
        if c % 3 == 0:
            x = x + pulse
        rows.append(x.astype(np.float32))
    respec = np.fft.fft(resp.astype(np.float32).astype(np.float64)).astype(np.complex64)
    return np.stack(rows), respec


def test_coherent_noise_sub_matches_oracle():
    """Three disjoint groups in interleaved rows: two with the response
    path (U-plane settings), one without (collection)."""
    torch = _torch()
    from wct_amd import noise as N
    rng = np.random.default_rng(77)
    nt = 2000
    blocks, respecs = [], []
    for k, resp in enumerate((True, True, False)):
        rows, rs = _coherent_case(rng, nt, 16 + 8 * k, resp)
        blocks.append(rows)
        respecs.append(rs)
    nrows = sum(len(b) for b in blocks)
    # frame row of each (group, channel), plus one row in no group; the
    # odd last row is in a response group (its complex row has no partner)
    order = rng.permutation(nrows + 1)
    j = int(np.flatnonzero(order == nrows)[0])
    order[[0, j]] = order[[j, 0]]              # into group 0 (response path)
    frame = np.zeros((nrows + 1, nt), np.float32)
    frame[order[nrows]] = rng.normal(0, 3, nt).astype(np.float32)
    groups, pos = [], 0
    spectra = np.stack([respecs[0], respecs[1]])
    for k, b in enumerate(blocks):
        idx = [int(order[pos + i]) for i in range(len(b))]
        pos += len(b)
        for i, r in enumerate(idx):
            frame[r] = b[i]
        p = dict(rows=idx, pad_f=10, pad_b=10, adc_limit=15.0, protection_factor=5.0, min_adc_limit=50.0,
                 decon_limit=0.02, decon_limit1=0.07, roi_min_max_ratio=3.0)
        if k < 2:
            p.update(respec=k, res_offset=120)
        groups.append(p)
    print(f"coherent {len(groups)} groups, {nrows} rows x {nt}", flush=True)
    got = N.coherent(frame, groups, spectra).cpu().numpy()
    for k, grp in enumerate(groups):
        want = NF.coherent_noise_sub([frame[r] for r in grp["rows"]],
                                     respec=spectra[k] if k < 2 else None, res_offset=120 if k < 2 else 0,
                                     pad_f=10, pad_b=10, decon_limit=0.02, decon_lf_cutoff=0.08, adc_limit=15.0,
                                     decon_limit1=0.07, roi_min_max_ratio=3.0, protection_factor=5.0,
                                     min_adc_limit=50.0)
        for i, r in enumerate(grp["rows"]):
            d = np.abs(got[r].astype(np.float64) - want[i])
            assert d.max() <= 1e-3, (k, i, d.max(), int(d.argmax()))
        # the common mode is gone and the pulse kept on the pulsed channels
        for i, r in enumerate(grp["rows"]):
            assert float(np.std(got[r][:600])) < 1.6
    # rows of no group are untouched
    assert np.array_equal(got[order[nrows]], frame[order[nrows]])


def test_grouped_filter_component():
    """OmnibusNoiseFilter with mbCoherentNoiseSub as a grouped filter over
    the DB's groups (U/V groups with an explicit response waveform and
    offset, W groups without), all groups in one batch: traces equal the
    oracle's coherent subtraction group by group; a group with a missing
    channel is left alone."""
    _torch()
    from wct_amd import components as CO
    rng = np.random.default_rng(91)
    nt = 2000
    resp = np.zeros(nt)
    resp[100:140] = np.sin(np.linspace(0, 2 * np.pi, 40))
    rows = np.zeros((24, nt), np.float32)
    for k in range(3):
        r, _ = _coherent_case(rng, nt, 8, k < 2)
        rows[8 * k:8 * (k + 1)] = r
    groups = [list(range(0, 8)), list(range(8, 16)), list(range(16, 24)), [0, 1, 99]]   # the last one incomplete
    common = {"pad_window_front": 10, "pad_window_back": 10, "adc_limit": 15, "protection_factor": 5.0,
              "min_adc_limit": 50, "decon_limit": 0.02, "decon_limit1": 0.07, "roi_min_max_ratio": 3.0,
              "min_rms_cut": 0.0, "max_rms_cut": 1000.0}
    cfg = [
        {"type": "AnodeChannels", "name": "ca", "data": {"ident": 0, "plane_ranges": [[0, 8], [8, 8], [16, 8]]}},
        {"type": "OmniChannelNoiseDB", "name": "cdb", "data": {
            "anode": "AnodeChannels:ca", "tick": TICK, "nsamples": nt, "bad": [], "groups": groups,
            "channel_info": [
                dict(common, channels={"first": 0, "last": 15}, response_offset=120,
                     response={"waveform": [float(v) for v in resp[:200]], "waveformid": 1}),
                dict(common, channels={"first": 16, "last": 23}),
            ]}},
        {"type": "mbCoherentNoiseSub", "name": "coh", "data": {"anode": "AnodeChannels:ca",
                                                               "noisedb": "OmniChannelNoiseDB:cdb"}},
        {"type": "OmnibusNoiseFilter", "name": "cnf", "data": {
            "nticks": 0, "maskmap": {}, "channel_filters": [], "grouped_filters": ["mbCoherentNoiseSub:coh"],
            "channel_status_filters": [], "noisedb": "OmniChannelNoiseDB:cdb", "intraces": "orig",
            "outtraces": "quiet"}},
    ]
    CO.reset()
    CO.configure(cfg)
    fr = CO.Frame(ident=5, time=0.0, tick=TICK)
    for ch in range(24):
        fr.add_trace(ch, rows[ch], 0)
    fr.tag_traces("orig", list(range(24)))
    out = CO.frame_filter("OmnibusNoiseFilter:cnf", fr)
    got = {ch: q for ch, tb, q in out.traces()}
    respec = np.fft.fft(np.concatenate([resp[:200], np.zeros(nt - 200)]).astype(np.float32).astype(np.float64))
    for k in range(3):
        want = NF.coherent_noise_sub([rows[c] for c in range(8 * k, 8 * k + 8)],
                                     respec=respec.astype(np.complex64) if k < 2 else None,
                                     res_offset=120 if k < 2 else 0, pad_f=10, pad_b=10, decon_limit=0.02,
                                     decon_lf_cutoff=0.08, adc_limit=15.0, decon_limit1=0.07, roi_min_max_ratio=3.0,
                                     protection_factor=5.0, min_adc_limit=50.0)
        for i in range(8):
            d = np.abs(got[8 * k + i].astype(np.float64) - want[i])
            assert d.max() <= 1e-3, (k, i, d.max())
    CO.reset()


def _status_rows(rng, nt, n=43):
    """Rows for OneChannelStatus: unit noise (a ratio near 5 at 2000
    ticks), plus a sine inside the 250 summed bins at amplitude 40 (ratio
    near 25) or 12 (near 11), a sine past them at 60, or a sine of 40 with
    five 500 ADC spikes; then a zero row and a constant row."""
    i = np.arange(nt)
    rows = []
    for k in range(n - 2):
        r = rng.normal(0, 1.0, nt)
        kind = k % 5
        ph = rng.uniform(0, 2 * np.pi)
        if kind in (1, 4):
            r += 40 * np.sin(2 * np.pi * rng.integers(1, 251) * i / nt + ph)
        elif kind == 2:
            r += 12 * np.sin(2 * np.pi * rng.integers(1, 251) * i / nt + ph)
        elif kind == 3:
            r += 60 * np.sin(2 * np.pi * rng.integers(300, 900) * i / nt + ph)
        if kind == 4:
            r[rng.integers(0, nt, 5)] += 500.0
        rows.append(r)
    rows.append(np.zeros(nt))
    rows.append(np.full(nt, 300.0))
    return np.asarray(rows, np.float32)


def test_lf_noisy_matches_oracle():
    """wct_nf_lf_noisy on an odd number of rows against the oracle's
    ID_lf_noisy: ratios within 1e-4 (FP32 transforms of two channels at
    once against the oracle's), flags equal away from the cut."""
    _torch()
    from wct_amd import noise as N
    rng = np.random.default_rng(1281)
    nt = 2000
    rows = _status_rows(rng, nt)
    print(f"lf_noisy {rows.shape}", flush=True)
    flags, ratio = N.lf_noisy(rows)
    nflag = 0
    for c in range(len(rows)):
        want, wr = NF.lf_noisy(rows[c])
        assert abs(ratio[c] - wr) <= 1e-4 * max(wr, 1.0), (c, ratio[c], wr)
        if abs(wr - 14.0) > 0.01:
            assert bool(flags[c]) == want, (c, flags[c], want, wr)
        nflag += want
    assert 10 <= nflag <= 30
    assert not flags[-2] and ratio[-2] == 0             # the zero row
    assert not flags[-1] and ratio[-1] < 1e-3           # the constant row: only bin 0
    # other parameters
    flags, ratio = N.lf_noisy(rows[:8], threshold=2.5, window=9, nbins=600, cut=8.0)
    for c in range(8):
        want, wr = NF.lf_noisy(rows[c], 2.5, 9, 600, 8.0)
        assert abs(ratio[c] - wr) <= 1e-4 * max(wr, 1.0), (c, ratio[c], wr)
        if abs(wr - 8.0) > 0.01:
            assert bool(flags[c]) == want


def test_status_filter_component():
    """OmnibusNoiseFilter with mbOneChannelStatus as its status filter:
    the induction channels whose oracle ID_lf_noisy holds are masked
    lf_noisy over the whole waveform, collection channels never; the
    samples pass unchanged."""
    _torch()
    from wct_amd import components as CO
    rng = np.random.default_rng(1263)
    nt = 2000
    rows = _status_rows(rng, nt, n=26)[:24]
    cfg = [
        {"type": "AnodeChannels", "name": "sa", "data": {"ident": 0, "plane_ranges": [[0, 8], [8, 8], [16, 8]]}},
        {"type": "OmniChannelNoiseDB", "name": "sdb", "data": {
            "anode": "AnodeChannels:sa", "tick": TICK, "nsamples": nt, "bad": [],
            "channel_info": [{"channels": {"first": 0, "last": 23}, "min_rms_cut": 0.0, "max_rms_cut": 1000.0}]}},
        {"type": "mbOneChannelStatus", "name": "st", "data": {"anode": "AnodeChannels:sa", "Threshold": 3.5,
                                                              "Window": 5, "Nbins": 250, "Cut": 14}},
        {"type": "OmnibusNoiseFilter", "name": "snf", "data": {
            "nticks": 0, "maskmap": {}, "channel_filters": [], "grouped_filters": [],
            "channel_status_filters": ["mbOneChannelStatus:st"], "noisedb": "OmniChannelNoiseDB:sdb",
            "intraces": "orig", "outtraces": "raw"}},
    ]
    CO.reset()
    CO.configure(cfg)
    fr = CO.Frame(ident=7, time=0.0, tick=TICK)
    for ch in range(24):
        fr.add_trace(ch, rows[ch], 0)
    fr.tag_traces("orig", list(range(24)))
    out = CO.frame_filter("OmnibusNoiseFilter:snf", fr)
    got = {ch: q for ch, tb, q in out.traces()}
    want = {ch for ch in range(16) if NF.one_channel_status(rows[ch], ch // 8)}
    assert want and any(NF.lf_noisy(rows[ch])[0] for ch in range(16, 24))   # a collection channel that would be
    m = out.masks("lf_noisy")
    assert set(m) == want, (set(m), want)
    assert all(v == [(0, nt)] for v in m.values())
    for ch in range(24):
        assert np.array_equal(got[ch], rows[ch])
    CO.reset()


def _rotl12(x, n):
    x = np.asarray(x, dtype=np.int64)
    return (((x << n) | (x >> (12 - n))) & 0xFFF).astype(np.int64)


def _bitshift_rows(rng, nt, nrows=21):
    """Baselines whose top bits hold still (600-900 +- 3 or 20: bits 9-11
    fixed; 1100 +- 3 for a 5-bit rotation: bits 7-11 fixed), rotated left by 0, 1, 2, 3 or 5 bits in 12,
    some with samples whose bit 11 is set (glitches the correction must
    catch), and a zero row."""
    rows = []
    for k in range(nrows - 1):
        n = (0, 1, 2, 3, 5)[k % 5]
        base, sd = (1100, 3.0) if n == 5 else (rng.uniform(600, 900), (3.0, 20.0)[k % 2])
        x = np.clip(np.rint(base + rng.normal(0, sd, nt)), 0, 2047).astype(np.int64)
        if k % 3 == 0:
            x[rng.integers(600, nt, 4)] += 2048
        rows.append(_rotl12(x, n) if n else x)
    rows.append(np.zeros(nt, np.int64))
    return np.asarray(rows, np.float32)


def test_adc_bit_shift_matches_oracle():
    """wct_nf_adc_bit_shift against the oracle: shift counts equal and
    every sample equal (integer arithmetic), at the default and at other
    thresholds."""
    _torch()
    from wct_amd import noise as N
    rng = np.random.default_rng(1071)
    nt = 3000
    rows = _bitshift_rows(rng, nt)
    print(f"adc bit shift {rows.shape}", flush=True)
    got, ns = N.adc_bit_shift(rows)
    got = got.cpu().numpy()
    seen = set()
    for c in range(len(rows)):
        want, wn = NF.adc_bit_shift(rows[c])
        assert ns[c] == wn, (c, ns[c], wn)
        assert np.array_equal(got[c], want), (c, np.flatnonzero(got[c] != want)[:5])
        seen.add(wn)
    assert {0, 1, 2, 3, 5} <= seen
    got, ns = N.adc_bit_shift(rows[:10], exam_nticks=800, threshold_sigma=5.0, threshold_fix=0.5)
    got = got.cpu().numpy()
    for c in range(10):
        want, wn = NF.adc_bit_shift(rows[c], exam_nticks=800, threshold_sigma=5.0, threshold_fix=0.5)
        assert ns[c] == wn and np.array_equal(got[c], want), c


def test_bit_shift_filter_component():
    """OmnibusNoiseFilter with mbADCBitShift as a channel filter: traces
    equal the oracle's, corrected channels masked ADCBitShift (n, n)."""
    _torch()
    from wct_amd import components as CO
    rng = np.random.default_rng(1106)
    nt = 3000
    rows = _bitshift_rows(rng, nt, nrows=12)
    cfg = [
        {"type": "AnodeChannels", "name": "ba", "data": {"ident": 0, "plane_ranges": [[0, 4], [4, 4], [8, 4]]}},
        {"type": "OmniChannelNoiseDB", "name": "bdb", "data": {
            "anode": "AnodeChannels:ba", "tick": TICK, "nsamples": nt, "bad": [],
            "channel_info": [{"channels": {"first": 0, "last": 11}, "min_rms_cut": 0.0, "max_rms_cut": 1000.0}]}},
        {"type": "mbADCBitShift", "name": "bs", "data": {"Number_of_ADC_bits": 12, "Exam_number_of_ticks_test": 500,
                                                          "Threshold_sigma_test": 7.5, "Threshold_fix": 0.8}},
        {"type": "OmnibusNoiseFilter", "name": "bnf", "data": {
            "nticks": 0, "maskmap": {}, "channel_filters": ["mbADCBitShift:bs"], "grouped_filters": [],
            "channel_status_filters": [], "noisedb": "OmniChannelNoiseDB:bdb", "intraces": "orig",
            "outtraces": "raw"}},
    ]
    CO.reset()
    CO.configure(cfg)
    fr = CO.Frame(ident=3, time=0.0, tick=TICK)
    for ch in range(12):
        fr.add_trace(ch, rows[ch], 0)
    fr.tag_traces("orig", list(range(12)))
    out = CO.frame_filter("OmnibusNoiseFilter:bnf", fr)
    got = {ch: q for ch, tb, q in out.traces()}
    want_mask = {}
    for ch in range(12):
        want, wn = NF.adc_bit_shift(rows[ch])
        assert np.array_equal(got[ch], want), ch
        if wn:
            want_mask[ch] = [(wn, wn)]
    assert want_mask
    assert out.masks("ADCBitShift") == want_mask
    CO.reset()
