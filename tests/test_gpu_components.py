"""The C++ components on the device, called through their interfaces.

HipDFT (IDFT): the aux/test/test_idft.cxx cases (impulse, flat, zero,
round trips, 1b both axes, 2d, transposes; tolerance 1e-8 absolute on
those exact cases) plus random data against the FP64 oracle DFT, and
calls from several threads at once.

HipOmnibusSigProc (IFrameFilter): a three-plane anode frame pushed
through the component must give exactly what the C ABI and the oracle
give on the same data: the wiener and gauss traces (the hits and charge
decons of the C ABI cut by the oracle's apply_roi to the ROIs of the
oracle's ROI refinement, which starts from the oracle's tight and loose
ROI lists on the C ABI's mask, loose and refine decons), with the
per-wire RMS summary on wiener, and the debug mode's tight_lf / loose_lf
/ decon_charge outputs; save_data's post-processing (negatives and bad
ranges zeroed; debug mode keeps negatives; sparse runs); tight ROI lists
equal to the oracle's run/extend/connect rules applied to the device
mask; and no traces without ROI refinement or the debug mode, as the
reference saves none then.
"""
import json
import os
import threading

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture
def K(dev):
    from wct_amd import components
    components.reset()
    yield components
    components.reset()


def _crand(rng, *shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


def test_hipdft_reference_cases(K):
    K.configure([{"type": "HipDFT", "name": ""}])
    for n in (1, 2, 8, 60, 6000):
        imp = np.zeros(n, np.complex64)
        imp[0] = 1
        np.testing.assert_allclose(K.dft("HipDFT", "fwd1d", imp), np.ones(n), atol=1e-8)
        flat = np.ones(n, np.complex64)
        back = K.dft("HipDFT", "inv1d", K.dft("HipDFT", "fwd1d", imp))
        np.testing.assert_allclose(back, imp, atol=1e-7)
        np.testing.assert_allclose(K.dft("HipDFT", "fwd1d", flat)[0], n, rtol=1e-7)
        assert not np.any(K.dft("HipDFT", "fwd1d", np.zeros(n, np.complex64)))
    for shape in [(4, 6), (16, 1), (1, 16), (64, 96)]:
        imp = np.zeros(shape, np.complex64)
        imp[0, 0] = 1
        for axis in (0, 1):
            np.testing.assert_allclose(K.dft("HipDFT", "fwd1b", imp, axis), np.fft.fft(imp, axis=axis), atol=1e-8)
        np.testing.assert_allclose(K.dft("HipDFT", "fwd2d", imp), np.ones(shape), atol=1e-8)
        np.testing.assert_allclose(K.dft("HipDFT", "inv2d", np.ones(shape, np.complex64)), imp, atol=1e-7)


@pytest.mark.parametrize("shape", [(48, 600), (960, 64), (77, 143)])
def test_hipdft_random_against_oracle(K, shape):
    K.configure([{"type": "HipDFT", "name": "gpu0", "data": {"device": 0}}])
    rng = np.random.default_rng(shape[0])
    x = _crand(rng, *shape)
    for axis in (0, 1):
        ref = np.fft.fft(x.astype(np.complex128), axis=axis)
        got = K.dft("HipDFT:gpu0", "fwd1b", x, axis)
        assert np.max(np.abs(got - ref)) <= 2e-6 * np.max(np.abs(ref))
        back = K.dft("HipDFT:gpu0", "inv1b", got, axis)
        assert np.max(np.abs(back - x)) <= 2e-6 * np.max(np.abs(x))
    ref2 = O.fwd2d(x, np.float64)
    got2 = K.dft("HipDFT:gpu0", "fwd2d", x)
    assert np.max(np.abs(got2 - ref2)) <= 2e-6 * np.max(np.abs(ref2))
    xr = rng.standard_normal(shape).astype(np.float32)
    np.testing.assert_array_equal(K.dft("HipDFT:gpu0", "transpose", xr), xr.T)
    np.testing.assert_array_equal(K.dft("HipDFT:gpu0", "transpose", x), x.T)


def test_hipdft_concurrent_threads(K):
    K.configure([{"type": "HipDFT", "name": ""}])
    rng = np.random.default_rng(5)
    xs = [_crand(rng, 8, 1000 + 100 * i) for i in range(6)]
    errs = [None] * len(xs)

    def work(i):
        try:
            for _ in range(5):
                got = K.dft("HipDFT", "fwd1b", xs[i], 1)
                ref = np.fft.fft(xs[i].astype(np.complex128), axis=1)
                assert np.max(np.abs(got - ref)) <= 2e-6 * np.max(np.abs(ref))
        except Exception as e:    # reported below
            errs[i] = e

    th = [threading.Thread(target=work, args=(i,)) for i in range(len(xs))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert errs == [None] * len(xs)


# ---- HipOmnibusSigProc -------------------------------------------------

NW = (48, 48, 56)
NT = 1200
CH0 = (1000, 2000, 3000)


def _sigproc_config(tmpdir, extra=None):
    from wct_amd import components as K
    from wct_amd import workload as W
    planes = []
    for ip, kind in enumerate(("ind", "ind", "col")):
        pos, cur = W.fr_paths(kind)
        planes.append((ip, 0.0, W.PITCH, pos, cur))
    frj = K.field_response_json(planes, W.ORIGIN, 0.0, W.FINE_PERIOD, W.SPEED)
    path = os.path.join(str(tmpdir), "fr.json")
    with open(path, "w") as f:
        json.dump(frj, f)
    data = {"anode": "AnodeChannels:apa0", "field_response": "FieldResponse:synth",
            "elecresponse": "ColdElecResponse:ce", "ctoffset": W.CTOFFSET, "ftoffset": W.FTOFFSET,
            "postgain": W.INTER_GAIN, "ADC_mV": W.ADC_MV, "troi_ind_th_factor": 3.0,
            "troi_col_th_factor": 5.0}
    data.update(extra or {})
    wire = dict(W.WIRE)
    return [
        {"type": "AnodeChannels", "name": "apa0",
         "data": {"ident": 0, "plane_ranges": [[CH0[i], NW[i]] for i in range(3)]}},
        {"type": "FieldResponse", "name": "synth", "data": {"filename": path}},
        {"type": "ColdElecResponse", "name": "ce",
         "data": {"gain": W.GAIN, "shaping": W.SHAPING, "postgain": 1.0}},
        {"type": "HfFilter", "name": "Wiener_tight_U", "data": W.WIENER},
        {"type": "HfFilter", "name": "Wiener_tight_V", "data": W.WIENER},
        {"type": "HfFilter", "name": "Wiener_tight_W", "data": W.WIENER},
        {"type": "HfFilter", "name": "Wire_ind", "data": wire},
        {"type": "HfFilter", "name": "Wire_col", "data": wire},
        {"type": "LfFilter", "name": "ROI_tight_lf", "data": W.LF_TIGHT},
        {"type": "LfFilter", "name": "ROI_tighter_lf", "data": W.LF_TIGHTER},
        {"type": "LfFilter", "name": "ROI_loose_lf", "data": W.LF_LOOSE},
        {"type": "HfFilter", "name": "Gaus_wide", "data": W.GAUS_WIDE},
        {"type": "HfFilter", "name": "Wiener_wide_U", "data": W.WIENER_WIDE},
        {"type": "HfFilter", "name": "Wiener_wide_V", "data": W.WIENER_WIDE},
        {"type": "HfFilter", "name": "Wiener_wide_W", "data": W.WIENER_WIDE},
        {"type": "HipDFT", "name": ""},
        {"type": "HipOmnibusSigProc", "name": "osp", "data": data},
    ]


def _adc_planes(seed):
    rng = np.random.default_rng(seed)
    out = []
    for ip in range(3):
        a = rng.normal(0.0, 1.5, (NW[ip], NT)).astype(np.float32)
        for _ in range(4):   # a few bipolar/unipolar pulses
            w, t = rng.integers(2, NW[ip] - 2), rng.integers(100, NT - 100)
            a[w - 1:w + 2, t:t + 20] += (40.0 if ip == 2 else 25.0) * np.hanning(20)
        out.append(a)
    return out


def _filters(kind, nt, debug=False, tight=True, refine=True, charge=True):
    """The component's output filters of one plane, in FP64: the base
    (tight, or all ones for an induction plane without ROI refinement),
    the tighter one; in debug mode the loose one (loose_dbg, no tight-LF
    rows) and the charge one; with refinement the loose one with its
    tight-LF alternative (alt), the refine, hits and charge ones."""
    from wct_amd import host, workload as W
    wiener = host.hf_filter(nt, **W.WIENER)
    ind = kind == "ind"
    f = {}
    if tight:
        f["base"] = wiener * host.lf_filter(nt, **W.LF_TIGHT) if ind else wiener
        if ind:
            f["tighter"] = wiener * host.lf_filter(nt, **W.LF_TIGHTER)
    else:
        f["base"] = np.ones(nt)
    if debug and ind:
        f["loose_dbg"] = wiener * host.lf_filter(nt, **W.LF_LOOSE)
    if refine or (debug and charge):
        f["charge"] = host.hf_filter(nt, **W.GAUS_WIDE)
    if refine:
        if ind:
            f["loose"] = wiener * host.lf_filter(nt, **W.LF_LOOSE)
            f["alt"] = wiener * host.lf_filter(nt, **W.LF_TIGHT)
            f["refine"] = wiener
        f["hits"] = host.hf_filter(nt, **W.WIENER_WIDE)
    return f


# the component's extra outputs, in its order; baseline restored on
# induction planes for the loose and refine decons, on the collection
# plane for every one (decon_2D_hits/charge, OmnibusSigProc.cxx:1344-1379)
_EXTRAS = ("loose_dbg", "charge", "loose", "refine", "hits")


def _alt_rows(bad, lfn, ip):
    """decon_2D_looseROI's tight-LF rows: a bad channel within 2 wires (1
    on the V plane) or an lf_noisy one within 2, in the same plane."""
    nw = NW[ip]
    isbad = [bool(bad.get((ip, w))) for w in range(nw)]
    isl = [(ip, w) in lfn for w in range(nw)]
    nb = 2 if ip == 0 else 1
    return np.array([int(any(isbad[max(0, w - nb):w + nb + 1]) or any(isl[max(0, w - 2):w + 3]))
                     for w in range(nw)], np.uint8)


def _expected(adc, bad, th, debug=False, refine=True, chan=None, charge=True, lfn=()):
    """The same planes through the C ABI (workload's spec, padded to
    plane 0's 21 response regions as the reference sizes them), with the
    component's unfiltered kernels and output filters.  chan is
    (responses, {channel: response row}, t0) for the per-channel
    correction; lfn the (plane, wire)s of lf_noisy channels.  Returns
    per plane (base, mask, rms, {extra name: frame})."""
    import torch
    from wct_amd import decon as D
    from wct_amd import workload as W
    dev = torch.device("cuda:0")
    kinds = ("ind", "ind", "col")
    jobs, keep, res = [], [], []
    nfine_rows = len(W.fr_paths("ind")[0])
    table = None
    if chan is not None:
        from wct_amd import host
        cresp, chmap, t0 = chan
        ewave = host.coldelec_samples(NT, t0, t0 + NT * W.TICK, W.GAIN, W.SHAPING, 1.0)
        table = D.channel_factors(cresp, ewave, NT, device=dev)
    for ip, kind in enumerate(kinds):
        sp = W.plane_spec(kind, NW[ip], NT, fold_lf=False)
        rows = np.zeros((nfine_rows, sp.fr_rows.shape[1]))
        rows[:sp.fr_rows.shape[0]] = sp.fr_rows
        sp.fr_rows = rows
        sp.wire_shift = (nfine_rows - 1) // 2
        tight = kind == "col" or refine
        flt = _filters(kind, NT, debug, tight, refine, charge)
        env = np.max(np.abs(np.stack(list(flt.values()))), axis=0)
        pk = D.PlaneKernel(sp, unfiltered=True, envelope=env)
        g = {k: torch.from_numpy(v[:pk.kc].astype(np.float32)).to(dev) for k, v in flt.items()}
        a = torch.from_numpy(adc[ip]).to(dev)
        y = torch.empty_like(a)
        m = torch.empty(a.shape, dtype=torch.uint8, device=dev) if tight else None
        r = torch.empty(NW[ip], dtype=torch.float32, device=dev) if tight else None
        ptr, lo, hi = [0], [], []
        for w in range(NW[ip]):
            for (b0, b1) in bad.get((ip, w), []):
                lo.append(b0)
                hi.append(b1)
            ptr.append(len(lo))
        badt = tuple(torch.tensor(v if v else [0], dtype=torch.int32, device=dev) for v in (ptr, lo, hi))
        extra = {k: torch.empty_like(a) for k in _EXTRAS if k in flt}
        alt = torch.from_numpy(_alt_rows(bad, lfn, ip)).to(dev) if "loose" in flt else None
        outs = [D.Output(g[k], out=extra[k], restore=1 if (k in ("loose_dbg", "loose", "refine") or kind == "col")
                         else 0, save_mode=0, filter_alt=g["alt"] if k == "loose" else None,
                         alt_rows=alt if k == "loose" else None) for k in extra]
        keep.append(alt)
        cj = None
        if table is not None:
            rows = torch.tensor([chmap[CH0[ip] + w] for w in range(NW[ip])], dtype=torch.int32, device=dev)
            cj = (table, rows)
            keep.append(rows)
        jobs.append(D.make_job(pk, a, y, mask=m, rms=r, th_factor=th[ip], restore=1 if tight else 0,
                               save_mode=0, lf_tight=g["base"], lf_tighter=g.get("tighter"),
                               bad=badt, outputs=outs, chan=cj))
        keep += [pk, a, y, m, r, badt, g, extra]
        res.append((y, m, r, extra))
    b = D.Batch(jobs, device=dev)
    b.run()
    torch.cuda.synchronize()
    out = [(y.cpu().numpy(), None if m is None else m.cpu().numpy(), None if r is None else r.cpu().numpy(),
            {k: v.cpu().numpy() for k, v in extra.items()}) for (y, m, r, extra) in res]
    b.close()
    return out


def _frame(K, adc, bad):
    f = K.Frame(17, 3.0, 500.0)
    for ip in range(3):
        for w in range(NW[ip]):
            f.add_trace(CH0[ip] + w, adc[ip][w])
    for (ip, w), ranges in bad.items():
        for (b0, b1) in ranges:
            f.add_mask("bad", CH0[ip] + w, b0, b1)
    f.add_mask("bad", 99999, 0, 10)   # a channel of another anode is ignored
    return f


def _refined(exp, bad, nt=NT):
    """Per plane the oracle's refined ROIs (ROI_refinement restated in
    oracle/roi_refine.py, from the oracle's tight ROI lists on the C ABI's
    mask and, on the induction planes, its loose ROI lists on the C ABI's
    loose decon, with the refine decon as the ROIs' data) and the wiener
    and gauss frames: the C ABI's hits and charge decons cut by the
    oracle's apply_roi, then save_data's zeroing of negatives and
    [lo, hi) bad ranges."""
    import roi_refine as RR
    res = []
    for ip in range(3):
        y, m, r, extra = exp[ip]
        nw = NW[ip]
        badl = [list(bad.get((ip, w), [])) for w in range(nw)]
        tight = O.roi_lists(m, pad=5, asy=0.1)
        if ip < 2:
            loose = O.loose_rois(extra["loose"], bad=badl, tight=tight)
            base = extra["refine"]
        else:
            loose, base = None, y
        R = RR.RoiRefinement(ip, nw)
        R.load_data(base, r, tight, loose, bad=badl)
        R.refine()
        rois = R.rois(True)
        frames = []
        for key in ("hits", "charge"):
            cut = O.apply_roi(extra[key], rois, ip < 2)
            cut = np.where(cut > 0, cut, 0).astype(np.float32)
            for w in range(nw):
                for (b0, b1) in badl[w]:
                    cut[w, max(0, b0):min(nt, b1)] = 0
            frames.append(cut)
        res.append((rois, frames[0], frames[1]))
    return res


def test_sigproc_matches_c_abi(K, tmp_path):
    print("sigproc vs C ABI + oracle ROI refinement, 3 planes", NW, "x", NT, flush=True)
    K.configure(_sigproc_config(tmp_path))
    adc = _adc_planes(3)
    bad = {(0, 5): [(100, 300)], (2, 7): [(0, NT)], (2, 8): [(500, 510), (900, 950)]}
    out = K.frame_filter("HipOmnibusSigProc:osp", _frame(K, adc, bad))
    exp = _expected(adc, bad, (3.0, 3.0, 5.0))
    ref = _refined(exp, bad)
    info = out.info()
    assert info["ident"] == 17 and info["time"] == 3.0 and info["tick"] == 500.0
    assert out.frame_tags() == ["sigproc"]
    assert sorted(out.trace_tags()) == ["gauss", "wiener"]
    tr = out.traces()
    assert len(tr) == 2 * sum(NW)
    widx, gidx = list(out.tagged_traces("wiener")), list(out.tagged_traces("gauss"))
    assert len(widx) == len(gidx) == sum(NW)
    summ = out.trace_summary("wiener")
    k = 0
    nroi = 0
    for ip in range(3):
        rois, wf, gf = ref[ip]
        r = exp[ip][2]
        nroi += sum(len(x) for x in rois)
        for w in range(NW[ip]):
            ch, tb, q = tr[widx[k]]
            assert ch == CH0[ip] + w and tb == 0
            np.testing.assert_array_equal(q, wf[w])
            assert summ[k] == r[w]
            ch, tb, q = tr[gidx[k]]
            assert ch == CH0[ip] + w and tb == 0
            np.testing.assert_array_equal(q, gf[w])
            k += 1
    print("refined ROIs", nroi, flush=True)
    assert nroi > 0 and any(np.any(tr[i][2] > 0) for i in widx)
    # bad ranges zeroed, input "bad" map passed through
    assert not np.any(tr[widx[5]][2][100:300])
    assert not np.any(tr[widx[NW[0] + NW[1] + 7]][2])
    assert out.masks("bad")[CH0[0] + 5] == [(100, 300)]
    # tight ROIs: the oracle's rules over the device mask
    rois = out.masks("tight_roi")
    for ip in range(3):
        y, m, r, _ = exp[ip]
        ref = O.roi_lists(m, pad=5, asy=0.1)
        for w in range(NW[ip]):
            assert rois.get(CH0[ip] + w, []) == [(a, b + 1) for (a, b) in ref[w]]
    assert any(rois.get(CH0[2] + w) for w in range(NW[2]))
    # a second frame of the same shape reuses the spectra and gives the same
    out2 = K.frame_filter("HipOmnibusSigProc:osp", _frame(K, adc, bad))
    for i in (widx[3], gidx[NW[0] + 2]):
        np.testing.assert_array_equal(out2.trace(i)[2], tr[i][2])


def _track_planes(seed, ntracks=(8, 8, 10)):
    """ADC planes with straight tracks through the forward response
    (workload's induction and collection models) plus noise."""
    import torch
    from wct_amd import decon as D
    from wct_amd import workload as W
    rng = np.random.default_rng(seed)
    dev = torch.device("cuda:0")
    out = []
    for ip, kind in enumerate(("ind", "ind", "col")):
        fk = D.ForwardKernel(W.plane_spec(kind, NW[ip], NT))
        q = torch.from_numpy(W.tracks_raster(rng, ntracks[ip], NW[ip], NT)).to(dev)
        sig = torch.empty_like(q)
        D.convolve(fk, q, sig)
        torch.cuda.synchronize()
        out.append((sig.cpu().numpy() + rng.normal(0.0, W.NOISE_ADC, (NW[ip], NT))).astype(np.float32))
    return out


def test_sigproc_refinement_on_tracks(K, tmp_path):
    """wiener and gauss on frames with straight tracks (many ROIs, linked
    across wires, broken at valleys and shrunk) equal the C ABI's decons cut
    by the oracle's refinement, with bad channels on every plane."""
    print("sigproc refinement on track frames", flush=True)
    K.configure(_sigproc_config(tmp_path))
    for seed in (21, 22):
        adc = _track_planes(seed)
        bad = {(0, 9): [(0, NT)], (1, 30): [(400, 700)], (2, 17): [(100, 120)]}
        out = K.frame_filter("HipOmnibusSigProc:osp", _frame(K, adc, bad))
        exp = _expected(adc, bad, (3.0, 3.0, 5.0))
        ref = _refined(exp, bad)
        tr = out.traces()
        widx, gidx = list(out.tagged_traces("wiener")), list(out.tagged_traces("gauss"))
        k = 0
        for ip in range(3):
            for w in range(NW[ip]):
                np.testing.assert_array_equal(tr[widx[k]][2], ref[ip][1][w])
                np.testing.assert_array_equal(tr[gidx[k]][2], ref[ip][2][w])
                k += 1
        nroi = [sum(len(x) for x in ref[ip][0]) for ip in range(3)]
        print("seed", seed, "refined ROIs per plane", nroi, flush=True)
        assert min(nroi) >= 5


def test_sigproc_loose_fallback_rows(K, tmp_path):
    """decon_2D_looseROI's tight-LF rows next to bad and lf_noisy channels
    (:1215-1231) in the component's refinement: an lf_noisy channel and a
    bad one give what the C ABI gives with those rows flagged."""
    print("sigproc loose decon fallback rows", flush=True)
    K.configure(_sigproc_config(tmp_path))
    adc = _adc_planes(11)
    bad = {(1, 20): [(0, 40)]}
    lfn = {(0, 10), (1, 33)}
    f = _frame(K, adc, bad)
    for (ip, w) in sorted(lfn):
        f.add_mask("lf_noisy", CH0[ip] + w, 0, NT)
    out = K.frame_filter("HipOmnibusSigProc:osp", f)
    exp = _expected(adc, bad, (3.0, 3.0, 5.0), lfn=lfn)
    assert _alt_rows(bad, lfn, 0)[8:13].all() and _alt_rows(bad, lfn, 1)[19:22].all()
    ref = _refined(exp, bad)
    tr = out.traces()
    widx = list(out.tagged_traces("wiener"))
    k = 0
    for ip in range(3):
        for w in range(NW[ip]):
            np.testing.assert_array_equal(tr[widx[k]][2], ref[ip][1][w])
            k += 1


def _chan_file(tmp_path, t0, tick=500.0, nresp=4, nsamp=150):
    """A PerChannelResponse file over the three planes' channels (groups
    of channels sharing a response) and its contents."""
    from test_gpu_chanresp import _channel_responses
    cr = _channel_responses(nresp, nsamp, 21)
    chmap = {CH0[ip] + w: (w + ip) % nresp for ip in range(3) for w in range(NW[ip])}
    doc = {"tick": tick, "t0": t0,
           "channels": [[[ch for ch, r in chmap.items() if r == k], cr[k].tolist()] for k in range(nresp)]}
    path = os.path.join(str(tmp_path), "chresp.json")
    with open(path, "w") as f:
        json.dump(doc, f)
    return path, cr, chmap


def test_sigproc_per_channel_response(K, tmp_path):
    """per_chan_resp: decon_2D_init's channel-by-channel correction
    (:949-982) in the component gives what the C ABI gives with the same
    factor table (the nominal response sampled from the responses' t0);
    a response binned at another tick is refused at the first frame."""
    print("sigproc per-channel correction", flush=True)
    t0 = -1000.0
    path, cr, chmap = _chan_file(tmp_path, t0)
    cfg = _sigproc_config(tmp_path, {"per_chan_resp": "PerChannelResponse:pcr", "use_roi_debug_mode": True})
    cfg.insert(-1, {"type": "PerChannelResponse", "name": "pcr", "data": {"filename": path}})
    K.configure(cfg)
    adc = _adc_planes(8)
    bad = {(0, 5): [(100, 300)]}
    out = K.frame_filter("HipOmnibusSigProc:osp", _frame(K, adc, bad))
    exp = _expected(adc, bad, (3.0, 3.0, 5.0), debug=True, chan=(cr, chmap, t0))
    plain = _expected(adc, bad, (3.0, 3.0, 5.0), debug=True)
    tr = out.traces()
    idx = list(out.tagged_traces("tight_lf"))
    k = 0
    for ip in range(3):
        y = exp[ip][0].copy()
        assert np.any(y != plain[ip][0])
        for (ipp, w), ranges in bad.items():
            if ipp == ip:
                for (b0, b1) in ranges:
                    y[w, b0:b1] = 0
        for w in range(NW[ip]):
            np.testing.assert_array_equal(tr[idx[k]][2], y[w])
            k += 1
    K.reset()
    path2, _, _ = _chan_file(tmp_path, t0, tick=400.0)
    cfg = _sigproc_config(tmp_path, {"per_chan_resp": "PerChannelResponse:pcr"})
    cfg.insert(-1, {"type": "PerChannelResponse", "name": "pcr", "data": {"filename": path2}})
    K.configure(cfg)
    from wct_amd.capi import WctError
    with pytest.raises(WctError, match="channel response size mismatch"):
        K.frame_filter("HipOmnibusSigProc:osp", _frame(K, adc, bad))


@pytest.mark.parametrize("refine", [True, False])
def test_sigproc_debug_outputs(K, tmp_path, refine):
    """use_roi_debug_mode: tight_lf (the unfiltered decon_2D_init frame on
    induction planes without refinement), loose_lf (the collection plane
    repeats its tight decon) and decon_charge, each plane in order,
    negatives kept and bad ranges zeroed, equal to the C ABI's outputs."""
    print(f"sigproc debug outputs, refinement {refine}", flush=True)
    K.configure(_sigproc_config(tmp_path, {"use_roi_debug_mode": True, "use_roi_refinement": refine}))
    adc = _adc_planes(6)
    bad = {(0, 5): [(100, 300)], (2, 8): [(500, 510)]}
    out = K.frame_filter("HipOmnibusSigProc:osp", _frame(K, adc, bad))
    exp = _expected(adc, bad, (3.0, 3.0, 5.0), debug=True, refine=refine)
    tags = ["decon_charge", "loose_lf", "tight_lf"] + (["gauss", "wiener"] if refine else [])
    assert sorted(out.trace_tags()) == sorted(tags)
    tr = out.traces()
    assert len(tr) == len(tags) * sum(NW)

    def zero_bad(ip, frame):
        f = frame.copy()
        for (ipp, w), ranges in bad.items():
            if ipp == ip:
                for (b0, b1) in ranges:
                    f[w, b0:b1] = 0
        return f

    for tag, key in (("tight_lf", "base"), ("loose_lf", "loose_dbg"), ("decon_charge", "charge")):
        idx = list(out.tagged_traces(tag))
        assert len(idx) == sum(NW)
        k = 0
        for ip in range(3):
            y, m, r, extra = exp[ip]
            want = y if (key == "base" or (key == "loose_dbg" and ip == 2)) else extra[key]
            want = zero_bad(ip, want)
            for w in range(NW[ip]):
                ch, tb, q = tr[idx[k]]
                assert ch == CH0[ip] + w
                np.testing.assert_array_equal(q, want[w])
                k += 1
    assert any(np.any(q < 0) for _, _, q in tr)
    if refine:
        assert out.masks("tight_roi")
    else:     # no tight ROIs on the induction planes without refinement
        assert all(CH0[2] <= ch < CH0[2] + NW[2] for ch in out.masks("tight_roi"))


def test_sigproc_without_refinement_saves_nothing(K, tmp_path):
    K.configure(_sigproc_config(tmp_path, {"use_roi_refinement": False}))
    out = K.frame_filter("HipOmnibusSigProc:osp", _frame(K, _adc_planes(3), {(0, 1): [(0, 10)]}))
    assert out.info()["ntraces"] == 0 and out.trace_tags() == []
    assert out.frame_tags() == ["sigproc"]
    assert out.masks("bad")[CH0[0] + 1] == [(0, 10)]


def test_sigproc_debug_and_sparse(K, tmp_path):
    print("sigproc debug + sparse", flush=True)
    # no decon_charge output: its filter would widen the kernel's kept
    # bins, and the dense tight decon is compared bit for bit below
    K.configure(_sigproc_config(tmp_path, {"use_roi_debug_mode": True, "process_planes": [2],
                                           "decon_charge_tag": ""}))
    adc = _adc_planes(4)
    bad = {(2, 3): [(10, 20)]}
    out = K.frame_filter("HipOmnibusSigProc:osp", _frame(K, adc, bad))
    assert sorted(out.trace_tags()) == ["gauss", "loose_lf", "tight_lf", "wiener"]
    tr = [out.trace(i) for i in out.tagged_traces("tight_lf")]
    assert len(tr) == NW[2] and tr[0][0] == CH0[2]
    assert any(np.any(q < 0) for _, _, q in tr)         # debug keeps negatives
    assert not np.any(tr[3][2][10:20])

    K.reset()
    K.configure(_sigproc_config(tmp_path, {"process_planes": [2]}))
    out = K.frame_filter("HipOmnibusSigProc:osp", _frame(K, adc, bad))
    dense = {ch: q for ch, _, q in (out.trace(i) for i in out.tagged_traces("wiener"))}
    assert len(dense) == NW[2] and any(np.any(q > 0) for q in dense.values())
    K.reset()
    K.configure(_sigproc_config(tmp_path, {"sparse": True, "process_planes": [2]}))
    out = K.frame_filter("HipOmnibusSigProc:osp", _frame(K, adc, bad))
    rebuilt = {}
    widx = list(out.tagged_traces("wiener"))
    summ = out.trace_summary("wiener")
    assert len(summ) == len(widx)
    for i in widx:
        ch, tb, q = out.trace(i)
        assert q.size > 0 and np.all(q > 0)
        rebuilt.setdefault(ch, np.zeros(NT, np.float32))[tb:tb + q.size] = q
    for ch, q in dense.items():
        np.testing.assert_array_equal(rebuilt.get(ch, np.zeros(NT, np.float32)), q)


def test_sigproc_between_frame_files(K, tmp_path):
    """FrameFileSource -> HipOmnibusSigProc -> FrameFileSink: an ADC
    archive in the reference's layout (int16 frame, column by column, a
    bad-channel mask) gives an output archive whose wiener frame, channels
    and RMS summary equal the component's in-memory output."""
    import io
    import tarfile
    print("frame files around sigproc, 3 planes", NW, "x", NT, flush=True)
    adc = _adc_planes(5)
    adc16 = [np.round(a * 8).astype(np.int16) for a in adc]   # integer ADC counts
    bad = {(1, 4): [(200, 260)]}
    src, dst = tmp_path / "adc.tar.gz", tmp_path / "sp.tar.gz"
    rows = np.concatenate(adc16)
    chans = np.concatenate([np.arange(CH0[ip], CH0[ip] + NW[ip], dtype=np.int32) for ip in range(3)])
    members = [("frame_orig_17.npy", np.asfortranarray(rows)), ("channels_orig_17.npy", chans),
               ("tickinfo_orig_17.npy", np.array([3.0, 500.0, 0.0])),
               ("chanmask_bad_17.npy", np.asfortranarray(np.array([[CH0[1] + 4, 200, 260]], np.int32)))]
    with tarfile.open(src, "w:gz") as tf:
        for name, arr in members:
            b = io.BytesIO()
            np.save(b, arr)
            ti = tarfile.TarInfo(name)
            ti.size = b.tell()
            b.seek(0)
            tf.addfile(ti, b)
    cfg = _sigproc_config(tmp_path) + [
        {"type": "FrameFileSource", "name": "in", "data": {"inname": str(src), "tags": ["orig"]}},
        {"type": "FrameFileSink", "name": "out", "data": {"outname": str(dst), "tags": ["wiener"]}}]
    K.configure(cfg)
    f = K.frame_source("FrameFileSource:in")
    out = K.frame_filter("HipOmnibusSigProc:osp", f)
    K.frame_sink("FrameFileSink:out", out)
    assert K.frame_source("FrameFileSource:in") is None
    K.frame_sink("FrameFileSink:out", None)
    # the same frame built in memory
    direct = K.frame_filter("HipOmnibusSigProc:osp", _frame(K, [a.astype(np.float32) for a in adc16], bad))
    want = np.stack([direct.trace(i)[2] for i in direct.tagged_traces("wiener")])
    got = {}
    with tarfile.open(dst) as tf:
        for m in tf.getmembers():
            got[m.name] = np.load(io.BytesIO(tf.extractfile(m).read()))
    assert sorted(got) == sorted(["channels_wiener_17.npy", "chanmask_bad_17.npy", "chanmask_tight_roi_17.npy",
                                  "frame_wiener_17.npy", "summary_wiener_17.npy", "tickinfo_wiener_17.npy"])
    np.testing.assert_array_equal(got["frame_wiener_17.npy"], want)
    np.testing.assert_array_equal(got["channels_wiener_17.npy"], chans)
    np.testing.assert_array_equal(got["summary_wiener_17.npy"], direct.trace_summary("wiener"))
    assert list(got["tickinfo_wiener_17.npy"]) == [3.0, 500.0, 0.0]
    assert not np.any(got["frame_wiener_17.npy"][NW[0] + 4, 200:260])
    rois = direct.masks("tight_roi")
    flat = sorted((c, a, b) for c, rl in rois.items() for (a, b) in rl)
    assert sorted(map(tuple, got["chanmask_tight_roi_17.npy"].tolist())) == flat and flat
