"""Device paths of the tensor-set deconvolution (sig/src/Decon2DResponse.cxx,
Decon2DFilter.cxx): the spectrum mode of the fused plan and
wct_spectrum_filter, against the FP64 oracle on seeded frames.

Tolerances:
  spectrum (Decon2DResponse)     exact Hermitian symmetry (real DC and
      Nyquist, bin nt - k the conjugate of bin k); weighted by the time
      filter the spectrum later meets in Decon2DFilter, relative RMS
      error <= 1e-4 against FP64.  Unweighted, the spectrum near Nyquist
      is the noise divided by a response that has fallen to ~1e-20 of
      its peak, below the FP64 transform's own rounding floor (1e-16 of
      the peak): there the FP64 oracle, this path and the reference's
      FP32 arithmetic all divide by rounding noise and cannot agree.
      Weighted, the bins past the fused path's spectral cut (filtered
      kernel below 1e-7 of its peak, where the response is near that
      floor) still leave a few 1e-5 of disagreement between any two
      evaluations; the fused path drops them (2-3e-7), this mode keeps
      them as the reference does.
  filter pass (Decon2DFilter)    relative RMS error <= 5e-5 against the
      FP64 oracle on the same spectrum (restore_baseline's median can
      land on a neighbouring sample, which shifts a row)
  the pair, with the Wiener and LF filters, against the FP64 fused
      decon + restore_baseline: relative RMS error <= 5e-5 (as for the
      filter pass; measured 3-5e-7)
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

CASES = [("ind", 800, 6000, 33), ("col", 960, 6000, 40), ("ind", 96, 1000, 6)]


def _rel_rms(a, b):
    return float(np.sqrt(np.sum(np.abs(a - b) ** 2) / np.sum(np.abs(b) ** 2)))


def _setup(kind, nw, nt, ntr):
    from wct_amd import workload as W
    spec = W.plane_spec(kind, nw, nt)
    resp = O.overall_response(spec.fr_rows, spec.elec, W.FINE_PERIOD, W.TICK, W.INTER_GAIN,
                              W.ADC_MV, W.FTOFFSET, nt, dtype=np.float64)
    seed = 3 * nw + nt
    q = O.tracks(seed, 1, ntr, nw, nt, W.Q_PER_TICK)
    adc = O.forward_conv(q, resp, spec.wire_shift, spec.time_shift)
    adc += O.normal(seed, 2, nw * nt, W.NOISE_ADC).reshape(nw, nt)
    return spec, resp, adc.astype(np.float32)


def _gpu_spectrum(spec, resp, adc, dev):
    import torch
    from wct_amd import decon
    pk = decon.PlaneKernel(spec, cut_tol=0.0, time_filter=np.ones(spec.nt), resp=resp)
    assert pk.kc == spec.nt // 2 + 1
    a = torch.from_numpy(adc).to(dev)
    sp = torch.empty((spec.nw, spec.nt), dtype=torch.complex64, device=dev)
    b = decon.Batch([decon.make_job(pk, a, None, spectrum=sp)])
    b.run()
    torch.cuda.synchronize()
    b.close()
    return sp


def _time_filter(spec, kind):
    # the tight Wiener filter, times the tight LF filter on induction
    # planes (plane_spec folds it in): a Decon2DFilter "filters" list of
    # HfFilter:Wiener_tight_* (+ LfFilter:ROI_tight_lf)
    return np.asarray(spec.time_filter, np.float64)


@pytest.mark.parametrize("kind,nw,nt,ntr", CASES)
def test_spectrum_mode_vs_fp64(dev, kind, nw, nt, ntr):
    print(f"tensor-set spectrum {kind} {nw}x{nt}", flush=True)
    spec, resp, adc = _setup(kind, nw, nt, ntr)
    sp = _gpu_spectrum(spec, resp, adc, dev).cpu().numpy()
    assert np.all(np.isfinite(sp))
    k = np.arange(nt)
    assert np.array_equal(sp[:, (-k) % nt], np.conj(sp))
    assert np.all(sp[:, 0].imag == 0) and np.all(sp[:, nt // 2].imag == 0)
    ref = O.decon2d_response(adc.astype(np.float64), resp, spec.wire_filter, spec.wire_shift, spec.time_shift)
    w = _time_filter(spec, kind)[None, :]
    e = _rel_rms(sp.astype(np.complex128) * w, ref * w)
    print(f"  filter-weighted spectrum rel_rms={e:.3e}")
    assert e <= 1e-4


@pytest.mark.parametrize("kind,nw,nt,ntr", CASES)
def test_spectrum_filter_vs_fp64(dev, kind, nw, nt, ntr):
    import torch
    from wct_amd import decon
    print(f"tensor-set filter {kind} {nw}x{nt}", flush=True)
    spec, resp, adc = _setup(kind, nw, nt, ntr)
    sp = _gpu_spectrum(spec, resp, adc, dev)
    tf = _time_filter(spec, kind)
    ft = torch.from_numpy(tf[: nt // 2 + 1].astype(np.float32)).to(dev)
    out = torch.empty((nw, nt), device=dev)
    decon.spectrum_filter(sp, ft, out, restore=1)
    torch.cuda.synchronize()
    y = out.cpu().numpy().astype(np.float64)
    # the same spectrum through the FP64 oracle
    ref_same = O.decon2d_filter(sp.cpu().numpy().astype(np.complex128), tf)
    e1 = _rel_rms(y, ref_same)
    # the pair against the FP64 fused decon + restore_baseline
    y64 = O.decon_fused(adc.astype(np.float64), resp, spec.wire_filter, tf, spec.wire_shift, spec.time_shift)
    ref_pair, _ = O.restore_baseline(y64, np.float64)
    e2 = _rel_rms(y, ref_pair)
    print(f"  filter rel_rms={e1:.3e} pair vs fused rel_rms={e2:.3e}")
    assert e1 <= 5e-5 and e2 <= 5e-5
    # without the baseline the pass is linear: a scaled filter scales the output
    out2 = torch.empty_like(out)
    decon.spectrum_filter(sp, ft * 2.0, out2, restore=0)
    decon.spectrum_filter(sp, ft, out, restore=0)
    torch.cuda.synchronize()
    assert torch.allclose(out2, 2.0 * out, rtol=1e-6, atol=1e-6 * float(out.abs().max()))


def test_spectrum_mode_rejects_mixed_jobs(dev):
    import torch
    from wct_amd import decon
    from wct_amd.capi import WctError
    spec, resp, adc = _setup("ind", 96, 1000, 2)
    pk = decon.PlaneKernel(spec, cut_tol=0.0, time_filter=np.ones(spec.nt), resp=resp)
    a = torch.from_numpy(adc).to(dev)
    sp = torch.empty((96, 1000), dtype=torch.complex64, device=dev)
    o = torch.empty((96, 1000), device=dev)
    with pytest.raises(WctError, match="spectrum mode"):
        decon.Batch([decon.make_job(pk, a, None, spectrum=sp), decon.make_job(pk, a, o)])
    bad = decon.make_job(pk, a, None, spectrum=sp)
    bad.out = o.data_ptr()
    with pytest.raises(WctError, match="outputs of the decon mode"):
        decon.Batch([bad])
    pk.close()


# ---- the components through the C front (column-by-column tensors) ------

def _components_setup(K, tmp_path, nch=96, nt=1000, ntr=6):
    from wct_amd import workload as W
    from test_components import _tensorset_config
    K.reset()
    K.configure(_tensorset_config(tmp_path, resp_extra={
        "gain": W.GAIN, "shaping_time": W.SHAPING, "inter_gain": W.INTER_GAIN, "ADC_mV": W.ADC_MV,
        "ftoffset": W.FTOFFSET, "ctoffset": W.CTOFFSET, "iplane": 0}))
    spec, resp, adc = _setup("ind", nch, nt, ntr)
    s = K.TensorSet(7, {"tick": W.TICK, "run": 3})
    s.add_array(adc, "f4", {"tag": "trace_tag", "type": "waveform", "tbin": 0.0, "pad": 0.0})
    s.add("i4", (nch,), np.arange(100, 100 + nch, dtype=np.int32), {"tag": "trace_tag", "type": "channels"})
    s.add("f8", (nch,), np.linspace(1.0, 2.0, nch), {"tag": "trace_tag", "type": "summary"})
    return spec, resp, adc, s


def test_tensorset_components_match_device_paths(dev, tmp_path):
    import torch
    from wct_amd import components as K
    from wct_amd import decon
    print("Decon2DResponse -> Decon2DFilter, 96 x 1000", flush=True)
    spec, resp, adc, s = _components_setup(K, tmp_path)
    r = K.tensorset_filter("Decon2DResponse:r", s)
    info = r.info()
    assert info["ident"] == 7 and info["metadata"]["tags"] == ["trace_tag"] and info["metadata"]["run"] == 3
    iw = r.find("trace_tag", "waveform")
    dt, shape, _, md = r.tensor(iw)
    assert dt == "c8" and shape == (96, 1000) and md["type"] == "waveform" and md["tag"] == "trace_tag"
    assert r.find("trace_tag", "channels") is not None and r.find("trace_tag", "summary") is not None
    sp = r.array(iw)
    # the device path on the same frame and response
    want = _gpu_spectrum(spec, decon.overall_response(spec), adc, dev).cpu().numpy()
    assert np.abs(sp - want).max() <= 1e-5 * np.abs(want).max()
    f = K.tensorset_filter("Decon2DFilter:f", r)
    iy = f.find("trace_tag", "waveform")
    assert f.tensor(iy)[0] == "f4" and f.find("trace_tag", "channels") is not None
    y = f.array(iy).astype(np.float64)
    tf = _time_filter(spec, "ind")
    ref_same = O.decon2d_filter(sp.astype(np.complex128), tf)
    y64 = O.decon_fused(adc.astype(np.float64), resp, spec.wire_filter, tf, spec.wire_shift, spec.time_shift)
    ref_pair, _ = O.restore_baseline(y64, np.float64)
    e1, e2 = _rel_rms(y, ref_same), _rel_rms(y, ref_pair)
    print(f"  component filter rel_rms={e1:.3e} pair vs fused rel_rms={e2:.3e}")
    assert e1 <= 5e-5 and e2 <= 5e-5
    assert K.tensorset_filter("Decon2DResponse:r", None) is None
    K.reset()


def test_decon2d_response_per_channel(dev, tmp_path):
    """per_chan_resp (Decon2DResponse.cxx:280-308): the component's
    spectrum equals the device path's with the factor table of the
    frame's channels (nominal response Response::ColdElec(gain,
    shaping_time) from the responses' t0); a tick mismatch is refused."""
    import json
    import torch
    from wct_amd import components as K
    from wct_amd import decon, host
    from wct_amd import workload as W
    from wct_amd.capi import WctError
    from test_components import _tensorset_config
    from test_gpu_chanresp import _channel_responses
    print("Decon2DResponse per-channel correction, 96 x 1000", flush=True)
    nch, nt, t0 = 96, 1000, 2000.0
    cr = _channel_responses(3, 120, 4)
    chans = np.arange(100, 100 + nch, dtype=np.int32)
    rows = (np.arange(nch) // 5) % 3
    for tick, ok in ((W.TICK, True), (400.0, False)):
        doc = {"tick": tick, "t0": t0,
               "channels": [[[int(c) for c, r in zip(chans, rows) if r == k], cr[k].tolist()] for k in range(3)]}
        path = tmp_path / "chresp.json"
        path.write_text(json.dumps(doc))
        K.reset()
        cfg = _tensorset_config(tmp_path, resp_extra={
            "gain": W.GAIN, "shaping_time": W.SHAPING, "inter_gain": W.INTER_GAIN, "ADC_mV": W.ADC_MV,
            "ftoffset": W.FTOFFSET, "ctoffset": W.CTOFFSET, "iplane": 0,
            "per_chan_resp": "PerChannelResponse:pcr"})
        cfg.insert(0, {"type": "PerChannelResponse", "name": "pcr", "data": {"filename": str(path)}})
        K.configure(cfg)
        spec, resp, adc = _setup("ind", nch, nt, 6)
        s = K.TensorSet(7, {"tick": W.TICK})
        s.add_array(adc, "f4", {"tag": "trace_tag", "type": "waveform"})
        s.add("i4", (nch,), chans, {"tag": "trace_tag", "type": "channels"})
        if not ok:
            with pytest.raises(WctError, match="channel response size mismatch"):
                K.tensorset_filter("Decon2DResponse:r", s)
            continue
        r = K.tensorset_filter("Decon2DResponse:r", s)
        sp = r.array(r.find("trace_tag", "waveform"))
        ewave = host.coldelec_samples(nt, t0, t0 + nt * W.TICK, W.GAIN, W.SHAPING, 1.0)
        table = decon.channel_factors(cr, ewave, nt, device=dev)
        pk = decon.PlaneKernel(spec, cut_tol=0.0, time_filter=np.ones(nt), resp=decon.overall_response(spec))
        a = torch.from_numpy(adc).to(dev)
        want = torch.empty((nch, nt), dtype=torch.complex64, device=dev)
        b = decon.Batch([decon.make_job(pk, a, None, spectrum=want,
                                        chan=(table, torch.from_numpy(rows.astype(np.int32)).to(dev)))])
        b.run()
        torch.cuda.synchronize()
        want = want.cpu().numpy()
        assert np.abs(sp - want).max() <= 1e-5 * np.abs(want).max()
        plain = _gpu_spectrum(spec, decon.overall_response(spec), adc, dev).cpu().numpy()
        assert np.abs(plain - want).max() > 1e-3 * np.abs(want).max()
    K.reset()


def test_decon2d_filter_cmm_ranges_and_empty_rows(dev, tmp_path):
    from wct_amd import components as K
    from wct_amd.capi import WctError
    print("Decon2DFilter cmm ranges", flush=True)
    _, _, _, s = _components_setup(K, tmp_path)
    r = K.tensorset_filter("Decon2DResponse:r", s)
    sp = r.array(r.find("trace_tag", "waveform"))
    base = K.tensorset_filter("Decon2DFilter:f", r)
    y0 = base.array(base.find("trace_tag", "waveform"))
    t = K.TensorSet(8, {})
    t.add_array(sp, "c8", {"tag": "trace_tag", "type": "waveform"})
    t.add("i4", (96,), np.arange(96, dtype=np.int32), {"tag": "trace_tag", "type": "channels"})
    ranges = np.array([[5, 40], [0, 1000], [10, 20]], np.float64)     # (tmin, tmax) per range
    chans = np.array([3.0, 95.0, 500.0])                               # 500: past the last row, skipped
    t.add_array(ranges, "f8", {"tag": "trace_tag", "type": "bad:cmm_range"})
    t.add("f8", (3,), chans, {"tag": "trace_tag", "type": "bad:cmm_channel"})
    o = K.tensorset_filter("Decon2DFilter:f", t)
    y = o.array(o.find("trace_tag", "waveform"))
    keep = np.ones_like(y, bool)
    keep[3, 5:40] = False
    keep[95, :] = False
    assert np.all(y[~keep] == 0) and np.array_equal(y[keep], y0[keep])
    # Sig::restore_baseline refuses a row with no non-zero sample
    z = sp.copy()
    z[17, :] = 0
    u = K.TensorSet(9, {})
    u.add_array(z, "c8", {"tag": "trace_tag", "type": "waveform"})
    u.add("i4", (96,), np.arange(96, dtype=np.int32), {"tag": "trace_tag", "type": "channels"})
    with pytest.raises(WctError, match="empty waveform"):
        K.tensorset_filter("Decon2DFilter:f", u)
    K.reset()
