"""The C++ component layer on the CPU: configuration, filters, the ROI
list builder against the oracle and brute-force cases, frame handles,
and the frame filter's end-of-stream / empty-frame / configuration error
behaviour (OmnibusSigProc.cxx:54-217, :1382-1400)."""
import numpy as np
import pytest

import oracle as O


@pytest.fixture(autouse=True)
def _fresh():
    from wct_amd import components as K
    K.reset()
    yield
    K.reset()


def test_filters_match_reference_formula():
    from wct_amd import components as K
    from wct_amd import workload as W
    K.configure([
        {"type": "HfFilter", "name": "Wiener_tight_W", "data": W.WIENER},
        {"type": "HfFilter", "name": "Wire_col", "data": W.WIRE},
        {"type": "LfFilter", "name": "ROI_tight_lf", "data": W.LF_TIGHT},
    ])
    n = 6000
    f = np.arange(n) * 2 * W.MHZ / n
    f = np.where(f > W.MHZ, f - 2 * W.MHZ, f)
    hf = np.exp(-0.5 * (np.abs(f) / W.WIENER["sigma"]) ** W.WIENER["power"])
    hf[0] = 0.0   # flag: zero at DC
    np.testing.assert_allclose(K.filter_waveform("HfFilter:Wiener_tight_W", n), hf.astype(np.float32),
                               rtol=1e-6, atol=1e-30)
    lf = 1 - np.exp(-(np.abs(f) / W.LF_TIGHT["tau"]) ** 2)
    np.testing.assert_allclose(K.filter_waveform("LfFilter:ROI_tight_lf", n), lf.astype(np.float32),
                               rtol=1e-6, atol=1e-7)
    nw = 480
    g = np.arange(nw) * 2.0 / nw
    g = np.where(g > 1.0, g - 2.0, g)
    np.testing.assert_allclose(K.filter_waveform("HfFilter:Wire_col", nw),
                               np.exp(-0.5 * (np.abs(g) / 0.75) ** 2).astype(np.float32), rtol=1e-6)


def test_defaults_follow_reference_headers():
    from wct_amd import components as K
    # HfFilter() = (1 MHz, 3 MHz, 2, true): at f = 0.5 MHz exp(-0.5 (1/6)^2)
    n = 8
    v = K.filter_waveform("HfFilter:dflt", n)
    assert v[0] == 0.0
    assert v[2] == pytest.approx(np.exp(-0.5 * (0.5 / 3.0) ** 2), rel=1e-6)
    # LfFilter() tau = 0.02 MHz
    v = K.filter_waveform("LfFilter:dflt", n)
    assert v[1] == pytest.approx(1 - np.exp(-(0.25 / 0.02) ** 2), rel=1e-6)


def _random_mask(rng, nw, nt, density, runlen):
    m = np.zeros((nw, nt), np.uint8)
    for w in range(nw):
        k = rng.poisson(density * nt / runlen)
        for _ in range(k):
            a = rng.integers(0, nt)
            m[w, a:a + rng.integers(1, 2 * runlen)] = rng.integers(1, 4)
    return m


@pytest.mark.parametrize("seed,nw,nt,density,runlen,pad,asy", [
    (1, 40, 300, 0.05, 6, 5, 0.1), (2, 64, 1000, 0.15, 12, 5, 0.1), (3, 30, 200, 0.4, 3, 2, 0.3),
    (4, 80, 500, 0.02, 20, 5, 0.1), (5, 16, 64, 0.9, 8, 0, 0.5)])
def test_roi_lists_match_oracle(seed, nw, nt, density, runlen, pad, asy):
    from wct_amd import components as K
    rng = np.random.default_rng(seed)
    m = _random_mask(rng, nw, nt, density, runlen)
    assert K.roi_lists(m, pad, asy) == O.roi_lists(m, pad, asy)


def test_roi_lists_edge_cases():
    from wct_amd import components as K
    nt = 50
    m = np.zeros((4, nt), np.uint8)
    m[0, nt - 1] = 1            # a run cannot start on the last tick
    m[1, :] = 1                 # whole row
    m[2, 0] = 1                 # clamped at 0 after padding
    m[3, 10:13] = 1
    m[3, 14] = 1                # a separate run: 13 is below threshold
    got = K.roi_lists(m, pad=5, asy=0.1)
    assert got == O.roi_lists(m, pad=5, asy=0.1)
    assert got[0] == []
    assert got[1] == [(0, nt - 1)]
    assert got[2][0] == (0, 5)
    assert K.roi_lists(np.zeros((0, 10), np.uint8)) == []
    assert K.roi_lists(np.zeros((3, 1), np.uint8)) == [[], [], []]


def test_roi_connect_by_brute_force():
    """Wire 1 gains the midpoint ROI of wires 0 and 2 (similar lengths),
    and not when lengths differ or it already overlaps."""
    from wct_amd import components as K
    nt = 200
    m = np.zeros((3, nt), np.uint8)
    m[0, 50:61] = 1
    m[2, 54:65] = 1
    rois = K.roi_lists(m, pad=0, asy=0.1)
    assert rois[1] == [(52, 62)]
    m[2, 54:100] = 1
    assert K.roi_lists(m, pad=0, asy=0.1)[1] == []
    m[2, 54:65] = 1
    m[2, 65:] = 0
    m[1, 55:58] = 1
    assert K.roi_lists(m, pad=0, asy=0.1)[1] == [(55, 57)]


def test_frame_handles_round_trip():
    from wct_amd import components as K
    f = K.Frame(7, 2.5, 500.0)
    f.add_trace(11, [1.0, -2.0, 3.0], tbin=4)
    f.add_trace(12, np.arange(5))
    f.add_mask("bad", 11, 0, 3)
    f.add_mask("bad", 11, 10, 20)
    assert f.info() == dict(ident=7, time=2.5, tick=500.0, ntraces=2)
    ch, tb, q = f.trace(0)
    assert (ch, tb) == (11, 4) and list(q) == [1.0, -2.0, 3.0]
    assert f.masks("bad") == {11: [(0, 3), (10, 20)]}
    assert f.masks("none") == {}


def _osp_config(extra=None):
    data = {"anode": "AnodeChannels:a0"}
    data.update(extra or {})
    return [{"type": "AnodeChannels", "name": "a0", "data": {"plane_ranges": [[0, 8], [100, 8], [200, 10]]}},
            {"type": "HipDFT", "name": ""},
            {"type": "HipOmnibusSigProc", "name": "osp", "data": data}]


def test_sigproc_eos_and_empty_frame():
    from wct_amd import components as K
    K.configure(_osp_config())
    assert K.frame_filter("HipOmnibusSigProc:osp", None) is None
    out = K.frame_filter("HipOmnibusSigProc:osp", K.Frame(3, 1.5, 500.0))
    assert out.info() == dict(ident=3, time=1.5, tick=500.0, ntraces=0)
    assert out.frame_tags() == []


@pytest.mark.parametrize("extra,msg", [({"fft_flag": 1}, "fft_flag"),
                                       ({"process_planes": [0, 3]}, "process_planes"),
                                       ({"anode": "AnodeChannels:missing_planes"}, "plane"),
                                       ({"use_multi_plane_protection": True}, "wire geometry"),
                                       ({"response_floor": 1.5}, "response_floor")])
def test_sigproc_configuration_errors(extra, msg):
    from wct_amd import components as K
    from wct_amd.capi import WctError
    with pytest.raises(WctError, match=msg):
        K.configure(_osp_config(extra))


def test_unknown_component_type():
    from wct_amd import components as K
    from wct_amd.capi import WctError
    with pytest.raises(WctError, match="unknown component type"):
        K.configure([{"type": "NoSuchThing", "name": "x"}])


def test_configuration_parse_error():
    from wct_amd.host import lib
    rc = lib().wcth_configure(b'[{"type": "HfFilter", "name": "x", "data": {"sigma": }}]')
    assert rc != 0
    assert "JSON parse error" in lib().wcth_last_error().decode()


def test_sigproc_full_configuration_parses_field_response(tmp_path):
    """The three-plane configuration of the device tests, including a
    field-response file whose currents reach subnormal values."""
    from wct_amd import components as K
    from test_gpu_components import _sigproc_config
    cfg = _sigproc_config(tmp_path)
    text = open(tmp_path / "fr.json").read()
    assert "e-310" in text or "e-309" in text
    K.configure(cfg)
    assert K.frame_filter("HipOmnibusSigProc:osp", None) is None


# ---- tensor-set deconvolution components (no device work) --------------

def _tensorset_config(tmpdir, resp_extra=None, filt_extra=None):
    import json as _json
    import os as _os
    from wct_amd import components as K
    from wct_amd import workload as W
    planes = []
    for ip, kind in enumerate(("ind", "ind", "col")):
        pos, cur = W.fr_paths(kind)
        planes.append((ip, 0.0, W.PITCH, pos, cur))
    path = _os.path.join(str(tmpdir), "fr.json")
    with open(path, "w") as f:
        _json.dump(K.field_response_json(planes, W.ORIGIN, 0.0, W.FINE_PERIOD, W.SPEED), f)
    rdata = {"anode": "AnodeChannels:apa0", "field_response": "FieldResponse:synth", "per_chan_resp": ""}
    rdata.update(resp_extra or {})
    fdata = {"filters": ["HfFilter:Wiener_tight_U", "LfFilter:ROI_tight_lf"]}
    fdata.update(filt_extra or {})
    return [
        {"type": "AnodeChannels", "name": "apa0", "data": {"ident": 0, "plane_ranges": [[0, 96], [96, 96], [192, 96]]}},
        {"type": "FieldResponse", "name": "synth", "data": {"filename": path}},
        {"type": "HfFilter", "name": "Wiener_tight_U", "data": W.WIENER},
        {"type": "HfFilter", "name": "Wire_ind", "data": dict(W.WIRE)},
        {"type": "HfFilter", "name": "Wire_col", "data": dict(W.WIRE)},
        {"type": "LfFilter", "name": "ROI_tight_lf", "data": W.LF_TIGHT},
        {"type": "Decon2DResponse", "name": "r", "data": rdata},
        {"type": "Decon2DFilter", "name": "f", "data": fdata},
    ]


def test_tensorset_handles_roundtrip():
    from wct_amd import components as K
    rng = np.random.default_rng(3)
    a = rng.standard_normal((5, 7)).astype(np.float32)
    s = K.TensorSet(ident=42, metadata={"tick": 500.0, "note": "x"})
    s.add_array(a, "f4", {"tag": "t", "type": "waveform"})
    s.add("i4", (5,), np.arange(5, dtype=np.int32), {"tag": "t", "type": "channels"})
    info = s.info()
    assert info["ident"] == 42 and info["ntensors"] == 2 and info["metadata"]["tick"] == 500.0
    dt, shape, flat, md = s.tensor(0)
    assert dt == "f4" and shape == (5, 7) and md == {"tag": "t", "type": "waveform"}
    assert np.array_equal(flat, a.ravel(order="F"))       # column by column
    assert np.array_equal(s.array(0), a)
    assert s.find("t", "channels") == 1 and s.find("t", "summary") is None
    # the C front refuses an unknown element type
    import ctypes
    from wct_amd.host import lib as hlib
    one = (ctypes.c_size_t * 1)(1)
    z = np.zeros(1, np.float32)
    assert hlib().wcth_tensorset_add(s.h, b"zz", 1, one, z.ctypes.data, b"{}") == -1


def test_decon2d_configuration_errors(tmp_path):
    from wct_amd import components as K
    from wct_amd.capi import WctError
    K.reset()
    cfg = _tensorset_config(tmp_path)
    bad = [dict(c) for c in cfg]
    bad[-2] = {"type": "Decon2DResponse", "name": "r", "data": {"per_chan_resp": ""}}   # anode "Anode"
    with pytest.raises(WctError, match="no anode"):
        K.configure(bad)
    K.reset()
    bad[-2] = {"type": "Decon2DResponse", "name": "r",
               "data": {"anode": "AnodeChannels:apa0", "per_chan_resp": "NoSuchResponse:x"}}
    with pytest.raises(WctError, match="m_cresp"):
        K.configure(bad)
    K.reset()
    bad[-2] = {"type": "Decon2DResponse", "name": "r",
               "data": {"anode": "AnodeChannels:apa0", "per_chan_resp": "", "field_response": "Nope:x"}}
    with pytest.raises(WctError, match="no field response"):
        K.configure(bad)
    K.reset()


def test_decon2d_eos_and_input_errors(tmp_path):
    from wct_amd import components as K
    from wct_amd.capi import WctError
    K.reset()
    K.configure(_tensorset_config(tmp_path))
    assert K.tensorset_filter("Decon2DResponse:r", None) is None
    assert K.tensorset_filter("Decon2DFilter:f", None) is None
    s = K.TensorSet(0, {"tick": 500.0})
    s.add("i4", (4,), np.arange(4, dtype=np.int32), {"tag": "trace_tag", "type": "channels"})
    with pytest.raises(WctError, match="no waveform tensor"):
        K.tensorset_filter("Decon2DResponse:r", s)
    s = K.TensorSet(0, {})
    s.add_array(np.ones((4, 10), np.float32), "f4", {"tag": "trace_tag", "type": "waveform"})
    with pytest.raises(WctError, match="no channels tensor"):
        K.tensorset_filter("Decon2DResponse:r", s)
    with pytest.raises(WctError, match="dtype f4, expected c8"):
        K.tensorset_filter("Decon2DFilter:f", s)
    s = K.TensorSet(0, {})
    s.add_array(np.ones((4, 11), np.complex64), "c8", {"tag": "trace_tag", "type": "waveform"})
    s.add("i4", (4,), np.arange(4, dtype=np.int32), {"tag": "trace_tag", "type": "channels"})
    with pytest.raises(WctError, match="odd tick count"):
        K.tensorset_filter("Decon2DFilter:f", s)
    K.reset()
    K.configure(_tensorset_config(tmp_path, filt_extra={"filters": ["NoSuchFilter:x"]}))
    s = K.TensorSet(0, {})
    s.add_array(np.ones((4, 10), np.complex64), "c8", {"tag": "trace_tag", "type": "waveform"})
    s.add("i4", (4,), np.arange(4, dtype=np.int32), {"tag": "trace_tag", "type": "channels"})
    with pytest.raises(WctError, match="no filter"):
        K.tensorset_filter("Decon2DFilter:f", s)
    K.reset()


def test_per_channel_response_file(tmp_path):
    """SigProc::PerChannelResponse (sigproc/src/PerChannelResponse.cxx:
    28-78): single channels and channel groups, float samples, the
    binning of the first entry, KeyError for a channel without a
    response, and its configuration errors."""
    import json
    from wct_amd import components as K
    from wct_amd.capi import WctError
    r1 = [0.0, 1.5, 2.25, -0.5]
    r2 = [1.0, 2.0, 3.0, 4.0, 0.1]
    path = tmp_path / "chresp.json"
    path.write_text(json.dumps({"tick": 500.0, "t0": 1000.0, "channels": [[7, r1], [[8, 9, 10], r2]]}))
    K.reset()
    K.configure([{"type": "PerChannelResponse", "name": "pcr", "data": {"filename": str(path)}}])
    v, t0, bs = K.channel_response("PerChannelResponse:pcr", 7)
    assert list(v) == r1 and t0 == 1000.0 and bs == 500.0
    for ch in (8, 9, 10):
        v, _, _ = K.channel_response("PerChannelResponse:pcr", ch)
        np.testing.assert_array_equal(v, np.array(r2, np.float32))
    with pytest.raises(WctError, match="no response for channel 11"):
        K.channel_response("PerChannelResponse:pcr", 11)
    bad = [({"tick": 500.0, "t0": 0.0}, "no channels given"),
           ({"tick": 500.0, "t0": 0.0, "channels": [[1, []]]}, "zero length response")]
    for i, (doc, msg) in enumerate(bad):
        p = tmp_path / f"bad{i}.json"
        p.write_text(json.dumps(doc))
        with pytest.raises(WctError, match=msg):
            K.configure([{"type": "PerChannelResponse", "name": f"b{i}", "data": {"filename": str(p)}}])
    with pytest.raises(WctError, match="must supply a PerChannelResponse filename"):
        K.configure([{"type": "PerChannelResponse", "name": "nofile", "data": {"filename": ""}}])
    K.reset()
