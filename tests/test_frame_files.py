"""Frame files on the CPU: FrameFileSink writes the reference's archive
layout (sio/src/FrameFileSink.cxx:123-258: per tag frame / channels /
tickinfo / summary .npy members, Eigen arrays column by column, then
chanmask_<name>_<ident>.npy), checked member by member with tarfile and
numpy; FrameFileSource reads archives written here by numpy in that
layout (sio/src/FrameFileSource.cxx:125-397: framelets per tag, tag
filtering, trace tags with summaries, channel masks, one end of stream
and then false).  The reference ships no frame archive, so the expected
arrays come from a direct restatement of Aux::fill
(aux/src/FrameTools.cxx:181-223) in numpy below."""
import io
import tarfile

import numpy as np
import pytest


@pytest.fixture(autouse=True)
def _fresh():
    from wct_amd import components as K
    K.reset()
    yield
    K.reset()


def _members(path):
    """[(name, array, header)] of an archive, in order."""
    out = []
    with tarfile.open(path) as tf:
        for m in tf.getmembers():
            raw = tf.extractfile(m).read()
            f = io.BytesIO(raw)
            ver = np.lib.format.read_magic(f)
            rd = np.lib.format.read_array_header_1_0 if ver == (1, 0) else np.lib.format.read_array_header_2_0
            shape, fortran, dt = rd(f)
            hdr = {"descr": dt.str, "fortran_order": fortran, "shape": shape}
            out.append((m.name, np.load(io.BytesIO(raw)), hdr))
    return out


def _dense(traces, channels, tb0, tb1, baseline=0.0, scale=1.0, offset=0.0):
    """Aux::fill: charges summed in trace order into rows by channel,
    clipped to [tb0, tb1), in float32 arithmetic."""
    arr = np.full((len(channels), tb1 - tb0), np.float32(baseline), np.float32)
    row = {c: i for i, c in enumerate(channels)}
    for ch, tb, q in traces:
        if ch not in row:
            continue
        for k, v in enumerate(np.asarray(q, np.float32)):
            col = tb + k - tb0
            if 0 <= col < tb1 - tb0:
                arr[row[ch], col] += v
    return arr * np.float32(scale) + np.float32(offset)


TRACES = [(3, 2, [1.0, 2.0, 3.0]), (1, 0, [4.0, 5.0]), (3, 4, [10.0]), (8, 1, [-2.7, 0.5])]


def _frame(K, ident=7, time=1.5, tick=500.0):
    f = K.Frame(ident, time, tick)
    for ch, tb, q in TRACES:
        f.add_trace(ch, q, tb)
    f.tag_traces("gauss", [0, 1], [0.5, 0.25])
    f.tag_frame("sigproc")
    f.add_mask("bad", 1, 0, 3)
    f.add_mask("bad", 5, 2, 4)
    f.add_mask("bad", 5, 7, 9)
    return f


def test_sink_writes_reference_layout(tmp_path):
    from wct_amd import components as K
    path = tmp_path / "frames.tar.gz"
    K.configure([{"type": "FrameFileSink", "name": "out",
                  "data": {"outname": str(path), "tags": ["gauss", "*"]}}])
    K.frame_sink("FrameFileSink:out", _frame(K))
    K.frame_sink("FrameFileSink:out", None)
    mem = _members(path)
    assert [m[0] for m in mem] == [
        "frame_gauss_7.npy", "channels_gauss_7.npy", "tickinfo_gauss_7.npy", "summary_gauss_7.npy",
        "frame_*_7.npy", "channels_*_7.npy", "tickinfo_*_7.npy", "chanmask_bad_7.npy"]
    d = {m[0]: (m[1], m[2]) for m in mem}
    # the tagged traces: channels 1 and 3, ticks [0, 5)
    fr, hdr = d["frame_gauss_7.npy"]
    assert hdr["descr"] == "<f4" and hdr["fortran_order"] is True and hdr["shape"] == (2, 5)
    np.testing.assert_array_equal(fr, _dense(TRACES[:2], [1, 3], 0, 5))
    ch, hdr = d["channels_gauss_7.npy"]
    assert hdr["descr"] == "<i4" and list(ch) == [1, 3]
    ti, hdr = d["tickinfo_gauss_7.npy"]
    assert hdr["descr"] == "<f8" and list(ti) == [1.5, 500.0, 0.0]
    sm, hdr = d["summary_gauss_7.npy"]
    assert hdr["descr"] == "<f8" and list(sm) == [0.5, 0.25]
    # every trace: channels 1, 3, 8, ticks [0, 5); the two channel-3
    # traces share the row and sum where they overlap (tick 4)
    fr, hdr = d["frame_*_7.npy"]
    assert hdr["shape"] == (3, 5)
    np.testing.assert_array_equal(fr, _dense(TRACES, [1, 3, 8], 0, 5))
    assert fr[1, 4] == 13.0 and fr[1, 2] == 1.0 and fr[2, 1] == np.float32(-2.7)
    assert list(d["tickinfo_*_7.npy"][0]) == [1.5, 500.0, 0.0]
    cm, hdr = d["chanmask_bad_7.npy"]
    assert hdr["descr"] == "<i4" and hdr["fortran_order"] is True and hdr["shape"] == (3, 3)
    assert sorted(map(tuple, cm.tolist())) == [(1, 0, 3), (5, 2, 4), (5, 7, 9)]


def test_sink_dense_digitize_scale(tmp_path):
    from wct_amd import components as K
    path = tmp_path / "adc.tar"
    cfg = {"outname": str(path), "tags": [], "digitize": True, "baseline": 2048.4, "scale": 1.5,
           "offset": -0.25, "masks": False, "dense": {"chbeg": 2, "chend": 6, "tbbeg": 1, "tbend": 4}}
    K.configure([{"type": "FrameFileSink", "name": "adc", "data": cfg}])
    K.frame_sink("FrameFileSink:adc", _frame(K, ident=3))
    K.frame_sink("FrameFileSink:adc", None)
    mem = _members(path)
    assert [m[0] for m in mem] == ["frame_*_3.npy", "channels_*_3.npy", "tickinfo_*_3.npy"]
    fr, hdr = mem[0][1], mem[0][2]
    assert hdr["descr"] == "<i2" and hdr["fortran_order"] is True and hdr["shape"] == (4, 3)
    want = _dense(TRACES, [2, 3, 4, 5], 1, 4, 2048.4, 1.5, -0.25).astype(np.int16)   # truncation
    np.testing.assert_array_equal(fr, want)
    assert list(mem[1][1]) == [2, 3, 4, 5]
    assert list(mem[2][1]) == [1.5, 500.0, 1.0]


def test_sink_digitize_truncates_toward_zero(tmp_path):
    from wct_amd import components as K
    path = tmp_path / "neg.tar"
    K.configure([{"type": "FrameFileSink", "name": "n",
                  "data": {"outname": str(path), "digitize": True, "masks": False}}])
    f = K.Frame(0, 0.0, 500.0)
    f.add_trace(0, [-2.7, -0.5, 0.5, 2.7, -32768.0, 32767.0])
    K.frame_sink("FrameFileSink:n", f)
    K.frame_sink("FrameFileSink:n", None)
    fr = _members(path)[0][1]
    assert fr.dtype == np.int16 and fr.tolist() == [[-2, 0, 0, 2, -32768, 32767]]


def _write_archive(path, members):
    """A reference-layout archive written by numpy: members is a list of
    (name, array); 2D arrays are saved column by column."""
    mode = "w:gz" if str(path).endswith(".gz") else "w"
    with tarfile.open(path, mode) as tf:
        for name, arr in members:
            b = io.BytesIO()
            np.save(b, np.asfortranarray(arr) if np.ndim(arr) == 2 else arr)
            ti = tarfile.TarInfo(name)
            ti.size = b.tell()
            b.seek(0)
            tf.addfile(ti, b)


def _written_frame(rng, ident, tag, nch, nt, ch0, tb0, dtype=np.float32):
    q = (rng.standard_normal((nch, nt)) * 100).astype(dtype)
    return q, [(f"frame_{tag}_{ident}.npy", q),
               (f"channels_{tag}_{ident}.npy", np.arange(ch0, ch0 + nch, dtype=np.int32)),
               (f"tickinfo_{tag}_{ident}.npy", np.array([0.25 * ident, 500.0, tb0], np.float64))]


@pytest.mark.parametrize("suffix", [".tar", ".tar.gz"])
def test_source_reads_numpy_written_archive(tmp_path, suffix):
    from wct_amd import components as K
    rng = np.random.default_rng(5)
    path = tmp_path / ("in" + suffix)
    qa, ma = _written_frame(rng, 4, "orig", 3, 7, 100, 0, np.int16)
    qb, mb = _written_frame(rng, 4, "gauss", 2, 5, 200, 2)
    qc, mc = _written_frame(rng, 9, "*", 4, 6, 0, 1)
    summ = [("summary_gauss_4.npy", np.array([1.5, 2.5]))]
    masks = [("chanmask_bad_4.npy", np.array([[100, 0, 2], [101, 3, 5], [100, 4, 6]], np.int32))]
    _write_archive(path, ma + mb + summ + masks + mc)
    K.configure([{"type": "FrameFileSource", "name": "in",
                  "data": {"inname": str(path), "tags": ["orig", "gauss"], "frame_tags": ["loaded"]}}])
    f = K.frame_source("FrameFileSource:in")
    info = f.info()
    assert info == dict(ident=4, time=1.0, tick=500.0, ntraces=5)
    tr = f.traces()
    for i in range(3):
        assert tr[i][0] == 100 + i and tr[i][1] == 0
        np.testing.assert_array_equal(tr[i][2], qa[i].astype(np.float32))
    for i in range(2):
        assert tr[3 + i][0] == 200 + i and tr[3 + i][1] == 2
        np.testing.assert_array_equal(tr[3 + i][2], qb[i])
    assert f.frame_tags() == ["loaded"]
    assert sorted(f.trace_tags()) == ["gauss", "orig"]
    assert list(f.tagged_traces("orig")) == [0, 1, 2]
    assert list(f.tagged_traces("gauss")) == [3, 4]
    assert list(f.trace_summary("gauss")) == [1.5, 2.5]
    assert f.trace_summary("orig").size == 0
    assert f.masks("bad") == {100: [(0, 2), (4, 6)], 101: [(3, 5)]}
    # second frame: the "*" framelet is read whatever the tags, untagged
    g = K.frame_source("FrameFileSource:in")
    assert g.info() == dict(ident=9, time=2.25, tick=500.0, ntraces=4)
    assert g.trace_tags() == []
    np.testing.assert_array_equal(np.stack([t[2] for t in g.traces()]), qc)
    assert K.frame_source("FrameFileSource:in") is None   # end of stream
    assert K.frame_source("FrameFileSource:in") is False  # nothing after it


def test_source_skips_tags_not_asked_for(tmp_path):
    from wct_amd import components as K
    rng = np.random.default_rng(6)
    path = tmp_path / "in.tar"
    _, ma = _written_frame(rng, 1, "orig", 3, 4, 0, 0)
    qb, mb = _written_frame(rng, 1, "gauss", 2, 4, 10, 0)
    _write_archive(path, ma + mb)
    K.configure([{"type": "FrameFileSource", "name": "g", "data": {"inname": str(path), "tags": ["gauss"]}}])
    f = K.frame_source("FrameFileSource:g")
    assert f.info()["ntraces"] == 2 and f.trace_tags() == ["gauss"]
    np.testing.assert_array_equal(np.stack([t[2] for t in f.traces()]), qb)
    # no tags asked for: only untagged framelets, so nothing here
    K.configure([{"type": "FrameFileSource", "name": "n", "data": {"inname": str(path)}}])
    assert K.frame_source("FrameFileSource:n") is None


def test_sink_source_round_trip(tmp_path):
    from wct_amd import components as K
    path = tmp_path / "rt.tar.gz"
    K.configure([{"type": "FrameFileSink", "name": "o", "data": {"outname": str(path), "tags": ["gauss"]}},
                 {"type": "FrameFileSource", "name": "i", "data": {"inname": str(path), "tags": ["gauss"]}}])
    for ident in (1, 2):
        K.frame_sink("FrameFileSink:o", _frame(K, ident=ident, time=ident * 0.5))
    K.frame_sink("FrameFileSink:o", None)
    for ident in (1, 2):
        f = K.frame_source("FrameFileSource:i")
        assert f.info() == dict(ident=ident, time=ident * 0.5, tick=500.0, ntraces=2)
        got = f.traces()
        want = _dense(TRACES[:2], [1, 3], 0, 5)
        assert [g[0] for g in got] == [1, 3] and all(g[1] == 0 for g in got)
        np.testing.assert_array_equal(np.stack([g[2] for g in got]), want)
        assert list(f.trace_summary("gauss")) == [0.5, 0.25]
        assert f.masks("bad") == {1: [(0, 3)], 5: [(2, 4), (7, 9)]}
    assert K.frame_source("FrameFileSource:i") is None
    assert K.frame_source("FrameFileSource:i") is False


def test_empty_archive_is_one_end_of_stream(tmp_path):
    from wct_amd import components as K
    path = tmp_path / "empty.tar"
    K.configure([{"type": "FrameFileSink", "name": "o", "data": {"outname": str(path)}}])
    K.frame_sink("FrameFileSink:o", None)
    assert _members(path) == []
    K.configure([{"type": "FrameFileSource", "name": "i", "data": {"inname": str(path)}}])
    assert K.frame_source("FrameFileSource:i") is None
    assert K.frame_source("FrameFileSource:i") is False


@pytest.mark.parametrize("members,msg", [
    ([("frame_a_b_1.npy", np.zeros((1, 1), np.float32))], "cannot parse"),
    ([("frame_x_1.npy", np.zeros((2, 3), np.float64))], "float32 or int16"),
    ([("frame_x_1.npy", np.zeros((2, 3), np.float32)), ("channels_x_1.npy", np.arange(3, dtype=np.int32)),
      ("tickinfo_x_1.npy", np.zeros(3))], "differ in size"),
    ([("chanmask_bad_1.npy", np.zeros((2, 2), np.int32))], "{n, 3}"),
])
def test_source_input_errors(tmp_path, members, msg):
    from wct_amd import components as K
    path = tmp_path / "bad.tar"
    _write_archive(path, members)
    K.configure([{"type": "FrameFileSource", "name": "i", "data": {"inname": str(path), "tags": ["x"]}}])
    with pytest.raises(K.WctError, match=msg.replace("{", r"\{").replace("}", r"\}")):
        K.frame_source("FrameFileSource:i")


def test_configuration_errors(tmp_path):
    from wct_amd import components as K
    with pytest.raises(K.WctError, match="unsupported outname"):
        K.configure([{"type": "FrameFileSink", "name": "o", "data": {"outname": str(tmp_path / "x.zip")}}])
    K.reset()
    with pytest.raises(K.WctError):
        K.configure([{"type": "FrameFileSource", "name": "i", "data": {"inname": str(tmp_path / "none.tar")}}])
