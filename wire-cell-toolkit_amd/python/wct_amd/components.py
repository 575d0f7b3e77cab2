"""The C++ component layer of libWireCellHip.so from Python.

configure() takes the same list of {"type", "name", "data"} objects the
reference's application reads (apps/src/Main.cxx:330-350); dft() calls
an IDFT component through its virtual interface; Frame / frame_filter()
push an IFrame through an IFrameFilter (IFunctionNode.h:50) and read
back the traces, tags, summaries and channel masks of the result.
"""
import ctypes as C
import json

import numpy as np

from .capi import WctError
from .host import lib

OPS = {"fwd1d": 0, "inv1d": 1, "fwd1b": 2, "inv1b": 3, "fwd2d": 4, "inv2d": 5,
       "transpose_real": 6, "transpose_complex": 7}


def _check(rc):
    if rc != 0:
        raise WctError(lib().wcth_last_error().decode())


def configure(cfg_list):
    _check(lib().wcth_configure(json.dumps(cfg_list).encode()))


def reset():
    _check(lib().wcth_reset())


def known_types():
    n = lib().wcth_known_types(None, 0)
    buf = C.create_string_buffer(n + 1)
    lib().wcth_known_types(buf, n + 1)
    return buf.value.decode().split()


def dft(tn, op, arr, axis=1):
    """Runs IDFT op on a host array: complex64 for the transforms (1-d or
    2-d) and complex64 or float32 for transpose.  Returns a new array."""
    a = np.ascontiguousarray(arr)
    if op == "transpose":
        op = "transpose_real" if a.dtype == np.float32 else "transpose_complex"
        if a.dtype not in (np.float32, np.complex64):
            raise TypeError("transpose takes float32 or complex64")
    elif a.dtype != np.complex64:
        raise TypeError("IDFT transforms take complex64")
    if a.ndim == 1:
        nrows, ncols = a.shape[0], 1
    else:
        nrows, ncols = a.shape
    out = np.empty_like(a)
    _check(lib().wcth_dft(tn.encode(), OPS[op], a.ctypes.data, out.ctypes.data, nrows, ncols, axis))
    if op.startswith("transpose"):
        out = out.reshape(ncols, nrows)
    return out


def filter_waveform(tn, n):
    out = np.zeros(n, np.float32)
    _check(lib().wcth_filter_waveform(tn.encode(), n, out.ctypes.data))
    return out


def channel_response(tn, channel):
    """(samples, tmin, binsize) of an IChannelResponse's channel."""
    import ctypes as C
    n, t0, bs = C.c_int(), C.c_double(), C.c_double()
    _check(lib().wcth_channel_response(tn.encode(), channel, None, 0, C.byref(n), C.byref(t0), C.byref(bs)))
    out = np.zeros(n.value, np.float32)
    _check(lib().wcth_channel_response(tn.encode(), channel, out.ctypes.data, n.value, None, None, None))
    return out, t0.value, bs.value


class Frame:
    """An IFrame handle (input under construction, or a filter result)."""

    def __init__(self, ident=0, time=0.0, tick=500.0, _handle=None):
        self.h = _handle if _handle is not None else lib().wcth_frame_new(ident, time, tick)

    def add_trace(self, channel, charge, tbin=0):
        q = np.ascontiguousarray(charge, np.float32)
        _check(lib().wcth_frame_add_trace(self.h, int(channel), int(tbin), q.size, q.ctypes.data))

    def add_mask(self, name, channel, first, second):
        _check(lib().wcth_frame_add_mask(self.h, name.encode(), int(channel), int(first), int(second)))

    def tag_frame(self, tag):
        _check(lib().wcth_frame_tag_frame(self.h, tag.encode()))

    def tag_traces(self, tag, indices, summary=None):
        idx = np.ascontiguousarray(indices, np.uint64)
        sm = np.ascontiguousarray(summary if summary is not None else [], np.float64)
        _check(lib().wcth_frame_tag_traces(self.h, tag.encode(), idx.ctypes.data if idx.size else None, idx.size,
                                           sm.ctypes.data if sm.size else None, sm.size))

    def info(self):
        ident, ntr = C.c_int(), C.c_int()
        time, tick = C.c_double(), C.c_double()
        _check(lib().wcth_frame_info(self.h, C.byref(ident), C.byref(time), C.byref(tick), C.byref(ntr)))
        return dict(ident=ident.value, time=time.value, tick=tick.value, ntraces=ntr.value)

    def trace(self, i):
        ch, tb, n = C.c_int(), C.c_int(), C.c_int()
        p = C.POINTER(C.c_float)()
        _check(lib().wcth_frame_trace(self.h, i, C.byref(ch), C.byref(tb), C.byref(n), C.byref(p)))
        q = np.ctypeslib.as_array(p, shape=(n.value,)).copy() if n.value else np.zeros(0, np.float32)
        return ch.value, tb.value, q

    def traces(self):
        return [self.trace(i) for i in range(self.info()["ntraces"])]

    def _tags(self, which):
        n = lib().wcth_frame_tags(self.h, which, -1, None)
        out = []
        for i in range(n):
            s = C.c_char_p()
            lib().wcth_frame_tags(self.h, which, i, C.byref(s))
            out.append(s.value.decode())
        return out

    def frame_tags(self):
        return self._tags(0)

    def trace_tags(self):
        return self._tags(1)

    def tagged_traces(self, tag):
        p = C.POINTER(C.c_size_t)()
        n = C.c_int()
        _check(lib().wcth_frame_tagged(self.h, tag.encode(), C.byref(p), C.byref(n)))
        return np.ctypeslib.as_array(p, shape=(n.value,)).copy() if n.value else np.zeros(0, np.uint64)

    def trace_summary(self, tag):
        p = C.POINTER(C.c_double)()
        n = C.c_int()
        _check(lib().wcth_frame_summary(self.h, tag.encode(), C.byref(p), C.byref(n)))
        return np.ctypeslib.as_array(p, shape=(n.value,)).copy() if n.value else np.zeros(0)

    def masks(self, name):
        """{channel: [(first, second), ...]} of one channel-mask-map entry."""
        n = lib().wcth_frame_masks(self.h, name.encode(), None, None, None, 0)
        if n < 0:
            raise WctError("bad frame handle")
        ch = np.zeros(max(n, 1), np.int32)
        lo = np.zeros(max(n, 1), np.int32)
        hi = np.zeros(max(n, 1), np.int32)
        lib().wcth_frame_masks(self.h, name.encode(), ch.ctypes.data, lo.ctypes.data, hi.ctypes.data, n)
        out = {}
        for c, a, b in zip(ch[:n], lo[:n], hi[:n]):
            out.setdefault(int(c), []).append((int(a), int(b)))
        return out

    def close(self):
        if self.h:
            lib().wcth_frame_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def frame_filter(tn, frame):
    """Calls IFrameFilter tn on frame (None = end of stream)."""
    out = C.c_void_p()
    _check(lib().wcth_frame_filter(tn.encode(), frame.h if frame is not None else None, C.byref(out)))
    return Frame(_handle=out.value) if out.value else None


class TensorSet:
    """An ITensorSet handle (input under construction, or a filter
    result).  2D waveform tensors of shape (nrows, ncols) keep their
    elements column by column, as the reference's tensor-set nodes map
    them (aux/inc/WireCellAux/Util.h:119-151): add_array / array take a
    numpy (nrows, ncols) array and do that layout."""

    _DT = {"f4": np.float32, "f8": np.float64, "c8": np.complex64, "c16": np.complex128, "i4": np.int32,
           "i8": np.int64, "u1": np.uint8, "i2": np.int16, "u4": np.uint32}

    def __init__(self, ident=0, metadata=None, _handle=None):
        if _handle is not None:
            self.h = _handle
        else:
            self.h = lib().wcth_tensorset_new(int(ident), json.dumps(metadata or {}).encode())
            if not self.h:
                _check(-1)

    def add(self, dtype, shape, data, metadata=None):
        """Adds a tensor whose store is the bytes of data (C order of the
        given array, copied)."""
        a = np.ascontiguousarray(data, self._DT[dtype])
        sh = (C.c_size_t * max(len(shape), 1))(*shape)
        _check(lib().wcth_tensorset_add(self.h, dtype.encode(), len(shape), sh, a.ctypes.data if a.size else None,
                                        json.dumps(metadata or {}).encode()))

    def add_array(self, arr, dtype, metadata=None):
        """A 2D (nrows, ncols) array stored column by column."""
        a = np.asarray(arr)
        self.add(dtype, a.shape, np.asfortranarray(a, self._DT[dtype]).ravel(order="F"), metadata)

    def info(self):
        ident, n = C.c_int(), C.c_int()
        ln = lib().wcth_tensorset_info(self.h, C.byref(ident), C.byref(n), None, 0)
        if ln < 0:
            raise WctError("bad tensor set handle")
        buf = C.create_string_buffer(ln + 1)
        lib().wcth_tensorset_info(self.h, None, None, buf, ln + 1)
        return dict(ident=ident.value, ntensors=n.value, metadata=json.loads(buf.value.decode() or "null"))

    def tensor(self, i):
        """(dtype, shape, flat data copy, metadata) of tensor i."""
        dt = C.create_string_buffer(8)
        nd = C.c_int()
        sh = (C.c_size_t * 8)()
        p, nb = C.c_void_p(), C.c_size_t()
        _check(lib().wcth_tensorset_tensor(self.h, int(i), dt, C.byref(nd), sh, C.byref(p), C.byref(nb)))
        dtype = dt.value.decode()
        shape = tuple(sh[k] for k in range(nd.value))
        raw = C.string_at(p.value, nb.value) if nb.value else b""
        flat = np.frombuffer(raw, self._DT[dtype]).copy()
        ln = lib().wcth_tensorset_tensor_md(self.h, int(i), None, 0)
        buf = C.create_string_buffer(ln + 1)
        lib().wcth_tensorset_tensor_md(self.h, int(i), buf, ln + 1)
        return dtype, shape, flat, json.loads(buf.value.decode() or "null")

    def array(self, i):
        """Tensor i as a (nrows, ncols) array (column-by-column store)."""
        dtype, shape, flat, _ = self.tensor(i)
        return flat.reshape(shape, order="F")

    def find(self, tag, type_):
        """Index of the first tensor with metadata tag and type, or None."""
        for i in range(self.info()["ntensors"]):
            md = self.tensor(i)[3] or {}
            if md.get("tag") == tag and md.get("type") == type_:
                return i
        return None

    def close(self):
        if self.h:
            lib().wcth_tensorset_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def tensorset_filter(tn, tset):
    """Calls ITensorSetFilter tn on tset (None = end of stream)."""
    out = C.c_void_p()
    _check(lib().wcth_tensorset_filter(tn.encode(), tset.h if tset is not None else None, C.byref(out)))
    return TensorSet(_handle=out.value) if out.value else None


def frame_sink(tn, frame):
    """Passes frame (None = end of stream) to IFrameSink tn."""
    _check(lib().wcth_frame_sink(tn.encode(), frame.h if frame is not None else None))


def frame_source(tn):
    """The next frame of IFrameSource tn: a Frame, None at end of stream,
    and False once the source has nothing more after that."""
    out = C.c_void_p()
    rc = lib().wcth_frame_source(tn.encode(), C.byref(out))
    if rc == 1:
        return False
    _check(rc)
    return Frame(_handle=out.value) if out.value else None


def roi_lists(mask, pad=5, asy=0.1, cap=1 << 22):
    """Tight ROI lists (inclusive tick pairs per wire) from a threshold mask."""
    m = np.ascontiguousarray(mask, np.uint8)
    nw, nt = m.shape
    cnt = np.zeros(nw, np.int32)
    pairs = np.zeros(2 * cap, np.int32)
    tot = lib().wcth_roi_lists(m.ctypes.data, nw, nt, nt, pad, asy, cnt.ctypes.data, pairs.ctypes.data, cap)
    if tot < 0:
        raise WctError("roi_lists failed" if tot == -1 else "ROI capacity exceeded")
    out, k = [], 0
    for w in range(nw):
        out.append([(int(pairs[2 * (k + i)]), int(pairs[2 * (k + i) + 1])) for i in range(cnt[w])])
        k += cnt[w]
    return out


def _csr(lists, nw):
    cnt = np.array([len(lists[w]) if w < len(lists) else 0 for w in range(nw)], np.int32)
    flat = [v for w in range(nw) for pair in (lists[w] if w < len(lists) else []) for v in pair]
    return cnt, np.ascontiguousarray(flat if flat else [0, 0], np.int32)


def _rows(cnt, pairs):
    out, k = [], 0
    for c in cnt:
        out.append([(int(pairs[2 * (k + i)]), int(pairs[2 * (k + i) + 1])) for i in range(c)])
        k += c
    return out


def loose_rois(r, bad=None, rms=None, tight=None, rebin=6, l_factor=3.5, l_max_th=10000.0, l_factor1=0.7,
               l_short_length=3, l_jump_one_bin=0, cap=1 << 22):
    """find_ROI_loose (+ extend_ROI_loose by the tight lists) of a plane's
    loose decon r (nw x nt float32).  bad: per-row lists of inclusive
    (lo, hi) tick ranges; rms: per-row RMS; tight: per-row inclusive ROIs
    or None."""
    a = np.ascontiguousarray(r, np.float32)
    nw, nt = a.shape
    keep = []
    bp = blo = bhi = None
    if bad is not None:
        ptr = np.zeros(nw + 1, np.int32)
        lo, hi = [], []
        for w in range(nw):
            for (x, y) in (bad[w] if w < len(bad) else []):
                lo.append(x)
                hi.append(y)
            ptr[w + 1] = len(lo)
        keep += [ptr, np.ascontiguousarray(lo or [0], np.int32), np.ascontiguousarray(hi or [0], np.int32)]
        bp, blo, bhi = (k.ctypes.data for k in keep[-3:])
    rp = None
    if rms is not None:
        keep.append(np.ascontiguousarray(rms, np.float32))
        rp = keep[-1].ctypes.data
    tc = tpairs = None
    if tight is not None:
        keep += list(_csr(tight, nw))
        tc, tpairs = keep[-2].ctypes.data, keep[-1].ctypes.data
    cnt = np.zeros(nw, np.int32)
    pairs = np.zeros(2 * cap, np.int32)
    tot = lib().wcth_loose_rois(a.ctypes.data, nw, nt, nt, bp, blo, bhi, rp, tc, tpairs, rebin, l_factor, l_max_th,
                                l_factor1, l_short_length, l_jump_one_bin, cnt.ctypes.data, pairs.ctypes.data, cap)
    if tot < 0:
        raise WctError(lib().wcth_last_error().decode() if tot == -1 else "ROI capacity exceeded")
    return _rows(cnt, pairs)


def apply_roi(r, rois, linear):
    """ROI_refinement::apply_roi on a copy of r (float32)."""
    a = np.array(r, np.float32, copy=True, order="C")
    nw, nt = a.shape
    cnt, flat = _csr(rois, nw)
    _check(lib().wcth_apply_roi(a.ctypes.data, nw, nt, nt, cnt.ctypes.data, flat.ctypes.data, int(linear)))
    return a


def find_peaks(signal, max_peaks=200, sigma=1.0, threshold=0.05):
    """PeakFinding::find_peak: (positions, heights) as float64 arrays."""
    a = np.ascontiguousarray(signal, np.float32)
    cap = max(1, a.size)
    x = np.zeros(cap)
    y = np.zeros(cap)
    n = lib().wcth_find_peaks(a.ctypes.data, a.size, max_peaks, sigma, threshold, x.ctypes.data, y.ctypes.data, cap)
    if n < 0:
        raise WctError(lib().wcth_last_error().decode() if n == -1 else "peak capacity exceeded")
    return x[:n].copy(), y[:n].copy()


REFINE_KEYS = ("r_th_factor", "r_fake_signal_low_th", "r_fake_signal_high_th", "r_fake_signal_low_th_ind_factor",
               "r_fake_signal_high_th_ind_factor", "r_pad", "r_break_roi_loop", "r_th_peak", "r_sep_peak",
               "r_low_peak_sep_threshold_pre", "r_max_npeaks", "r_sigma", "r_th_percent")
REFINE_DEFAULTS = (3.0, 500.0, 1000.0, 1.0, 1.0, 5, 2, 3.0, 6.0, 1200.0, 200, 2.0, 0.1)


class RoiRefinement:
    """ROI_refinement of one plane through the C front (wcth_refine_*).
    plane 0/1 induction, 2 collection; params: dict of OmnibusSigProc's
    r_* keys (defaults for the rest)."""

    def __init__(self, plane, nw, nt, params=None, wrapped=False):
        vals = dict(zip(REFINE_KEYS, REFINE_DEFAULTS))
        vals.update(params or {})
        self._par = np.array([vals[k] for k in REFINE_KEYS], np.float64)
        self.nw, self.nt = nw, nt
        h = lib().wcth_refine_new(plane, nw, nt, self._par.ctypes.data, int(bool(wrapped)))
        if not h:
            raise WctError(lib().wcth_last_error().decode())
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().wcth_refine_free(h)
            self._h = None

    def load(self, r, rms, tight, loose=None, bad=None):
        a = np.ascontiguousarray(r, np.float32)
        assert a.shape == (self.nw, self.nt)
        rm = np.ascontiguousarray(rms, np.float32)
        tc, tp = _csr(tight, self.nw)
        lc = lp = None
        keep = []
        if loose is not None:
            lc, lp = _csr(loose, self.nw)
            keep += [lc, lp]
        bp = blo = bhi = None
        if bad is not None:
            ptr = np.zeros(self.nw + 1, np.int32)
            lo, hi = [], []
            for w in range(self.nw):
                for (x, y) in bad[w]:
                    lo.append(x)
                    hi.append(y)
                ptr[w + 1] = len(lo)
            keep += [ptr, np.ascontiguousarray(lo or [0], np.int32), np.ascontiguousarray(hi or [0], np.int32)]
            bp, blo, bhi = (k.ctypes.data for k in keep[-3:])
        rc = lib().wcth_refine_load(self._h, a.ctypes.data, self.nt, bp, blo, bhi, rm.ctypes.data, tc.ctypes.data,
                                    tp.ctypes.data, None if lc is None else lc.ctypes.data,
                                    None if lp is None else lp.ctypes.data)
        _check(rc)

    def step(self, name):
        _check(lib().wcth_refine_step(self._h, name.encode()))

    def rois(self, extended=True, cap=1 << 22):
        cnt = np.zeros(self.nw, np.int32)
        pairs = np.zeros(2 * cap, np.int32)
        tot = lib().wcth_refine_rois(self._h, int(bool(extended)), cnt.ctypes.data, pairs.ctypes.data, cap)
        if tot < 0:
            raise WctError(lib().wcth_last_error().decode() if tot == -1 else "ROI capacity exceeded")
        return _rows(cnt, pairs)

    def contents(self, wire):
        """The contents of the wire's ROIs, concatenated in list order."""
        cap = 4 * self.nt
        while True:
            out = np.zeros(cap, np.float32)
            n = lib().wcth_refine_contents(self._h, wire, out.ctypes.data, cap)
            if n == -1:
                raise WctError(lib().wcth_last_error().decode())
            if n >= 0:
                return out[:n].copy()
            cap *= 4


def field_response_json(planes, origin, tstart, period, speed):
    """A field response in the reference's JSON schema
    (util/src/Response.cxx:32-72).  planes: list of (planeid, location,
    pitch, pitchpos[], current[npaths, nsamp])."""
    pl = []
    for planeid, location, pitch, pos, cur in planes:
        paths = [{"PathResponse": {"current": {"array": {"shape": [len(c)], "elements": [float(x) for x in c]}},
                                   "pitchpos": float(p), "wirepos": 0.0}}
                 for p, c in zip(pos, cur)]
        pl.append({"PlaneResponse": {"paths": paths, "planeid": planeid, "location": location,
                                     "pitch": pitch}})
    return {"FieldResponse": {"planes": pl, "axis": [1.0, 0.0, 0.0], "origin": origin, "tstart": tstart,
                              "period": period, "speed": speed}}
