"""ctypes/numpy front end of the CPU oracle (oracle/wct_oracle.c).

TEST INFRASTRUCTURE ONLY.  Only tests/, the smoke check and the
CPU-baseline leg of bench.py may import this module, and only as the
checker.  Every function cites the reference code it restates.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


def build():
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libwct_oracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        I, D, F = C.c_int, C.c_double, C.c_float
        sig = {
            "wo_dft_naive": (None, [I, f64p, f64p, I]),
            "wo_fwd1d_f": (I, [f32p, f32p, I]), "wo_inv1d_f": (I, [f32p, f32p, I]),
            "wo_fwd1b_f": (I, [f32p, f32p, I, I, I]), "wo_inv1b_f": (I, [f32p, f32p, I, I, I]),
            "wo_fwd2d_f": (I, [f32p, f32p, I, I]), "wo_inv2d_f": (I, [f32p, f32p, I, I]),
            "wo_fwd1d_d": (I, [f64p, f64p, I]), "wo_inv1d_d": (I, [f64p, f64p, I]),
            "wo_fwd1b_d": (I, [f64p, f64p, I, I, I]), "wo_inv1b_d": (I, [f64p, f64p, I, I, I]),
            "wo_fwd2d_d": (I, [f64p, f64p, I, I]), "wo_inv2d_d": (I, [f64p, f64p, I, I]),
            "wo_hermitian_mirror_f": (None, [f32p, I]), "wo_hermitian_mirror_d": (None, [f64p, I]),
            "wo_percentile_f": (F, [f32p, C.c_size_t, F]),
            "wo_percentile_d": (D, [f64p, C.c_size_t, F]),
            "wo_restore_baseline_f": (None, [f32p, I, I, f32p]),
            "wo_restore_baseline_d": (None, [f64p, I, I, f64p]),
            "wo_overall_response_f": (I, [f32p, I, I, f32p, I, D, D, D, D, D, I, f32p]),
            "wo_overall_response_d": (I, [f64p, I, I, f64p, I, D, D, D, D, D, I, f64p]),
            "wo_decon_init_f": (I, [f32p, I, I, f32p, I, f32p, I, I, f32p]),
            "wo_decon_init_d": (I, [f64p, I, I, f64p, I, f64p, I, I, f64p]),
            "wo_decon_init_cr_f": (I, [f32p, I, I, f32p, I, f32p, I, I, C.c_void_p, I, C.c_void_p, f32p]),
            "wo_decon_init_cr_d": (I, [f64p, I, I, f64p, I, f64p, I, I, C.c_void_p, I, C.c_void_p, f64p]),
            "wo_decon_apply_filter_f": (I, [f32p, I, I, f32p, f32p]),
            "wo_decon_apply_filter_d": (I, [f64p, I, I, f64p, f64p]),
            "wo_decon_fused_d": (I, [f64p, I, I, f64p, I, f64p, f64p, I, I, f64p]),
            "wo_forward_conv_d": (I, [f64p, I, I, f64p, I, I, I, f64p]),
            "wo_cal_rms_f": (D, [f32p, C.c_size_t]),
            "wo_cal_rms_d": (D, [f64p, C.c_size_t]),
            "wo_roi_threshold_f": (None, [f32p, C.c_void_p, I, I, F, i32p, i32p, i32p, I, f64p, f64p, u8p]),
            "wo_roi_threshold_d": (None, [f64p, C.c_void_p, I, I, F, i32p, i32p, i32p, I, f64p, f64p, u8p]),
            "wo_roi_lists": (I, [u8p, I, I, I, F, i32p, i32p, I]),
            "wo_loose_rois": (I, [f32p, I, I, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p, I, D, D, D, I, I, i32p, i32p, I]),
            "wo_apply_roi": (None, [f32p, I, I, i32p, i32p, I]),
            "wo_hf_filter": (None, [I, D, D, D, I, f64p]),
            "wo_lf_filter": (None, [I, D, D, f64p]),
            "wo_coldelec": (D, [D, D, D]),
            "wo_elec_samples": (None, [I, D, D, D, D, D, C.c_void_p, C.c_void_p]),
            "wo_fft_best_length": (C.c_long, [C.c_long, I]),
            "wo_wire_region_average": (I, [I, f64p, f32p, I, D, I, i32p, f32p]),
            "wo_uniform": (D, [C.c_uint64, C.c_uint64, C.c_uint64]),
            "wo_normal": (None, [C.c_uint64, C.c_uint64, C.c_size_t, D, f64p]),
            "wo_tracks": (None, [C.c_uint64, C.c_uint64, I, I, I, D, f64p]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def _cplx(a, dtype):
    a = np.ascontiguousarray(a, dtype=np.complex128 if dtype == np.float64 else np.complex64)
    return a.view(dtype).copy()


def _uncplx(a):
    return a.view(np.complex128 if a.dtype == np.float64 else np.complex64)


def _sfx(dtype):
    return "_d" if dtype == np.float64 else "_f"


# ---- FFT / IDFT semantics (aux/src/FftwDFT.cxx:88-217) ----

def dft_naive(x, sign=-1):
    x = _cplx(x, np.float64)
    n = x.size // 2
    out = np.zeros_like(x)
    lib().wo_dft_naive(n, x, out, sign)
    return _uncplx(out)


def fwd1d(x, dtype=np.float32):
    a = _cplx(x, dtype)
    getattr(lib(), "wo_fwd1d" + _sfx(dtype))(a, a, a.size // 2)
    return _uncplx(a)


def inv1d(x, dtype=np.float32):
    a = _cplx(x, dtype)
    getattr(lib(), "wo_inv1d" + _sfx(dtype))(a, a, a.size // 2)
    return _uncplx(a)


def fwd1b(x, axis, dtype=np.float32):
    nr, nc = x.shape
    a = _cplx(x, dtype)
    getattr(lib(), "wo_fwd1b" + _sfx(dtype))(a, a, nr, nc, axis)
    return _uncplx(a).reshape(nr, nc)


def inv1b(x, axis, dtype=np.float32):
    nr, nc = x.shape
    a = _cplx(x, dtype)
    getattr(lib(), "wo_inv1b" + _sfx(dtype))(a, a, nr, nc, axis)
    return _uncplx(a).reshape(nr, nc)


def fwd2d(x, dtype=np.float32):
    nr, nc = x.shape
    a = _cplx(x, dtype)
    getattr(lib(), "wo_fwd2d" + _sfx(dtype))(a, a, nr, nc)
    return _uncplx(a).reshape(nr, nc)


def inv2d(x, dtype=np.float32):
    nr, nc = x.shape
    a = _cplx(x, dtype)
    getattr(lib(), "wo_inv2d" + _sfx(dtype))(a, a, nr, nc)
    return _uncplx(a).reshape(nr, nc)


def hermitian_mirror(x, dtype=np.float64):
    a = _cplx(x, dtype)
    getattr(lib(), "wo_hermitian_mirror" + _sfx(dtype))(a, a.size // 2)
    return _uncplx(a)


# ---- order statistics / baseline (util/src/Waveform.cxx:85-106,
#      sigproc/src/OmnibusSigProc.cxx:867-906) ----

def percentile(x, p, dtype=np.float32):
    a = np.ascontiguousarray(x, dtype=dtype)
    return getattr(lib(), "wo_percentile" + _sfx(dtype))(a, a.size, p)


def restore_baseline(arr, dtype=np.float32):
    a = np.ascontiguousarray(arr, dtype=dtype).copy()
    base = np.zeros(a.shape[0], dtype=dtype)
    getattr(lib(), "wo_restore_baseline" + _sfx(dtype))(a, a.shape[0], a.shape[1], base)
    return a, base


def cal_rms(x, dtype=np.float32):
    a = np.ascontiguousarray(x, dtype=dtype)
    return getattr(lib(), "wo_cal_rms" + _sfx(dtype))(a, a.size)


def roi_threshold(r_data, th_factor, r_tight=None, bad=(), dtype=np.float32):
    """ROI_formation::find_ROI_by_decon_itself threshold part.
    bad: iterable of (row, lo, hi) inclusive tick ranges."""
    a = np.ascontiguousarray(r_data, dtype=dtype)
    nw, nt = a.shape
    rt = None
    if r_tight is not None:
        rt = np.ascontiguousarray(r_tight, dtype=dtype)
    bad = list(bad)
    br = np.array([b[0] for b in bad] or [0], np.int32)
    bl = np.array([b[1] for b in bad] or [0], np.int32)
    bh = np.array([b[2] for b in bad] or [0], np.int32)
    rms = np.zeros(nw)
    th = np.zeros(nw)
    mask = np.zeros((nw, nt), np.uint8)
    getattr(lib(), "wo_roi_threshold" + _sfx(dtype))(
        a, None if rt is None else rt.ctypes.data, nw, nt, th_factor, br, bl, bh, len(bad),
        rms, th, mask)
    return rms, th, mask


def roi_lists(mask, pad=5, asy=0.1, cap=1 << 22):
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    nw, nt = m.shape
    cnt = np.zeros(nw, np.int32)
    pairs = np.zeros(2 * cap, np.int32)
    tot = lib().wo_roi_lists(m, nw, nt, pad, asy, cnt, pairs, cap)
    if tot < 0:
        raise RuntimeError("ROI capacity exceeded")
    out, k = [], 0
    for w in range(nw):
        out.append([(int(pairs[2 * (k + i)]), int(pairs[2 * (k + i) + 1])) for i in range(cnt[w])])
        k += cnt[w]
    return out


def _csr(lists, nw):
    """Per-row lists of (a, b) pairs as CSR int32 arrays (ptr, flat pairs)."""
    ptr = np.zeros(nw + 1, np.int32)
    flat = []
    for w in range(nw):
        rows = lists[w] if w < len(lists) else []
        ptr[w + 1] = ptr[w] + len(rows)
        for a, b in rows:
            flat += [a, b]
    return ptr, np.ascontiguousarray(flat if flat else [0, 0], np.int32)


def _unflat(cnt, pairs):
    out, k = [], 0
    for c in cnt:
        out.append([(int(pairs[2 * (k + i)]), int(pairs[2 * (k + i) + 1])) for i in range(c)])
        k += c
    return out


def loose_rois(r_data, bad=None, rms=None, tight=None, rebin=6, l_factor=3.5, l_max_th=10000.0,
               l_factor1=0.7, l_short_length=3, l_jump_one_bin=0, cap=1 << 22):
    """find_ROI_loose (+ extend_ROI_loose with the tight lists) of one
    plane.  bad: per-row lists of inclusive (lo, hi) tick ranges; rms:
    per-row RMS replacing cal_RMS; tight: per-row inclusive tight ROIs or
    None (no extension, the collection plane)."""
    r = np.ascontiguousarray(r_data, np.float32)
    nw, nt = r.shape
    keep = []

    def ptr(a):
        if a is None:
            return None
        keep.append(a)
        return a.ctypes.data
    bp = blo = bhi = None
    if bad is not None:
        p, flat = _csr(bad, nw)
        bp, blo, bhi = ptr(p), ptr(np.ascontiguousarray(flat[0::2])), ptr(np.ascontiguousarray(flat[1::2]))
    rp = ptr(None if rms is None else np.ascontiguousarray(rms, np.float64))
    tp = tpairs = None
    if tight is not None:
        p, flat = _csr(tight, nw)
        tp, tpairs = ptr(p), ptr(flat)
    cnt = np.zeros(nw, np.int32)
    pairs = np.zeros(2 * cap, np.int32)
    tot = lib().wo_loose_rois(r, nw, nt, bp, blo, bhi, rp, tp, tpairs, rebin, l_factor, l_max_th, l_factor1,
                              l_short_length, l_jump_one_bin, cnt, pairs, cap)
    if tot < 0:
        raise RuntimeError("ROI capacity exceeded")
    return _unflat(cnt, pairs)


def apply_roi(r_data, rois, linear):
    """ROI_refinement::apply_roi on a copy of r_data (float32)."""
    r = np.array(r_data, np.float32, copy=True, order="C")
    nw, nt = r.shape
    p, flat = _csr(rois, nw)
    lib().wo_apply_roi(r, nw, nt, p, flat, int(linear))
    return r


# ---- response (util/src/Response.cxx, OmnibusSigProc.cxx:707-865) ----

def hf_filter(n, max_freq, sigma, power, flag):
    out = np.zeros(n)
    lib().wo_hf_filter(n, max_freq, sigma, power, int(bool(flag)), out)
    return out


def lf_filter(n, max_freq, tau):
    out = np.zeros(n)
    lib().wo_lf_filter(n, max_freq, tau, out)
    return out


def coldelec(t, gain, shaping):
    return lib().wo_coldelec(t, gain, shaping)


def elec_samples(n, tmin, tmax, gain, shaping, postgain):
    of = np.zeros(n, np.float32)
    od = np.zeros(n, np.float64)
    lib().wo_elec_samples(n, tmin, tmax, gain, shaping, postgain, of.ctypes.data, od.ctypes.data)
    return of, od


def fft_best_length(n, keep_odd_even=False):
    return lib().wo_fft_best_length(n, int(keep_odd_even))


def wire_region_average(pitchpos, current, pitch, max_regions=256):
    pp = np.ascontiguousarray(pitchpos, np.float64)
    cur = np.ascontiguousarray(current, np.float32)
    npaths, nsamp = cur.shape
    regs = np.zeros(max_regions, np.int32)
    avg = np.zeros((max_regions, nsamp), np.float32)
    n = lib().wo_wire_region_average(npaths, pp, cur, nsamp, pitch, max_regions, regs, avg)
    if n < 0:
        raise RuntimeError("too many regions")
    return regs[:n].copy(), avg[:n].copy()


def overall_response(avg, elec, fine_period, coarse_period, inter_gain, adc_mv,
                     fine_time_offset, nticks, dtype=np.float32):
    a = np.ascontiguousarray(avg, dtype=dtype)
    e = np.ascontiguousarray(elec, dtype=dtype)
    nr, nin = a.shape
    out = np.zeros((nr, nticks), dtype=dtype)
    rc = getattr(lib(), "wo_overall_response" + _sfx(dtype))(
        a, nr, nin, e, e.size, fine_period, coarse_period, inter_gain, adc_mv,
        fine_time_offset, nticks, out)
    if rc:
        raise RuntimeError("overall_response failed")
    return out


# ---- decon (OmnibusSigProc.cxx:944-1117) ----

def decon_init(data, resp, wire_filter, wire_shift, time_shift, dtype=np.float32,
               ch_resp=None, ewave=None):
    """decon_2D_init's m_c_data (before the time filters).  With ch_resp
    (nw x nsamp channel responses) and ewave (nt electronics samples) the
    per-channel correction (:949-982) runs after the time transform."""
    d = np.ascontiguousarray(data, dtype=dtype)
    r = np.ascontiguousarray(resp, dtype=dtype)
    wf = np.ascontiguousarray(wire_filter, dtype=dtype)
    nw, nt = d.shape
    c = np.zeros(2 * nw * nt, dtype=dtype)
    if ch_resp is None:
        rc = getattr(lib(), "wo_decon_init" + _sfx(dtype))(d, nw, nt, r, r.shape[0], wf,
                                                           wire_shift, time_shift, c)
    else:
        cr = np.ascontiguousarray(ch_resp, dtype=dtype)
        ew = np.ascontiguousarray(ewave, dtype=dtype)
        if cr.ndim != 2 or cr.shape[0] != nw or ew.shape != (nt,):
            raise ValueError("ch_resp must be nw x nsamp and ewave nt long")
        rc = getattr(lib(), "wo_decon_init_cr" + _sfx(dtype))(
            d, nw, nt, r, r.shape[0], wf, wire_shift, time_shift,
            cr.ctypes.data, cr.shape[1], ew.ctypes.data, c)
    if rc:
        raise RuntimeError("decon_init failed")
    return _uncplx(c).reshape(nw, nt)


def decon_apply_filter(c_data, time_filter, dtype=np.float32):
    nw, nt = c_data.shape
    c = _cplx(c_data, dtype)
    tf = np.ascontiguousarray(time_filter, dtype=dtype)
    out = np.zeros((nw, nt), dtype=dtype)
    getattr(lib(), "wo_decon_apply_filter" + _sfx(dtype))(c, nw, nt, tf, out)
    return out


def decon_stepwise(data, resp, wire_filter, time_filter, wire_shift, time_shift,
                   dtype=np.float32, ch_resp=None, ewave=None):
    """decon_2D_init followed by one decon_2D_*ROI variant, before
    restore_baseline, in the reference's step order and precision."""
    c = decon_init(data, resp, wire_filter, wire_shift, time_shift, dtype, ch_resp, ewave)
    return decon_apply_filter(c, time_filter, dtype)


def decon_fused(data, resp, wire_filter, time_filter, wire_shift, time_shift):
    """FP64 fused decon (the same linear map, no intermediate round trips)."""
    d = np.ascontiguousarray(data, np.float64)
    r = np.ascontiguousarray(resp, np.float64)
    nw, nt = d.shape
    out = np.zeros((nw, nt))
    rc = lib().wo_decon_fused_d(d, nw, nt, r, r.shape[0],
                                np.ascontiguousarray(wire_filter, np.float64),
                                np.ascontiguousarray(time_filter, np.float64),
                                wire_shift, time_shift, out)
    if rc:
        raise RuntimeError("decon_fused failed")
    return out


# ---- tensor-set variant (sig/src/Decon2DResponse.cxx, Decon2DFilter.cxx) ----

def decon2d_response(data, resp, wire_filter, wire_shift, time_shift, dtype=np.float64):
    """Sig::Decon2DResponse::operator() (sig/src/Decon2DResponse.cxx:203-390):
    fwd_r2c along ticks, fwd along wires, divide by the response
    spectrum, NaN -> -0 and Inf -> 0, wire filter, inverse along wires,
    inv_c2r along ticks, roll by wire_shift (and time_shift when > 0),
    fwd_r2c along ticks: the steps of decon_2D_init
    (OmnibusSigProc.cxx:944-1051) on the same response, so the complex
    (nw, nt) time spectrum of the rolled frame."""
    return decon_init(data, resp, wire_filter, wire_shift, time_shift, dtype)


def decon2d_filter(spec, time_filter, cmm=(), restore=True, dtype=np.float64):
    """Sig::Decon2DFilter::operator() (sig/src/Decon2DFilter.cxx:90-145):
    each spectrum column times the filter product, inv_c2r along ticks,
    Sig::restore_baseline (sig/src/Util.cxx:5-35), then row ch zeroed
    over [tmin, tmax) for every (ch, tmin, tmax) of bad:cmm_range with
    ch below the row count."""
    r = decon_apply_filter(spec, time_filter, dtype)
    if restore:
        r, _ = restore_baseline(r, dtype)
    r = np.array(r)
    for ch, tmin, tmax in cmm:
        if ch < r.shape[0]:
            r[int(ch), int(tmin):int(tmax)] = 0.0
    return r


def forward_conv(q, resp, wire_shift, time_shift):
    qq = np.ascontiguousarray(q, np.float64)
    r = np.ascontiguousarray(resp, np.float64)
    nw, nt = qq.shape
    out = np.zeros((nw, nt))
    lib().wo_forward_conv_d(qq, nw, nt, r, r.shape[0], wire_shift, time_shift, out)
    return out


# ---- synthetic inputs ----

def normal(seed, stream, n, sigma):
    out = np.zeros(n)
    lib().wo_normal(seed, stream, n, sigma, out)
    return out


def tracks(seed, stream, ntracks, nw, nt, q_per_tick):
    q = np.zeros((nw, nt))
    lib().wo_tracks(seed, stream, ntracks, nw, nt, q_per_tick, q)
    return q
