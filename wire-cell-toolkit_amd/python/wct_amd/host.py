"""ctypes declarations of include/wct_host.h (libWireCellHip.so)."""
import ctypes as C
import os

import numpy as np

from .capi import WctError, _PKG

HOST_LIB_PATH = os.path.join(_PKG, "libWireCellHip.so")
_LIB = None
_f64 = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32 = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_SIGS = {
    "wcth_coldelec_samples": (C.c_int, [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                        C.c_double, _f64]),
    "wcth_hf_filter": (C.c_int, [C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, _f64]),
    "wcth_lf_filter": (C.c_int, [C.c_int, C.c_double, C.c_double, _f64]),
    "wcth_fft_best_length": (C.c_long, [C.c_long, C.c_int]),
    "wcth_wire_region_average": (C.c_int, [C.c_int, _f64, _f64, C.c_int, C.c_double, C.c_int,
                                           _i32, _f64]),
    "wcth_last_error": (C.c_char_p, []),
    "wcth_configure": (C.c_int, [C.c_char_p]),
    "wcth_reset": (C.c_int, []),
    "wcth_known_types": (C.c_int, [C.c_char_p, C.c_size_t]),
    "wcth_dft": (C.c_int, [C.c_char_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "wcth_filter_waveform": (C.c_int, [C.c_char_p, C.c_int, C.c_void_p]),
    "wcth_channel_response": (C.c_int, [C.c_char_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                        C.c_void_p]),
    "wcth_frame_new": (C.c_void_p, [C.c_int, C.c_double, C.c_double]),
    "wcth_frame_free": (None, [C.c_void_p]),
    "wcth_frame_add_trace": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p]),
    "wcth_frame_add_mask": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, C.c_int, C.c_int]),
    "wcth_frame_filter": (C.c_int, [C.c_char_p, C.c_void_p, C.POINTER(C.c_void_p)]),
    "wcth_frame_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                  C.POINTER(C.c_double), C.POINTER(C.c_int)]),
    "wcth_frame_trace": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                   C.POINTER(C.c_int), C.POINTER(C.POINTER(C.c_float))]),
    "wcth_frame_tags": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_char_p)]),
    "wcth_frame_tagged": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.POINTER(C.c_size_t)),
                                    C.POINTER(C.c_int)]),
    "wcth_frame_summary": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.POINTER(C.c_double)),
                                     C.POINTER(C.c_int)]),
    "wcth_frame_masks": (C.c_int, [C.c_void_p, C.c_char_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_int]),
    "wcth_loose_rois": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_long, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double,
                                  C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int]),
    "wcth_apply_roi": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_long, C.c_void_p, C.c_void_p, C.c_int]),
    "wcth_roi_lists": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_long, C.c_int, C.c_float,
                                 C.c_void_p, C.c_void_p, C.c_int]),
    "wcth_find_peaks": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_void_p,
                                  C.c_void_p, C.c_int]),
    "wcth_refine_new": (C.c_void_p, [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int]),
    "wcth_refine_free": (None, [C.c_void_p]),
    "wcth_refine_load": (C.c_int, [C.c_void_p, C.c_void_p, C.c_long, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "wcth_refine_step": (C.c_int, [C.c_void_p, C.c_char_p]),
    "wcth_refine_rois": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int]),
    "wcth_refine_contents": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int]),
    "wcth_frame_tag_frame": (C.c_int, [C.c_void_p, C.c_char_p]),
    "wcth_frame_tag_traces": (C.c_int, [C.c_void_p, C.c_char_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int]),
    "wcth_frame_sink": (C.c_int, [C.c_char_p, C.c_void_p]),
    "wcth_frame_source": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "wcth_tensorset_new": (C.c_void_p, [C.c_int, C.c_char_p]),
    "wcth_tensorset_free": (None, [C.c_void_p]),
    "wcth_tensorset_add": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, C.c_void_p, C.c_void_p, C.c_char_p]),
    "wcth_tensorset_filter": (C.c_int, [C.c_char_p, C.c_void_p, C.POINTER(C.c_void_p)]),
    "wcth_tensorset_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_char_p,
                                      C.c_size_t]),
    "wcth_tensorset_tensor": (C.c_int, [C.c_void_p, C.c_int, C.c_char_p, C.POINTER(C.c_int), C.c_void_p,
                                        C.POINTER(C.c_void_p), C.POINTER(C.c_size_t)]),
    "wcth_tensorset_tensor_md": (C.c_int, [C.c_void_p, C.c_int, C.c_char_p, C.c_size_t]),
}
EXPORTED = sorted(_SIGS)
MISSING = []


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(HOST_LIB_PATH):
            raise WctError(f"host library not built: {HOST_LIB_PATH}")
        L = C.CDLL(HOST_LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name, None)
            if fn is None:
                MISSING.append(name)
                continue
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def coldelec_samples(n, tmin, tmax, gain, shaping, postgain):
    out = np.zeros(n)
    if lib().wcth_coldelec_samples(n, tmin, tmax, gain, shaping, postgain, out):
        raise WctError("coldelec_samples: bad arguments")
    return out


def hf_filter(n, max_freq, sigma, power, flag):
    out = np.zeros(n)
    lib().wcth_hf_filter(n, max_freq, sigma, power, int(bool(flag)), out)
    return out


def lf_filter(n, max_freq, tau):
    out = np.zeros(n)
    lib().wcth_lf_filter(n, max_freq, tau, out)
    return out


def fft_best_length(n, keep_odd_even=False):
    return int(lib().wcth_fft_best_length(n, int(keep_odd_even)))


def wire_region_average(pitchpos, current, pitch, max_regions=256):
    pp = np.ascontiguousarray(pitchpos, np.float64)
    cur = np.ascontiguousarray(current, np.float64)
    npaths, nsamp = cur.shape
    regs = np.zeros(max_regions, np.int32)
    avg = np.zeros((max_regions, nsamp))
    n = lib().wcth_wire_region_average(npaths, pp, cur, nsamp, pitch, max_regions, regs, avg)
    if n < 0:
        raise WctError("wire_region_average failed")
    return regs[:n].copy(), avg[:n].copy()
