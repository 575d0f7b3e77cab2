"""ctypes declarations of include/wct_hip.h."""
import ctypes as C
import os

_PKG = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
LIB_PATH = os.environ.get("WCT_HIP_LIBRARY") or os.path.join(_PKG, "libwct_hip.so")
_LIB = None


class WctError(RuntimeError):
    pass


class DeconTuning(C.Structure):
    _fields_ = [("wire_tile", C.c_int), ("row_grid_pct", C.c_int), ("row_plan", C.c_int),
                ("r2c_blocks", C.c_int), ("c2r_blocks", C.c_int)]


class DeconOutput(C.Structure):
    _fields_ = [
        ("filter", C.c_void_p), ("filter_alt", C.c_void_p), ("alt_rows", C.c_void_p),
        ("out", C.c_void_p), ("ld_out", C.c_int),
        ("rms", C.c_void_p),
        ("restore_baseline", C.c_int),
        ("save_mode", C.c_int),
    ]


MAX_OUTPUTS = 6


class DeconJob(C.Structure):
    _fields_ = [
        ("kernel", C.c_void_p),
        ("adc", C.c_void_p), ("ld_adc", C.c_int),
        ("out", C.c_void_p), ("ld_out", C.c_int),
        ("mask", C.c_void_p), ("ld_mask", C.c_int),
        ("rms", C.c_void_p),
        ("bad_ptr", C.c_void_p), ("bad_lo", C.c_void_p), ("bad_hi", C.c_void_p),
        ("lf_tight", C.c_void_p), ("lf_tighter", C.c_void_p),
        ("th_factor", C.c_float),
        ("restore_baseline", C.c_int),
        ("save_mode", C.c_int),
        ("adjoint", C.c_int),
        ("spectrum", C.c_void_p), ("ld_spectrum", C.c_int),
        ("outputs", C.POINTER(DeconOutput)), ("noutputs", C.c_int),
        ("chan_factors", C.c_void_p), ("ld_chan_factors", C.c_int),
        ("chan_factor_row", C.c_void_p),
    ]


class ImpactFrame(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("start_ch", "end_ch", "npad_wire", "frame_start_ch", "frame_end_ch",
                                        "frame_start_tick", "frame_end_tick", "nw", "nt")]


class NfChannel(C.Structure):
    _fields_ = [("plane", C.c_int), ("rcrc", C.c_int), ("bins_begin", C.c_int), ("bins_end", C.c_int),
                ("min_rms", C.c_float), ("max_rms", C.c_float), ("mode", C.c_int), ("baseline", C.c_float),
                ("gain", C.c_float), ("config", C.c_int), ("noise", C.c_int)]


class NfGroup(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("first", "count", "respec", "res_offset", "pad_f", "pad_b")] + \
               [(n, C.c_float) for n in ("decon_limit", "decon_lf_cutoff", "adc_limit", "decon_limit1",
                                         "roi_min_max_ratio", "protection_factor", "min_adc_limit")]


NF_PARTIAL, NF_NOISY, NF_HARMONIC, NF_CHIRP = 1, 2, 4, 8
NF_PROTODUNE, NF_MICROBOONE = 0, 1


# name -> (restype, argtypes)
_V, _I, _D, _F, _Z = C.c_void_p, C.c_int, C.c_double, C.c_float, C.c_size_t
_SIGS = {
    "wct_last_error": (C.c_char_p, []),
    "wct_abi_version": (_I, []),
    "wct_channel_factors": (_I, [_V, _I, _V, _I, _I, _V, _I, _V]),
    "wct_debug_inject": (_I, [C.c_char_p, _I, _I]),
    "wct_debug_alloc_cap": (_I, [_Z]),
    "wct_device_count": (_I, [C.POINTER(C.c_int)]),
    "wct_set_device": (_I, [_I]),
    "wct_malloc": (_I, [C.POINTER(C.c_void_p), _Z]),
    "wct_free": (_I, [_V]),
    "wct_memcpy_htod": (_I, [_V, _V, _Z, _V]),
    "wct_memcpy_dtoh": (_I, [_V, _V, _Z, _V]),
    "wct_memcpy_dtod": (_I, [_V, _V, _Z, _V]),
    "wct_memset": (_I, [_V, _I, _Z, _V]),
    "wct_host_malloc": (_I, [C.POINTER(C.c_void_p), _Z]),
    "wct_host_free": (_I, [_V]),
    "wct_memcpy_htod_async": (_I, [_V, _V, _Z, _V]),
    "wct_memcpy_dtoh_async": (_I, [_V, _V, _Z, _V]),
    "wct_stream_create": (_I, [C.POINTER(C.c_void_p)]),
    "wct_stream_destroy": (_I, [_V]),
    "wct_stream_sync": (_I, [_V]),
    "wct_dft_fwd1d": (_I, [_V, _V, _I, _V]),
    "wct_dft_inv1d": (_I, [_V, _V, _I, _V]),
    "wct_dft_fwd1b": (_I, [_V, _V, _I, _I, _I, _V]),
    "wct_dft_inv1b": (_I, [_V, _V, _I, _I, _I, _V]),
    "wct_dft_fwd2d": (_I, [_V, _V, _I, _I, _V]),
    "wct_dft_inv2d": (_I, [_V, _V, _I, _I, _V]),
    "wct_dft_transpose_real": (_I, [_V, _V, _I, _I, _V]),
    "wct_dft_transpose_complex": (_I, [_V, _V, _I, _I, _V]),
    "wct_dft_fwd1b_f64": (_I, [_V, _V, _I, _I, _I, _V]),
    "wct_dft_inv1b_f64": (_I, [_V, _V, _I, _I, _I, _V]),
    "wct_dft_supported": (_I, [_I]),
    "wct_dft_native_length": (_I, [_I]),
    "wct_overall_response": (_I, [_V, _I, _I, _V, _I, _D, _D, _D, _D, _D, _I, _V, _V]),
    "wct_decon_kernel_create": (_I, [C.POINTER(C.c_void_p), _V, _I, _I, _I, _V, _V, _I, _I, _D, _V]),
    "wct_forward_kernel_create": (_I, [C.POINTER(C.c_void_p), _V, _I, _I, _I, _I, _I, _V]),
    "wct_decon_kernel_create_unfiltered": (_I, [C.POINTER(C.c_void_p), _V, _I, _I, _I, _V, _V, _I, _I, _D,
                                                _D, _V]),
    "wct_decon_kernel_info": (_I, [_V, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "wct_decon_kernel_download": (_I, [_V, _V]),
    "wct_decon_kernel_destroy": (_I, [_V]),
    "wct_decon_workspace_size": (_Z, [C.POINTER(DeconJob), _I]),
    "wct_decon_run": (_I, [C.POINTER(DeconJob), _I, _V, _Z, _V]),
    "wct_spectrum_filter": (_I, [_V, _I, _I, _I, _V, _I, _V, _I, _V]),
    "wct_noalloc_prepare": (_I, [_I]),
    "wct_spectrum_filter_noalloc": (_I, [_V, _I, _I, _I, _V, _I, _V, _I, _V]),
    "wct_decon_plan_create": (_I, [C.POINTER(C.c_void_p), C.POINTER(DeconJob), _I, _V, _Z]),
    "wct_decon_plan_run": (_I, [_V, _V]),
    "wct_decon_plan_run_stage": (_I, [_V, _I, _V]),
    "wct_decon_plan_destroy": (_I, [_V]),
    "wct_decon_plan_create_tuned": (_I, [C.POINTER(C.c_void_p), C.POINTER(DeconJob), _I, _V, _Z,
                                         C.POINTER(DeconTuning)]),
    "wct_decon_graph_create": (_I, [C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), _I, _V]),
    "wct_decon_graph_launch": (_I, [_V, _V]),
    "wct_decon_graph_destroy": (_I, [_V]),
    "wct_plain_workspace_size": (_Z, [_I, _I]),
    "wct_plain_decon_plane": (_I, [_V, _I, _I, _I, _V, _I, _V, _V, _V, _I, _I, _F, _V, _V, _V, _V, _I, _V,
                                   _I, _V, _I, _V, _Z, _V]),
    "wct_fft_best_length": (C.c_long, [C.c_long, _I]),
    "wct_impact_geometry": (_I, [_I, _I, _I, _I, _I, _I, _I, C.POINTER(ImpactFrame)]),
    "wct_impact_create": (_I, [C.POINTER(C.c_void_p), _I, _I, _V, _I, _V]),
    "wct_impact_destroy": (_I, [_V]),
    "wct_impact_run": (_I, [_V, C.POINTER(ImpactFrame), C.POINTER(C.c_long), _V, _V, _V, _V, _V]),
    "wct_impact_waveforms": (_I, [_V, C.POINTER(ImpactFrame), _V, _I, _V, _I, _I, _V, _V]),
    "wct_nf_rcrc_spectrum": (_I, [_D, _D, _I, _I, _V, _V]),
    "wct_nf_workspace_size": (_Z, [_I, _I]),
    "wct_nf_one_channel": (_I, [_V, _I, _V, _I, _I, _I, _V, _V, _V, _V, _V, _Z, _V]),
    "wct_nf_rel_gain": (_I, [_V, _I, _I, _I, _V, _F, _F, _V]),
    "wct_nf_adc_bit_shift": (_I, [_V, _I, _I, _I, _I, _I, _D, _D, _V, _V]),
    "wct_nf_status_workspace_size": (_Z, [_I, _I]),
    "wct_nf_lf_noisy": (_I, [_V, _I, _I, _I, _D, _I, _I, _D, _V, _V, _V, _Z, _V]),
    "wct_nf_coherent_workspace_size": (_Z, [_I, _I, _I]),
    "wct_nf_coherent": (_I, [_V, _I, _I, _I, _V, _I, _V, _V, _V, _I, _V, _Z, _V]),
}

# Every symbol include/wct_hip.h declares (checked by the CPU tests).
EXPORTED = sorted(_SIGS)
MISSING = []


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise WctError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name, None)
            if fn is None:
                MISSING.append(name)
                continue
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def check(rc):
    if rc != 0:
        msg = lib().wct_last_error().decode(errors="replace")
        raise WctError(f"wct status {rc}: {msg}")
    return rc
