"""The per-channel noise filters (Protodune::OneChannelNoise,
sigproc/src/Protodune.cxx:753-918; Microboone::OneChannelNoise,
sigproc/src/Microboone.cxx:902-1064; Protodune::RelGainCalib,
Protodune.cxx:975-1009) through the C ABI, for every channel of a frame at
once.  Device buffers are torch tensors used as plain device memory.
"""
import ctypes as C

import numpy as np

from .capi import NF_CHIRP, NF_HARMONIC, NF_MICROBOONE, NF_NOISY, NF_PARTIAL, NF_PROTODUNE, NfChannel, NfGroup, check, lib


def _stream():
    import torch
    return torch.cuda.current_stream().cuda_stream


def rcrc_spectrum(width, tick, nt, layers=1, stream=None):
    """OmniChannelNoiseDB's rcrc spectrum (nt complex64, numpy)."""
    import torch
    out = torch.empty(nt, dtype=torch.complex64, device="cuda")
    check(lib().wct_nf_rcrc_spectrum(float(width), float(tick), int(nt), int(layers), C.c_void_p(out.data_ptr()),
                                     _stream() if stream is None else stream))
    return out.cpu().numpy()


def flagged_bins(noise):
    """Runs of bins whose real part is below 0.5, as (first, last) pairs
    (Protodune.cxx:791-806)."""
    runs = []
    for i, v in enumerate(np.real(np.asarray(noise))):
        if v < 0.5:
            if runs and runs[-1][1] == i - 1:
                runs[-1][1] = i
            else:
                runs.append([i, i])
    return [tuple(r) for r in runs]


def one_channel(frame, channels, rcrc=None, stream=None):
    """frame: (nch, nt) float32 samples (numpy or CUDA tensor); channels:
    one dict per row with plane, mode ("protodune" or "microboone"), rcrc
    (row of the spectra table or -1), bins (ProtoDUNE: flagged (first,
    last) runs), min_rms, max_rms, and for MicroBooNE baseline, gain,
    config and noise_row (rows of the spectra table or -1); rcrc: the
    (nr, nt) complex64 spectra table.  Returns (filtered CUDA tensor,
    flags (nch, 3) numpy: bits, chirp begin, chirp end)."""
    import torch
    x = torch.as_tensor(frame, dtype=torch.float32, device="cuda").contiguous()
    nch, nt = x.shape
    assert len(channels) == nch
    entries = (NfChannel * nch)()
    pairs = []
    for i, c in enumerate(channels):
        b = list(c.get("bins", []))
        mode = c.get("mode", "protodune")
        mode = NF_MICROBOONE if mode in ("microboone", NF_MICROBOONE) else NF_PROTODUNE
        entries[i] = NfChannel(int(c.get("plane", 0)), int(c.get("rcrc", -1)), len(pairs), len(pairs) + len(b),
                               float(c.get("min_rms", 0.5)), float(c.get("max_rms", 10.0)), mode,
                               float(c.get("baseline", 0.0)), float(c.get("gain", 1.0)), int(c.get("config", -1)),
                               int(c.get("noise_row", -1)))
        pairs.extend(b)
    dch = torch.frombuffer(bytearray(bytes(entries)), dtype=torch.uint8).to("cuda")
    dbins = torch.as_tensor(np.asarray(pairs, np.int32).reshape(-1), device="cuda") if pairs else None
    drc = None
    if rcrc is not None and len(rcrc):
        drc = torch.as_tensor(np.ascontiguousarray(rcrc, np.complex64), device="cuda")
    out = torch.empty_like(x)
    flags = torch.zeros((nch, 3), dtype=torch.int32, device="cuda")
    ws_bytes = lib().wct_nf_workspace_size(nch, nt)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda")
    check(lib().wct_nf_one_channel(
        C.c_void_p(x.data_ptr()), nt, C.c_void_p(out.data_ptr()), nt, nch, nt, C.c_void_p(dch.data_ptr()),
        C.c_void_p(drc.data_ptr() if drc is not None else 0), C.c_void_p(dbins.data_ptr() if dbins is not None else 0),
        C.c_void_p(flags.data_ptr()), C.c_void_p(ws.data_ptr()), ws_bytes, _stream() if stream is None else stream))
    return out, flags.cpu().numpy()


def masks(flag, plane, nt, mode="protodune"):
    """The mask map one channel's flags (bits, chirp begin, chirp end)
    stand for: {name: [(first, second), ...]}, as Protodune.cxx:846-915
    and Microboone.cxx:924-1061 fill it."""
    bits, beg, end = (int(v) for v in flag)
    out = {}
    if mode in ("microboone", NF_MICROBOONE):
        if bits & NF_CHIRP:
            out["chirp"] = [(beg, end)]
            if plane != 2 and (beg > 0 or end < nt):
                out.setdefault("lf_noisy", []).append((beg, end))
        if bits & NF_PARTIAL and plane != 2:
            out.setdefault("lf_noisy", []).append((0, nt))
        if bits & NF_NOISY:
            out["noisy"] = [(0, nt)]
            out.pop("lf_noisy", None)
        return out
    if bits & NF_HARMONIC:
        out["harmonic"] = [(0, nt)]
    if bits & NF_PARTIAL and plane != 2:
        out["lf_noisy"] = [(0, nt)]
    if bits & NF_NOISY:
        out["noisy"] = [(0, nt)]
    return out


def mask_names(flag, plane):
    """Names of the masks of a ProtoDUNE channel's flag bits."""
    return sorted(masks((flag, -1, -1), plane, 0))


def rel_gain(frame, gains, gain_min=0.8, gain_max=1.25, stream=None):
    """Protodune::RelGainCalib on every row (in place on a CUDA tensor)."""
    import torch
    x = torch.as_tensor(frame, dtype=torch.float32, device="cuda").contiguous()
    g = torch.as_tensor(np.asarray(gains, np.float32), device="cuda")
    nch, nt = x.shape
    check(lib().wct_nf_rel_gain(C.c_void_p(x.data_ptr()), nt, nch, nt, C.c_void_p(g.data_ptr()), float(gain_min),
                                float(gain_max), _stream() if stream is None else stream))
    return x


def adc_bit_shift(frame, nbits=12, exam_nticks=500, threshold_sigma=7.5, threshold_fix=0.8, stream=None):
    """Microboone::ADCBitShift on every row (a corrected CUDA copy of
    frame and the per-row shift count, numpy int32)."""
    import torch
    x = torch.as_tensor(frame, dtype=torch.float32, device="cuda").contiguous().clone()
    nch, nt = x.shape
    ns = torch.empty(nch, dtype=torch.int32, device="cuda")
    check(lib().wct_nf_adc_bit_shift(C.c_void_p(x.data_ptr()), nt, nch, nt, int(nbits), int(exam_nticks),
                                     float(threshold_sigma), float(threshold_fix), C.c_void_p(ns.data_ptr()),
                                     _stream() if stream is None else stream))
    torch.cuda.synchronize()
    return x, ns.cpu().numpy()


def lf_noisy(frame, threshold=3.5, window=5, nbins=250, cut=14.0, stream=None):
    """Microboone::OneChannelStatus::ID_lf_noisy on every row of frame:
    (flags, ratios) as numpy arrays (bool, float64)."""
    import torch
    x = torch.as_tensor(frame, dtype=torch.float32, device="cuda").contiguous()
    nch, nt = x.shape
    flags = torch.empty(nch, dtype=torch.int32, device="cuda")
    ratio = torch.empty(nch, dtype=torch.float64, device="cuda")
    ws_bytes = lib().wct_nf_status_workspace_size(nch, nt)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda")
    check(lib().wct_nf_lf_noisy(C.c_void_p(x.data_ptr()), nt, nch, nt, float(threshold), int(window), int(nbins),
                                float(cut), C.c_void_p(flags.data_ptr()), C.c_void_p(ratio.data_ptr()),
                                C.c_void_p(ws.data_ptr()), ws_bytes, _stream() if stream is None else stream))
    torch.cuda.synchronize()
    return flags.cpu().numpy().astype(bool), ratio.cpu().numpy()


GROUP_DEFAULTS = dict(respec=-1, res_offset=0, pad_f=0, pad_b=0, decon_limit=0.02, decon_lf_cutoff=0.08,
                      adc_limit=0.0, decon_limit1=0.08, roi_min_max_ratio=0.8, protection_factor=5.0,
                      min_adc_limit=50.0)


def coherent(frame, groups, spectra=None, stream=None):
    """Microboone::CoherentNoiseSub on every group (in place on a CUDA copy
    of frame, returned).  groups: dicts with rows (row indices in channel
    order) and the GROUP_DEFAULTS keys (respec: row of spectra or -1)."""
    import torch
    x = torch.as_tensor(frame, dtype=torch.float32, device="cuda").contiguous().clone()
    nrows, nt = x.shape
    entries = (NfGroup * len(groups))()
    grows, row_group = [], np.full(nrows, -1, np.int32)
    any_resp = 0
    for g, grp in enumerate(groups):
        p = dict(GROUP_DEFAULTS, **{k: v for k, v in grp.items() if k != "rows"})
        entries[g] = NfGroup(len(grows), len(grp["rows"]), int(p["respec"]), int(p["res_offset"]), int(p["pad_f"]),
                             int(p["pad_b"]), *(float(p[k]) for k in ("decon_limit", "decon_lf_cutoff", "adc_limit",
                                                                     "decon_limit1", "roi_min_max_ratio",
                                                                     "protection_factor", "min_adc_limit")))
        any_resp |= int(p["respec"]) >= 0
        if not 1 <= len(grp["rows"]) <= 128:
            raise ValueError(f"coherent: group {g} has {len(grp['rows'])} channels (1 to 128 supported)")
        for r in grp["rows"]:
            if row_group[r] >= 0:
                raise ValueError(f"coherent: row {r} is in groups {row_group[r]} and {g}")
            row_group[r] = g
        grows.extend(grp["rows"])
    dg = torch.frombuffer(bytearray(bytes(entries)), dtype=torch.uint8).to("cuda")
    dgr = torch.as_tensor(np.asarray(grows, np.int32), device="cuda")
    drg = torch.as_tensor(row_group, device="cuda")
    dsp = torch.as_tensor(np.ascontiguousarray(spectra, np.complex64), device="cuda") if spectra is not None else None
    ws_bytes = lib().wct_nf_coherent_workspace_size(len(groups), nrows, nt)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda")
    check(lib().wct_nf_coherent(C.c_void_p(x.data_ptr()), nt, nrows, nt, C.c_void_p(dg.data_ptr()), len(groups),
                                C.c_void_p(dgr.data_ptr()), C.c_void_p(drg.data_ptr()),
                                C.c_void_p(dsp.data_ptr() if dsp is not None else 0), any_resp,
                                C.c_void_p(ws.data_ptr()), ws_bytes, _stream() if stream is None else stream))
    return x
