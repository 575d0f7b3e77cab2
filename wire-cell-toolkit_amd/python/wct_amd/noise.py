"""The ProtoDUNE per-channel noise filter (Protodune::OneChannelNoise,
sigproc/src/Protodune.cxx:753-918) through the C ABI, for every channel
of a frame at once.  Device buffers are torch tensors used as plain
device memory.
"""
import ctypes as C

import numpy as np

from .capi import NF_HARMONIC, NF_NOISY, NF_PARTIAL, NfChannel, check, lib


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
    one dict per row with plane, rcrc (row of the rcrc table or -1), bins
    (flagged (first, last) runs), min_rms, max_rms; rcrc: (nr, nt)
    complex64 table.  Returns (filtered CUDA tensor, flags numpy)."""
    import torch
    x = torch.as_tensor(frame, dtype=torch.float32, device="cuda").contiguous()
    nch, nt = x.shape
    assert len(channels) == nch
    entries = (NfChannel * nch)()
    pairs = []
    for i, c in enumerate(channels):
        b = list(c.get("bins", []))
        entries[i] = NfChannel(int(c.get("plane", 0)), int(c.get("rcrc", -1)), len(pairs), len(pairs) + len(b),
                               float(c.get("min_rms", 0.5)), float(c.get("max_rms", 10.0)))
        pairs.extend(b)
    dch = torch.frombuffer(bytearray(bytes(entries)), dtype=torch.uint8).to("cuda")
    dbins = torch.as_tensor(np.asarray(pairs, np.int32).reshape(-1), device="cuda") if pairs else None
    drc = None
    if rcrc is not None and len(rcrc):
        drc = torch.as_tensor(np.ascontiguousarray(rcrc, np.complex64), device="cuda")
    out = torch.empty_like(x)
    flags = torch.zeros(nch, dtype=torch.int32, device="cuda")
    ws_bytes = lib().wct_nf_workspace_size(nch, nt)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda")
    check(lib().wct_nf_one_channel(
        C.c_void_p(x.data_ptr()), nt, C.c_void_p(out.data_ptr()), nt, nch, nt, C.c_void_p(dch.data_ptr()),
        C.c_void_p(drc.data_ptr() if drc is not None else 0), C.c_void_p(dbins.data_ptr() if dbins is not None else 0),
        C.c_void_p(flags.data_ptr()), C.c_void_p(ws.data_ptr()), ws_bytes, _stream() if stream is None else stream))
    return out, flags.cpu().numpy()


def mask_names(flag, plane):
    """The ChannelMaskMap names one channel's flags stand for."""
    names = []
    if flag & NF_HARMONIC:
        names.append("harmonic")
    if flag & NF_PARTIAL and plane != 2:
        names.append("lf_noisy")
    if flag & NF_NOISY:
        names.append("noisy")
    return names
