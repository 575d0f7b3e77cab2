"""Gen::ImpactTransform (gen/src/ImpactTransform.cxx:19-347) through the C
ABI: the forward simulation convolution of impact-group charges with the
plane's impact responses.  Device buffers are torch tensors used as plain
device memory.
"""
import ctypes as C

import numpy as np

from .capi import ImpactFrame, check, lib

FRAME_KEYS = ("start_ch", "end_ch", "npad_wire", "frame_start_ch", "frame_end_ch", "frame_start_tick",
              "frame_end_tick", "nw", "nt")


def _stream():
    import torch
    return torch.cuda.current_stream().cuda_stream


def fft_best_length(n, keep_odd_even=False):
    """WireCell::fft_best_length (util/src/FFTBestLength.cxx:12-62)."""
    return int(lib().wct_fft_best_length(int(n), int(bool(keep_odd_even))))


def geometry(num_group, nwires_resp, impact_range, time_range, waveform_pad):
    """The transform's frame (ImpactTransform.cxx:84-115) as a dict."""
    fr = ImpactFrame()
    check(lib().wct_impact_geometry(int(num_group), int(nwires_resp), int(impact_range[0]), int(impact_range[1]),
                                    int(time_range[0]), int(time_range[1]), int(waveform_pad), C.byref(fr)))
    return {k: getattr(fr, k) for k in FRAME_KEYS}


def _frame(d):
    return ImpactFrame(*[int(d[k]) for k in FRAME_KEYS])


class ImpactTransform:
    """resp: (num_group, nwires_resp, nresp) response waveforms, row
    (i, j) that of impact group i at wire offset j - (nwires_resp - 1) / 2
    (numpy array or CUDA tensor)."""

    def __init__(self, resp, stream=None):
        import torch
        r = torch.as_tensor(resp, dtype=torch.float32, device="cuda").contiguous()
        self.num_group, self.nwires_resp, self.nresp = (int(x) for x in r.shape)
        self._h = C.c_void_p()
        check(lib().wct_impact_create(C.byref(self._h), self.num_group, self.nwires_resp, C.c_void_p(r.data_ptr()),
                                      self.nresp, _stream() if stream is None else stream))

    def close(self):
        if self._h:
            check(lib().wct_impact_destroy(self._h))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def geometry(self, impact_range, time_range, waveform_pad):
        return geometry(self.num_group, self.nwires_resp, impact_range, time_range, waveform_pad)

    @staticmethod
    def pack_charges(charges):
        """charges[g]: (ch, tick, q) tuples or an (n, 3) array per group ->
        (offsets, ch, tick, q) as numpy arrays."""
        offs = [0]
        cols = [[], [], []]
        for grp in charges:
            a = np.asarray(grp, dtype=np.float64).reshape(-1, 3)
            offs.append(offs[-1] + len(a))
            for c in range(3):
                cols[c].append(a[:, c])
        cat = [np.concatenate(c) if c else np.zeros(0) for c in cols]
        return (np.asarray(offs, np.int64), cat[0].astype(np.int32), cat[1].astype(np.int32),
                cat[2].astype(np.float64))

    def run(self, frame, offsets, ch, tick, q, out=None, stream=None):
        """m_decon_data (nw x nt float32 tensor).  offsets: host array of
        num_group + 1; ch, tick, q: per-charge arrays (numpy or CUDA)."""
        import torch
        offs = np.ascontiguousarray(offsets, np.int64)
        assert offs.size == self.num_group + 1
        dch = torch.as_tensor(ch, dtype=torch.int32, device="cuda").contiguous()
        dtk = torch.as_tensor(tick, dtype=torch.int32, device="cuda").contiguous()
        dq = torch.as_tensor(q, dtype=torch.float64, device="cuda").contiguous()
        if out is None:
            out = torch.empty((frame["nw"], frame["nt"]), dtype=torch.float32, device="cuda")
        fr = _frame(frame)
        check(lib().wct_impact_run(self._h, C.byref(fr), offs.ctypes.data_as(C.POINTER(C.c_long)),
                                   C.c_void_p(dch.data_ptr()), C.c_void_p(dtk.data_ptr()),
                                   C.c_void_p(dq.data_ptr()), C.c_void_p(out.data_ptr()),
                                   _stream() if stream is None else stream))
        return out

    def waveforms(self, frame, decon, nsamples, long_aux=None, long_pad=0, stream=None):
        """waveform(iwire) for the frame's wires (nw x nsamples tensor)."""
        import torch
        out = torch.empty((frame["nw"], nsamples), dtype=torch.float32, device="cuda")
        la = None
        nlong = 0
        if long_aux is not None and len(long_aux) > 0:
            la = torch.as_tensor(long_aux, dtype=torch.float32, device="cuda").contiguous()
            nlong = la.numel()
        fr = _frame(frame)
        check(lib().wct_impact_waveforms(self._h, C.byref(fr), C.c_void_p(decon.data_ptr()), int(nsamples),
                                         C.c_void_p(la.data_ptr() if la is not None else 0), nlong, int(long_pad),
                                         C.c_void_p(out.data_ptr()), _stream() if stream is None else stream))
        return out
