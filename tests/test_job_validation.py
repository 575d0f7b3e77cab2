"""make_job checks every tensor against the kernel's plane shape before
anything reaches the C side (the C ABI trusts pointers and strides, so a
short or strided tensor would be read or written out of bounds on the
device).  CPU only: no launch happens."""
import pytest
import torch

from wct_amd import decon


class _Kernel:
    """Stands in for a PlaneKernel: only the shape fields make_job reads."""
    handle = None

    def __init__(self, nw, nt, kc):
        self.nw, self.nt, self.kc = nw, nt, kc


NW, NT, KC = 8, 20, 6


def _frames():
    return torch.zeros((NW, NT)), torch.zeros((NW, NT))


@pytest.mark.parametrize("what,msg", [
    ("short_adc", "adc: shape"),
    ("f64_adc", "adc: dtype"),
    ("strided_out", "out: inner stride"),
    ("small_out", "out: shape"),
    ("mask_dtype", "mask: dtype"),
    ("rms_len", "rms: shape"),
    ("lf_short", "lf_tight: 5 values"),
    ("bad_ptr_len", "bad_ptr: shape"),
    ("bad_hi_len", "bad_hi: shape"),
    ("spectrum_dtype", "spectrum: dtype"),
    ("cpu", "expected a GPU device"),
])
def test_make_job_rejects_bad_tensors(what, msg):
    k = _Kernel(NW, NT, KC)
    adc, out = _frames()
    kw = {}
    if what == "short_adc":
        adc = torch.zeros((NW - 1, NT))
    elif what == "f64_adc":
        adc = torch.zeros((NW, NT), dtype=torch.float64)
    elif what == "strided_out":
        out = torch.zeros((NW, 2 * NT))[:, ::2]
    elif what == "small_out":
        out = torch.zeros((NW, NT - 2))
    elif what == "mask_dtype":
        kw["mask"] = torch.zeros((NW, NT), dtype=torch.int32)
    elif what == "rms_len":
        kw["rms"] = torch.zeros(NW + 1)
    elif what == "lf_short":
        kw["lf_tight"] = torch.ones(KC - 1)
    elif what == "bad_ptr_len":
        kw["bad"] = tuple(torch.zeros(n, dtype=torch.int32) for n in (NW, 1, 1))
    elif what == "bad_hi_len":
        kw["bad"] = tuple(torch.zeros(n, dtype=torch.int32) for n in (NW + 1, 2, 1))
    elif what == "spectrum_dtype":
        with pytest.raises(ValueError, match=msg):
            decon.make_job(k, adc, None, spectrum=torch.zeros((NW, NT), dtype=torch.complex128))
        return
    with pytest.raises(ValueError, match=msg):
        decon.make_job(k, adc, out, **kw)

