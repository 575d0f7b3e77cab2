"""wct_spectrum_filter_noalloc: the spectrum filter with its job table and
scaled filter in the library's static device scratch instead of
allocated memory.  Two calls in a row, the second smaller, each equal
bit for bit to wct_spectrum_filter on the same inputs (the same kernels
run on the same data); an unprepared length and a length beyond the
largest transform are refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(dev, nw, nt, seed):
    import torch
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((nw, nt)).astype(np.float32)
    spec = torch.from_numpy(np.fft.fft(x, axis=1).astype(np.complex64)).to(dev)
    filt = torch.from_numpy(rng.uniform(0.0, 1.0, nt // 2 + 1).astype(np.float32)).to(dev)
    return spec, filt


def _run(fn, spec, filt, out, restore):
    import torch
    from wct_amd.capi import check, lib
    nw, nt = out.shape
    check(getattr(lib(), fn)(spec.data_ptr(), spec.stride(0), nw, nt, filt.data_ptr(), restore, out.data_ptr(),
                             out.stride(0), torch.cuda.current_stream().cuda_stream))


def test_noalloc_matches_allocating_entry_point(dev):
    import torch
    from wct_amd.capi import WctError, check, lib
    print("noalloc spectrum filter: refusals", flush=True)
    # a length whose tables do not exist yet is refused, nothing allocated
    spec, filt = _case(dev, 8, 2002, 1)
    out = torch.empty((8, 2002), device=dev)
    with pytest.raises(WctError, match="no transform tables"):
        _run("wct_spectrum_filter_noalloc", spec, filt, out, 1)
    assert lib().wct_noalloc_prepare(20000) != 0
    for nw, nt, seed in ((96, 6000, 2), (40, 1000, 3)):     # the second call smaller
        print(f"noalloc spectrum filter: {nw} x {nt}", flush=True)
        spec, filt = _case(dev, nw, nt, seed)
        for restore in (1, 0):
            ref = torch.empty((nw, nt), device=dev)
            _run("wct_spectrum_filter", spec, filt, ref, restore)
            check(lib().wct_noalloc_prepare(nt))
            got = torch.full((nw, nt), float("nan"), device=dev)
            _run("wct_spectrum_filter_noalloc", spec, filt, got, restore)
            torch.cuda.synchronize()
            assert torch.equal(got, ref), (nw, nt, restore)
