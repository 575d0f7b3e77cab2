"""Failure paths forced from outside (the reference's contract: errors
are exceptions, util/inc/WireCellUtil/Exceptions.h:43), with the C ABI's
hooks: an injected status and an allocation cap.

  - the C ABI: a kernel build under the cap fails as out of memory and
    succeeds once the cap is lifted; the spectrum filter's allocating
    entry point fails under a 1-byte cap while the no-alloc one still
    runs and gives the same values as before;
  - the components: HipOmnibusSigProc (a failed build under the cap, an
    injected failure of the batched run), HipDFT and Decon2DResponse
    turn the status into an exception, and the next call on the same
    instance gives exactly what a fresh instance gives.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture
def hooks(dev):
    from wct_amd.capi import lib
    L = lib()
    yield L
    L.wct_debug_inject(None, 0, 0)
    L.wct_debug_alloc_cap(0)


def test_kernel_build_under_allocation_cap(hooks):
    from wct_amd import decon, workload as W
    from wct_amd.capi import WctError
    spec = W.plane_spec("col", 96, 2000)
    hooks.wct_debug_alloc_cap(1 << 16)
    with pytest.raises(WctError, match="wct status -4: .*above the 65536-byte cap"):
        decon.PlaneKernel(spec)
    hooks.wct_debug_alloc_cap(0)
    pk = decon.PlaneKernel(spec)
    assert pk.kc > 0
    pk.close()


def test_noalloc_entry_point_runs_under_a_one_byte_cap(hooks, dev):
    import torch
    from wct_amd.capi import WctError, check
    from test_zz_no_alloc import _case, _run
    nw, nt = 64, 2000
    spec, filt = _case(dev, nw, nt, 9)
    ref = torch.empty((nw, nt), device=dev)
    _run("wct_spectrum_filter", spec, filt, ref, 1)
    check(hooks.wct_noalloc_prepare(nt))
    hooks.wct_debug_alloc_cap(1)
    out = torch.empty((nw, nt), device=dev)
    with pytest.raises(WctError, match="cap"):
        _run("wct_spectrum_filter", spec, filt, out, 1)
    got = torch.full((nw, nt), float("nan"), device=dev)
    _run("wct_spectrum_filter_noalloc", spec, filt, got, 1)
    torch.cuda.synchronize()
    assert torch.equal(got, ref)


def _sigproc(K, tmp_path):
    from test_gpu_components import _sigproc_config
    K.reset()
    K.configure(_sigproc_config(tmp_path))


def _traces(out):
    return [(ch, tb, q.copy()) for ch, tb, q in out.traces()]


def test_sigproc_recovers_from_failures(hooks, tmp_path):
    from wct_amd import components as K
    from wct_amd.capi import WctError
    from test_gpu_components import _adc_planes, _frame
    print("sigproc failure paths", flush=True)
    adc = _adc_planes(12)
    bad = {(0, 5): [(100, 300)]}
    _sigproc(K, tmp_path)
    want = _traces(K.frame_filter("HipOmnibusSigProc:osp", _frame(K, adc, bad)))
    # a build under the allocation cap: the first frame fails, the next
    # one (cap lifted) builds and matches
    _sigproc(K, tmp_path)
    hooks.wct_debug_alloc_cap(1 << 16)
    with pytest.raises(WctError, match="cap"):
        K.frame_filter("HipOmnibusSigProc:osp", _frame(K, adc, bad))
    hooks.wct_debug_alloc_cap(0)
    got = _traces(K.frame_filter("HipOmnibusSigProc:osp", _frame(K, adc, bad)))
    assert len(got) == len(want)
    for (c1, t1, q1), (c2, t2, q2) in zip(got, want):
        assert (c1, t1) == (c2, t2)
        np.testing.assert_array_equal(q1, q2)
    # an injected failure of the batched run, then a good frame
    hooks.wct_debug_inject(b"wct_decon_plan_run", 0, -2)
    with pytest.raises(WctError, match="injected failure"):
        K.frame_filter("HipOmnibusSigProc:osp", _frame(K, adc, bad))
    got = _traces(K.frame_filter("HipOmnibusSigProc:osp", _frame(K, adc, bad)))
    for (c1, t1, q1), (c2, t2, q2) in zip(got, want):
        np.testing.assert_array_equal(q1, q2)
    K.reset()


def test_hipdft_and_decon2d_turn_statuses_into_exceptions(hooks, tmp_path):
    from wct_amd import components as K
    from wct_amd.capi import WctError
    from test_gpu_tensorset import _components_setup
    K.reset()
    K.configure([{"type": "HipDFT", "name": ""}])
    x = (np.arange(64 * 100) % 7).astype(np.complex64).reshape(64, 100)
    want = K.dft("HipDFT", "fwd1b", x, 1)
    hooks.wct_debug_inject(b"wct_dft_fwd1b", 0, -2)
    with pytest.raises(WctError, match="injected failure"):
        K.dft("HipDFT", "fwd1b", x, 1)
    np.testing.assert_array_equal(K.dft("HipDFT", "fwd1b", x, 1), want)
    _, _, _, s = _components_setup(K, tmp_path)
    r0 = K.tensorset_filter("Decon2DResponse:r", s)
    want = r0.array(r0.find("trace_tag", "waveform"))
    _, _, _, s = _components_setup(K, tmp_path)
    hooks.wct_debug_inject(b"wct_decon_kernel_create", 0, -4)
    with pytest.raises(WctError, match="injected failure"):
        K.tensorset_filter("Decon2DResponse:r", s)
    r1 = K.tensorset_filter("Decon2DResponse:r", s)
    np.testing.assert_array_equal(r1.array(r1.find("trace_tag", "waveform")), want)
    K.reset()
