"""IDFT lengths with a prime factor above 13 (Bluestein's chirp-z
algorithm on 5-smooth transforms), as FftwDFT takes any length: the
reference's own benchmark sizes 9503, 9592, 9595 and its 2D shapes
(2400, 9595) and (3456, 9595) (aux/test/check_idft_bench.cxx:90,105),
small primes, both 1b axes, 2D with FftwDFT's swapped extents, the
test_idft.cxx impulse/zero cases at those sizes (aux/test/
test_idft.cxx:28-131), FP64 transforms, and the plain per-plane decon
of a 9595-tick frame.

Tolerance: against numpy FP64, max error <= 1e-5 of the largest output
(the chirp is formed in FP64 and the two length-L transforms are FP32:
about 3e-7 typical); impulse -> flat within 2e-6 absolute (the
reference's own 1e-8 holds for the native lengths, where no rounding
enters an impulse's transform; a Bluestein transform rounds in its
chirp products).
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _dev_cplx(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.complex64)).to(dev)


def _call(name, *args):
    import torch
    from wct_amd import lib, check
    s = torch.cuda.current_stream().cuda_stream
    check(getattr(lib(), name)(*args, s))
    torch.cuda.synchronize()


def _relerr(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)


@pytest.mark.parametrize("n", [7 * 17, 17, 19, 97, 101, 2 * 17, 1009, 9503, 9592, 9595, 65537])
def test_fwd_inv_1d_any_length(dev, n):
    from wct_amd import lib
    assert lib().wct_dft_supported(n) == 1 and lib().wct_dft_native_length(n) == 0
    rng = np.random.default_rng(n)
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    ref = np.fft.fft(x.astype(np.complex128))
    t = _dev_cplx(x, dev)
    _call("wct_dft_fwd1d", t.data_ptr(), t.data_ptr(), n)
    e = _relerr(t.cpu().numpy(), ref)
    print(f"  n={n} fwd max rel {e:.2e}", flush=True)
    assert e < 1e-5, n
    _call("wct_dft_inv1d", t.data_ptr(), t.data_ptr(), n)
    assert np.abs(t.cpu().numpy() - x).max() < 1e-5 * np.abs(x).max()


@pytest.mark.parametrize("n", [17, 9503, 9595])
def test_impulse_and_zero_any_length(dev, n):
    import torch
    z = torch.zeros(n, dtype=torch.complex64, device=dev)
    _call("wct_dft_fwd1d", z.data_ptr(), z.data_ptr(), n)
    assert z.abs().max().item() == 0
    for at in (0, 3):
        x = torch.zeros(n, dtype=torch.complex64, device=dev)
        x[at] = 1
        f = torch.empty_like(x)
        _call("wct_dft_fwd1d", x.data_ptr(), f.data_ptr(), n)
        want = np.exp(-2j * np.pi * at * np.arange(n) / n)
        assert np.abs(f.cpu().numpy() - want).max() < 2e-6
        _call("wct_dft_inv1d", f.data_ptr(), f.data_ptr(), n)
        ref = np.zeros(n, np.complex64)
        ref[at] = 1
        assert np.abs(f.cpu().numpy() - ref).max() < 2e-6


@pytest.mark.parametrize("shape,axis", [((2400, 9595), 1), ((3456, 9595), 1), ((101, 64), 0), ((9595, 4), 0),
                                        ((17, 19), 0), ((17, 19), 1)])
def test_1b_any_length(dev, shape, axis):
    nr, nc = shape
    print(f"  1b {shape} axis {axis}", flush=True)
    rng = np.random.default_rng(nr + nc)
    x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    ref = np.fft.fft(x.astype(np.complex128), axis=axis)
    t = _dev_cplx(x, dev)
    out = t.clone()
    _call("wct_dft_fwd1b", t.data_ptr(), out.data_ptr(), nr, nc, axis)
    assert _relerr(out.cpu().numpy(), ref) < 1e-5
    del ref
    _call("wct_dft_inv1b", out.data_ptr(), out.data_ptr(), nr, nc, axis)
    assert np.abs(out.cpu().numpy() - x).max() < 1e-5 * np.abs(x).max()


@pytest.mark.parametrize("shape", [(97, 101), (17, 64), (64, 17)])
def test_2d_any_length(dev, shape):
    nr, nc = shape
    rng = np.random.default_rng(nr * nc)
    x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    ref = O.fwd2d(x, np.float64)     # FftwDFT's (ncols x nrows) view
    t = _dev_cplx(x, dev)
    _call("wct_dft_fwd2d", t.data_ptr(), t.data_ptr(), nr, nc)
    assert _relerr(t.cpu().numpy(), ref) < 1e-5
    _call("wct_dft_inv2d", t.data_ptr(), t.data_ptr(), nr, nc)
    assert np.abs(t.cpu().numpy() - x).max() < 1e-5 * np.abs(x).max()


def test_f64_any_length(dev):
    import torch
    rng = np.random.default_rng(9)
    x = rng.standard_normal((6, 9595)) + 1j * rng.standard_normal((6, 9595))
    for axis in (1, 0):
        t = torch.from_numpy(x.copy()).to(dev)
        _call("wct_dft_fwd1b_f64", t.data_ptr(), t.data_ptr(), 6, 9595, axis)
        assert _relerr(t.cpu().numpy(), np.fft.fft(x, axis=axis)) < 1e-12
        _call("wct_dft_inv1b_f64", t.data_ptr(), t.data_ptr(), 6, 9595, axis)
        assert np.abs(t.cpu().numpy() - x).max() < 1e-12 * np.abs(x).max()


def test_plain_decon_on_a_9595_tick_frame(dev):
    """The plain per-plane path (the reference's per-step control flow)
    on MicroBooNE-length frames: 9595 ticks = 5 x 19 x 101, against the
    FP64 oracle on a well-conditioned response."""
    import torch
    from wct_amd import decon, workload as W
    spec = W.plane_spec("col", 64, 9595, fold_lf=True)
    nr = spec.fr_rows.shape[0]
    c = (nr - 1) // 2
    resp = np.zeros((nr, spec.nt))
    for r in range(nr):
        a = 1.0 / (1.0 + (r - c) ** 2)
        resp[r, 3] = 50 * a
        resp[r, 4] = 15 * a
    q = O.tracks(5, 1, 3, spec.nw, spec.nt, W.Q_PER_TICK)
    adc = O.forward_conv(q, resp, spec.wire_shift, spec.time_shift)
    adc = (adc + O.normal(5, 2, spec.nw * spec.nt, W.NOISE_ADC).reshape(spec.nw, spec.nt)).astype(np.float32)
    pp = decon.PlainPlane(spec, resp=resp, device=dev)
    a = torch.from_numpy(adc).to(dev)
    y = torch.empty_like(a)
    pp.run(a, y)
    torch.cuda.synchronize()
    y64 = O.decon_fused(adc.astype(np.float64), resp, spec.wire_filter, spec.time_filter, spec.wire_shift,
                        spec.time_shift)
    ybl, _ = O.restore_baseline(y64, dtype=np.float64)
    got = y.cpu().numpy().astype(np.float64)
    e = float(np.sqrt(np.sum((got - ybl) ** 2) / np.sum(ybl ** 2)))
    print(f"  plain 64x9595 rel_rms {e:.2e}", flush=True)
    assert e <= 5e-5
