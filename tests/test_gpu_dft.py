"""IDFT drop-in (include/wct_hip.h, wct_dft_*) against the oracle and
numpy FP64, with the cases of the reference's IDFT conformance test
(aux/test/test_idft.cxx:28-237)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _dev_cplx(a, dev):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a, np.complex64)).to(dev)
    return t


def _call(name, *args):
    import torch
    from wct_amd import lib, check
    s = torch.cuda.current_stream().cuda_stream
    check(getattr(lib(), name)(*args, s))
    torch.cuda.synchronize()


def _relerr(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)


SIZES = [1, 2, 3, 5, 6, 8, 15, 16, 25, 30, 60, 64, 77, 120, 121, 128, 143, 169, 480, 800,
         960, 1000, 1024, 3000, 4096, 6000, 6272, 8000]


@pytest.mark.parametrize("n", SIZES)
def test_fwd_inv_1d(dev, n):
    rng = np.random.default_rng(n)
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    ref = np.fft.fft(x.astype(np.complex128))
    t = _dev_cplx(x, dev)
    _call("wct_dft_fwd1d", t.data_ptr(), t.data_ptr(), n)
    got = t.cpu().numpy()
    assert _relerr(got, ref) < 5e-6, n
    _call("wct_dft_inv1d", t.data_ptr(), t.data_ptr(), n)
    back = t.cpu().numpy()
    assert np.abs(back - x).max() < 5e-6 * max(1, np.abs(x).max())


def test_impulse_and_zero(dev):
    # test_idft.cxx:28-76: zero -> zero, impulse -> flat -> impulse
    import torch
    n = 1024
    z = torch.zeros(n, dtype=torch.complex64, device=dev)
    _call("wct_dft_fwd1d", z.data_ptr(), z.data_ptr(), n)
    assert z.abs().max().item() == 0
    x = torch.zeros(n, dtype=torch.complex64, device=dev)
    x[0] = 1
    f = torch.empty_like(x)
    _call("wct_dft_fwd1d", x.data_ptr(), f.data_ptr(), n)
    assert torch.allclose(f, torch.ones_like(f), atol=1e-6)
    _call("wct_dft_inv1d", f.data_ptr(), f.data_ptr(), n)
    ref = np.zeros(n, np.complex64)
    ref[0] = 1
    assert np.abs(f.cpu().numpy() - ref).max() < 1e-6
    # 2d 128x128
    x2 = torch.zeros(128 * 128, dtype=torch.complex64, device=dev)
    x2[0] = 1
    _call("wct_dft_fwd2d", x2.data_ptr(), x2.data_ptr(), 128, 128)
    assert torch.allclose(x2, torch.ones_like(x2), atol=1e-6)
    _call("wct_dft_inv2d", x2.data_ptr(), x2.data_ptr(), 128, 128)
    r2 = np.zeros(128 * 128, np.complex64)
    r2[0] = 1
    assert np.abs(x2.cpu().numpy() - r2).max() < 1e-6


@pytest.mark.parametrize("shape", [(2, 8), (8, 2), (3, 5), (800, 6000), (960, 1000), (33, 77), (9, 6272), (5, 8000)])
@pytest.mark.parametrize("axis", [0, 1])
def test_1b(dev, shape, axis):
    # test_idft.cxx:79-131 impulse on (2,8)/(8,2) plus random data
    nr, nc = shape
    rng = np.random.default_rng(nr * 7 + nc)
    x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    ref = np.fft.fft(x.astype(np.complex128), axis=axis)
    t = _dev_cplx(x, dev)
    out = t.clone()
    _call("wct_dft_fwd1b", t.data_ptr(), out.data_ptr(), nr, nc, axis)
    assert _relerr(out.cpu().numpy(), ref) < 5e-6
    # oracle (FftwDFT semantics) agrees with numpy here
    assert _relerr(O.fwd1b(x, axis, np.float64), ref) < 1e-12
    _call("wct_dft_inv1b", out.data_ptr(), out.data_ptr(), nr, nc, axis)
    assert np.abs(out.cpu().numpy() - x).max() < 5e-6 * np.abs(x).max()


@pytest.mark.parametrize("shape", [(2, 8), (8, 2), (5, 3), (128, 128), (96, 40)])
def test_2d_fftw_semantics(dev, shape):
    nr, nc = shape
    rng = np.random.default_rng(nr + 100 * nc)
    x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    ref = O.fwd2d(x, np.float64)     # FftwDFT's (ncols x nrows) view
    t = _dev_cplx(x, dev)
    _call("wct_dft_fwd2d", t.data_ptr(), t.data_ptr(), nr, nc)
    assert _relerr(t.cpu().numpy(), ref) < 5e-6
    _call("wct_dft_inv2d", t.data_ptr(), t.data_ptr(), nr, nc)
    assert np.abs(t.cpu().numpy() - x).max() < 5e-6 * np.abs(x).max()


@pytest.mark.parametrize("shape", [(2, 8), (8, 2), (37, 65), (800, 600)])
def test_transpose(dev, shape):
    # test_idft.cxx:181-202 (copy and in place, real and complex)
    import torch
    nr, nc = shape
    a = np.arange(nr * nc, dtype=np.float32).reshape(nr, nc)
    t = torch.from_numpy(a).to(dev)
    o = torch.empty_like(t)
    _call("wct_dft_transpose_real", t.data_ptr(), o.data_ptr(), nr, nc)
    assert np.array_equal(o.cpu().numpy().reshape(nc, nr), a.T)
    _call("wct_dft_transpose_real", t.data_ptr(), t.data_ptr(), nr, nc)
    assert np.array_equal(t.cpu().numpy().reshape(nc, nr), a.T)
    c = (a + 1j * (a + 0.5)).astype(np.complex64)
    tc = _dev_cplx(c, dev)
    _call("wct_dft_transpose_complex", tc.data_ptr(), tc.data_ptr(), nr, nc)
    assert np.array_equal(tc.cpu().numpy().reshape(nc, nr), c.T)


@pytest.mark.parametrize("shape", [(21, 1024), (800, 6000), (960, 6000)])
def test_f64_1b(dev, shape):
    import torch
    nr, nc = shape
    rng = np.random.default_rng(3)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    for axis in (0, 1):
        t = torch.from_numpy(x.copy()).to(dev)
        _call("wct_dft_fwd1b_f64", t.data_ptr(), t.data_ptr(), nr, nc, axis)
        ref = np.fft.fft(x, axis=axis)
        assert _relerr(t.cpu().numpy(), ref) < 1e-13
        _call("wct_dft_inv1b_f64", t.data_ptr(), t.data_ptr(), nr, nc, axis)
        assert np.abs(t.cpu().numpy() - x).max() < 1e-13 * np.abs(x).max()
