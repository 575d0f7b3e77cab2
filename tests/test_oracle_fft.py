"""Pins the oracle's FFT, IDFT semantics and Hermitian mirror.

Checks: brute-force long-double DFT on small sizes (all radices incl.
7/11/13 and a large prime), numpy FP64 at full sizes, the IDFT
conformance cases of aux/test/test_idft.cxx, the mirror cases of
aux/test/test_dfttools.cxx and the zero-padding known answer of
sigproc/test/test_zero_padding.cxx (tests/golden/reference_kats.json).
"""
import json
import os

import numpy as np
import pytest

import oracle as O

KATS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 11, 13, 15, 16, 25, 30, 49, 60, 64, 77, 120,
                               121, 143, 169, 480, 101])
def test_fft_vs_brute_force(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = O.dft_naive(x, -1)
    # the brute force itself against the textbook definition on tiny n
    if n <= 8:
        k = np.arange(n)
        mat = np.exp(-2j * np.pi * np.outer(k, k) / n)
        assert np.abs(mat @ x - ref).max() < 1e-13 * max(1, np.abs(ref).max())
    d = O.fwd1d(x, np.float64)
    f = O.fwd1d(x, np.float32)
    scale = np.abs(ref).max()
    assert np.abs(d - ref).max() < 1e-14 * scale * max(1, np.log2(n))
    assert np.abs(f - ref).max() < 3e-7 * scale * max(1, np.log2(n))
    back = O.inv1d(ref, np.float64)
    assert np.abs(back - x).max() < 1e-14 * max(1, np.log2(n)) * np.abs(x).max()
    bn = O.dft_naive(ref, +1) / n
    assert np.abs(bn - x).max() < 1e-13 * np.abs(x).max()


@pytest.mark.parametrize("n", [800, 960, 1000, 1024, 3000, 6000, 8000, 10000, 9595])
def test_fft_vs_numpy(n):
    rng = np.random.default_rng(n + 1)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = np.fft.fft(x)
    assert np.abs(O.fwd1d(x, np.float64) - ref).max() < 5e-15 * np.abs(ref).max() * np.log2(n)
    assert np.abs(O.fwd1d(x, np.float32) - ref).max() < 3e-7 * np.abs(ref).max() * np.log2(n)


def test_parseval_and_linearity():
    rng = np.random.default_rng(5)
    x = rng.standard_normal(6000) + 1j * rng.standard_normal(6000)
    y = rng.standard_normal(6000) + 1j * rng.standard_normal(6000)
    X, Y = O.fwd1d(x, np.float64), O.fwd1d(y, np.float64)
    assert abs(np.sum(np.abs(X) ** 2) / 6000 - np.sum(np.abs(x) ** 2)) < 1e-10 * np.sum(np.abs(x) ** 2)
    Z = O.fwd1d(2 * x - 3j * y, np.float64)
    assert np.abs(Z - (2 * X - 3j * Y)).max() < 1e-11 * np.abs(Z).max()


def test_idft_zero_impulse():
    # aux/test/test_idft.cxx:28-76
    n = KATS["idft"]["oned_size"]
    tol = KATS["idft"]["tolerance"]
    for dt in (np.float32, np.float64):
        assert np.abs(O.fwd1d(np.zeros(n, complex), dt)).max() == 0
        imp = np.zeros(n, complex)
        imp[0] = 1
        f = O.fwd1d(imp, dt)
        assert np.abs(f - 1).max() <= tol
        b = O.inv1d(f, dt)
        assert np.abs(b - imp).max() <= (tol if dt == np.float64 else 1e-7)
    nr, nc = KATS["idft"]["twod_shape"]
    imp2 = np.zeros((nr, nc), complex)
    imp2[0, 0] = 1
    f2 = O.fwd2d(imp2, np.float32)
    assert np.abs(f2 - 1).max() <= tol
    assert np.abs(O.inv2d(f2, np.float32) - imp2).max() <= 1e-7


@pytest.mark.parametrize("shape", [(2, 8), (8, 2)])
@pytest.mark.parametrize("axis", [0, 1])
def test_idft_1b_impulse(shape, axis):
    # aux/test/test_idft.cxx:79-131: impulse -> one unit row (axis 1) or column (axis 0)
    nr, nc = shape
    imp = np.zeros(shape, complex)
    imp[0, 0] = 1
    f = O.fwd1b(imp, axis, np.float32)
    expect = np.zeros(shape)
    if axis:
        expect[0, :] = 1
    else:
        expect[:, 0] = 1
    assert np.abs(np.abs(f) - expect).max() < 1e-7
    assert np.abs(O.inv1b(f, axis, np.float32) - imp).max() < 1e-7


def test_2d_uses_fftw_swapped_extents():
    # FftwDFT::fwd2d plans (ncols, nrows): on a non-square buffer this is
    # the 2D DFT of the row-major (ncols x nrows) view (aux/src/FftwDFT.cxx:186-197).
    rng = np.random.default_rng(2)
    a = rng.standard_normal((3, 5)) + 1j * rng.standard_normal((3, 5))
    got = O.fwd2d(a, np.float64)
    assert np.abs(got - np.fft.fft2(a.reshape(5, 3)).reshape(3, 5)).max() < 1e-12
    assert np.abs(got - np.fft.fft2(a)).max() > 1.0    # not the naive row-major 2D DFT
    back = O.inv2d(got, np.float64)
    assert np.abs(back - a).max() < 1e-13


def test_hermitian_mirror_kats():
    h = KATS["hermitian_mirror"]
    for key in ("even", "odd"):
        x = np.array([complex(*v) for v in h[key + "_in"]])
        want = np.array([complex(*v) for v in h[key + "_out"]])
        got = O.hermitian_mirror(x, np.float32)
        assert np.array_equal(got, want.astype(np.complex64)), key


def test_c2r_ignores_bins_above_nyquist():
    # aux/test/test_dfttools.cxx:267-285 (unreachable in the reference's main)
    size = 64
    imp = np.zeros(size)
    imp[0] = 1
    spec = O.fwd1d(imp, np.float32)
    spec[3 * size // 4] = 42
    wave = O.inv1d(O.hermitian_mirror(spec, np.float32), np.float32).real
    want = np.zeros(size)
    want[0] = 1
    assert np.abs(wave - want).max() < 1e-7


def test_zero_padding_kat():
    # sigproc/test/test_zero_padding.cxx:25-66 with the DftTools r2c/c2r
    z = KATS["zero_padding"]
    a = np.array(z["input"], dtype=np.float64)
    tran = O.fwd1d(a, np.float32)
    n = tran.size
    tran = np.concatenate([tran, np.zeros(n, np.complex64)])
    # odd size: rotate so bins (n+1)/2..n-1 move to the end
    k = (n + 1) // 2
    tran = np.concatenate([tran[:k], tran[n:], tran[k:n]])
    b = O.inv1d(O.hermitian_mirror(tran, np.float32), np.float32).real * (tran.size / n)
    assert np.abs(b - np.array(z["output"])).max() < 5e-6 * 10
