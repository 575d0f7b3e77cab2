"""Fused deconvolution (wct_overall_response, wct_decon_kernel_create,
wct_decon_run) against the FP64 oracle on seeded frames.

Tolerances (FP32 device path vs the FP64 oracle):
  decon output       relative RMS error <= 2e-6, max abs error <= 1e-4 * max|y|
  baseline/RMS       relative error <= 1e-4
  ROI mask           agreement >= 0.9999, and every disagreeing sample lies
                     within 1e-3 * rms of the threshold
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _spec(kind, nw, nt, fold_lf=True):
    from wct_amd import workload as W
    return W.plane_spec(kind, nw, nt, fold_lf=fold_lf)


def _oracle_resp(spec):
    from wct_amd import workload as W
    return O.overall_response(spec.fr_rows, spec.elec, W.FINE_PERIOD, W.TICK, W.INTER_GAIN,
                              W.ADC_MV, W.FTOFFSET, spec.nt, dtype=np.float64)


def _frame(spec, resp, seed, ntracks):
    from wct_amd import workload as W
    q = O.tracks(seed, 1, ntracks, spec.nw, spec.nt, W.Q_PER_TICK)
    adc = O.forward_conv(q, resp, spec.wire_shift, spec.time_shift)
    adc += O.normal(seed, 2, spec.nw * spec.nt, W.NOISE_ADC).reshape(spec.nw, spec.nt)
    return adc.astype(np.float32)


def _rel_rms(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


@pytest.mark.parametrize("kind,nt", [("col", 6000), ("ind", 6000), ("col", 1000)])
def test_overall_response_fp64(dev, kind, nt):
    from wct_amd import decon
    spec = _spec(kind, 64, nt)
    got = decon.overall_response(spec)
    ref = _oracle_resp(spec)
    assert got.shape == ref.shape
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_kernel_matches_fp64_definition(dev):
    from wct_amd import decon
    spec = _spec("ind", 96, 1000)
    pk = decon.PlaneKernel(spec, cut_tol=0.0)
    K = pk.download()                              # kc x nw
    resp = pk.resp
    R = np.zeros((spec.nw, spec.nt), complex)
    R[:resp.shape[0]] = resp
    R = np.fft.fft2(R)
    m = np.arange(spec.nw)[:, None]
    k = np.arange(spec.nt // 2 + 1)[None, :]
    ph = np.exp(-2j * np.pi * (m * spec.wire_shift / spec.nw + k * spec.time_shift / spec.nt))
    Rh = R[:, :spec.nt // 2 + 1]
    with np.errstate(divide="ignore", invalid="ignore"):
        ref = np.where(Rh != 0, spec.wire_filter[:, None] * spec.time_filter[None, :spec.nt // 2 + 1]
                       * ph / (Rh * spec.nw * spec.nt), 0)
    assert pk.kc == spec.nt // 2 + 1
    err = np.abs(K.T - ref).max() / np.abs(ref).max()
    assert err < 2e-7


CASES = [("col", 480, 6000, 20), ("ind", 800, 6000, 33), ("col", 64, 1000, 3), ("ind", 96, 2000, 4)]


@pytest.mark.parametrize("kind,nw,nt,ntr", CASES)
def test_decon_vs_fp64_oracle(dev, kind, nw, nt, ntr):
    import torch
    from wct_amd import decon
    spec = _spec(kind, nw, nt)
    resp = _oracle_resp(spec)
    adc = _frame(spec, resp, seed=nw + nt, ntracks=ntr)
    ref = O.decon_fused(adc.astype(np.float64), resp, spec.wire_filter, spec.time_filter,
                        spec.wire_shift, spec.time_shift)
    pk = decon.PlaneKernel(spec, cut_tol=1e-7)
    d_adc = torch.from_numpy(adc).to(dev)
    out = torch.empty_like(d_adc)
    # raw decon (no baseline) first
    b = decon.Batch([decon.make_job(pk, d_adc, out, restore=0)])
    b.run()
    torch.cuda.synchronize()
    y = out.cpu().numpy().astype(np.float64)
    e = _rel_rms(y, ref)
    print(f"{kind} {nw}x{nt}: kc={pk.kc} rel_rms={e:.3e} max={np.abs(y-ref).max()/np.abs(ref).max():.3e}")
    assert e <= 2e-6
    assert np.abs(y - ref).max() <= 1e-4 * np.abs(ref).max()


@pytest.mark.parametrize("kind,nw,nt,ntr", CASES[:2] + CASES[3:])
def test_baseline_rms_mask(dev, kind, nw, nt, ntr):
    import torch
    from wct_amd import decon, workload as W
    spec = _spec(kind, nw, nt)
    resp = _oracle_resp(spec)
    adc = _frame(spec, resp, seed=7 * nw + nt, ntracks=ntr)
    y64 = O.decon_fused(adc.astype(np.float64), resp, spec.wire_filter, spec.time_filter,
                        spec.wire_shift, spec.time_shift)
    ybl, base = O.restore_baseline(y64, dtype=np.float64)
    rms_ref, th_ref, mask_ref = O.roi_threshold(ybl, W.TH_FACTOR, dtype=np.float64)
    pk = decon.PlaneKernel(spec, cut_tol=1e-7)
    d_adc = torch.from_numpy(adc).to(dev)
    out = torch.empty_like(d_adc)
    mask = torch.empty(adc.shape, dtype=torch.uint8, device=dev)
    rms = torch.empty(nw, dtype=torch.float32, device=dev)
    b = decon.Batch([decon.make_job(pk, d_adc, out, mask=mask, rms=rms, restore=1)])
    b.run()
    torch.cuda.synchronize()
    y = out.cpu().numpy().astype(np.float64)
    assert _rel_rms(y, ybl) <= 5e-6
    r = rms.cpu().numpy()
    assert np.abs(r - rms_ref).max() <= 1e-4 * np.abs(rms_ref).max()
    m = mask.cpu().numpy() & 1
    agree = np.mean(m == mask_ref)
    bad = np.argwhere(m != mask_ref)
    for (w, t) in bad[:50]:
        assert abs(ybl[w, t] - th_ref[w]) <= 1e-3 * rms_ref[w] + 1e-3
    print(f"mask agreement {agree:.6f}, mismatches {len(bad)}")
    assert agree >= 0.9999


def test_batch_of_uneven_planes(dev):
    """Planes of different wire and tick counts in one run equal single runs."""
    import torch
    from wct_amd import decon
    shapes = [("col", 96, 2000), ("ind", 80, 1000), ("col", 96, 1200)]
    kernels, adcs, outs, jobs, singles = [], [], [], [], []
    for i, (kind, nw, nt) in enumerate(shapes):
        spec = _spec(kind, nw, nt)
        resp = _oracle_resp(spec)
        adc = torch.from_numpy(_frame(spec, resp, seed=100 + i, ntracks=3)).to(dev)
        pk = decon.PlaneKernel(spec, cut_tol=1e-7)
        out = torch.empty_like(adc)
        kernels.append(pk)
        adcs.append(adc)
        outs.append(out)
        jobs.append(decon.make_job(pk, adc, out))
        o1 = torch.empty_like(adc)
        decon.Batch([decon.make_job(pk, adc, o1)]).run()
        singles.append(o1)
    decon.Batch(jobs).run()
    torch.cuda.synchronize()
    for o, s in zip(outs, singles):
        assert torch.equal(o, s)
