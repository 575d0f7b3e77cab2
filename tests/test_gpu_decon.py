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


def _run(pk, adc, dev, restore=1, mask=True, bad=None, lf_tight=None, lf_tighter=None,
         save_mode=0):
    import torch
    from wct_amd import decon
    nw, nt = adc.shape
    d_adc = torch.from_numpy(adc).to(dev)
    out = torch.empty_like(d_adc)
    m = torch.full(adc.shape, 255, dtype=torch.uint8, device=dev) if mask else None
    rms = torch.empty(nw, dtype=torch.float32, device=dev)
    dbad = None
    if bad:
        ptr = np.zeros(nw + 1, np.int32)
        rows = sorted(bad)
        for (w, lo, hi) in rows:
            ptr[w + 1:] += 1
        dbad = tuple(torch.from_numpy(np.ascontiguousarray(a, np.int32)).to(dev) for a in
                     (ptr, [b[1] for b in rows], [b[2] for b in rows]))
    lt = None if lf_tight is None else torch.from_numpy(np.ascontiguousarray(lf_tight, np.float32)).to(dev)
    ltt = None if lf_tighter is None else torch.from_numpy(np.ascontiguousarray(lf_tighter, np.float32)).to(dev)
    job = decon.make_job(pk, d_adc, out, mask=m, rms=rms, restore=restore, bad=dbad,
                         lf_tight=lt, lf_tighter=ltt, save_mode=save_mode)
    decon.Batch([job]).run()
    torch.cuda.synchronize()
    return (out.cpu().numpy(), None if m is None else m.cpu().numpy(), rms.cpu().numpy())


@pytest.mark.parametrize("kind,nw,nt,ntr", CASES[:2] + CASES[3:])
def test_baseline_rms_mask(dev, kind, nw, nt, ntr):
    """restore_baseline / cal_RMS / threshold on the device are exact
    order statistics of the device's own decon (bit-identical to the
    oracle's float path on the same data); against the FP64 chain they
    agree to the order-statistic spacing."""
    from wct_amd import decon, workload as W
    spec = _spec(kind, nw, nt)
    resp = _oracle_resp(spec)
    adc = _frame(spec, resp, seed=7 * nw + nt, ntracks=ntr)
    pk = decon.PlaneKernel(spec, cut_tol=1e-7)
    raw, _, _ = _run(pk, adc, dev, restore=0, mask=False)
    y, m, r = _run(pk, adc, dev, restore=1)
    # exactness on identical data
    ybl_f, _ = O.restore_baseline(raw, dtype=np.float32)
    assert np.array_equal(y, ybl_f)
    # the device sums x^2 in FP32 per thread (at most 24 terms, relative
    # error <= 24 * 2^-24) and in double across threads: compare with the
    # oracle's double accumulation on the same data (the RMS, a square
    # root, halves that bound), and loosely with the reference's float
    # accumulation (which drifts ~1e-6 on 6000 samples)
    rms_d, _, _ = O.roi_threshold(ybl_f.astype(np.float64), W.TH_FACTOR, dtype=np.float64)
    rms_f, th_f, mask_f = O.roi_threshold(ybl_f, W.TH_FACTOR, dtype=np.float32)
    assert np.all(np.abs(r - rms_d) <= 1e-6 * rms_d + 1e-30)
    assert np.all(np.abs(r - rms_f) <= 2e-5 * rms_f + 1e-30)
    mism = np.argwhere((m & 1) != mask_f)
    for (w, t) in mism:
        assert abs(float(ybl_f[w, t]) - th_f[w]) <= 1e-5 * th_f[w]
    assert np.all(m <= 1)
    # against the FP64 chain
    y64 = O.decon_fused(adc.astype(np.float64), resp, spec.wire_filter, spec.time_filter,
                        spec.wire_shift, spec.time_shift)
    ybl, base = O.restore_baseline(y64, dtype=np.float64)
    rms_ref, th_ref, mask_ref = O.roi_threshold(ybl, W.TH_FACTOR, dtype=np.float64)
    e = _rel_rms(y.astype(np.float64), ybl)
    agree = np.mean((m & 1) == mask_ref)
    print(f"{kind} {nw}x{nt}: after baseline rel_rms={e:.3e} rms max rel="
          f"{np.abs(r - rms_ref).max() / np.abs(rms_ref).max():.3e} mask agreement={agree:.7f}")
    assert e <= 5e-5
    assert np.abs(r - rms_ref).max() <= 1e-4 * np.abs(rms_ref).max()
    assert agree >= 0.9999


def test_bad_channels_and_save_mode(dev):
    """Bad ranges: [lo, hi) zeroed on input (load_data), [lo, hi] left out
    of the RMS and the mask (find_ROI_by_decon_itself), and save_data's
    nominal mode zeroes negatives and [lo, hi)."""
    from wct_amd import decon, workload as W
    spec = _spec("col", 96, 2000)
    resp = _oracle_resp(spec)
    adc = _frame(spec, resp, seed=31, ntracks=4)
    bad = [(3, 100, 400), (3, 1500, 1600), (50, 0, 2000), (95, 1990, 1999)]
    pk = decon.PlaneKernel(spec, cut_tol=1e-7)
    raw, _, _ = _run(pk, adc, dev, restore=0, mask=False, bad=bad)
    adc_z = adc.copy()
    for (w, lo, hi) in bad:
        adc_z[w, lo:hi] = 0
    ref = O.decon_fused(adc_z.astype(np.float64), resp, spec.wire_filter, spec.time_filter,
                        spec.wire_shift, spec.time_shift)
    assert _rel_rms(raw.astype(np.float64), ref) <= 2e-6
    y, m, r = _run(pk, adc, dev, restore=1, bad=bad)
    ybl_f, _ = O.restore_baseline(raw, dtype=np.float32)
    assert np.array_equal(y, ybl_f)
    rms_f, th_f, mask_f = O.roi_threshold(ybl_f, W.TH_FACTOR, dtype=np.float32,
                                          bad=[(w, lo, hi) for (w, lo, hi) in bad])
    assert np.all(np.abs(r - rms_f) <= 2e-5 * rms_f + 1e-30)
    assert np.mean((m & 1) == mask_f) >= 0.99999
    for (w, lo, hi) in bad:
        assert np.all(m[w, lo:min(hi + 1, spec.nt)] == 0)
    ys, _, _ = _run(pk, adc, dev, restore=1, bad=bad, save_mode=1)
    want = np.where(ybl_f > 0, ybl_f, 0).astype(np.float32)
    for (w, lo, hi) in bad:
        want[w, lo:hi] = 0
    assert np.array_equal(ys, want)


def test_tight_and_tighter_outputs(dev):
    """Induction plane with the LF filters applied per output: the tight
    output uses Wiener x ROI_tight_lf; bit 1 of the mask is the tighter
    decon (Wiener x ROI_tighter_lf) above the tight threshold."""
    from wct_amd import decon, workload as W
    spec = _spec("ind", 128, 2000, fold_lf=False)
    resp = _oracle_resp(spec)
    adc = _frame(spec, resp, seed=41, ntracks=6)
    pk = decon.PlaneKernel(spec, cut_tol=1e-7)
    lt, ltt = spec.lf_tight[:pk.kc], spec.lf_tighter[:pk.kc]
    y, m, r = _run(pk, adc, dev, restore=1, lf_tight=lt, lf_tighter=ltt)
    y_t = O.decon_fused(adc.astype(np.float64), resp, spec.wire_filter,
                        spec.time_filter * spec.lf_tight, spec.wire_shift, spec.time_shift)
    y_tt = O.decon_fused(adc.astype(np.float64), resp, spec.wire_filter,
                         spec.time_filter * spec.lf_tighter, spec.wire_shift, spec.time_shift)
    bl_t, _ = O.restore_baseline(y_t, dtype=np.float64)
    bl_tt, _ = O.restore_baseline(y_tt, dtype=np.float64)
    rms_ref, th_ref, mask_ref = O.roi_threshold(bl_t, W.TH_FACTOR, r_tight=bl_tt, dtype=np.float64)
    assert _rel_rms(y.astype(np.float64), bl_t) <= 5e-5
    assert np.abs(r - rms_ref).max() <= 1e-4 * np.abs(rms_ref).max()
    combined = ((m & 3) != 0).astype(np.uint8)
    assert np.mean(combined == mask_ref) >= 0.9999
    assert np.any(m & 2)


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
