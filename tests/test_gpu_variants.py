"""The other decon variants of a plane as extra outputs of one job
(OmnibusSigProc.cxx: decon_2D_looseROI :1162-1238, decon_2D_ROI_refine
:1053-1075, decon_2D_hits :1316-1347, decon_2D_charge :1349-1380), on a
kernel without a time filter (wct_decon_kernel_create_unfiltered): each
output applies its whole filter to the one deconvolved spectrum.

Checks, per output: before restore_baseline, against the FP64 oracle
(the fused decon with that output's filter) within 2e-6 relative RMS,
including the loose output's alternative filter on the flagged rows (its
tight-LF fallback next to bad or lf_noisy channels); restore_baseline
exact on the device's own raw output (the order statistics are
discontinuous in the data: with tracks of 1e5 ADC a row's median can move
to a neighbouring sample between FP32 and FP64, so the baselined outputs
are compared with FP64 for the record only); the loose output's per-wire
cal_RMS as the oracle's rule on the same data.  Gaus_wide over a response that falls below FP64 rounding at
high frequency is not a defined quantity (the oracle and the device
divide different rounding noise), so the charge output is checked on the
well-conditioned response, and the regularised inverse (r_floor) on the
benchmark response.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _oracle_resp(spec):
    from wct_amd import workload as W
    return O.overall_response(spec.fr_rows, spec.elec, W.FINE_PERIOD, W.TICK, W.INTER_GAIN,
                              W.ADC_MV, W.FTOFFSET, spec.nt, dtype=np.float64)


def _wellcond_resp(spec):
    nr = spec.fr_rows.shape[0]
    c = (nr - 1) // 2
    resp = np.zeros((nr, spec.nt))
    for r in range(nr):
        a = 1.0 / (1.0 + (r - c) ** 2)
        resp[r, 3] = a
        resp[r, 4] = 0.3 * a
    return resp * 50.0


def _frame(spec, resp, seed, ntracks):
    from wct_amd import workload as W
    q = O.tracks(seed, 1, ntracks, spec.nw, spec.nt, W.Q_PER_TICK)
    adc = O.forward_conv(q, resp, spec.wire_shift, spec.time_shift)
    adc += O.normal(seed, 2, spec.nw * spec.nt, W.NOISE_ADC).reshape(spec.nw, spec.nt)
    return adc.astype(np.float32)


def _rel_rms(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


def _filters(spec, with_charge):
    from wct_amd import host, workload as W
    nt = spec.nt
    wiener = host.hf_filter(nt, **W.WIENER)
    f = {"refine": wiener, "hits": host.hf_filter(nt, **W.WIENER_WIDE)}
    if spec.kind == "ind":
        f["tight"] = wiener * host.lf_filter(nt, **W.LF_TIGHT)
        f["loose"] = wiener * host.lf_filter(nt, **W.LF_LOOSE)
    else:
        f["tight"] = wiener
        f["loose"] = wiener
    if with_charge:
        f["charge"] = host.hf_filter(nt, **W.GAUS_WIDE)
    return f


CASES = [("ind", 800, 6000, "bench", 33), ("col", 480, 6000, "bench", 20),
         ("ind", 96, 2000, "wellcond", 4), ("col", 96, 2000, "wellcond", 4)]


@pytest.mark.parametrize("kind,nw,nt,rk,ntr", CASES)
def test_variant_outputs_vs_fp64(dev, kind, nw, nt, rk, ntr):
    """Each output before restore_baseline against the FP64 fused decon
    with its filter (<= 2e-6 relative RMS); restore_baseline of each
    output exact on the device's own raw output; the loose output's
    cal_RMS as the oracle's rule on the same data."""
    import torch
    from wct_amd import decon, workload as W
    print(f"variants vs FP64: {kind} {nw}x{nt} {rk} response", flush=True)
    spec = W.plane_spec(kind, nw, nt, fold_lf=False)
    resp = _oracle_resp(spec) if rk == "bench" else _wellcond_resp(spec)
    adc = _frame(spec, resp, seed=5 * nw + nt, ntracks=ntr)
    flt = _filters(spec, with_charge=(rk == "wellcond"))
    env = np.max(np.stack(list(flt.values())), axis=0)
    pk = decon.PlaneKernel(spec, resp=resp, unfiltered=True, envelope=env, cut_tol=1e-7)
    print(f"  kept bins {pk.kc} of {nt // 2 + 1}", flush=True)
    g = {k: torch.from_numpy(v[:pk.kc].astype(np.float32)).to(dev) for k, v in flt.items()}
    a = torch.from_numpy(adc).to(dev)
    flags = np.zeros(nw, np.uint8)
    flags[[1, 2, nw // 2, nw - 1]] = 1
    d_flags = torch.from_numpy(flags).to(dev)
    extra = [k for k in ("loose", "refine", "hits", "charge") if k in flt]
    raw = {k: torch.empty_like(a) for k in flt}
    based = {k: torch.empty_like(a) for k in extra}
    mask = torch.empty(a.shape, dtype=torch.uint8, device=dev)
    rms = torch.empty(nw, dtype=torch.float32, device=dev)
    rms_loose = torch.empty(nw, dtype=torch.float32, device=dev)
    restore_hc = 1 if kind == "col" else 0      # hits/charge restore only on collection (:1344-1346)
    restores = {"loose": 1, "refine": 1, "hits": restore_hc, "charge": restore_hc}

    def out(k, dst, restore, rms_=None):
        alt = dict(filter_alt=g["tight"], alt_rows=d_flags) if k == "loose" else {}
        return decon.Output(g[k], out=dst, restore=restore, rms=rms_, **alt)

    # job 1: every output raw; job 2: as the reference restores them
    j1 = decon.make_job(pk, a, raw["tight"], lf_tight=g["tight"], restore=0,
                        outputs=[out(k, raw[k], 0) for k in extra])
    j2 = decon.make_job(pk, a, torch.empty_like(a), mask=mask, rms=rms, lf_tight=g["tight"],
                        outputs=[out(k, based[k], restores[k], rms_loose if k == "loose" else None)
                                 for k in extra])
    decon.Batch([j1]).run()
    decon.Batch([j2]).run()
    torch.cuda.synchronize()
    a64 = adc.astype(np.float64)
    errs = {}
    for k in flt:
        ref = O.decon_fused(a64, resp, spec.wire_filter, flt[k], spec.wire_shift, spec.time_shift)
        if k == "loose":
            alt = O.decon_fused(a64, resp, spec.wire_filter, flt["tight"], spec.wire_shift, spec.time_shift)
            ref[flags == 1] = alt[flags == 1]
        got = raw[k].cpu().numpy()
        errs[k] = _rel_rms(got.astype(np.float64), ref)
        line = f"  {k:6s} raw rel_rms={errs[k]:.3e} max_rel={np.abs(got - ref).max() / np.abs(ref).max():.3e}"
        if k in restores and restores[k]:
            ref_b, _ = O.restore_baseline(ref, dtype=np.float64)
            want, _ = O.restore_baseline(got, dtype=np.float32)
            b = based[k].cpu().numpy()
            assert np.array_equal(b, want), k      # exact on the device's own output
            line += f"  baselined vs FP64 rel_rms={_rel_rms(b.astype(np.float64), ref_b):.3e}"
        elif k in restores:
            assert np.array_equal(based[k].cpu().numpy(), got), k
        print(line, flush=True)
    for k, e in errs.items():
        assert e <= 2e-6, k
    # the loose output's cal_RMS: the oracle's rule (double sums) on the
    # device's own baselined loose output
    yl = based["loose"].cpu().numpy()
    r_ref, _, _ = O.roi_threshold(yl.astype(np.float64), W.TH_FACTOR, dtype=np.float64)
    rl = rms_loose.cpu().numpy()
    assert np.all(np.abs(rl - r_ref) <= 1e-6 * r_ref + 1e-30)
    pk.close()


def test_response_floor_zeroes_small_response_bins(dev):
    """r_floor: K = 0 exactly where |R| <= r_floor max|R| (R from the
    FP64 2D transform of the padded response), nonzero elsewhere."""
    from wct_amd import decon, workload as W
    spec = W.plane_spec("col", 96, 2000, fold_lf=False)
    resp = _oracle_resp(spec)
    floor = 1e-6
    pk = decon.PlaneKernel(spec, resp=resp, unfiltered=True, envelope=np.ones(spec.nt), cut_tol=0.0,
                           r_floor=floor)
    K = pk.download()                                  # (kc, nw), k-major
    rp = np.zeros((spec.nw, spec.nt))
    rp[:resp.shape[0]] = resp
    R = np.fft.fft2(rp)[:, :pk.kc].T                   # (kc, nw)
    ar = np.abs(R) / np.abs(R).max()
    assert np.all(K[ar < 0.5 * floor] == 0)
    assert np.all(K[ar > 2.0 * floor] != 0)
    assert np.any(ar < 0.5 * floor)
    pk.close()


def test_too_many_outputs_refused(dev):
    import torch
    from wct_amd import decon, workload as W
    from wct_amd.capi import MAX_OUTPUTS
    spec = W.plane_spec("col", 96, 2000, fold_lf=False)
    pk = decon.PlaneKernel(spec, resp=_wellcond_resp(spec), unfiltered=True, cut_tol=0.0)
    a = torch.zeros((96, 2000), device=dev)
    g = torch.ones(pk.kc, device=dev)
    with pytest.raises(ValueError, match="exceed"):
        decon.make_job(pk, a, torch.empty_like(a), outputs=[decon.Output(g)] * (MAX_OUTPUTS + 1))
    pk.close()
