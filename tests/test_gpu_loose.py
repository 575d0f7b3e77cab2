"""Loose ROIs and apply_roi on the device's own outputs (a one-APA
induction plane, 800 x 6000, tracks, bad ranges): the loose decon
(Wiener x ROI_loose_lf, the tight-LF fallback on flagged rows) with its
per-wire cal_RMS comes from the decon's extra output, the tight ROIs from
the device mask; the host's find_ROI_loose + extend_ROI_loose
(sigproc/src/ROI_formation.cxx:538-937) must give exactly the oracle's
lists on that data and RMS, and the host's apply_roi
(sigproc/src/ROI_refinement.cxx:118-218) exactly the oracle's values
(induction with the loose ROIs, collection rule with the tight ones).
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def test_loose_rois_and_apply_roi_on_device_outputs(dev):
    import torch
    from wct_amd import components as K
    from wct_amd import decon, host, workload as W
    from test_gpu_variants import _frame, _oracle_resp
    nw, nt = 800, 6000
    print(f"loose ROIs on device outputs, ind {nw}x{nt}", flush=True)
    spec = W.plane_spec("ind", nw, nt, fold_lf=False)
    resp = _oracle_resp(spec)
    adc = _frame(spec, resp, seed=91, ntracks=33)
    wiener = host.hf_filter(nt, **W.WIENER)
    tight_f = wiener * host.lf_filter(nt, **W.LF_TIGHT)
    loose_f = wiener * host.lf_filter(nt, **W.LF_LOOSE)
    env = np.maximum(tight_f, loose_f)
    pk = decon.PlaneKernel(spec, resp=resp, unfiltered=True, envelope=env, cut_tol=1e-7)
    g_t = torch.from_numpy(tight_f[:pk.kc].astype(np.float32)).to(dev)
    g_l = torch.from_numpy(loose_f[:pk.kc].astype(np.float32)).to(dev)
    flags = np.zeros(nw, np.uint8)
    flags[[3, 4, 400]] = 1
    # bad ranges: [lo, hi) zeroed on input, [lo, hi] left out of the RMS
    bad_rows = {7: [(0, nt)], 120: [(2000, 2600)], 121: [(100, 150), (3000, 3100)]}
    ptr, lo, hi = [0], [], []
    for w in range(nw):
        for (a, b) in bad_rows.get(w, []):
            lo.append(a)
            hi.append(b)
        ptr.append(len(lo))
    badt = tuple(torch.tensor(v, dtype=torch.int32, device=dev) for v in (ptr, lo, hi))
    a = torch.from_numpy(adc).to(dev)
    y_t = torch.empty_like(a)
    y_l = torch.empty_like(a)
    mask = torch.empty(a.shape, dtype=torch.uint8, device=dev)
    rms_t = torch.empty(nw, dtype=torch.float32, device=dev)
    rms_l = torch.empty(nw, dtype=torch.float32, device=dev)
    out_l = decon.Output(g_l, out=y_l, restore=1, rms=rms_l, filter_alt=g_t,
                         alt_rows=torch.from_numpy(flags).to(dev))
    job = decon.make_job(pk, a, y_t, mask=mask, rms=rms_t, lf_tight=g_t, restore=1, bad=badt, outputs=[out_l])
    decon.Batch([job]).run()
    torch.cuda.synchronize()
    yl, rl = y_l.cpu().numpy(), rms_l.cpu().numpy()
    m = mask.cpu().numpy()
    tight = K.roi_lists(m)
    assert tight == O.roi_lists(m)
    bad = [[(x, y) for (x, y) in bad_rows.get(w, [])] for w in range(nw)]
    got = K.loose_rois(yl, bad=bad, rms=rl, tight=tight)
    want = O.loose_rois(yl, bad=bad, rms=rl.astype(np.float64), tight=tight)
    assert got == want
    n = sum(len(x) for x in got)
    cover = sum(e - s + 1 for row in got for (s, e) in row) / float(nw * nt)
    print(f"  {n} loose ROIs, {cover:.3f} of the plane covered", flush=True)
    assert n > 0 and cover < 0.9
    # the device RMS against the reference's float rule on the same rows
    for w in (0, 120, 121, 400):
        keep = np.ones(nt, bool)
        for (x, y) in bad[w]:
            keep[x:y + 1] = False
        ref = O.cal_rms(yl[w][keep])
        assert abs(rl[w] - ref) <= 1e-5 * ref, w
    # apply_roi: induction rule with the loose ROIs, collection rule with the tight ones
    for rois, linear in ((got, 1), (tight, 0)):
        np.testing.assert_array_equal(K.apply_roi(yl, rois, linear), O.apply_roi(yl, rois, linear))
    pk.close()
