"""The batched configuration of 24 plane frames in one plan: 8 APAs of
U/V/W 800/800/960 wires with tick counts cycling 6000/8000/10000, half
noise-only and half with tracks, plus a hipGraph replay.  Every plane of
the batch equals its own single-plane run bit for bit, and one plane of
each tick count matches the FP64 oracle (raw decon, relative RMS error
<= 2e-6)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _rel_rms(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


def test_batch24_mixed_ticks(dev):
    import torch
    from wct_amd import decon as D, workload as W
    planes = W.config_planes("batch24")
    assert len(planes) == 24 and sorted(set(nt for _, _, nt in planes)) == [6000, 8000, 10000]
    print("batch24: 24 planes, ticks 6000/8000/10000", flush=True)
    kern = {t: D.PlaneKernel(W.plane_spec(*t)) for t in sorted(set(planes))}
    g = torch.Generator(device=dev)
    g.manual_seed(24)
    rng = np.random.default_rng(24)
    adc, q = [], {}
    for i, (kind, nw, nt) in enumerate(planes):
        a = torch.randn((nw, nt), generator=g, device=dev) * W.NOISE_ADC
        if (i // 3) % 2 == 1:   # tracks on every other APA
            qr = W.tracks_raster(rng, 50, nw, nt)
            fk = D.ForwardKernel(W.plane_spec(kind, nw, nt), resp=kern[(kind, nw, nt)].resp)
            sig = torch.empty_like(a)
            D.convolve(fk, torch.from_numpy(qr).to(dev), sig)
            a += sig
            fk.close()
        adc.append(a)
    out = [torch.empty_like(a) for a in adc]
    mask = [torch.empty(a.shape, dtype=torch.uint8, device=dev) for a in adc]
    rms = [torch.empty(a.shape[0], device=dev) for a in adc]
    jobs = [D.make_job(kern[planes[i]], adc[i], out[i], mask=mask[i], rms=rms[i]) for i in range(24)]
    b = D.Batch(jobs)
    graph = D.Graph([b])
    graph.launch()
    torch.cuda.synchronize()
    for i in range(24):
        o1 = torch.empty_like(adc[i])
        m1 = torch.empty_like(mask[i])
        r1 = torch.empty_like(rms[i])
        D.Batch([D.make_job(kern[planes[i]], adc[i], o1, mask=m1, rms=r1)]).run()
        torch.cuda.synchronize()
        assert torch.equal(out[i], o1) and torch.equal(mask[i], m1) and torch.equal(rms[i], r1), i
    # raw decon of one plane per tick count against the FP64 oracle
    for i in (0, 4, 8):
        kind, nw, nt = planes[i]
        pk = kern[planes[i]]
        raw = torch.empty_like(adc[i])
        D.Batch([D.make_job(pk, adc[i], raw, restore=0)]).run()
        torch.cuda.synchronize()
        sp = pk.spec
        ref = O.decon_fused(adc[i].cpu().numpy().astype(np.float64), pk.resp, sp.wire_filter, sp.time_filter,
                            sp.wire_shift, sp.time_shift)
        e = _rel_rms(raw.cpu().numpy().astype(np.float64), ref)
        print(f"  plane {i} {kind} {nw}x{nt}: kc={pk.kc} rel_rms={e:.3e}")
        assert e <= 2e-6
    graph.close()
    b.close()
