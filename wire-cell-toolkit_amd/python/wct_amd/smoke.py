"""One small deconvolution on cuda:0 checked against the oracle (used by
__graft_entry__.smoke(); the oracle is imported here only as the checker)."""
import numpy as np


def run():
    import torch
    import oracle as O
    from . import decon, workload as W

    assert torch.cuda.is_available(), "smoke needs a GPU"
    dev = torch.device("cuda:0")
    spec = W.plane_spec("col", 96, 1000)
    resp = O.overall_response(spec.fr_rows, spec.elec, W.FINE_PERIOD, W.TICK, W.INTER_GAIN,
                              W.ADC_MV, W.FTOFFSET, spec.nt, dtype=np.float64)
    q = O.tracks(1, 1, 3, spec.nw, spec.nt, W.Q_PER_TICK)
    adc = O.forward_conv(q, resp, spec.wire_shift, spec.time_shift)
    adc += O.normal(1, 2, spec.nw * spec.nt, W.NOISE_ADC).reshape(spec.nw, spec.nt)
    adc = adc.astype(np.float32)
    ref = O.decon_fused(adc.astype(np.float64), resp, spec.wire_filter, spec.time_filter,
                        spec.wire_shift, spec.time_shift)
    pk = decon.PlaneKernel(spec, cut_tol=1e-7)
    d_adc = torch.from_numpy(adc).to(dev)
    out = torch.empty_like(d_adc)
    decon.Batch([decon.make_job(pk, d_adc, out, restore=0)]).run()
    torch.cuda.synchronize()
    y = out.cpu().numpy().astype(np.float64)
    err = float(np.sqrt(np.sum((y - ref) ** 2) / np.sum(ref ** 2)))
    print(f"smoke: 96x1000 collection plane, rel RMS error vs FP64 oracle = {err:.3e}")
    assert err < 2e-6, err
