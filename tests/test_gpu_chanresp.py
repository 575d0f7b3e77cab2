"""Channel-by-channel electronics correction (decon_2D_init,
sigproc/src/OmnibusSigProc.cxx:949-982) on the device.

  - wct_channel_factors against numpy's FP64 transforms: each row is
    fft(elec) / fft(response cut or zero-padded to nt) over the half
    spectrum, 0 where the response's transform is 0 (float rounding:
    max relative error <= 1e-6 of the row's largest factor);
  - the corrected decon against FP64 truth, on the benchmark response:
    the correction multiplies each row's time spectrum before the linear
    2D decon, so the reference's output equals the FP64 fused decon
    (the oracle) of the rows corrected in FP64 on the host (relative RMS
    <= 2e-6, max <= 1e-4 of the largest value, the decon's own bar);
  - on a well-conditioned response, against the oracle's step-by-step
    FP64 restatement with the correction (same bar);
  - channel responses equal to the nominal one change nothing beyond
    float rounding, and a row index of -1 leaves a wire uncorrected.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _elec(nt, gain=None, shaping=None, t0=0.0):
    from wct_amd import host, workload as W
    g = W.GAIN if gain is None else gain
    s = W.SHAPING if shaping is None else shaping
    return host.coldelec_samples(nt, t0, t0 + nt * W.TICK, g, s, 1.0)


def _channel_responses(nresp, nsamp, seed):
    """Measured-like responses: cold electronics with gains within 10 %
    and shaping times of 1.8-2.2 us, float samples at the tick."""
    from wct_amd import workload as W
    rng = np.random.default_rng(seed)
    out = np.zeros((nresp, nsamp), np.float32)
    for r in range(nresp):
        g = W.GAIN * rng.uniform(0.9, 1.1)
        s = rng.uniform(1.8, 2.2) * 1000.0
        out[r] = _elec(nsamp, g, s).astype(np.float32)
    return out


def _ratio_fp64(elec, resp, nt):
    r = np.zeros(nt)
    n = min(nt, resp.shape[0])
    r[:n] = resp[:n]
    E = np.fft.fft(elec)
    C = np.fft.fft(r)
    out = np.zeros(nt, np.complex128)
    nz = C != 0
    out[nz] = E[nz] / C[nz]
    return out


def _rel_rms(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


@pytest.mark.parametrize("nt,nsamp", [(6000, 200), (1000, 1400), (2000, 2000)])
def test_factor_table_against_numpy(dev, nt, nsamp):
    from wct_amd import decon
    resp = _channel_responses(5, nsamp, nt)
    resp[3] = 0.0                      # |ch_elec| == 0 everywhere: factor 0
    elec = _elec(nt)
    tab = decon.channel_factors(resp, elec, nt, device=dev).cpu().numpy()
    assert tab.shape == (5, nt // 2 + 1)
    for r in range(5):
        ref = _ratio_fp64(elec, resp[r].astype(np.float64), nt)[: nt // 2 + 1]
        if r == 3:
            assert not np.any(tab[r])
            continue
        err = np.max(np.abs(tab[r] - ref)) / np.max(np.abs(ref))
        assert err <= 1e-6, (r, err)


def _corrected_rows(adc, ratios, rows):
    """FP64 on the host: row w's time spectrum times its factor row."""
    out = adc.astype(np.float64).copy()
    for w, ri in enumerate(rows):
        if ri >= 0:
            out[w] = np.fft.ifft(np.fft.fft(out[w]) * ratios[ri]).real
    return out


def _run(pk, adc, dev, chan=None):
    import torch
    from wct_amd import decon
    d_adc = torch.from_numpy(adc).to(dev)
    out = torch.empty_like(d_adc)
    b = decon.Batch([decon.make_job(pk, d_adc, out, restore=0, chan=chan)])
    b.run()
    torch.cuda.synchronize()
    return out.cpu().numpy().astype(np.float64)


@pytest.mark.parametrize("kind,nw,nt", [("col", 480, 6000), ("ind", 96, 2000)])
def test_corrected_decon_against_fp64(dev, kind, nw, nt):
    import torch
    from wct_amd import decon
    from test_gpu_decon import _frame, _oracle_resp, _spec
    print(f"per-channel correction {kind} {nw}x{nt}", flush=True)
    spec = _spec(kind, nw, nt)
    resp = _oracle_resp(spec)
    adc = _frame(spec, resp, seed=nw + 7, ntracks=max(2, nw // 24))
    cresp = _channel_responses(7, 200, nw)
    elec = _elec(nt)
    rows = np.arange(nw, dtype=np.int32) % 7
    rows[5] = -1                        # one wire left uncorrected
    ratios = [_ratio_fp64(elec, cresp[r].astype(np.float64), nt) for r in range(7)]
    ref = O.decon_fused(_corrected_rows(adc, ratios, rows), resp, spec.wire_filter, spec.time_filter,
                        spec.wire_shift, spec.time_shift)
    pk = decon.PlaneKernel(spec, cut_tol=1e-7)
    tab = decon.channel_factors(cresp, elec, nt, device=dev)
    y = _run(pk, adc, dev, chan=(tab, torch.from_numpy(rows).to(dev)))
    e = _rel_rms(y, ref)
    print(f"  rel_rms={e:.3e} max={np.abs(y - ref).max() / np.abs(ref).max():.3e}", flush=True)
    assert e <= 2e-6
    assert np.abs(y - ref).max() <= 1e-4 * np.abs(ref).max()
    # the correction is not a no-op on these responses
    plain = _run(pk, adc, dev)
    assert _rel_rms(plain, ref) > 1e-3


def test_corrected_decon_against_oracle_steps(dev):
    """The oracle's step-by-step FP64 decon_2D_init with the correction,
    on a well-conditioned response (the step order in FP64 is then exact
    to rounding)."""
    import torch
    from wct_amd import decon
    from test_gpu_variants import _wellcond_resp
    from test_gpu_decon import _frame, _spec
    nw, nt = 96, 2000
    spec = _spec("ind", nw, nt)
    resp = _wellcond_resp(spec)
    adc = _frame(spec, resp, seed=11, ntracks=4)
    cresp = _channel_responses(3, 300, 5)
    elec = _elec(nt)
    rows = (np.arange(nw, dtype=np.int32) // 7) % 3
    ch = np.stack([cresp[r] for r in rows]).astype(np.float64)
    c = O.decon_init(adc.astype(np.float64), resp, spec.wire_filter, spec.wire_shift, spec.time_shift,
                     np.float64, ch_resp=ch, ewave=elec)
    ref = O.decon_apply_filter(c, spec.time_filter, np.float64)
    pk = decon.PlaneKernel(spec, resp=resp, cut_tol=0.0)
    tab = decon.channel_factors(cresp, elec, nt, device=dev)
    y = _run(pk, adc, dev, chan=(tab, torch.from_numpy(rows).to(dev)))
    e = _rel_rms(y, ref)
    print(f"wellcond 96x2000 vs oracle steps: rel_rms={e:.3e}", flush=True)
    assert e <= 2e-6
    assert np.abs(y - ref).max() <= 1e-4 * np.abs(ref).max()


def test_nominal_responses_change_nothing(dev):
    import torch
    from wct_amd import decon
    from test_gpu_decon import _frame, _oracle_resp, _spec
    nw, nt = 64, 1000
    spec = _spec("col", nw, nt)
    resp = _oracle_resp(spec)
    adc = _frame(spec, resp, seed=4, ntracks=3)
    e32 = _elec(nt).astype(np.float32)   # the same samples on both sides: factors 1 to rounding
    pk = decon.PlaneKernel(spec, cut_tol=1e-7)
    tab = decon.channel_factors(e32[None, :], e32.astype(np.float64), nt, device=dev)
    assert np.max(np.abs(tab.cpu().numpy() - 1.0)) <= 1e-7
    rows = torch.zeros(nw, dtype=torch.int32, device=dev)
    y = _run(pk, adc, dev, chan=(tab, rows))
    plain = _run(pk, adc, dev)
    assert _rel_rms(y, plain) <= 1e-7
    none = _run(pk, adc, dev, chan=(tab, torch.full((nw,), -1, dtype=torch.int32, device=dev)))
    np.testing.assert_array_equal(none, plain)
