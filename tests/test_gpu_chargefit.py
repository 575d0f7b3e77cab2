"""The charge fit of the 150-APA workload's second half
(wct_amd.chargefit): steepest descent on L(q) = 1/2 |G (F q - d)|^2
through the hand-written adjoints.

Checks: the gradient at the first step against autograd through an FP64
torch.fft restatement of the same operators (within 1e-4 relative), the
loss non-increasing over 20 steps (exact line search on a quadratic:
each step can only lower it, up to FP32 rounding of 1e-6 relative), and
a substantial decrease overall.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _setup(kind, nw, nt, B, seed):
    from wct_amd import autograd as A, workload as W
    m = A.Model(W.plane_spec(kind, nw, nt))
    rng = np.random.default_rng(seed)
    q_true = np.stack([W.charge_raster(rng, nw, nt, occupancy=0.02, ntracks=2) for _ in range(B)])
    adc = np.stack([O.forward_conv(q.astype(np.float64), m.resp, m.spec.wire_shift, m.spec.time_shift)
                    for q in q_true]) + rng.normal(0, W.NOISE_ADC, q_true.shape)
    return m, adc.astype(np.float32)


def _fp64_grad(m, q, adc, st, sw):
    """L and dL/dq of 1/2 |G (F q - d)|^2 by autograd in FP64 torch.fft."""
    import torch
    from wct_amd import autograd as A
    sp = m.spec
    nw, nt = sp.nw, sp.nt
    Rr = np.zeros((nw, nt))
    Rr[:m.resp.shape[0]] = m.resp
    R = np.fft.fft2(Rr)
    mm = np.arange(nw)[:, None]
    kk = np.arange(nt)[None, :]
    ts = sp.time_shift if sp.time_shift > 0 else 0
    ph = np.exp(2j * np.pi * (mm * sp.wire_shift / nw + kk * ts / nt))
    with np.errstate(divide="ignore", invalid="ignore"):
        G = np.where(R != 0, np.conj(ph) / R, 0)
    Ft = A.hf_value(nt, m.tf, st, m.tp, m.tflag)
    Fw = A.hf_value(nw, m.wf, sw, m.wp, False)
    Gf = torch.from_numpy(G * Fw[:, None] * Ft[None, :])
    Kf = torch.from_numpy(R * ph)
    qt = torch.tensor(q, dtype=torch.float64, requires_grad=True)
    d = torch.from_numpy(adc.astype(np.float64))
    sim = torch.fft.ifft2(torch.fft.fft2(qt) * Kf).real
    y = torch.fft.ifft2(torch.fft.fft2(sim - d) * Gf).real
    L = 0.5 * (y ** 2).sum()
    L.backward()
    return float(L.detach()), qt.grad.numpy()


@pytest.mark.parametrize("kind,nw,nt", [("ind", 64, 600), ("col", 48, 1000)])
def test_first_step_gradient_vs_fp64(dev, kind, nw, nt):
    import torch
    from wct_amd import chargefit as CF
    print(f"charge fit gradient vs FP64 {kind} {nw}x{nt}", flush=True)
    m, adc = _setup(kind, nw, nt, 2, seed=nw)
    st, sw = 0.1e-3, 0.75
    rng = np.random.default_rng(7)
    q = rng.exponential(100.0, adc.shape).astype(np.float32)
    L, g = CF.gradient(m, torch.from_numpy(adc).to(dev), torch.from_numpy(q).to(dev), st, sw)
    Lr, gr = _fp64_grad(m, q, adc, st, sw)
    e = float(np.linalg.norm(g.cpu().numpy() - gr) / np.linalg.norm(gr))
    print(f"  L rel={abs(L - Lr) / Lr:.2e} grad rel={e:.2e}", flush=True)
    assert abs(L - Lr) <= 1e-4 * Lr
    assert e <= 1e-4
    m.close()


def test_twenty_steps_never_increase_the_loss(dev):
    import torch
    from wct_amd import chargefit as CF
    print("charge fit 20 steps", flush=True)
    m, adc = _setup("ind", 96, 1000, 3, seed=11)
    q, losses = CF.fit(m, torch.from_numpy(adc).to(dev), 20, 0.1e-3, 0.75)
    print("  losses", [f"{v:.4e}" for v in losses[::5]], flush=True)
    assert len(losses) == 21
    for a, b in zip(losses, losses[1:]):
        assert b <= a * (1 + 1e-6)
    assert losses[-1] < 0.5 * losses[0]
    m.close()
