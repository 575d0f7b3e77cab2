"""Charge fit through the hand-written adjoints (the second half of the
150-APA workload: gradient steps on a few selected APAs).

For one plane type (an autograd.Model) and a batch of ADC frames d
(B, nw, nt) the fit minimises, over the charge raster q,

    L(q) = 1/2 | G (F q - d) |^2

with F the forward response convolution and G the filtered
deconvolution at widths (s_t, s_w): the frame simulated from q,
deconvolved, against the deconvolved data.  The gradient
g = F^T G^T G (F q - d) comes from backward() through the libtorch
autograd Functions (their adjoints are the conjugate-kernel fused
passes).  L is quadratic in q, so each step is steepest descent with the
exact line search alpha = |g|^2 / |G F g|^2, and L never increases.
One step is six fused passes over the batch: G F q, the two adjoints,
and G F g.
"""
import numpy as np

from . import autograd as A


def _dot(a, b):
    return float((a.double() * b.double()).sum())


def fit(model, adc, steps, sigma_t, sigma_w, q0=None, log=None):
    """Runs `steps` steps from q0 (zeros by default).  Returns (q, losses):
    losses[k] is L before step k, losses[steps] after the last one."""
    import torch
    st = torch.tensor(float(sigma_t), dtype=torch.float64)
    sw = torch.tensor(float(sigma_w), dtype=torch.float64)
    target = model.decon(adc, sigma_t, sigma_w)          # G d, fixed
    q = torch.zeros_like(adc) if q0 is None else q0.clone()
    losses = []
    for k in range(steps + 1):
        qv = q.detach().requires_grad_(True)
        y = A.decon(A.forward_conv(qv, model), st, sw, model)
        loss = 0.5 * ((y - target) ** 2).sum()
        losses.append(float(loss.detach()))
        if k == steps:
            break
        loss.backward()
        g = qv.grad
        ag = model.decon(model.forward_conv(g), sigma_t, sigma_w)
        den = _dot(ag, ag)
        alpha = _dot(g, g) / den if den > 0 else 0.0
        q = (qv.detach() - alpha * g).detach()
        if log:
            log(f"chargefit: step {k}: L={losses[-1]:.6e} alpha={alpha:.3e}")
    return q.detach(), losses


def gradient(model, adc, q, sigma_t, sigma_w):
    """(L, dL/dq) at q, as the fit computes them."""
    import torch
    st = torch.tensor(float(sigma_t), dtype=torch.float64)
    sw = torch.tensor(float(sigma_w), dtype=torch.float64)
    target = model.decon(adc, sigma_t, sigma_w)
    qv = q.detach().clone().requires_grad_(True)
    y = A.decon(A.forward_conv(qv, model), st, sw, model)
    loss = 0.5 * ((y - target) ** 2).sum()
    loss.backward()
    return float(loss.detach()), qv.grad.detach()
