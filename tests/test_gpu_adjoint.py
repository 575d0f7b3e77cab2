"""Hand-written adjoints of the forward response convolution and of the
filtered deconvolution (wct_amd.autograd), checked three ways:
  - the dot-product identity <A x, y> = <x, A^T y> for both operators;
  - gradients of L = 1/2 |decon(fwd(q); s_t, s_w) - T|^2 (T = fwd(q_true)
    + noise) with respect to q, s_t and s_w against autograd through an
    FP64 torch.fft restatement of the same operators;
  - central finite differences in FP64 through the oracle.
Tolerances: dot products within 1e-5 |A x| |y|; gradients within 1e-4
relative (2-norm) of the FP64 references; finite differences within
1e-3 relative.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _model(kind, nw, nt):
    from wct_amd import autograd as A
    from wct_amd import workload as W
    return A.Model(W.plane_spec(kind, nw, nt))


def _dot(a, b):
    return float((a.double() * b.double()).sum())


@pytest.mark.parametrize("kind,nw,nt", [("ind", 128, 1000), ("col", 96, 2000)])
def test_dot_product_identity(dev, kind, nw, nt):
    import torch
    print(f"adjoint dot-product identity {kind} {nw}x{nt}", flush=True)
    m = _model(kind, nw, nt)
    g = torch.Generator(device=dev)
    g.manual_seed(5)
    x = torch.randn((2, nw, nt), generator=g, device=dev)
    y = torch.randn((2, nw, nt), generator=g, device=dev)
    for fwd, adj in [(lambda v: m.forward_conv(v), lambda v: m.forward_conv(v, adjoint=True)),
                     (lambda v: m.decon(v, 1e-4, 0.75), lambda v: m.decon(v, 1e-4, 0.75, adjoint=True))]:
        ax = fwd(x)
        aty = adj(y)
        torch.cuda.synchronize()
        lhs, rhs = _dot(ax, y), _dot(x, aty)
        scale = float(ax.double().norm() * y.double().norm())
        assert abs(lhs - rhs) <= 1e-5 * scale, (lhs, rhs, scale)
    m.close()


def _problem(kind, nw, nt, B, seed):
    from wct_amd import workload as W
    rng = np.random.default_rng(seed)
    q_true = np.stack([W.charge_raster(rng, nw, nt, occupancy=0.02, ntracks=2) for _ in range(B)])
    q0 = np.stack([W.charge_raster(rng, nw, nt, occupancy=0.02, ntracks=2) for _ in range(B)])
    noise = rng.normal(0, W.NOISE_ADC, q_true.shape)
    return q_true.astype(np.float32), q0.astype(np.float32), noise


def _torch_fp64_reference(m, q, T, st, sw):
    """FP64 torch.fft restatement: returns (L, dL/dq, dL/ds_t, dL/ds_w)."""
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
    Kf = torch.from_numpy(R * ph)
    with np.errstate(divide="ignore", invalid="ignore"):
        G = np.where(R != 0, np.conj(ph) / R, 0)
    G = torch.from_numpy(G)
    qt = torch.tensor(q, dtype=torch.float64, requires_grad=True)
    s_t = torch.tensor(st, dtype=torch.float64, requires_grad=True)
    s_w = torch.tensor(sw, dtype=torch.float64, requires_grad=True)
    at = torch.from_numpy(A.freq_grid(nt, m.tf))
    aw = torch.from_numpy(A.freq_grid(nw, m.wf))
    Ft = torch.exp(-0.5 * (at / s_t) ** m.tp)
    Ft = torch.cat([torch.zeros(1, dtype=torch.float64), Ft[1:]]) if m.tflag else Ft
    Fw = torch.exp(-0.5 * (aw / s_w) ** m.wp)
    d = torch.fft.ifft2(torch.fft.fft2(qt) * Kf).real
    y = torch.fft.ifft2(torch.fft.fft2(d) * (Fw[:, None] * Ft[None, :] * G)).real
    L = 0.5 * ((y - torch.from_numpy(T)) ** 2).sum()
    L.backward()
    return float(L.detach()), qt.grad.numpy(), float(s_t.grad), float(s_w.grad)


def _gpu_grads(m, q, T, st, sw, dev):
    import torch
    from wct_amd import autograd as A
    qg = torch.tensor(q, device=dev, requires_grad=True)
    s_t = torch.tensor(st, dtype=torch.float64, requires_grad=True)
    s_w = torch.tensor(sw, dtype=torch.float64, requires_grad=True)
    y = A.decon(A.forward_conv(qg, m), s_t, s_w, m)
    L = 0.5 * ((y.double() - torch.from_numpy(T).to(dev)) ** 2).sum()
    L.backward()
    torch.cuda.synchronize()
    return float(L.detach()), qg.grad.cpu().numpy().astype(np.float64), float(s_t.grad), float(s_w.grad)


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b)))


@pytest.mark.parametrize("kind,nw,nt", [("ind", 64, 600), ("col", 48, 1000)])
def test_gradients_vs_torch_fft_fp64(dev, kind, nw, nt):
    import torch
    print(f"adjoint vs torch.fft FP64 {kind} {nw}x{nt}", flush=True)
    m = _model(kind, nw, nt)
    q_true, q0, noise = _problem(kind, nw, nt, 2, seed=nw + nt)
    # target in ADC space: forward of the true raster plus noise
    T = O.forward_conv(q_true[0].astype(np.float64), m.resp, m.spec.wire_shift, m.spec.time_shift)
    T = np.stack([O.forward_conv(q.astype(np.float64), m.resp, m.spec.wire_shift, m.spec.time_shift)
                  for q in q_true]) + noise
    st, sw = 0.12e-3, 0.8
    Lr, gq_r, gt_r, gw_r = _torch_fp64_reference(m, q0, T, st, sw)
    L, gq, gt, gw = _gpu_grads(m, q0, T, st, sw, dev)
    print(f"  L rel={abs(L - Lr) / abs(Lr):.2e} dq rel={_rel(gq, gq_r):.2e} "
          f"ds_t rel={abs(gt - gt_r) / abs(gt_r):.2e} ds_w rel={abs(gw - gw_r) / abs(gw_r):.2e}")
    assert abs(L - Lr) <= 1e-4 * abs(Lr)
    assert _rel(gq, gq_r) <= 1e-4
    assert abs(gt - gt_r) <= 1e-4 * abs(gt_r)
    assert abs(gw - gw_r) <= 1e-4 * abs(gw_r)
    m.close()


def test_gradients_vs_finite_differences_fp64(dev):
    """Central differences of the FP64 oracle loss in s_t, s_w and a few
    raster pixels."""
    from wct_amd import autograd as A
    kind, nw, nt = "ind", 48, 600
    m = _model(kind, nw, nt)
    sp = m.spec
    q_true, q0, noise = _problem(kind, nw, nt, 1, seed=3)
    T = O.forward_conv(q_true[0].astype(np.float64), m.resp, sp.wire_shift, sp.time_shift) + noise[0]
    st, sw = 0.1e-3, 0.75

    def loss(q, s_t, s_w):
        d = O.forward_conv(q, m.resp, sp.wire_shift, sp.time_shift)
        y = O.decon_fused(d, m.resp, m.wire_filter(s_w), m.time_filter(s_t), sp.wire_shift, sp.time_shift)
        return 0.5 * np.sum((y - T) ** 2)

    q64 = q0[0].astype(np.float64)
    _, gq, gt, gw = _gpu_grads(m, q0, T[None], st, sw, dev)
    ht, hw = 1e-4 * st, 1e-4 * sw
    fd_t = (loss(q64, st + ht, sw) - loss(q64, st - ht, sw)) / (2 * ht)
    fd_w = (loss(q64, st, sw + hw) - loss(q64, st, sw - hw)) / (2 * hw)
    print(f"  FD s_t {fd_t:.6e} vs {gt:.6e}; s_w {fd_w:.6e} vs {gw:.6e}")
    assert abs(gt - fd_t) <= 1e-3 * abs(fd_t)
    assert abs(gw - fd_w) <= 1e-3 * abs(fd_w)
    rng = np.random.default_rng(9)
    gmax = np.abs(gq).max()
    for _ in range(6):
        w, t = int(rng.integers(0, nw)), int(rng.integers(0, nt))
        h = 10.0
        qp, qm = q64.copy(), q64.copy()
        qp[w, t] += h
        qm[w, t] -= h
        fd = (loss(qp, st, sw) - loss(qm, st, sw)) / (2 * h)
        assert abs(gq[0, w, t] - fd) <= 1e-3 * gmax, (w, t, gq[0, w, t], fd)
    m.close()


def test_batch16_induction_frames(dev):
    """The benchmark's adjoint configuration: 16 frames of 800 x 6000,
    one gradient step; the gradient passes the dot-product identity of
    the composed operator and is finite."""
    import time
    import torch
    from wct_amd import autograd as A
    print("adjoint batch 16 x 800 x 6000", flush=True)
    m = _model("ind", 800, 6000)
    B = 16
    rng = np.random.default_rng(1)
    from wct_amd import workload as W
    q_true = torch.from_numpy(np.stack([W.charge_raster(rng, 800, 6000) for _ in range(B)])).to(dev)
    T = m.forward_conv(q_true) + torch.randn_like(q_true) * W.NOISE_ADC
    q = torch.zeros_like(q_true, requires_grad=True)
    s_t = torch.tensor(1e-4, dtype=torch.float64, requires_grad=True)
    s_w = torch.tensor(0.75, dtype=torch.float64, requires_grad=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    y = A.decon(A.forward_conv(q, m), s_t, s_w, m)
    L = 0.5 * ((y - T) ** 2).double().sum()
    L.backward()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    print(f"  one gradient step on {B} frames: {dt * 1e3:.1f} ms (first call, kernels built)")
    gq = q.grad
    assert torch.isfinite(gq).all() and torch.isfinite(s_t.grad) and torch.isfinite(s_w.grad)
    # q = 0: y = 0, so dL/dq = -F^T G^T T; check <F^T G^T T, z> = <T, G F z>
    z = torch.randn_like(q_true)
    gfz = m.decon(m.forward_conv(z), 1e-4, 0.75)
    lhs = -_dot(gq, z)
    rhs = _dot(T, gfz)
    assert abs(lhs - rhs) <= 1e-4 * float(T.double().norm() * gfz.double().norm())
    m.close()
