"""The FP64 torch.fft restatement used as the adjoint's reference
gradient, pinned on the CPU: its loss equals the oracle's FP64 loss,
its filter-width gradients equal central differences of the oracle's
loss, and the closed-form filter derivatives equal differences of the
filters."""
import numpy as np

import oracle as O


class _CpuModel:
    """The attributes of wct_amd.autograd.Model the reference reads, with
    the response from the oracle instead of the device."""

    def __init__(self, kind, nw, nt):
        from wct_amd import workload as W
        self.spec = W.plane_spec(kind, nw, nt)
        sp = self.spec
        self.resp = O.overall_response(sp.fr_rows, sp.elec, W.FINE_PERIOD, W.TICK, W.INTER_GAIN,
                                       W.ADC_MV, W.FTOFFSET, nt, dtype=np.float64)
        self.tp, self.tf, self.tflag = W.WIENER["power"], W.WIENER["max_freq"], True
        self.wp, self.wf = W.WIRE["power"], W.WIRE["max_freq"]


def test_filter_derivatives_closed_form():
    from wct_amd import autograd as A
    for (n, fmax, s, p, flag) in [(600, 1e-3, 1e-4, 3.35324, True), (96, 1.0, 0.75, 2, False)]:
        h = 1e-6 * s
        fd = (A.hf_value(n, fmax, s + h, p, flag) - A.hf_value(n, fmax, s - h, p, flag)) / (2 * h)
        an = A.hf_dsigma(n, fmax, s, p, flag)
        assert np.abs(fd - an).max() <= 1e-6 * np.abs(an).max()


def test_torch_reference_matches_oracle_and_fd():
    from wct_amd import autograd as A
    from test_gpu_adjoint import _problem, _torch_fp64_reference
    nw, nt = 48, 600
    m = _CpuModel("ind", nw, nt)
    sp = m.spec
    q_true, q0, noise = _problem("ind", nw, nt, 1, seed=3)
    T = O.forward_conv(q_true[0].astype(np.float64), m.resp, sp.wire_shift, sp.time_shift) + noise[0]
    st, sw = 0.1e-3, 0.75

    def loss(q, s_t, s_w):
        d = O.forward_conv(q, m.resp, sp.wire_shift, sp.time_shift)
        y = O.decon_fused(d, m.resp, A.hf_value(nw, m.wf, s_w, m.wp, False),
                          A.hf_value(nt, m.tf, s_t, m.tp, True), sp.wire_shift, sp.time_shift)
        return 0.5 * np.sum((y - T) ** 2)

    Lr, gq, gt, gw = _torch_fp64_reference(m, q0, T[None], st, sw)
    q64 = q0[0].astype(np.float64)
    assert abs(Lr - loss(q64, st, sw)) <= 1e-12 * abs(Lr)
    ht, hw = 1e-4 * st, 1e-4 * sw
    fd_t = (loss(q64, st + ht, sw) - loss(q64, st - ht, sw)) / (2 * ht)
    fd_w = (loss(q64, st, sw + hw) - loss(q64, st, sw - hw)) / (2 * hw)
    assert abs(gt - fd_t) <= 1e-6 * abs(fd_t)
    assert abs(gw - fd_w) <= 1e-6 * abs(fd_w)
    # one pixel: the loss is quadratic in q, so a central difference is exact
    w, t = 10, 300
    qp, qm = q64.copy(), q64.copy()
    qp[w, t] += 1.0
    qm[w, t] -= 1.0
    fd = (loss(qp, st, sw) - loss(qm, st, sw)) / 2.0
    assert abs(gq[0, w, t] - fd) <= 1e-6 * np.abs(gq).max()
