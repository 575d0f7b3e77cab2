"""Differentiable forward response convolution and filtered deconvolution.

Both operators are real 2D circular filters applied by the fused kernels
(row r2c, wire pass, row c2r):
  forward  D = F q,      kernel R exp(+i phase) / (nw nt)
  decon    y = G_s d,    kernel Fw(s_w) Ft(s_t) exp(-i phase) / (R nw nt)
Their adjoints are the same pipeline with the conjugate kernel (the
wire pass multiplies by conj(K), wct_decon_job.adjoint), so
  dL/dd = G_s^T (dL/dy),   dL/dq = F^T (dL/dD).
The filter-width gradients use the filters' closed-form derivatives:
  dL/ds_t = < G_{dFt/ds_t} d, dL/dy >,  dL/ds_w = < G_{dFw/ds_w} d, dL/dy >
with one more fused pass each, the kernels built in FP64 on the device
from the derivative filters.  Nothing is stored between forward and
backward but the input frames and the filter widths: the response
spectra are rebuilt (and cached by width) when needed.

The filters are SigProc::HfFilter (sigproc/src/HfFilter.cxx:38-52):
g(f) = exp(-(|f| / sigma)^p / 2), 0 at f = 0 when flag is set, on the
frequency grid f_i = 2 f_max i / n folded above f_max.
"""
import numpy as np

from . import decon as D
from . import workload as W


def freq_grid(n, max_freq):
    """|f| of the HfFilter grid (HfFilter.cxx:45-49)."""
    f = np.arange(n) * 1.0 / n * 2 * max_freq
    f = np.where(f > max_freq, f - 2 * max_freq, f)
    return np.abs(f)


def hf_value(n, max_freq, sigma, power, flag):
    a = freq_grid(n, max_freq)
    g = np.exp(-0.5 * (a / sigma) ** power)
    if flag:
        g[0] = 0.0
    return g


def hf_dsigma(n, max_freq, sigma, power, flag):
    """d g / d sigma = g * (p / 2) (|f| / sigma)^p / sigma."""
    a = freq_grid(n, max_freq)
    x = (a / sigma) ** power
    d = np.exp(-0.5 * x) * 0.5 * power * x / sigma
    if flag:
        d[0] = 0.0
    return d


class Model:
    """One plane type: its FP64 response, the forward kernel, and decon
    kernels built on demand for filter widths (time filter: Wiener of
    power p_t with flag, wire filter: Gaussian of power 2)."""

    def __init__(self, spec, device=None, cut_tol=1e-7, time_power=W.WIENER["power"],
                 time_max_freq=W.WIENER["max_freq"], time_flag=True, wire_power=W.WIRE["power"],
                 wire_max_freq=W.WIRE["max_freq"], cache=8):
        import torch
        self.spec = spec
        self.device = device or torch.device("cuda")
        self.resp = D.overall_response(spec)
        self.fwd_kernel = D.ForwardKernel(spec, resp=self.resp)
        self.cut_tol = cut_tol
        self.tp, self.tf, self.tflag = time_power, time_max_freq, time_flag
        self.wp, self.wf = wire_power, wire_max_freq
        self._cache = {}
        self._order = []
        self._cap = cache

    def time_filter(self, sigma_t, deriv=False):
        f = hf_dsigma if deriv else hf_value
        return f(self.spec.nt, self.tf, sigma_t, self.tp, self.tflag)

    def wire_filter(self, sigma_w, deriv=False):
        f = hf_dsigma if deriv else hf_value
        return f(self.spec.nw, self.wf, sigma_w, self.wp, False)

    def kernel(self, sigma_t, sigma_w, which="value"):
        """Decon kernel for the widths: which = 'value', 'dt' (d/ds_t) or
        'dw' (d/ds_w).  Rebuilt in FP64 on the device, kept in a small
        cache keyed by the widths."""
        key = (float(sigma_t), float(sigma_w), which)
        k = self._cache.get(key)
        if k is None:
            tf = self.time_filter(sigma_t, which == "dt")
            wf = self.wire_filter(sigma_w, which == "dw")
            sp = _SpecView(self.spec, wf, tf)
            k = D.PlaneKernel(sp, cut_tol=self.cut_tol, resp=self.resp)
            self._cache[key] = k
            self._order.append(key)
            while len(self._order) > self._cap:
                old = self._order.pop(0)
                self._cache.pop(old).close()
        return k

    def apply(self, kernel, x, adjoint=False):
        """kernel applied to every frame of x (B, nw, nt) float32."""
        import torch
        x = x.contiguous()
        out = torch.empty_like(x)
        jobs = [D.make_job(kernel, x[b], out[b], restore=0, adjoint=adjoint) for b in range(x.shape[0])]
        D.Batch(jobs, device=x.device).run()
        return out

    def forward_conv(self, q, adjoint=False):
        return self.apply(self.fwd_kernel, q, adjoint)

    def decon(self, d, sigma_t, sigma_w, adjoint=False, which="value"):
        return self.apply(self.kernel(sigma_t, sigma_w, which), d, adjoint)

    def close(self):
        for k in self._cache.values():
            k.close()
        self._cache.clear()
        self._order.clear()
        self.fwd_kernel.close()


class _SpecView:
    """A PlaneSpec with replaced filters (the kernel builder reads nw,
    nt, the shifts and the two filters)."""

    def __init__(self, spec, wire_filter, time_filter):
        self.nw, self.nt = spec.nw, spec.nt
        self.wire_shift, self.time_shift = spec.wire_shift, spec.time_shift
        self.wire_filter, self.time_filter = wire_filter, time_filter
        self.kind = spec.kind


def _functions():
    import torch

    class ForwardConv(torch.autograd.Function):
        """D = F q on a batch (B, nw, nt); backward applies F^T."""

        @staticmethod
        def forward(ctx, q, model):
            ctx.model = model
            return model.forward_conv(q)

        @staticmethod
        def backward(ctx, g):
            return ctx.model.forward_conv(g, adjoint=True), None

    class Decon(torch.autograd.Function):
        """y = G_{s_t, s_w} d on a batch; backward gives dL/dd through
        G^T and dL/ds_t, dL/ds_w through the derivative-filter passes."""

        @staticmethod
        def forward(ctx, d, sigma_t, sigma_w, model):
            st, sw = float(sigma_t), float(sigma_w)
            ctx.model, ctx.st, ctx.sw = model, st, sw
            ctx.places = tuple((getattr(v, "dtype", torch.float64), getattr(v, "device", torch.device("cpu")))
                               for v in (sigma_t, sigma_w))
            ctx.save_for_backward(d)
            return model.decon(d, st, sw)

        @staticmethod
        def backward(ctx, g):
            (d,) = ctx.saved_tensors
            m, st, sw = ctx.model, ctx.st, ctx.sw
            g = g.contiguous()
            gd = gst = gsw = None
            if ctx.needs_input_grad[0]:
                gd = m.decon(g, st, sw, adjoint=True)
            if ctx.needs_input_grad[1]:
                yt = m.decon(d, st, sw, which="dt")
                dt_, dv_ = ctx.places[0]
                gst = torch.sum(yt.double() * g.double()).to(device=dv_, dtype=dt_).reshape(())
            if ctx.needs_input_grad[2]:
                yw = m.decon(d, st, sw, which="dw")
                dt_, dv_ = ctx.places[1]
                gsw = torch.sum(yw.double() * g.double()).to(device=dv_, dtype=dt_).reshape(())
            return gd, gst, gsw, None

    return ForwardConv, Decon


_FN = None


def functions():
    """(ForwardConv, Decon) torch.autograd.Function classes."""
    global _FN
    if _FN is None:
        _FN = _functions()
    return _FN


def forward_conv(q, model):
    return functions()[0].apply(q, model)


def decon(d, sigma_t, sigma_w, model):
    return functions()[1].apply(d, sigma_t, sigma_w, model)
