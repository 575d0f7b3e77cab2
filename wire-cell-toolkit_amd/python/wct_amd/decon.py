"""Host-side driver of the fused deconvolution through the C ABI.

PlaneKernel builds one plane type's FP64 response spectrum and its
regularised inverse on the device (wct_overall_response +
wct_decon_kernel_create); Batch runs any number of plane jobs with one
launch per stage (wct_decon_run).  Buffers are torch tensors on the
current device, used as plain device memory.
"""
import ctypes as C

import numpy as np

from .capi import MAX_OUTPUTS, DeconJob, DeconOutput, DeconTuning, check, lib
from . import workload as W


def _stream():
    import torch
    return torch.cuda.current_stream().cuda_stream


def overall_response(spec, nt=None, stream=None):
    """OmnibusSigProc::init_overall_response on the device in FP64."""
    nt = spec.nt if nt is None else nt
    rows = np.ascontiguousarray(spec.fr_rows, np.float64)
    nr, nin = rows.shape
    elec = np.ascontiguousarray(spec.elec, np.float64)
    out = np.zeros((nr, nt))
    check(lib().wct_overall_response(
        rows.ctypes.data, nr, nin, elec.ctypes.data, elec.size, W.FINE_PERIOD, W.TICK,
        W.INTER_GAIN, W.ADC_MV, W.FTOFFSET, nt, out.ctypes.data,
        _stream() if stream is None else stream))
    return out


class PlaneKernel:
    """Device-resident K(m, k) for one plane type (see include/wct_hip.h).

    unfiltered=True builds K without a time filter
    (wct_decon_kernel_create_unfiltered): each output of a job then
    applies its whole time filter, envelope (nt values, default the
    spec's time filter) sets the spectral cut and r_floor regularises
    the inverse."""

    def __init__(self, spec, cut_tol=1e-7, time_filter=None, resp=None, stream=None, unfiltered=False,
                 envelope=None, r_floor=0.0):
        self.spec = spec
        self.unfiltered = bool(unfiltered)
        self.resp = overall_response(spec) if resp is None else np.ascontiguousarray(resp, np.float64)
        tf = spec.time_filter if time_filter is None else time_filter
        tf = np.ascontiguousarray(tf, np.float64)
        wf = np.ascontiguousarray(spec.wire_filter, np.float64)
        h = C.c_void_p()
        st = _stream() if stream is None else stream
        if self.unfiltered:
            env = tf if envelope is None else np.ascontiguousarray(envelope, np.float64)
            check(lib().wct_decon_kernel_create_unfiltered(
                C.byref(h), self.resp.ctypes.data, self.resp.shape[0], spec.nw, spec.nt,
                wf.ctypes.data, env.ctypes.data, spec.wire_shift, spec.time_shift, cut_tol,
                float(r_floor), st))
        else:
            check(lib().wct_decon_kernel_create(
                C.byref(h), self.resp.ctypes.data, self.resp.shape[0], spec.nw, spec.nt,
                wf.ctypes.data, tf.ctypes.data, spec.wire_shift, spec.time_shift, cut_tol, st))
        self.handle = h
        nw, nt, kc = C.c_int(), C.c_int(), C.c_int()
        check(lib().wct_decon_kernel_info(h, C.byref(nw), C.byref(nt), C.byref(kc)))
        self.nw, self.nt, self.kc = nw.value, nt.value, kc.value

    def download(self):
        buf = np.zeros(2 * self.kc * self.nw, np.float32)
        check(lib().wct_decon_kernel_download(self.handle, buf.ctypes.data))
        return buf.view(np.complex64).reshape(self.kc, self.nw)

    def close(self):
        if getattr(self, "handle", None):
            lib().wct_decon_kernel_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _ptr(t):
    return None if t is None else t.data_ptr()


def _check_tensor(name, t, shape, dtypes, device):
    """Raises ValueError unless t is a tensor of the given shape (None
    entries match any extent), one of dtypes, on device, with unit inner
    stride: the C side trusts the pointers and strides it is given."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name}: expected a torch tensor, got {type(t).__name__}")
    if t.dim() != len(shape) or any(w is not None and w != g for w, g in zip(shape, t.shape)):
        raise ValueError(f"{name}: shape {tuple(t.shape)}, expected {tuple(shape)}")
    if t.dtype not in dtypes:
        raise ValueError(f"{name}: dtype {t.dtype}, expected one of {dtypes}")
    if t.device != device:
        raise ValueError(f"{name}: on {t.device}, the job's frames are on {device}")
    if t.dim() and t.stride(-1) != 1:
        raise ValueError(f"{name}: inner stride {t.stride(-1)}, expected 1")
    if t.dim() == 2 and t.shape[0] > 1 and t.stride(0) < t.shape[1]:
        raise ValueError(f"{name}: row stride {t.stride(0)} below the row length {t.shape[1]}")


class Output:
    """One extra output of a job (wct_decon_output): the job's spectrum
    through filter (device float32, >= kc values; rows flagged in
    alt_rows, a device uint8 (nw,) tensor, take filter_alt), optionally
    baselined, with its per-wire cal_RMS and save_data's zeroing."""

    def __init__(self, filter, out=None, rms=None, restore=1, save_mode=0, filter_alt=None, alt_rows=None):
        self.filter, self.out, self.rms = filter, out, rms
        self.restore, self.save_mode = int(restore), int(save_mode)
        self.filter_alt, self.alt_rows = filter_alt, alt_rows


def _check_gpu(dev):
    if dev.type != "cuda":
        raise ValueError(f"the job's tensors are on {dev}, expected a GPU device")


def make_job(kernel, adc, out, mask=None, rms=None, th_factor=W.TH_FACTOR, restore=1,
             save_mode=0, lf_tight=None, lf_tighter=None, bad=None, adjoint=False, spectrum=None,
             outputs=(), chan=None):
    """One wct_decon_job.  adc/out/mask/rms are torch tensors (2D for the
    frames); bad is (ptr, lo, hi) int32 device tensors or None.  With
    spectrum (a complex64 (nw, nt) tensor) the job writes the time
    spectrum of the deconvolved frame instead (out must be None).  chan
    is (table, rows) for the per-channel electronics correction: table a
    complex64 (nresp, >= kc) tensor from channel_factors(), rows an
    int32 (nw,) tensor of table rows (-1: no correction).

    Every tensor is checked against the kernel's (nw, nt) before any
    launch (ValueError on a mismatch), and the job keeps references to
    them: a Batch built from it keeps them alive as long as its plan."""
    import torch
    nw, nt, kc = kernel.nw, kernel.nt, kernel.kc
    _check_tensor("adc", adc, (nw, nt), (torch.float32,), adc.device)
    dev = adc.device
    j = DeconJob()
    j.kernel = kernel.handle
    j.adc = adc.data_ptr()
    j.ld_adc = adc.stride(0)
    keep = [kernel, adc]
    if chan is not None:
        table, rows = chan
        _check_tensor("chan table", table, (None, None), (torch.complex64,), dev)
        if table.shape[1] < kc or table.stride(1) != 1:
            raise ValueError(f"chan table: rows of {table.shape[1]} bins, the kernel keeps {kc}")
        _check_tensor("chan rows", rows, (nw,), (torch.int32,), dev)
        j.chan_factors = table.data_ptr()
        j.ld_chan_factors = table.stride(0)
        j.chan_factor_row = rows.data_ptr()
        keep += [table, rows]
    if spectrum is not None:
        if out is not None:
            raise ValueError("a spectrum-mode job has no out frame")
        _check_tensor("spectrum", spectrum, (nw, nt), (torch.complex64,), dev)
        j.spectrum = spectrum.data_ptr()
        j.ld_spectrum = spectrum.stride(0)
        keep.append(spectrum)
        _check_gpu(dev)
        j._keep = keep
        return j
    _check_tensor("out", out, (nw, nt), (torch.float32,), dev)
    keep.append(out)
    if mask is not None:
        _check_tensor("mask", mask, (nw, nt), (torch.uint8,), dev)
        keep.append(mask)
    if rms is not None:
        _check_tensor("rms", rms, (nw,), (torch.float32,), dev)
        keep.append(rms)
    for name, f in (("lf_tight", lf_tight), ("lf_tighter", lf_tighter)):
        if f is not None:
            _check_tensor(name, f, (None,), (torch.float32,), dev)
            if f.numel() < kc:
                raise ValueError(f"{name}: {f.numel()} values, the kernel keeps {kc} bins")
            keep.append(f)
    if bad is not None:
        bp, blo, bhi = bad
        _check_tensor("bad_ptr", bp, (nw + 1,), (torch.int32,), dev)
        _check_tensor("bad_lo", blo, (None,), (torch.int32,), dev)
        _check_tensor("bad_hi", bhi, (blo.shape[0],), (torch.int32,), dev)
        keep.extend(bad)
    j.out = out.data_ptr()
    j.ld_out = out.stride(0)
    j.mask = _ptr(mask)
    j.ld_mask = mask.stride(0) if mask is not None else 0
    j.rms = _ptr(rms)
    if bad is not None:
        j.bad_ptr, j.bad_lo, j.bad_hi = (_ptr(b) for b in bad)
    j.lf_tight = _ptr(lf_tight)
    j.lf_tighter = _ptr(lf_tighter)
    j.th_factor = th_factor
    j.restore_baseline = restore
    j.save_mode = save_mode
    j.adjoint = int(bool(adjoint))
    outputs = list(outputs)
    if len(outputs) > MAX_OUTPUTS:
        raise ValueError(f"{len(outputs)} outputs exceed {MAX_OUTPUTS}")
    if outputs:
        arr = (DeconOutput * len(outputs))()
        for q, o in enumerate(outputs):
            _check_tensor(f"outputs[{q}].filter", o.filter, (None,), (torch.float32,), dev)
            if o.filter.numel() < kc:
                raise ValueError(f"outputs[{q}].filter: {o.filter.numel()} values, the kernel keeps {kc} bins")
            d = arr[q]
            d.filter = o.filter.data_ptr()
            keep.append(o.filter)
            if o.filter_alt is not None or o.alt_rows is not None:
                if o.filter_alt is None or o.alt_rows is None:
                    raise ValueError(f"outputs[{q}]: filter_alt and alt_rows go together")
                _check_tensor(f"outputs[{q}].filter_alt", o.filter_alt, (None,), (torch.float32,), dev)
                if o.filter_alt.numel() < kc:
                    raise ValueError(f"outputs[{q}].filter_alt: {o.filter_alt.numel()} values, the kernel "
                                     f"keeps {kc} bins")
                _check_tensor(f"outputs[{q}].alt_rows", o.alt_rows, (nw,), (torch.uint8,), dev)
                d.filter_alt = o.filter_alt.data_ptr()
                d.alt_rows = o.alt_rows.data_ptr()
                keep += [o.filter_alt, o.alt_rows]
            if o.out is not None:
                _check_tensor(f"outputs[{q}].out", o.out, (nw, nt), (torch.float32,), dev)
                d.out = o.out.data_ptr()
                d.ld_out = o.out.stride(0)
                keep.append(o.out)
            if o.rms is not None:
                _check_tensor(f"outputs[{q}].rms", o.rms, (nw,), (torch.float32,), dev)
                d.rms = o.rms.data_ptr()
                keep.append(o.rms)
            d.restore_baseline, d.save_mode = o.restore, o.save_mode
        j.outputs = arr
        j.noutputs = len(outputs)
        keep.append(arr)
    _check_gpu(dev)
    j._keep = keep
    return j


def channel_factors(responses, elec, nt, device="cuda:0", stream=None):
    """Per-channel correction table (wct_channel_factors): responses is
    (nresp, nsamp) float32 channel responses, elec nt float64 samples of
    the nominal electronics response; returns complex64 (nresp, nt/2+1)
    rows of elec(k) / C_r(k) on the device (0 where C_r(k) == 0)."""
    import numpy as np
    import torch
    r = np.ascontiguousarray(responses, np.float32)
    e = np.ascontiguousarray(elec, np.float64)
    if r.ndim != 2 or e.shape != (nt,):
        raise ValueError("responses must be (nresp, nsamp) and elec nt long")
    out = torch.empty((r.shape[0], nt // 2 + 1), dtype=torch.complex64, device=device)
    check(lib().wct_channel_factors(out.data_ptr(), out.stride(0), r.ctypes.data, r.shape[0], r.shape[1],
                                    e.ctypes.data, nt, _stream() if stream is None else stream))
    return out


class Batch:
    """A fixed list of plane jobs and its workspace; run() launches the
    three stages once over all of them."""

    def __init__(self, jobs, device=None, workspace=None, wire_tile=0, row_grid_pct=0, row_plan=0, r2c_blocks=0,
                 c2r_blocks=0):
        import torch
        # the jobs' tensors must outlive the plan (and any graph of it)
        self._keep = [getattr(j, "_keep", None) for j in jobs]
        self.jobs = (DeconJob * len(jobs))(*jobs)
        self.n = len(jobs)
        self.ws_bytes = lib().wct_decon_workspace_size(self.jobs, self.n)
        if workspace is None:
            self.ws = torch.empty(self.ws_bytes, dtype=torch.uint8, device=device or "cuda")
        else:   # a shared workspace (runs on one stream reuse it in turn)
            if workspace.numel() < self.ws_bytes:
                raise ValueError("shared workspace too small")
            self.ws = workspace

        h = C.c_void_p()
        tun = DeconTuning(int(wire_tile), int(row_grid_pct), int(row_plan), int(r2c_blocks), int(c2r_blocks))
        check(lib().wct_decon_plan_create_tuned(C.byref(h), self.jobs, self.n, self.ws.data_ptr(),
                                                self.ws.numel(), C.byref(tun)))
        self.plan = h

    def run(self, stream=None):
        check(lib().wct_decon_plan_run(self.plan, _stream() if stream is None else stream))

    def run_stage(self, stage, stream=None):
        """One of the three launches of run() (0 r2c, 1 wire pass, 2 c2r);
        run them in order."""
        check(lib().wct_decon_plan_run_stage(self.plan, int(stage), _stream() if stream is None else stream))

    def close(self):
        if getattr(self, "plan", None):
            import torch
            torch.cuda.synchronize()
            lib().wct_decon_plan_destroy(self.plan)
            self.plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """A hipGraph replaying the runs of several batches in order
    (wct_decon_graph_create), captured on a side stream; launch() replays
    it on the current torch stream."""

    def __init__(self, batches, stream=None):
        import torch
        self.batches = list(batches)     # the plans must outlive the graph
        arr = (C.c_void_p * len(self.batches))(*[b.plan for b in self.batches])
        h = C.c_void_p()
        if stream is None:
            stream = torch.cuda.Stream()
        self._capture_stream = stream
        torch.cuda.synchronize()
        check(lib().wct_decon_graph_create(C.byref(h), arr, len(self.batches), stream.cuda_stream))
        self.graph = h

    def launch(self, stream=None):
        check(lib().wct_decon_graph_launch(self.graph, _stream() if stream is None else stream))

    def close(self):
        if getattr(self, "graph", None):
            import torch
            torch.cuda.synchronize()
            lib().wct_decon_graph_destroy(self.graph)
            self.graph = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ForwardKernel(PlaneKernel):
    """The forward response convolution (wct_forward_kernel_create): a job
    run with it turns a charge raster into the ADC frame whose decon with
    unit filters returns the raster."""

    def __init__(self, spec, resp=None, stream=None):
        self.spec = spec
        self.resp = overall_response(spec) if resp is None else np.ascontiguousarray(resp, np.float64)
        h = C.c_void_p()
        check(lib().wct_forward_kernel_create(
            C.byref(h), self.resp.ctypes.data, self.resp.shape[0], spec.nw, spec.nt,
            spec.wire_shift, spec.time_shift, _stream() if stream is None else stream))
        self.handle = h
        nw, nt, kc = C.c_int(), C.c_int(), C.c_int()
        check(lib().wct_decon_kernel_info(h, C.byref(nw), C.byref(nt), C.byref(kc)))
        self.nw, self.nt, self.kc = nw.value, nt.value, kc.value


def convolve(fwd_kernel, q, out):
    """out = forward response convolution of q (torch tensors on device)."""
    Batch([make_job(fwd_kernel, q, out, restore=0)]).run()


class PlainPlane:
    """The plain per-plane path (wct_plain_decon_plane): the reference's
    per-step control flow, FP32, one launch per step.  Holds the plane's
    FP32 response rows and filters on the device; the tight filter is
    Wiener (x ROI_tight_lf on induction planes) and the tighter filter
    Wiener x ROI_tighter_lf (induction only), multiplied in float as the
    reference multiplies its filter waveforms."""

    def __init__(self, spec, resp=None, device=None):
        import torch
        dev = device or "cuda"
        self.spec = spec
        resp = overall_response(spec) if resp is None else np.asarray(resp, np.float64)
        f32 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(dev)
        self.resp = f32(resp)
        self.nr = resp.shape[0]
        self.wire = f32(spec.wire_filter)
        tf = np.asarray(spec.time_filter, np.float32)
        self.tighter = None
        if spec.kind == "ind" and spec.lf_tight is not None and not spec.lf_folded:
            self.tight = f32(tf * np.asarray(spec.lf_tight, np.float32))
            self.tighter = f32(tf * np.asarray(spec.lf_tighter, np.float32))
        else:   # collection, or the tight LF already folded into time_filter
            self.tight = f32(tf)
        self.ws_bytes = lib().wct_plain_workspace_size(spec.nw, spec.nt)
        self.ws = torch.empty(self.ws_bytes, dtype=torch.uint8, device=dev)

    def run(self, adc, out, mask=None, rms=None, th_factor=W.TH_FACTOR, save_mode=0, bad=None,
            tighter=True, stream=None):
        sp = self.spec
        bp = blo = bhi = None
        if bad is not None:
            bp, blo, bhi = (b.data_ptr() for b in bad)
        check(lib().wct_plain_decon_plane(
            adc.data_ptr(), adc.stride(0), sp.nw, sp.nt, self.resp.data_ptr(), self.nr,
            self.wire.data_ptr(), self.tight.data_ptr(),
            self.tighter.data_ptr() if (tighter and self.tighter is not None) else None,
            sp.wire_shift, sp.time_shift, th_factor, bp, blo, bhi, out.data_ptr(), out.stride(0),
            _ptr(mask), mask.stride(0) if mask is not None else 0, _ptr(rms), save_mode,
            self.ws.data_ptr(), self.ws_bytes, _stream() if stream is None else stream))


def spectrum_filter(spec, filt, out, restore=1, stream=None):
    """wct_spectrum_filter: Decon2DFilter's inverse pass on a complex64
    (nw, nt) device spectrum; filt holds nt/2 + 1 float32 device values;
    out is a float32 (nw, nt) device tensor."""
    nw, nt = out.shape
    check(lib().wct_spectrum_filter(spec.data_ptr(), spec.stride(0), nw, nt, filt.data_ptr(), int(restore),
                                    out.data_ptr(), out.stride(0), _stream() if stream is None else stream))
