"""Benchmark: deconvolved channel-ticks/s of the batched fused FP32
deconvolution on one MI355X.

Workload (default "apa150"): 150 APAs x (800 + 800 + 960) channels x
6000 ticks, float32 ADC resident in HBM (9.2 GB), half the APAs
noise-only and half with tracks at 20 tracks per 480 wires.  One step
deconvolves every plane (tight Wiener filter, wire filter, response
inverse), restores the baseline, computes the per-wire robust RMS and
writes the decon frame (fp32) and the 3-sigma ROI mask (u8).  The planes
are processed in chunks of --chunk APAs whose spectra share one
workspace, three launches per chunk.

Prints one JSON line (metric, value, config) with the roofline of the
dominant kernel (the stage with the longest launch, timed on its own
with HIP events after the timed steps; the whole pipeline's figures sit
under roofline.pipeline) and the CPU baseline (the oracle's FP32
step-by-step restatement of the reference path, one plane per thread)
beside it.
"""
import argparse
import json
import math
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "wire-cell-toolkit_amd", "python"))

HBM_PEAK_GBS = 8000.0
B_ALG = 9.0     # bytes per channel-tick: 4 read + 4 write + 1 mask


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="apa150")
    ap.add_argument("--chunk", type=int, default=8, help="APAs per launch group")
    ap.add_argument("--cut", type=float, default=1e-7, help="spectral cut tolerance")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="CPU baseline threads (0 = the CPUs this process may use, at most 16)")
    ap.add_argument("--chargefit-apas", type=int, default=8,
                    help="signal APAs in the timed 20-step charge fit (0 = skip)")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the FP64 accuracy check")
    ap.add_argument("--plain-apas", type=int, default=4,
                    help="APAs timed through the plain per-plane path (0 = skip)")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner (use --chunk and --wire-tile, full row grid)")
    ap.add_argument("--wire-tile", type=int, default=0,
                    help="bins per wire-pass tile with --no-tune (0 = the plan's default)")
    ap.add_argument("--row-plan", type=int, nargs=3, default=(0, 0, 0), metavar=("PLAN", "R2C", "C2R"),
                    help="with --no-tune: row plan (-1 runtime, 0 default = 1 compile-time) and the r2c / "
                         "c2r workgroups per CU of the compile-time plan (0 = its default)")
    ap.add_argument("--eager", action="store_true", help="time plan runs without the hipGraph")
    ap.add_argument("--impact", type=int, default=1,
                    help="1: time the forward simulation convolution (one 896 x 6272 plane frame) as a side leg")
    ap.add_argument("--impact-only", action="store_true",
                    help="run only the forward simulation leg and print its own JSON line")
    ap.add_argument("--noise-apas", type=int, default=4,
                    help="APAs (2560 channels x 6000 ticks) in the timed noise-filter side leg (0 = skip)")
    ap.add_argument("--noise-only", action="store_true",
                    help="run only the noise-filter leg and print its own JSON line")
    ap.add_argument("--adjoint-frames", type=int, default=16,
                    help="frames of 800 x 6000 in the timed adjoint gradient step (0 = skip)")
    return ap.parse_args()


def make_data(planes, kernels_fwd, dev, seed=1234):
    """Synthetic frames on the device: noise for every plane, the forward
    response convolution of a track raster for every other APA (a pool of
    6 distinct signal APAs reused cyclically)."""
    import numpy as np
    import torch
    from wct_amd import decon, workload as W
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    rng = np.random.default_rng(seed)
    frames = []
    pool = {}
    for i, (kind, nw, nt) in enumerate(planes):
        apa, p = divmod(i, 3)
        adc = torch.randn((nw, nt), generator=g, device=dev, dtype=torch.float32) * W.NOISE_ADC
        if apa % 2 == 1:
            key = ((apa // 2) % 6, p)
            if key not in pool:
                q = torch.from_numpy(W.tracks_raster(rng, W.tracks_per_plane(nw), nw, nt)).to(dev)
                sig = torch.empty_like(q)
                decon.convolve(kernels_fwd[(kind, nw, nt)], q, sig)
                pool[key] = sig
            adc += pool[key]
        frames.append(adc)
        if i % 90 == 0:
            log(f"bench: generated plane {i}/{len(planes)}")
    torch.cuda.synchronize()
    return frames


def cpu_info():
    """CPU model (/proc/cpuinfo), logical CPU count (nproc) and the CPUs
    this process may run on."""
    model, family, number = "", "", ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and not model:
                    model = v
                elif k == "cpu family" and not family:
                    family = v
                elif k == "model" and not number:
                    number = v
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    return {"model": model, "family": family, "model_number": number, "nproc": os.cpu_count(),
            "usable": usable}


def cpu_baseline(planes, nthreads):
    """The oracle's FP32 restatement of the reference CPU path (decon_2D_init,
    tight filter, restore_baseline, cal_RMS/threshold), one plane per
    thread as the reference runs one OmnibusSigProc per APA thread, on
    bounded samples: the CPU-runnable collection plane (480 x 6000), one
    APA (800/800/960 x 6000, three threads) and nthreads planes of the
    workload (one per thread)."""
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O
    from wct_amd import workload as W
    specs = {}

    def spec(t):
        if t not in specs:
            sp = W.plane_spec(*t)
            resp = O.overall_response(sp.fr_rows, sp.elec, W.FINE_PERIOD, W.TICK, W.INTER_GAIN,
                                      W.ADC_MV, W.FTOFFSET, t[2], dtype=np.float32)
            specs[t] = (sp, resp)
        return specs[t]

    rng = np.random.default_rng(7)

    def run(sample):
        inputs = [(rng.standard_normal((nw, nt)) * W.NOISE_ADC).astype(np.float32) for (_, nw, nt) in sample]
        for t in set(sample):
            spec(t)

        def one(i):
            sp, resp = spec(sample[i])
            c = O.decon_init(inputs[i], resp, sp.wire_filter.astype(np.float32), sp.wire_shift,
                             sp.time_shift, np.float32)
            y = O.decon_apply_filter(c, sp.time_filter.astype(np.float32), np.float32)
            yb, _ = O.restore_baseline(y)
            O.roi_threshold(yb, W.TH_FACTOR)
            return y.size

        t0 = time.perf_counter()
        with ThreadPoolExecutor(len(sample)) as ex:
            tot = sum(ex.map(one, range(len(sample))))
        dt = time.perf_counter() - t0
        return tot / dt, dt

    info = cpu_info()
    v0, d0 = run([("col", 480, 6000)])
    v1, d1 = run(W.config_planes("apa"))
    sample = [planes[i % len(planes)] for i in range(nthreads)]
    v4, d4 = run(sample)
    return {"value": v4, "unit": "channel-ticks/s", "cores": nthreads, "kind": "port",
            "sample": f"{nthreads} planes of the 150-APA workload ({'/'.join(f'{k}{nw}x{nt}' for k, nw, nt in sample[:3])}"
                      f"...) one per thread, {d4:.1f} s wall; own mixed-radix FFT (no FFTW); "
                      f"CPU: {info['model']} (family {info['family']} model {info['model_number']}), "
                      f"nproc {info['nproc']}, {info['usable']} usable",
            "configs": {"col480x6000 (1 thread)": {"value": v0, "wall_s": round(d0, 2)},
                        "apa 800/800/960x6000 (3 threads)": {"value": v1, "wall_s": round(d1, 2)},
                        f"apa150 sample ({nthreads} threads)": {"value": v4, "wall_s": round(d4, 2)}},
            "cpu": info}


STAGES = ("k_row_r2c (forward row transforms)", "k_wire_pass (wire transforms x K)",
          "k_row_c2r_roi (inverse rows, baseline, RMS, ROI mask, outputs)")


def stage_times(batches, reps):
    """ms per step of each of the three launches of a step, timed with HIP
    events on the stream the launches run on (each chunk's stages in
    order, as a step runs them)."""
    import torch
    stream = torch.cuda.current_stream()
    acc = [0.0, 0.0, 0.0]
    for _ in range(reps):
        for b in batches:
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
            ev[0].record(stream)
            for st in range(3):
                b.run_stage(st)
                ev[st + 1].record(stream)
            torch.cuda.synchronize()
            for st in range(3):
                acc[st] += ev[st].elapsed_time(ev[st + 1])
    return [a / reps for a in acc]


def signal_split_leg(planes, make_jobs, tuning, reps=5):
    """The three stages timed on one chunk of noise-only APAs and on one
    chunk of APAs with tracks (every other APA carries them), each as its
    own batch of the tuned shape: the c2r stage skips its +-500 window pass
    on rows whose histogram proves every sample inside the window, which
    every noise-only row does and rows crossed by tracks may not."""
    from wct_amd import decon
    per = 3 * tuning["chunk_apas"]
    sets = {"noise_only": [i for i in range(len(planes)) if (i // 3) % 2 == 0][:per],
            "tracks": [i for i in range(len(planes)) if (i // 3) % 2 == 1][:per]}
    if min(len(v) for v in sets.values()) < per:
        return None
    out = {}
    for name, idx in sets.items():
        rp = tuning["row_plan"]
        b = decon.Batch(make_jobs(idx), wire_tile=tuning["wire_tile"], row_grid_pct=tuning["row_grid_pct"],
                        row_plan=rp[0], r2c_blocks=rp[1], c2r_blocks=rp[2])
        b.run()
        ms = stage_times([b], reps)
        out[name] = {"apas": per // 3, "stage_ms": [round(m, 4) for m in ms]}
        b.close()
    out["c2r_ratio"] = round(out["tracks"]["stage_ms"][2] / out["noise_only"]["stage_ms"][2], 4)
    return out


def stage_alg_bytes(planes, kern):
    """Algorithmic bytes of each stage per step: r2c reads the frame (4 B
    per sample) and writes the kept bins (8 B each); the wire pass reads
    and writes the kept bins and reads K; c2r reads the kept bins and
    writes the output (4 B) and the mask (1 B) per sample."""
    b = [0.0, 0.0, 0.0]
    for t in planes:
        kind, nw, nt = t
        kc = kern[t].kc
        b[0] += 4.0 * nw * nt + 8.0 * kc * nw
        b[1] += 24.0 * kc * nw
        b[2] += 8.0 * kc * nw + 5.0 * nw * nt
    return b


def calibration_copy(dev, mib=1024):
    """One elementwise copy of known size (mib MiB read, mib MiB written):
    its FETCH_SIZE / WRITE_SIZE counters scale the decon kernels' to
    bytes."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    del src, dst


def hbm_traffic():
    """HBM bytes of the fused pipeline from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE of the launches of one 150-APA step run as
    one plan, scaled by the calibration copy), as (bytes of the step,
    [bytes of each stage], channel-ticks), or None when absent."""
    path = os.path.join(REPO, "profiles", "decon_fused_apa150_hbm_traffic.md")
    try:
        vals = {}
        with open(path) as f:
            for line in f:
                if line.startswith("- ") and ":" in line:
                    k, v = line[2:].split(":", 1)
                    vals[k.strip()] = v.strip()
        stages = [float(x) for x in vals["stage_bytes_per_chunk"].split(",")]
        return float(vals["bytes_per_chunk"]), stages, int(vals["channel_ticks_per_chunk"])
    except (OSError, KeyError, ValueError):
        return None


def adjoint_leg(dev, nframes, reps=3):
    """One gradient step of the charge fit on nframes induction frames of
    800 x 6000: forward response convolution, filtered decon, least-squares
    loss, backward through the hand-written adjoints to the raster and
    both filter widths (six fused passes)."""
    import numpy as np
    import torch
    from wct_amd import autograd as A, workload as W
    m = A.Model(W.plane_spec("ind", 800, 6000))
    rng = np.random.default_rng(3)
    q_true = torch.from_numpy(np.stack([W.charge_raster(rng, 800, 6000) for _ in range(nframes)])).to(dev)
    T = m.forward_conv(q_true) + torch.randn_like(q_true) * W.NOISE_ADC
    s_t = torch.tensor(1e-4, dtype=torch.float64, requires_grad=True)
    s_w = torch.tensor(0.75, dtype=torch.float64, requires_grad=True)

    def one():
        q = torch.zeros_like(q_true, requires_grad=True)
        y = A.decon(A.forward_conv(q, m), s_t, s_w, m)
        L = 0.5 * ((y - T) ** 2).sum()
        L.backward()
        return L

    one()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        one()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    m.close()
    ct = nframes * 800 * 6000
    return {"ms": round(ms, 3), "frames": nframes, "shape": "800x6000",
            "passes": 6, "channel_ticks_per_s": ct / (ms / 1e3)}


def impact_leg(dev, reps=20, cpu=True):
    """Gen::ImpactTransform on one plane: 20000 depositions (1.44 M charge
    entries in 11 impact groups) convolved with 21-wire responses into an
    896 x 6272 frame, charges and responses resident on the device.  The
    dominant kernel, the forward time transform of the six packed frames
    (6 x 896 rows of 6272), is timed on its own through wct_dft_fwd1b,
    which launches the same kernel on the same shape.  The CPU baseline is
    the oracle's FP64 numpy restatement (single-threaded transforms) on
    the same inputs, which also gives the relative error."""
    import ctypes as C
    import numpy as np
    import torch
    from wct_amd import workload as W
    from wct_amd.capi import check, lib
    from wct_amd.impact import ImpactTransform
    P = W.IMPACT
    G = P["num_group"]
    resp = W.impact_responses(G, P["nwires_resp"], P["nresp"]).astype(np.float32)
    it = ImpactTransform(resp)
    fr = it.geometry(P["impact_range"], P["time_range"], P["waveform_pad"])
    offs, ch, tk, q = W.impact_charges(fr, G, P["ndepo"], P["time_range"])
    dch = torch.as_tensor(ch, device=dev)
    dtk = torch.as_tensor(tk, device=dev)
    dq = torch.as_tensor(q, device=dev)
    nw, nt = fr["nw"], fr["nt"]
    out = torch.empty((nw, nt), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream()

    def timed(fn, n):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(n):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / n

    run = lambda: it.run(fr, offs, dch, dtk, dq, out=out)
    for _ in range(3):
        run()
    ms = timed(run, reps)
    nframes = (G - 1) // 2 + 1
    rows = torch.empty((nframes * nw, nt), dtype=torch.complex64, device=dev)
    rows.normal_()
    fwd = lambda: check(lib().wct_dft_fwd1b(C.c_void_p(rows.data_ptr()), C.c_void_p(rows.data_ptr()),
                                            nframes * nw, nt, 1, stream.cuda_stream))
    fwd()
    row_ms = timed(fwd, reps)
    row_bytes = 2 * 8 * nframes * nw * nt
    res = {
        "workload": f"ImpactTransform, {G} impact groups x {P['nwires_resp']}-wire responses, "
                    f"{P['ndepo']} depositions, frame {nw} x {nt}",
        "charges": int(offs[-1]), "frame": [nw, nt], "ms_per_plane": round(ms, 4),
        "samples_per_s": nw * nt / (ms / 1e3),
        "dominant_kernel": {"kernel": "k_fft_rows_lin<6272> (forward time transform of the packed frames)",
                            "launch_ms": round(row_ms, 4), "alg_bytes": row_bytes,
                            "achieved_gbs": row_bytes / (row_ms / 1e3) / 1e9,
                            "frac": row_bytes / (row_ms / 1e3) / 1e9 / HBM_PEAK_GBS},
        "dtype": "f32",
    }
    if cpu:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import impact_transform as IT
        geo = IT.geometry(G, P["nwires_resp"], P["impact_range"], P["time_range"], P["waveform_pad"])
        groups = [np.stack([ch[a:b], tk[a:b], q[a:b]], axis=1) for a, b in zip(offs[:-1], offs[1:])]
        t0 = time.perf_counter()
        want = IT.transform(groups, resp.astype(np.float64), geo)
        cpu_s = time.perf_counter() - t0
        got = out.cpu().numpy().astype(np.float64)
        d = got - want
        res["rel_rms_vs_fp64"] = float(np.sqrt((d * d).sum() / (want * want).sum()))
        res["cpu_baseline"] = {"value": nw * nt / cpu_s, "unit": "samples/s", "cores": 1, "kind": "port",
                               "sample": "the whole plane through the oracle's numpy FP64 restatement",
                               "seconds": round(cpu_s, 3)}
    it.close()
    return res


def noise_leg(dev, napas=4, reps=10, cpu=True):
    """ProtoDUNE's per-channel noise filter (pdOneChannelNoise as
    OmnibusNoiseFilter runs it) on napas APAs of 2560 channels x 6000
    ticks resident on the device: rcrc 1.1 ms on every channel (rc_layers
    1), RMS cuts 1-30 ADC, seeded noise with RC-shaped pulses on a tenth of
    the channels and a slow partial waveform on one channel in 200.  The
    whole call is timed with HIP events on its stream, and one row
    transform of the same shape (wct_dft_fwd1b, the kernel of both
    transform stages) beside it.  CPU baseline: the oracle's one-channel
    restatement (pure Python and numpy) on the first 100 channels (about
    10 s)."""
    import ctypes as C
    import numpy as np
    import torch
    from wct_amd import noise as N
    from wct_amd.capi import NfChannel, check, lib
    nt, per_apa = 6000, 2560
    nch = napas * per_apa
    rng = np.random.default_rng(17)
    g = torch.Generator(device=dev)
    g.manual_seed(5)
    x = torch.randn((nch, nt), generator=g, device=dev) * 3.0 + 900.0
    t = torch.arange(nt, device=dev, dtype=torch.float32)
    rc = N.rcrc_spectrum(1.1e6, 500.0, nt)
    pulse_rows = rng.choice(nch, size=nch // 10, replace=False)
    q = np.zeros(nt)
    q[[1000, 3000, 4500]] = [600.0, 300.0, 900.0]
    q = np.convolve(q, np.exp(-0.5 * (np.arange(-20, 21) / 4.0) ** 2), "same")
    qrc = torch.as_tensor(np.fft.ifft(np.fft.fft(q) * rc.astype(np.complex128)).real, dtype=torch.float32,
                          device=dev)
    x[torch.as_tensor(pulse_rows, device=dev)] += qrc
    x[::200] += 3000.0 * torch.exp(-t / 300.0)
    planes = [(c % per_apa) // 800 if (c % per_apa) < 1600 else 2 for c in range(nch)]
    entries = (NfChannel * nch)(*[NfChannel(p, 0, 0, 0, 1.0, 30.0, 0, 0.0, 1.0, -1, -1) for p in planes])
    dch = torch.frombuffer(bytearray(bytes(entries)), dtype=torch.uint8).to(dev)
    drc = torch.as_tensor(rc, device=dev)
    out = torch.empty_like(x)
    flags = torch.zeros((nch, 3), dtype=torch.int32, device=dev)
    ws_bytes = lib().wct_nf_workspace_size(nch, nt)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream()
    run = lambda: check(lib().wct_nf_one_channel(
        C.c_void_p(x.data_ptr()), nt, C.c_void_p(out.data_ptr()), nt, nch, nt, C.c_void_p(dch.data_ptr()),
        C.c_void_p(drc.data_ptr()), None, C.c_void_p(flags.data_ptr()), C.c_void_p(ws.data_ptr()), ws_bytes,
        stream.cuda_stream))
    rows = torch.empty((nch, nt), dtype=torch.complex64, device=dev)
    rows.normal_()
    fwd = lambda: check(lib().wct_dft_fwd1b(C.c_void_p(rows.data_ptr()), C.c_void_p(rows.data_ptr()), nch, nt, 1,
                                            stream.cuda_stream))

    def timed(fn, n):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(n):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / n

    run()
    ms = timed(run, reps)
    fwd()
    fft_ms = timed(fwd, reps)
    fl = flags.cpu().numpy()[:, 0]
    # the coherent subtraction (mbCoherentNoiseSub) on the filtered rows:
    # ProtoDUNE's groups (20 x 40 U, 20 x 40 V, 20 x 48 W channels per APA),
    # U and V groups through the response path (a 40-tick bipolar response
    # at offset 120)
    from wct_amd.capi import NfGroup
    groups, grows = [], []
    row_group = np.full(nch, -1, np.int32)
    for a in range(napas):
        for first, size, count in ((0, 40, 20), (800, 40, 20), (1600, 48, 20)):
            for k in range(count):
                rr = [a * per_apa + first + k * size + i for i in range(size)]
                resp = first < 1600
                groups.append(NfGroup(len(grows), size, 0 if resp else -1, 120 if resp else 0, 10, 10, 0.02, 0.08,
                                      15.0, 0.07, 3.0, 5.0, 50.0))
                row_group[rr] = len(groups) - 1
                grows.extend(rr)
    resp = np.zeros(nt)
    resp[100:140] = np.sin(np.linspace(0, 2 * np.pi, 40))
    dspec = torch.as_tensor(np.fft.fft(resp).astype(np.complex64), device=dev)
    dg = torch.frombuffer(bytearray(bytes((NfGroup * len(groups))(*groups))), dtype=torch.uint8).to(dev)
    dgr = torch.as_tensor(np.asarray(grows, np.int32), device=dev)
    drg = torch.as_tensor(row_group, device=dev)
    cws_bytes = lib().wct_nf_coherent_workspace_size(len(groups), nch, nt)
    cws = torch.empty(cws_bytes, dtype=torch.uint8, device=dev)
    work = out.clone()
    coh = lambda: check(lib().wct_nf_coherent(
        C.c_void_p(work.data_ptr()), nt, nch, nt, C.c_void_p(dg.data_ptr()), len(groups), C.c_void_p(dgr.data_ptr()),
        C.c_void_p(drg.data_ptr()), C.c_void_p(dspec.data_ptr()), 1, C.c_void_p(cws.data_ptr()), cws_bytes,
        stream.cuda_stream))
    coh()
    coh_ms = timed(coh, reps)
    # the status filter (mbOneChannelStatus) on the induction channels of
    # the filtered rows: Threshold 3.5, Window 5, Nbins 250, Cut 14
    ind = torch.cat([out[a * per_apa:a * per_apa + 1600] for a in range(napas)]).contiguous()
    nind = ind.shape[0]
    sws_bytes = lib().wct_nf_status_workspace_size(nind, nt)
    sws = torch.empty(sws_bytes, dtype=torch.uint8, device=dev)
    sflags = torch.empty(nind, dtype=torch.int32, device=dev)
    stat = lambda: check(lib().wct_nf_lf_noisy(
        C.c_void_p(ind.data_ptr()), nt, nind, nt, 3.5, 5, 250, 14.0, C.c_void_p(sflags.data_ptr()), None,
        C.c_void_p(sws.data_ptr()), sws_bytes, stream.cuda_stream))
    stat()
    stat_ms = timed(stat, reps)
    nlf = int(sflags.sum().item())
    # the bit-shift test (mbADCBitShift, 12 bits, 500 ticks examined) on
    # every raw row: none is shifted, so this is the detection alone
    bsx = x.clone()
    bns = torch.empty(nch, dtype=torch.int32, device=dev)
    bshift = lambda: check(lib().wct_nf_adc_bit_shift(
        C.c_void_p(bsx.data_ptr()), nt, nch, nt, 12, 500, 7.5, 0.8, C.c_void_p(bns.data_ptr()), stream.cuda_stream))
    bshift()
    bs_ms = timed(bshift, reps)
    nbs = int((bns > 0).sum().item())
    ct = nch * nt
    res = {
        "workload": f"pdOneChannelNoise on {napas} APAs x {per_apa} channels x {nt} ticks",
        "channels": nch, "ms": round(ms, 4), "channel_ticks_per_s": ct / (ms / 1e3),
        "alg_bytes_per_channel_tick": 8, "achieved_gbs": 8 * ct / (ms / 1e3) / 1e9,
        "frac": 8 * ct / (ms / 1e3) / 1e9 / HBM_PEAK_GBS,
        "row_transform_ms": round(fft_ms, 4),
        "coherent": {"groups": len(groups), "ms": round(coh_ms, 4),
                     "channel_ticks_per_s": ct / (coh_ms / 1e3)},
        "status": {"channels": nind, "ms": round(stat_ms, 4), "lf_noisy": nlf,
                   "channel_ticks_per_s": nind * nt / (stat_ms / 1e3)},
        "bit_shift": {"channels": nch, "ms": round(bs_ms, 4), "shifted": nbs},
        "partial": int(((fl & 1) != 0).sum()), "noisy": int(((fl & 2) != 0).sum()), "dtype": "f32",
    }
    if cpu:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import noise_filter as NFO
        sample = list(range(100))
        xs = x[sample].cpu().numpy()
        t0 = time.perf_counter()
        for i, c in enumerate(sample):
            NFO.one_channel_noise(xs[i], planes[c], rcrc=rc, min_rms=1.0, max_rms=30.0)
        sec = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": len(sample) * nt / sec, "unit": "channel-ticks/s", "cores": 1,
                               "kind": "port", "sample": f"{len(sample)} channels of 6000 ticks through the "
                                                         "oracle's one-channel restatement",
                               "seconds": round(sec, 3)}
    return res


def plain_leg(planes, frames, outs, masks, rmss, kern, napas, reps=2):
    """The plain per-plane device path (one launch per reference step,
    FP32, response spectrum rebuilt per call) over the first napas APAs
    of the same resident frames, timed on the same stream."""
    import torch
    from wct_amd import decon
    n = min(len(planes), 3 * napas)
    pp = {t: decon.PlainPlane(kern[t].spec, resp=kern[t].resp) for t in sorted(set(planes[:n]))}

    def run():
        for i in range(n):
            pp[planes[i]].run(frames[i], outs[i], mask=masks[i], rms=rmss[i])

    run()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        run()
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    ct = sum(nw * nt for _, nw, nt in planes[:n])
    return ct / (ms / 1e3), ms, n


def hbm_copy_peak(dev, gib=2, reps=5):
    """Device-to-device copy of gib GiB (read + write) timed with events:
    the HBM rate a plain stream copy reaches on this box, in GB/s."""
    import torch
    n = gib * (1 << 30) // 4
    src = torch.ones(n, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del src, dst
    torch.cuda.empty_cache()
    return 2.0 * n * 4 / (ms / 1e3) / 1e9


def fp64_decon(spec, resp, adc, dev):
    """The fused decon's definition in FP64 torch.fft on the device:
    Re IFFT2( FFT2(adc) Fw Ft conj(phase) / R ), 0 where R == 0."""
    import numpy as np
    import torch
    nw, nt = spec.nw, spec.nt
    Rr = np.zeros((nw, nt))
    Rr[:resp.shape[0]] = resp
    R = np.fft.fft2(Rr)
    mm = np.arange(nw)[:, None]
    kk = np.arange(nt)[None, :]
    ts = spec.time_shift if spec.time_shift > 0 else 0
    ph = np.exp(2j * np.pi * (mm * spec.wire_shift / nw + kk * ts / nt))
    with np.errstate(divide="ignore", invalid="ignore"):
        G = np.where(R != 0, np.conj(ph) / R, 0) * spec.wire_filter[:, None] * spec.time_filter[None, :]
    g = torch.from_numpy(G).to(dev)
    x = torch.fft.fft2(adc.double())
    return torch.fft.ifft2(x * g).real


def accuracy_leg(planes, frames, kern, dev, apa=1):
    """Per plane of one signal APA of the workload: the device decon before
    restore_baseline (the same kernels and kept bins as the timed step)
    against fp64_decon, as relative RMS and max error."""
    import torch
    from wct_amd import decon
    out = {}
    for i in range(3 * apa, 3 * apa + 3):
        t = planes[i]
        y = torch.empty_like(frames[i])
        decon.Batch([decon.make_job(kern[t], frames[i], y, restore=0)]).run()
        y64 = fp64_decon(kern[t].spec, kern[t].resp, frames[i], dev)
        torch.cuda.synchronize()
        d = y.double() - y64
        rel = float(torch.linalg.norm(d) / torch.linalg.norm(y64))
        mx = float(d.abs().max() / y64.abs().max())
        out[f"plane{i - 3 * apa}_{t[0]}{t[1]}x{t[2]}"] = {"rel_rms": rel, "max_rel": mx}
        del y, y64, d
    return out


def response_build_leg(types):
    """The FP64 response-spectrum build per plane type, timed on its own:
    init_overall_response on the device (wct_overall_response) and the
    regularised inverse with the spectral cut (wct_decon_kernel_create)."""
    import torch
    from wct_amd import decon, workload as W
    res = {}
    for t in types:
        spec = W.plane_spec(*t)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        resp = decon.overall_response(spec)
        t1 = time.perf_counter()
        k = decon.PlaneKernel(spec, resp=resp)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        k.close()
        res[f"{t[0]}{t[1]}x{t[2]}"] = {"overall_response_ms": round((t1 - t0) * 1e3, 2),
                                       "kernel_ms": round((t2 - t1) * 1e3, 2)}
    return res


def chargefit_leg(frames, planes, dev, napas=8, steps=20):
    """20 steepest-descent steps of the charge fit (wct_amd.chargefit) on
    the first napas signal APAs of the resident frames: 2 napas induction
    frames and napas collection frames per step, six fused passes each."""
    import torch
    from wct_amd import autograd as A, chargefit as CF, workload as W
    sel = [a for a in range(len(planes) // 3) if a % 2 == 1][:napas]
    groups = {}
    for a in sel:
        for p in range(3):
            groups.setdefault(planes[3 * a + p], []).append(frames[3 * a + p])
    models = {t: A.Model(W.plane_spec(*t)) for t in groups}
    adc = {t: torch.stack(v) for t, v in groups.items()}
    st, sw = W.WIENER["sigma"], W.WIRE["sigma"]
    for t in groups:     # warm the kernel caches
        CF.fit(models[t], adc[t], 1, st, sw)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    losses = {}
    for t in groups:
        _, losses[t] = CF.fit(models[t], adc[t], steps, st, sw)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ct = sum(v.numel() for v in adc.values())
    for m in models.values():
        m.close()
    return {"apas": len(sel), "steps": steps, "ms_per_step": round(dt * 1e3 / steps, 3),
            "channel_ticks_per_step": ct, "channel_ticks_per_s": ct * steps / dt,
            "loss_first_last": {f"{t[0]}{t[1]}x{t[2]}": [v[0], v[-1]] for t, v in losses.items()}}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; run with --gpus 1"}))
        sys.exit(2)
    import torch
    from wct_amd import decon, workload as W
    dev = torch.device("cuda:0")
    if args.noise_only:
        print(json.dumps({"metric": "noise-filtered channel-ticks/s (pdOneChannelNoise, 1 GPU)",
                          **noise_leg(dev, napas=max(1, args.noise_apas), cpu=not args.no_cpu)}), flush=True)
        return
    if args.impact_only:
        print(json.dumps({"metric": "forward simulation samples/s (ImpactTransform, 1 GPU)",
                          **impact_leg(dev, cpu=not args.no_cpu)}), flush=True)
        return
    planes = W.config_planes(args.workload)
    types = sorted(set(planes))
    calibration_copy(dev)
    copy_gbs = hbm_copy_peak(dev)
    log(f"bench: device copy {copy_gbs:.0f} GB/s (read + write)")
    resp_build = response_build_leg(types)
    log(f"bench: FP64 response build {resp_build}")
    log(f"bench: {args.workload}: {len(planes)} planes, plane types {types}")
    kern, fwd = {}, {}
    for t in types:
        spec = W.plane_spec(*t)
        kern[t] = decon.PlaneKernel(spec, cut_tol=args.cut)
        fwd[t] = decon.ForwardKernel(spec, resp=kern[t].resp)
    frames = make_data(planes, fwd, dev)
    outs = [torch.empty_like(f) for f in frames]
    masks = [torch.empty(f.shape, dtype=torch.uint8, device=dev) for f in frames]
    rmss = [torch.empty(f.shape[0], dtype=torch.float32, device=dev) for f in frames]
    def make_jobs(idx):
        return [decon.make_job(kern[planes[i]], frames[i], outs[i], mask=masks[i], rms=rmss[i],
                               restore=1) for i in idx]

    tuning = {"chunk_apas": args.chunk, "wire_tile": args.wire_tile, "row_grid_pct": 0,
              "row_plan": list(args.row_plan), "tuned": False}
    if not args.no_tune:
        from wct_amd import tune
        t0 = time.perf_counter()
        best, res = tune.tune(make_jobs, len(planes), log=log)
        tuning = {"chunk_apas": best[0], "wire_tile": best[1], "row_grid_pct": best[2],
                  "row_plan": list(best[3:]), "tuned": True,
                  "candidates": len(res), "tuning_s": round(time.perf_counter() - t0, 2),
                  "best_ms": round(res[best], 3),
                  "worst_ms": round(max(v for v in res.values() if isinstance(v, float)), 3)}
        log(f"bench: tuned {tuning}")
    from wct_amd import tune
    batches = tune.build_step(make_jobs, len(planes), tuning["chunk_apas"], tuning["wire_tile"],
                              tuning["row_grid_pct"], *tuning["row_plan"])
    graph = None if args.eager else decon.Graph(batches)
    ch_ticks = sum(nw * nt for _, nw, nt in planes)
    log(f"bench: {ch_ticks / 1e9:.3f} G channel-ticks per step, {len(batches)} chunks of "
        f"{tuning['chunk_apas']} APAs, kc={ {str(t): kern[t].kc for t in types} }, "
        f"{'eager' if graph is None else 'one hipGraph'} per step")

    def step():
        if graph is not None:
            graph.launch()
        else:
            for b in batches:
                b.run()

    for i in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log("bench: warmup done")
    stream = torch.cuda.current_stream()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for i in range(args.steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = ev0.elapsed_time(ev1) / args.steps
    ms = wall * 1e3 / args.steps
    value = ch_ticks / (ms / 1e3)
    achieved = B_ALG * ch_ticks / (gpu_ms / 1e3) / 1e9
    log(f"bench: {ms:.3f} ms/step wall, {gpu_ms:.3f} ms/step on the stream")

    stage_ms = stage_times(batches, reps=max(2, min(args.steps, 5)))
    log("bench: stages " + ", ".join(f"{m:.3f} ms" for m in stage_ms))

    plain = None
    if args.plain_apas > 0:
        pv, pms, pn = plain_leg(planes, frames, outs, masks, rmss, kern, args.plain_apas)
        log(f"bench: plain path {pms:.2f} ms for {pn} planes -> {pv:.3e} channel-ticks/s")
        plain = {"value": pv, "unit": "channel-ticks/s", "ratio": value / pv,
                 "sample": f"{pn} planes ({pn // 3} APAs) of the same frames, one launch per "
                           f"reference step, {pms:.2f} ms per pass"}

    split = signal_split_leg(planes, make_jobs, tuning)
    log(f"bench: stages on noise-only and track APAs {split}")

    accuracy, acc_ok = None, True
    if not args.no_accuracy:
        accuracy = accuracy_leg(planes, frames, kern, dev, apa=1 if len(planes) >= 6 else 0)
        acc_ok = all(v["rel_rms"] <= 1e-4 for v in accuracy.values())
        log(f"bench: FP64 accuracy {accuracy}")

    impact = None
    if args.impact:
        log("bench: forward simulation (ImpactTransform) leg")
        impact = impact_leg(dev, cpu=not args.no_cpu)
    noise = None
    if args.noise_apas:
        log("bench: noise-filter leg")
        noise = noise_leg(dev, napas=args.noise_apas, cpu=not args.no_cpu)
    adjoint = None
    if args.adjoint_frames > 0:
        adjoint = adjoint_leg(dev, args.adjoint_frames)
        log(f"bench: adjoint step {adjoint}")

    chargefit = None
    if args.chargefit_apas > 0 and len(planes) >= 6:
        chargefit = chargefit_leg(frames, planes, dev, napas=args.chargefit_apas)
        log(f"bench: charge fit {chargefit}")

    cpu = None
    if not args.no_cpu:
        n = args.cpu_threads or min(16, cpu_info()["usable"])
        log(f"bench: CPU baseline on {n} threads")
        cpu = cpu_baseline(planes, n)

    traffic = stage_traffic = None
    tr = hbm_traffic()
    if tr is not None:
        traffic = tr[0] * ch_ticks / tr[2]     # bytes per step
        stage_traffic = [b * ch_ticks / tr[2] for b in tr[1]]
    # the dominant kernel: the stage with the longest launch time
    alg = stage_alg_bytes(planes, kern)
    dom = max(range(3), key=lambda i: stage_ms[i])
    dom_gbs = alg[dom] / (stage_ms[dom] / 1e3) / 1e9

    res = {
        "metric": "deconvolved channel-ticks/s (batched vs plain, 1 GPU) at <=1e-4 rel RMS err",
        "value": value,
        "unit": "channel-ticks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded noise + forward-convolved straight tracks)",
        "config": {
            "workload": f"{args.workload}: " + (
                "150 APAs x (800+800+960) channels x 6000 ticks" if args.workload == "apa150"
                else f"{len(planes)} planes"),
            "planes": len(planes),
            "channel_ticks": ch_ticks,
            "chunk_apas": tuning["chunk_apas"],
            "tuning": tuning,
            "hipgraph": graph is not None,
            "spectral_cut": args.cut,
            "kept_bins": {f"{k}{nw}x{nt}": kern[(k, nw, nt)].kc for (k, nw, nt) in types},
        },
        "roofline": {
            "bound": "hbm",
            "achieved": dom_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": dom_gbs / HBM_PEAK_GBS,
            "traffic": None if stage_traffic is None else stage_traffic[dom],
            "kernel": STAGES[dom],
            "launch_ms": stage_ms[dom],
            "alg_bytes": alg[dom],
            "stages_ms": stage_ms,
            "stages_alg_bytes": alg,
            "stages_traffic": stage_traffic,
            "pipeline": {"achieved": achieved, "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                         "alg_bytes_per_channel_tick": B_ALG},
            "measured_copy_gbs": copy_gbs,
        },
        "accuracy_fp64": accuracy,
        "response_build_fp64": resp_build,
        "plain_gpu": plain,
        "stats_split": split,
        "adjoint_step": adjoint,
        "forward_sim": impact,
        "noise_filter": noise,
        "chargefit": chargefit,
        "batched_vs_plain": None if plain is None else plain["ratio"],
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)
    if not acc_ok:
        log("bench: FP64 accuracy above 1e-4 relative RMS")
        sys.exit(3)


if __name__ == "__main__":
    main()
