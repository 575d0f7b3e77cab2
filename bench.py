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

Prints one JSON line (the driver's contract) with the roofline of the
fused pipeline and the CPU baseline (the oracle's FP32 step-by-step
restatement of the reference path, one plane per thread) beside it.
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
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--plain-apas", type=int, default=4,
                    help="APAs timed through the plain per-plane path (0 = skip)")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner (use --chunk and --wire-tile, full row grid)")
    ap.add_argument("--wire-tile", type=int, default=0,
                    help="bins per wire-pass tile with --no-tune (0 = the plan's default)")
    ap.add_argument("--row-plan", type=int, nargs=3, default=(0, 0, 0), metavar=("PLAN", "R2C", "C2R"),
                    help="with --no-tune: row plan (0 runtime, 1 compile-time) and the r2c / c2r "
                         "workgroups per CU of the compile-time plan")
    ap.add_argument("--eager", action="store_true", help="time plan runs without the hipGraph")
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


def cpu_baseline(planes, nthreads):
    """The oracle's FP32 restatement of the reference CPU path (decon_2D_init,
    tight filter, restore_baseline, cal_RMS/threshold), one plane per
    thread, on a bounded sample of the workload's planes."""
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O
    from wct_amd import workload as W
    sample = [planes[i % len(planes)] for i in range(nthreads)]
    specs = {}
    for kind, nw, nt in set(sample):
        sp = W.plane_spec(kind, nw, nt)
        resp = O.overall_response(sp.fr_rows, sp.elec, W.FINE_PERIOD, W.TICK, W.INTER_GAIN,
                                  W.ADC_MV, W.FTOFFSET, nt, dtype=np.float32)
        specs[(kind, nw, nt)] = (sp, resp)
    rng = np.random.default_rng(7)
    inputs = [(rng.standard_normal((nw, nt)) * W.NOISE_ADC).astype(np.float32)
              for (kind, nw, nt) in sample]

    def one(i):
        sp, resp = specs[sample[i]]
        c = O.decon_init(inputs[i], resp, sp.wire_filter.astype(np.float32), sp.wire_shift,
                         sp.time_shift, np.float32)
        y = O.decon_apply_filter(c, sp.time_filter.astype(np.float32), np.float32)
        yb, _ = O.restore_baseline(y)
        O.roi_threshold(yb, W.TH_FACTOR)
        return y.size

    t0 = time.perf_counter()
    with ThreadPoolExecutor(nthreads) as ex:
        tot = sum(ex.map(one, range(nthreads)))
    dt = time.perf_counter() - t0
    cpu = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": tot / dt, "unit": "channel-ticks/s", "cores": nthreads, "kind": "port",
            "sample": f"{nthreads} planes ({'/'.join(f'{k}{nw}x{nt}' for k, nw, nt in sample[:3])}"
                      f"...) one per thread, {dt:.1f} s wall; own mixed-radix FFT (no FFTW); "
                      f"CPU: {cpu}"}


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
    """HBM bytes per channel-tick of the fused pipeline from the
    committed counter summary (FETCH_SIZE + WRITE_SIZE of the launches of
    one 150-APA step run as one plan, scaled by the calibration copy), as
    (bytes, channel-ticks), or None when absent."""
    path = os.path.join(REPO, "profiles", "decon_fused_apa150_hbm_traffic.md")
    try:
        vals = {}
        with open(path) as f:
            for line in f:
                if line.startswith("- ") and ":" in line:
                    k, v = line[2:].split(":", 1)
                    vals[k.strip()] = v.strip().split()[0]
        return float(vals["bytes_per_chunk"]), int(vals["channel_ticks_per_chunk"])
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


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; run with --gpus 1"}))
        sys.exit(2)
    import torch
    from wct_amd import decon, workload as W
    dev = torch.device("cuda:0")
    planes = W.config_planes(args.workload)
    types = sorted(set(planes))
    calibration_copy(dev)
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

    plain = None
    if args.plain_apas > 0:
        pv, pms, pn = plain_leg(planes, frames, outs, masks, rmss, kern, args.plain_apas)
        log(f"bench: plain path {pms:.2f} ms for {pn} planes -> {pv:.3e} channel-ticks/s")
        plain = {"value": pv, "unit": "channel-ticks/s", "ratio": value / pv,
                 "sample": f"{pn} planes ({pn // 3} APAs) of the same frames, one launch per "
                           f"reference step, {pms:.2f} ms per pass"}

    adjoint = None
    if args.adjoint_frames > 0:
        adjoint = adjoint_leg(dev, args.adjoint_frames)
        log(f"bench: adjoint step {adjoint}")

    cpu = None
    if not args.no_cpu:
        n = args.cpu_threads or min(16, os.cpu_count() or 1)
        log(f"bench: CPU baseline on {n} threads")
        cpu = cpu_baseline(planes, n)

    traffic = None
    tr = hbm_traffic()
    if tr is not None:
        traffic = tr[0] * ch_ticks / tr[1]     # bytes per step

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
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": "fused decon pipeline (row r2c, wire pass, row c2r+ROI) per step",
        },
        "plain_gpu": plain,
        "adjoint_step": adjoint,
        "batched_vs_plain": None if plain is None else plain["ratio"],
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
