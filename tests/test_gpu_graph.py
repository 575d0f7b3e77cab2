"""hipGraph replay of batched plans and the tuner's launch shapes.

A graph of two batches (uneven planes) is replayed several times with
eager work in between (outputs cleared, inputs changed) and every replay
must equal an eager run on the same inputs bit for bit; every tuning
candidate must give bit-identical outputs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PLANES = [("ind", 128, 2000), ("ind", 128, 2000), ("col", 96, 2000), ("ind", 200, 3000),
          ("col", 240, 3000), ("col", 240, 3000)]


def _setup(dev, planes=PLANES):
    import torch
    from wct_amd import decon as D, workload as W
    kern = {}
    for t in sorted(set(planes)):
        kern[t] = D.PlaneKernel(W.plane_spec(*t))
    g = torch.Generator(device=dev)
    g.manual_seed(11)
    adc = [torch.randn((nw, nt), generator=g, device=dev) * 1.5 for _, nw, nt in planes]
    out = [torch.empty_like(a) for a in adc]
    mask = [torch.empty(a.shape, dtype=torch.uint8, device=dev) for a in adc]
    rms = [torch.empty(a.shape[0], device=dev) for a in adc]

    def make_jobs(idx):
        return [D.make_job(kern[planes[i]], adc[i], out[i], mask=mask[i], rms=rms[i]) for i in idx]

    return kern, adc, out, mask, rms, make_jobs


def _snapshot(out, mask, rms):
    return [o.clone() for o in out], [m.clone() for m in mask], [r.clone() for r in rms]


def test_graph_replays_equal_eager(dev):
    import torch
    from wct_amd import decon as D
    print("graph replay vs eager, 6 uneven planes in 2 batches", flush=True)
    kern, adc, out, mask, rms, make_jobs = _setup(dev)
    batches = [D.Batch(make_jobs(range(0, 3))), D.Batch(make_jobs(range(3, 6)))]
    graph = D.Graph(batches)
    for rep in range(3):
        for a in adc:        # eager work between replays: new inputs
            a.mul_(1.0 + 0.25 * rep).add_(0.5)
        for b in batches:
            b.run()
        torch.cuda.synchronize()
        ref = _snapshot(out, mask, rms)
        for o in out:
            o.zero_()
        for m in mask:
            m.zero_()
        graph.launch()
        torch.cuda.synchronize()
        for x, y in zip(out, ref[0]):
            assert torch.equal(x, y)
        for x, y in zip(mask, ref[1]):
            assert torch.equal(x, y)
        for x, y in zip(rms, ref[2]):
            assert torch.equal(x, y)
    graph.close()
    for b in batches:
        b.close()


def test_tuning_candidates_bit_identical(dev):
    import torch
    from wct_amd import decon as D, tune as T
    print("tuning candidates give identical outputs", flush=True)
    kern, adc, out, mask, rms, make_jobs = _setup(dev)
    b = D.Batch(make_jobs(range(len(PLANES))))
    b.run()
    torch.cuda.synchronize()
    ref = _snapshot(out, mask, rms)
    b.close()
    for wt in (2, 4, 8, 16):
        for pct in (100, 50, 25):
            for o in out:
                o.zero_()
            b = D.Batch(make_jobs(range(len(PLANES))), wire_tile=wt, row_grid_pct=pct)
            b.run()
            torch.cuda.synchronize()
            b.close()
            for x, y in zip(out, ref[0]):
                assert torch.equal(x, y), (wt, pct)
            for x, y in zip(mask, ref[1]):
                assert torch.equal(x, y), (wt, pct)
    # the tuner over a small space returns a timed best candidate
    best, res = T.tune(make_jobs, len(PLANES), {"chunk_apas": (1, 2), "wire_tile": (8, 16),
                                                 "row_grid_pct": (100,)}, reps=1)
    assert best in res and isinstance(res[best], float)


def test_static_wire_lengths_bit_identical_across_tiles(dev):
    """800 and 960 wires take compile-time wire-pass plans at 2-, 4- and
    8-bin tiles: those give the same bits as the default (8), which the
    decon tests hold to the oracle.  16-bin tiles run the runtime plan,
    whose code the compiler contracts into FMAs differently: it agrees to
    FP32 rounding (relative RMS 1e-6; a mask bit may flip where a sample
    sits on the threshold)."""
    import torch
    from wct_amd import decon as D
    planes = [("ind", 800, 1200), ("col", 960, 1200)]
    print("static wire lengths, tiles 2/4/8/16", planes, flush=True)
    kern, adc, out, mask, rms, make_jobs = _setup(dev, planes)
    b = D.Batch(make_jobs(range(2)))
    b.run()
    torch.cuda.synchronize()
    ref = _snapshot(out, mask, rms)
    b.close()
    for wt in (2, 4, 16):
        for o in out:
            o.zero_()
        b = D.Batch(make_jobs(range(2)), wire_tile=wt)
        b.run()
        torch.cuda.synchronize()
        b.close()
        if wt != 16:
            for got, want in zip((out, mask, rms), ref):
                for x, y in zip(got, want):
                    assert torch.equal(x, y), wt
        else:
            for x, y in zip(out, ref[0]):
                assert (torch.linalg.norm(x - y) / torch.linalg.norm(y)).item() <= 1e-6
            for x, y in zip(mask, ref[1]):
                assert (x == y).float().mean().item() >= 0.9999
            for x, y in zip(rms, ref[2]):
                assert torch.allclose(x, y, rtol=1e-5, atol=0)


def test_compile_time_row_plan_agrees(dev):
    """Rows of 6000 ticks on the transforms compiled for 3000 points
    (row_plan 1, at each occupancy choice) against the runtime radix
    plan (row_plan -1): other pass orders and multiply-adds, so FP32
    rounding apart (relative RMS <= 1e-6; a mask bit may flip where a
    sample sits on the threshold)."""
    import torch
    from wct_amd import decon as D
    planes = [("ind", 800, 6000), ("col", 960, 6000), ("ind", 800, 8000)]
    print("compile-time row plan vs runtime plan", planes, flush=True)
    kern, adc, out, mask, rms, make_jobs = _setup(dev, planes)
    b = D.Batch(make_jobs(range(3)), row_plan=-1)
    b.run()
    torch.cuda.synchronize()
    ref = _snapshot(out, mask, rms)
    b.close()
    for r2c, c2r in ((3, 2), (4, 3), (5, 4), (0, 0)):
        for o in out:
            o.zero_()
        b = D.Batch(make_jobs(range(3)), row_plan=1, r2c_blocks=r2c, c2r_blocks=c2r)
        b.run()
        torch.cuda.synchronize()
        b.close()
        for x, y in zip(out, ref[0]):
            assert (torch.linalg.norm(x - y) / torch.linalg.norm(y)).item() <= 1e-6, (r2c, c2r)
        for x, y in zip(mask, ref[1]):
            assert (x == y).float().mean().item() >= 0.9999
        for x, y in zip(rms, ref[2]):
            assert torch.allclose(x, y, rtol=1e-5, atol=0)
        # the 8000-tick plane has no compile-time plan: the same bits
        assert torch.equal(out[2], ref[0][2])


def test_bad_tuning_is_refused(dev):
    from wct_amd import decon as D
    from wct_amd.capi import WctError
    kern, adc, out, mask, rms, make_jobs = _setup(dev)
    with pytest.raises(WctError, match="wire tile"):
        D.Batch(make_jobs(range(2)), wire_tile=32)
    with pytest.raises(WctError, match="row grid"):
        D.Batch(make_jobs(range(2)), row_grid_pct=5)
    with pytest.raises(WctError, match="r2c blocks"):
        D.Batch(make_jobs(range(2)), row_plan=-1, r2c_blocks=3)
    with pytest.raises(WctError, match="r2c blocks"):
        D.Batch(make_jobs(range(2)), r2c_blocks=2)
    with pytest.raises(WctError, match="c2r blocks"):
        D.Batch(make_jobs(range(2)), row_plan=1, c2r_blocks=5)
    with pytest.raises(WctError, match="row plan"):
        D.Batch(make_jobs(range(2)), row_plan=2)
