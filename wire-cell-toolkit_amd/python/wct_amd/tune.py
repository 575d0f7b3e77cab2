"""Autotuner of the batched deconvolution's launch shapes.

For a workload (a list of plane jobs over resident frames) it times each
candidate (APAs per launch group, wire-pass tile width, persistent grid
fraction of the row passes) on the device, replaying the whole step as
one hipGraph, and returns the fastest with every measurement.  The
candidates are launch shapes only: every one gives bit-identical
outputs (the per-element arithmetic of the transforms does not depend on
how rows and tiles are spread over workgroups).
"""
import itertools
import time

from . import decon as D

# A half-size persistent row grid measured 50-60 % slower than the full one
# on the 150-APA workload, so the default space keeps the full grid and
# spends the candidates on plan size and wire tile instead.
DEFAULT_SPACE = {"chunk_apas": (8, 16, 32, 64, 150), "wire_tile": (2, 4, 8, 16), "row_grid_pct": (100,)}


def build_step(make_jobs, nplanes, chunk_apas, wire_tile, row_grid_pct, planes_per_apa=3):
    """Batches of chunk_apas APAs sharing one workspace, and their graph."""
    per = planes_per_apa * chunk_apas
    batches, ws = [], None
    for c0 in range(0, nplanes, per):
        jobs = make_jobs(range(c0, min(c0 + per, nplanes)))
        b = D.Batch(jobs, workspace=ws, wire_tile=wire_tile, row_grid_pct=row_grid_pct)
        ws = b.ws
        batches.append(b)
    return batches


def time_graph(graph, reps):
    import torch
    stream = torch.cuda.current_stream()
    graph.launch()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        graph.launch()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def tune(make_jobs, nplanes, space=None, reps=2, log=None):
    """Returns (best_config, results): results maps each candidate
    (chunk_apas, wire_tile, row_grid_pct) to ms per step, or to the
    error string of a candidate the device refused."""
    space = space or DEFAULT_SPACE
    results = {}
    best, best_ms = None, float("inf")
    for cfg in itertools.product(space["chunk_apas"], space["wire_tile"], space["row_grid_pct"]):
        batches = graph = None
        try:
            batches = build_step(make_jobs, nplanes, *cfg)
            graph = D.Graph(batches)
            ms = time_graph(graph, reps)
        except Exception as e:      # a shape the device cannot take (LDS, wire count)
            results[cfg] = str(e)
            if log:
                log(f"tune: {cfg} refused: {e}")
            continue
        finally:
            if graph is not None:
                graph.close()
            for b in batches or []:
                b.close()
        results[cfg] = ms
        if log:
            log(f"tune: chunk {cfg[0]} APAs, wire tile {cfg[1]}, row grid {cfg[2]}%: {ms:.3f} ms")
        if ms < best_ms:
            best, best_ms = cfg, ms
    return best, results
