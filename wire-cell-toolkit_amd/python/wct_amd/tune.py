"""Autotuner of the batched deconvolution's launch shapes and row plans.

For a workload (a list of plane jobs over resident frames) it times
candidates on the device, replaying the whole step as one hipGraph, and
returns the fastest with every measurement.  A candidate is
(APAs per launch group, wire-pass tile width, persistent grid fraction
of the row passes, row plan, r2c workgroups per CU, c2r workgroups per
CU): the row plan chooses between the runtime radix plan (-1) and the
transforms compiled for 3000-point rows (radix 15, 20, 10 on an unpadded
row, 1), the latter at its occupancy choices.  Launch shapes give bit-identical
outputs; the two row plans order their multiply-adds differently, so
they agree to FP32 rounding.

The search is staged: the row-plan candidates at the whole workload as
one group and the default wire tile first, then group size and wire
tile with the best row plan.
"""
import itertools

from . import decon as D

# A half-size persistent row grid measured 50-60 % slower than the full one
# on the 150-APA workload, so the default space keeps the full grid and
# spends the candidates on plan size, wire tile and row plan instead.
ROW_PLANS = ((-1, 0, 0), (1, 4, 3), (1, 3, 3), (1, 5, 3), (1, 4, 2))
DEFAULT_SPACE = {"chunk_apas": (8, 16, 32, 64, 150), "wire_tile": (2, 4, 8, 16), "row_grid_pct": (100,),
                 "row_plan": ROW_PLANS}


def build_step(make_jobs, nplanes, chunk_apas, wire_tile, row_grid_pct, row_plan=0, r2c_blocks=0, c2r_blocks=0,
               planes_per_apa=3):
    """Batches of chunk_apas APAs sharing one workspace."""
    per = planes_per_apa * chunk_apas
    batches, ws = [], None
    for c0 in range(0, nplanes, per):
        jobs = make_jobs(range(c0, min(c0 + per, nplanes)))
        b = D.Batch(jobs, workspace=ws, wire_tile=wire_tile, row_grid_pct=row_grid_pct, row_plan=row_plan,
                    r2c_blocks=r2c_blocks, c2r_blocks=c2r_blocks)
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


def _time(make_jobs, nplanes, cfg, reps):
    batches = graph = None
    try:
        batches = build_step(make_jobs, nplanes, *cfg)
        graph = D.Graph(batches)
        return time_graph(graph, reps)
    finally:
        if graph is not None:
            graph.close()
        for b in batches or []:
            b.close()


def tune(make_jobs, nplanes, space=None, reps=2, log=None, planes_per_apa=3):
    """Returns (best_config, results): results maps each candidate
    (chunk_apas, wire_tile, row_grid_pct, row_plan, r2c_blocks,
    c2r_blocks) to ms per step, or to the error string of a candidate
    the device refused."""
    space = space or DEFAULT_SPACE
    results = {}
    napas = max(1, nplanes // planes_per_apa)
    whole = min([c for c in space["chunk_apas"] if c >= napas] or [max(space["chunk_apas"])])
    tile0 = 4 if 4 in space["wire_tile"] else space["wire_tile"][0]
    grid0 = space["row_grid_pct"][0]

    def run(cfg):
        if cfg in results:
            return results[cfg]
        try:
            ms = _time(make_jobs, nplanes, cfg, reps)
        except Exception as e:      # a shape the device cannot take (LDS, wire count)
            results[cfg] = str(e)
            if log:
                log(f"tune: {cfg} refused: {e}")
            return None
        results[cfg] = ms
        if log:
            log(f"tune: chunk {cfg[0]} APAs, wire tile {cfg[1]}, row grid {cfg[2]}%, row plan {cfg[3:]}: "
                f"{ms:.3f} ms")
        return ms

    # stage 1: row plans
    for rp in space.get("row_plan", ((1, 0, 0),)):
        run((whole, tile0, grid0) + tuple(rp))
    ok = {c: v for c, v in results.items() if isinstance(v, float)}
    best_rp = min(ok, key=ok.get)[3:] if ok else (1, 0, 0)
    # stage 2: group size, wire tile and grid with that row plan
    for chunk, tile, grid in itertools.product(space["chunk_apas"], space["wire_tile"], space["row_grid_pct"]):
        run((chunk, tile, grid) + tuple(best_rp))
    ok = {c: v for c, v in results.items() if isinstance(v, float)}
    best = min(ok, key=ok.get) if ok else None
    return best, results
