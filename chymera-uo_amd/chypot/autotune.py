"""Autotuner for the per-grid launch shapes of the potential solve.

For a grid class (its JMAX x KMAX x LMAX) it times, on the GPU, every
candidate of each tunable the library exposes (chy_grid_set_tune) and keeps
the fastest:

  * cr_lanes        lanes per r system of the cyclic reduction, 64 or 32
                    (the plain FP64 solve), timed on batched launches of the
                    largest batch size asked for;
  * refine_fused    the largest number of fields per launch for which the
                    refined solve runs all mode work of a plane in one kernel
                    (the binary32 plane in LDS) instead of one kernel per
                    stage: both forms are timed at every batch size asked
                    for, and the threshold is the largest size up to which
                    the one-kernel form was the faster at every size.

Timings use synthetic fields of the grid's shape, resident on the GPU, and
HIP events around whole batched calls on the solve stream, after warm-up
calls, the calls queued back to back.  Every candidate computes the same solve (the choices change rounding
only), so tuning never changes what a caller gets beyond that.

The result is a table {"JxKxL": {"cr_lanes": .., "refine_fused": ..,
"timings_ms": {...}}} that apply() puts on prepared grids and save()/load()
keep as JSON.  chypot.serve.Server applies a table to the grids it prepares.
measured_table() is the table this tuner made on an MI355X for 30 grid
classes.
"""
from __future__ import annotations

import json

import numpy as np

from . import MODES_CR, Context, PreparedGrid, fields


def grid_key(g) -> str:
    return f"{g.jmax}x{g.kmax}x{g.lmax}"


def fused_threshold(times_fused: dict, times_stages: dict) -> int:
    """The largest batch size b such that the one-kernel form was faster at
    every timed size <= b (0 if it lost at the smallest)."""
    best = 0
    for b in sorted(times_fused):
        if b in times_stages and times_fused[b] < times_stages[b]:
            best = b
        else:
            break
    return best


def _time_batch(ctx, jobs, steps, reps, warmup, stream):
    """Median milliseconds of a batched call on the GPU: events around each
    call on the stream the context issues on (the batch joins back to it),
    the calls queued back to back with one wait at the end, as a stream of
    batches runs.  (A wait after every call would add each call's launch
    time to its span and favour the form with fewer launches.)"""
    import torch
    for _ in range(warmup):
        ctx.solve_batch(jobs, refine_steps=steps)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    for t0, t1 in ev:
        t0.record(stream)
        ctx.solve_batch(jobs, refine_steps=steps)
        t1.record(stream)
    ev[-1][1].synchronize()
    return float(np.median([t0.elapsed_time(t1) for t0, t1 in ev]))


def tune_grid(ctx: Context, g, batch_sizes=(1, 32), refine_steps: int = 2, reps: int = 5, warmup: int = 2):
    """Time the candidates on grid g (a chypot.fields.Grid); returns the
    table entry for its class.  The context issues on a stream of the
    tuner's own meanwhile and on its own stream afterwards."""
    import torch
    stream = torch.cuda.Stream()
    ctx.set_stream(stream)
    nmax = max(batch_sizes)
    grids = [PreparedGrid(ctx, g) for _ in range(nmax)]
    jobs = []
    for i, pg in enumerate(grids):
        rho = fields.torus_device(g, 0.45, 0.25, 1.5, 0.05, 5000 + i)
        jobs.append(dict(grid=pg, rho=rho, phi=torch.zeros(pg.field_shape, dtype=torch.float64, device="cuda")))
    torch.cuda.synchronize()  # the fields were made on torch's stream
    entry = {"cr_lanes": grids[0].cr_lanes, "refine_fused": grids[0].refine_fused, "timings_ms": {}}
    try:
        # lanes per r system of the cyclic reduction (plain solve)
        if grids[0].mode_solver == MODES_CR:
            lane_ms = {}
            for lanes in (64, 32):
                try:
                    for pg in grids:
                        pg.cr_lanes = lanes
                except Exception:
                    continue  # not available on this grid
                lane_ms[lanes] = _time_batch(ctx, jobs, 0, reps, warmup, stream)
            entry["timings_ms"]["cr_lanes"] = {str(k): v for k, v in lane_ms.items()}
            entry["cr_lanes"] = min(lane_ms, key=lane_ms.get)
            for pg in grids:
                pg.cr_lanes = entry["cr_lanes"]
        # one kernel per plane or one per stage, per batch size (refined solve)
        if g.kmax <= 128 and refine_steps in (1, 2):
            tf, ts = {}, {}
            for b in sorted(batch_sizes):
                sub = jobs[:b]
                for fused, out in ((True, tf), (False, ts)):
                    for j in sub:
                        j["grid"].refine_fused = fused
                    out[b] = _time_batch(ctx, sub, refine_steps, reps, warmup, stream)
            entry["timings_ms"]["refine_fused"] = {str(b): {"one_kernel": tf[b], "per_stage": ts[b]} for b in tf}
            entry["refine_fused"] = fused_threshold(tf, ts)
    finally:
        stream.synchronize()
        for pg in grids:
            pg.close()
        ctx.set_stream(None)
    return entry


def tune(ctx: Context, grids, batch_sizes=(1, 32), refine_steps: int = 2, reps: int = 5):
    """A table for several grids (chypot.fields.Grid objects)."""
    return {grid_key(g): tune_grid(ctx, g, batch_sizes, refine_steps, reps) for g in grids}


def apply(pg: PreparedGrid, table: dict) -> bool:
    """Put the table's choices for pg's class on pg; False if the table has
    no entry for it."""
    e = table.get(grid_key(pg.grid))
    if e is None:
        return False
    pg.cr_lanes = int(e["cr_lanes"])
    pg.refine_fused = int(e["refine_fused"])
    return True


def measured_table() -> dict:
    """The table this tuner produced on an MI355X (tuned_mi355x.json beside
    this module), refined solves of 2 steps, calls queued back to back: nine
    grid classes with batch sizes 1, 8, 32 (1, 8 for JMAX = 512), medians of
    3 calls; and the other
    21 classes of the heterogeneous set JMAX in {64,128,256}, KMAX in
    {32,64,128}, LMAX in {64,128,256} with batch sizes 1 and 4, medians of 5
    calls."""
    import pathlib
    return load(pathlib.Path(__file__).with_name("tuned_mi355x.json"))


def save(table: dict, path) -> None:
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)


def load(path) -> dict:
    with open(path) as f:
        return json.load(f)
