"""Autotuner for the launch knobs of the overlap pipeline.

Knobs (flyehip_set_tuning): the DP lookback tile width, i.e. lanes that
chain one target group (1, 4, 8, 16, 32, 64: G lanes score G predecessors
per step, 64 / G groups share a wave), the number of query records per GPU
batch, and the batches in flight (streams the batches are dealt to).  The
tuner times the pipeline on a sample of the workload for every candidate of
the (batch, lanes) grid at two batches in flight, then the batches in flight
at the winner; it keeps the fastest and checks that every candidate produced
the same rows (the knobs change scheduling only, never results).
"""
from __future__ import annotations

import json
import time
from pathlib import Path

import numpy as np

from .overlap import FlyeHip, OverlapConfig

DP_LANES = (1, 4, 8, 16, 32, 64)


def tune(ctx: FlyeHip, query_ids: np.ndarray, config: OverlapConfig, batches=(1024, 4096),
         lanes=DP_LANES, flights=(1, 2, 3), repeats: int = 1, log=None) -> dict:
    results = []
    reference = [None]

    def run(g, qpb, fl):
        ctx.set_tuning(dp_lanes_per_group=g, queries_per_batch=qpb, batches_in_flight=fl)
        ctx.get_seq_overlaps(query_ids, config)  # warm buffers
        best = None
        for _ in range(repeats):
            t = time.perf_counter()
            rows, roff = ctx.get_seq_overlaps(query_ids, config)
            dt = time.perf_counter() - t
            best = dt if best is None else min(best, dt)
        st = ctx.stats()
        if reference[0] is None:
            reference[0] = rows.tobytes()
        elif rows.tobytes() != reference[0]:
            raise RuntimeError(f"tuning candidate lanes={g} batch={qpb} flights={fl} changed the results")
        r = {"dp_lanes": g, "queries_per_batch": qpb, "batches_in_flight": fl, "s": best, "ms_chain": st["ms_chain"]}
        results.append(r)
        if log:
            log(f"tune lanes={g} batch={qpb} flights={fl}: {best * 1e3:.1f} ms (dp {st['ms_chain']:.1f} ms)")
        return r

    stage = [run(g, qpb, 2) for qpb in batches for g in lanes]
    win = min(stage, key=lambda r: r["s"])
    for fl in flights:
        if fl != 2:
            run(win["dp_lanes"], win["queries_per_batch"], fl)
    win = min(results, key=lambda r: r["s"])
    ctx.set_tuning(dp_lanes_per_group=win["dp_lanes"], queries_per_batch=win["queries_per_batch"],
                   batches_in_flight=win["batches_in_flight"])
    return {"best": win, "candidates": results}


def save(result: dict, path: Path) -> None:
    path.parent.mkdir(parents=True, exist_ok=True)
    path.write_text(json.dumps(result, indent=2) + "\n")
