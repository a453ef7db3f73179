"""Autotuner for the launch knobs of the overlap pipeline.

Knobs (flyehip_set_tuning): the DP lookback tile width, i.e. lanes that
chain one target group (1, 4, 8, 16, 32, 64: G lanes score G predecessors
per step, 64 / G groups share a wave), the number of query records per GPU
batch, and the batches in flight (streams the batches are dealt to).  Large
batches pay the fixed per-batch costs (host round trips of the sort levels,
the serial heap-sort phase) fewer times, but every batch in flight holds its
own buffers, so the device memory budget bounds batch x flights.  The tuner
times the whole workload for every (batch, flights) candidate in order of
growing memory, skipping what the budget refuses, then the DP lanes at the
winner; it keeps the fastest and checks that every candidate produced the
same rows (the knobs change scheduling only, never results).
"""
from __future__ import annotations

import json
import time
from pathlib import Path

import numpy as np

from .overlap import FlyeHip, FlyeHipError, OverlapConfig

DP_LANES = (1, 4, 8, 16, 32, 64)
ERR_BUDGET = 3  # FLYEHIP_ERR_BUDGET


BATCHES = (2048, 3072, 4096, 5120, 6144)


def tune(ctx: FlyeHip, query_ids: np.ndarray, config: OverlapConfig, batches=None,
         lanes=(8, 4, 16, 32, 64), flights=(2, 3), repeats: int = 1, log=None) -> dict:
    """batches=None tries BATCHES plus, for every flight count, the run cut
    into that many equal batches (one batch per stream, no lone last batch)."""
    if batches is None:
        n = len(query_ids)
        batches = tuple(sorted(set(BATCHES) | {-(-n // fl) for fl in flights if n > 0}))
    results = []
    refused = []
    reference = [None]

    def run(g, qpb, fl):
        ctx.set_tuning(dp_lanes_per_group=g, queries_per_batch=qpb, batches_in_flight=fl)
        try:
            ctx.get_seq_overlaps(query_ids, config)  # warm buffers
        except FlyeHipError as e:
            if e.code != ERR_BUDGET:
                raise
            refused.append({"dp_lanes": g, "queries_per_batch": qpb, "batches_in_flight": fl})
            if log:
                log(f"tune lanes={g} batch={qpb} flights={fl}: over the memory budget, skipped")
            return None
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

    # (batch, flights) in order of growing memory; once a size is refused,
    # larger batches at as many flights are not tried
    cap = {}
    for qpb in sorted(batches):
        for fl in sorted(flights):
            if any(qpb >= b and fl >= f for f, b in cap.items()):
                continue
            if run(lanes[0], qpb, fl) is None:
                cap[fl] = qpb
    if not results:
        raise RuntimeError("no tuning candidate fits the device memory budget")
    win = min(results, key=lambda r: r["s"])
    for g in lanes[1:]:
        run(g, win["queries_per_batch"], win["batches_in_flight"])
    win = min(results, key=lambda r: r["s"])
    ctx.set_tuning(dp_lanes_per_group=win["dp_lanes"], queries_per_batch=win["queries_per_batch"],
                   batches_in_flight=win["batches_in_flight"])
    return {"best": win, "candidates": results, "refused": refused}


def save(result: dict, path: Path) -> None:
    path.parent.mkdir(parents=True, exist_ok=True)
    path.write_text(json.dumps(result, indent=2) + "\n")
