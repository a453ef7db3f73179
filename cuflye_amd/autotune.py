"""Autotuner for the launch knobs of the overlap pipeline.

Knobs (flyehip_set_tuning): the DP lookback tile width, i.e. lanes that
chain one target group (1, 4, 8, 16, 32, 64: G lanes score G predecessors
per step, 64 / G groups share a wave), and the number of query records per
GPU batch.  The tuner times the pipeline on a sample of the workload for
every candidate, keeps the fastest, and checks that every candidate produced
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
         lanes=DP_LANES, repeats: int = 1, log=None) -> dict:
    results = []
    reference = None
    for qpb in batches:
        for g in lanes:
            ctx.set_tuning(dp_lanes_per_group=g, queries_per_batch=qpb)
            ctx.get_seq_overlaps(query_ids, config)  # warm buffers
            best = None
            for _ in range(repeats):
                t = time.perf_counter()
                rows, roff = ctx.get_seq_overlaps(query_ids, config)
                dt = time.perf_counter() - t
                best = dt if best is None else min(best, dt)
            st = ctx.stats()
            if reference is None:
                reference = rows.tobytes()
            elif rows.tobytes() != reference:
                raise RuntimeError(f"tuning candidate lanes={g} batch={qpb} changed the results")
            results.append({"dp_lanes": g, "queries_per_batch": qpb, "s": best, "ms_chain": st["ms_chain"]})
            if log:
                log(f"tune lanes={g} batch={qpb}: {best * 1e3:.1f} ms (dp {st['ms_chain']:.1f} ms)")
    win = min(results, key=lambda r: r["s"])
    ctx.set_tuning(dp_lanes_per_group=win["dp_lanes"], queries_per_batch=win["queries_per_batch"])
    return {"best": win, "candidates": results}


def save(result: dict, path: Path) -> None:
    path.parent.mkdir(parents=True, exist_ok=True)
    path.write_text(json.dumps(result, indent=2) + "\n")
