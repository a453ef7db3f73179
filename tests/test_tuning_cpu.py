"""Host logic of the launch tuning and of the benchmark's reporting, without a
GPU: the autotuner's search order and default sizes (with the equal splits),
its memory-budget skipping and its check
that every candidate returns the same rows; the bench's refusal of several
GPUs and its lookup of the committed HBM-traffic table."""
import json
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from cuflye_amd import autotune
from cuflye_amd.overlap import FlyeHipError, OverlapConfig

ROOT = Path(__file__).resolve().parent.parent


class FakeCtx:
    """Stands in for FlyeHip: refuses batch x flights above a memory budget
    with FLYEHIP_ERR_BUDGET (3) and returns fixed rows otherwise."""

    def __init__(self, budget, rows=None, bad=None):
        self.budget = budget
        self.tuning = None
        self.calls = []
        self.rows = rows if rows is not None else np.arange(12, dtype=np.int32)
        self.bad = bad  # (lanes, batch, flights) whose rows differ

    def set_tuning(self, dp_lanes_per_group=0, queries_per_batch=0, batches_in_flight=0):
        self.tuning = (dp_lanes_per_group, queries_per_batch, batches_in_flight)

    def get_seq_overlaps(self, ids, cfg):
        g, qpb, fl = self.tuning
        self.calls.append(self.tuning)
        if qpb * fl > self.budget:
            raise FlyeHipError(3, "device memory budget exceeded")
        rows = self.rows.copy()
        if self.bad == self.tuning:
            rows[0] += 1
        return rows, np.array([0, len(rows)], np.int64)

    def stats(self):
        return {"ms_chain": 1.0}


def test_autotune_skips_refused_sizes_and_keeps_a_fitting_winner():
    ctx = FakeCtx(budget=4096 * 2)
    res = autotune.tune(ctx, np.arange(10), OverlapConfig(), batches=(2048, 4096, 8192), lanes=(8, 16),
                        flights=(2, 3))
    refused = {(r["queries_per_batch"], r["batches_in_flight"]) for r in res["refused"]}
    assert (4096, 3) in refused
    # once 4096 x 3 is refused, 8192 at 3 flights is not tried; 8192 x 2 is refused on its own
    assert (8192, 3) not in {(t[1], t[2]) for t in ctx.calls}
    assert (8192, 2) in refused
    best = res["best"]
    assert best["queries_per_batch"] * best["batches_in_flight"] <= ctx.budget
    assert ctx.tuning == (best["dp_lanes"], best["queries_per_batch"], best["batches_in_flight"])
    # the lanes sweep runs at the winner only
    assert {c["dp_lanes"] for c in res["candidates"]} == {8, 16}


def test_autotune_rejects_a_candidate_that_changes_results():
    ctx = FakeCtx(budget=1 << 30, bad=(8, 4096, 2))
    with pytest.raises(RuntimeError, match="changed the results"):
        autotune.tune(ctx, np.arange(10), OverlapConfig(), batches=(2048, 4096), lanes=(8,), flights=(2,))


def test_autotune_default_sizes_add_equal_splits():
    # 20,418 query records: the fixed sizes plus one batch per stream
    # (ceil(20418 / 2) = 10209, ceil(20418 / 3) = 6806), in growing order
    ctx = FakeCtx(budget=1 << 30)
    res = autotune.tune(ctx, np.arange(20418), OverlapConfig(), lanes=(8,), flights=(2, 3))
    sizes = [c["queries_per_batch"] for c in res["candidates"]]
    assert sorted(set(sizes)) == sorted(set(autotune.BATCHES) | {10209, 6806})
    assert sizes == sorted(sizes)
    # an explicit list is taken as it is
    ctx = FakeCtx(budget=1 << 30)
    res = autotune.tune(ctx, np.arange(20418), OverlapConfig(), batches=(4096,), lanes=(8,), flights=(2, 3))
    assert {c["queries_per_batch"] for c in res["candidates"]} == {4096}


def test_autotune_raises_when_nothing_fits():
    ctx = FakeCtx(budget=1)
    with pytest.raises(RuntimeError, match="no tuning candidate fits"):
        autotune.tune(ctx, np.arange(10), OverlapConfig(), batches=(2048,), lanes=(8,), flights=(2,))


def test_bench_refuses_several_gpus():
    out = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 2
    msg = json.loads(out.stdout.strip().splitlines()[-1])
    assert msg["n_gpus"] == 2 and "one GPU" in msg["error"]


def test_bench_traffic_table_matches_its_tuning_only():
    import bench
    table = json.loads((ROOT / "profiles" / "pmc_traffic.json").read_text())
    t = table["measured_with"]["tuning"]
    fam = next(iter(table["families"]))
    got = bench.pmc_traffic(fam, dict(t))
    assert got == pytest.approx(table["families"][fam]["hbm_bytes_per_launch"])
    other = dict(t, queries_per_batch=t["queries_per_batch"] + 1)
    assert bench.pmc_traffic(fam, other) is None
    # fetch and write parts add up, scaled by the copy probe
    f = table["families"][fam]
    assert f["hbm_bytes_per_launch"] == pytest.approx(f["fetch_bytes_per_launch"] + f["write_bytes_per_launch"])
    assert table["copy_probe"]["write_factor"] == pytest.approx(
        table["copy_probe"]["bytes"] / 1024 / table["copy_probe"]["write_kib"])
