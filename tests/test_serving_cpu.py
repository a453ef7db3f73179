"""Host logic of the serving loop without a GPU: iterations take reads in
arrival order up to the per-iteration budget, every request gets its reads'
rows back in order, several contexts run iterations side by side, admission
follows the library's measured device bytes, and the report's per-iteration
time comes from the library's own time per call."""
import threading
import time

import numpy as np

from cuflye_amd import native as N
from cuflye_amd.overlap import OverlapConfig
from cuflye_amd.serve import OverlapServer, Request


class FakeCtx:
    """Stands in for FlyeHip: one row per query id, 10 ms per call (reported
    through stats() as library time, as flyehip_get_stats does), and a
    device-memory high-water mark of `per_base` bytes per base of the largest
    call above a resident index of `index_bytes`."""

    def __init__(self, per_base=100.0, index_bytes=1000, sleep=0.0):
        self.calls = []
        self.bases = []
        self.per_base = per_base
        self.sleep = sleep
        self.peak = index_bytes
        self.index_info = N.IndexInfo()
        self.index_info.device_bytes = index_bytes
        self.active = 0
        self.max_active = 0
        self._mu = threading.Lock()

    def get_seq_overlaps_external(self, bases, offs, ids, cfg):
        with self._mu:
            self.active += 1
            self.max_active = max(self.max_active, self.active)
        if self.sleep:
            time.sleep(self.sleep)
        self.calls.append(len(ids))
        self.bases.append(int(offs[-1]))
        self.peak = max(self.peak, self.index_info.device_bytes + int(self.per_base * offs[-1]))
        rows = np.zeros(len(ids), N.OVERLAP_DTYPE)
        rows["cur_id"] = ids
        rows["cur_len"] = np.diff(offs)
        with self._mu:
            self.active -= 1
        return rows, np.arange(len(ids) + 1, dtype=np.int64)

    def stats(self):
        return {"ms_total": 10.0, "device_bytes_peak": self.peak}


def _request(rid, n, first_id, rng):
    lens = rng.integers(50, 200, size=n)
    offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    bases = rng.choice(np.frombuffer(b"ACGT", np.uint8), size=int(offs[-1]))
    return Request(rid, bases, offs, np.arange(first_id, first_id + n, dtype=np.int64))


def test_serving_iterations_rows_and_report():
    rng = np.random.default_rng(5)
    ctx = FakeCtx()
    srv = OverlapServer(ctx, OverlapConfig(), max_iteration_reads=16)
    reqs = [_request(0, 40, 1000, rng), _request(1, 7, 2000, rng), _request(2, 25, 3000, rng)]
    for r in reqs:
        srv.submit(r)
    srv.start()
    for r in reqs:
        assert r.done_event.wait(30)
    srv.stop()
    # 72 reads at up to 16 per iteration
    assert sum(ctx.calls) == 72 and max(ctx.calls) <= 16 and len(ctx.calls) == srv.iterations >= 5
    for r in reqs:
        got = np.concatenate(r.rows)
        assert np.array_equal(got["cur_id"], r.query_ids)
        assert np.array_equal(got["cur_len"], np.diff(r.offsets))
    rep = srv.latency_report()
    assert rep["requests"] == 3 and rep["reads"] == 72
    assert rep["iteration_ms_mean"] == 10.0
    assert abs(rep["iteration_reads_mean"] - 72 / srv.iterations) < 1e-9
    assert rep["capacity_reads_per_s"] > 0 and rep["busy_s"] > 0
    assert rep["p50_ms"] <= rep["p99_ms"] <= rep["max_ms"]
    # the admission estimate is the measured bytes per base of the largest call
    assert abs(rep["admission_bytes_per_base"] - 100.0) < 1e-9


def test_serving_contexts_run_side_by_side():
    rng = np.random.default_rng(6)
    ctxs = [FakeCtx(sleep=0.02), FakeCtx(sleep=0.02)]
    srv = OverlapServer(ctxs, OverlapConfig(), max_iteration_reads=8)
    reqs = [_request(k, 30, 1000 * (k + 1), rng) for k in range(4)]
    for r in reqs:
        srv.submit(r)
    srv.start()
    for r in reqs:
        assert r.done_event.wait(30)
    srv.stop()
    calls = ctxs[0].calls + ctxs[1].calls
    assert sum(calls) == 120 and max(calls) <= 8
    assert ctxs[0].calls and ctxs[1].calls
    for r in reqs:
        got = np.concatenate(r.rows)
        assert np.array_equal(got["cur_id"], r.query_ids)
    rep = srv.latency_report()
    assert rep["iterations_in_flight"] == 2 and rep["requests"] == 4
    # two iterations overlapped in wall clock: busy time below the sum of spans
    spans = sum(b - a for a, b in srv.iteration_spans)
    assert rep["busy_s"] < spans


def test_serving_admission_limits_iteration_bytes():
    rng = np.random.default_rng(7)
    ctx = FakeCtx(per_base=100.0)
    # reads of 50-199 bases at 100 B/base measured: 20 kB admits at most
    # 200 bases, i.e. at most 4 reads per iteration once measured
    srv = OverlapServer(ctx, OverlapConfig(), admission_bytes=20_000, max_iteration_reads=64)
    srv.bytes_per_base = 100.0
    req = _request(0, 30, 1, rng)
    srv.submit(req)
    srv.start()
    assert req.done_event.wait(30)
    srv.stop()
    assert sum(ctx.calls) == 30
    # every call fits the budget, except a lone read that alone exceeds it
    for n, b in zip(ctx.calls, ctx.bases):
        assert b * 100.0 <= 20_000 or n == 1
