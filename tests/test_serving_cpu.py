"""Host logic of the serving loop without a GPU: iterations take reads in
arrival order up to the per-iteration budget, every request gets its reads'
rows back in order, and the report's per-iteration time and capacity come
from the library's own time per call."""
import numpy as np

from cuflye_amd import native as N
from cuflye_amd.overlap import OverlapConfig
from cuflye_amd.serve import OverlapServer, Request


class FakeCtx:
    """Stands in for FlyeHip: one row per query id, 10 ms of library time per
    call (reported through stats(), as flyehip_get_stats does)."""

    def __init__(self):
        self.calls = []

    def get_seq_overlaps_external(self, bases, offs, ids, cfg):
        self.calls.append(len(ids))
        rows = np.zeros(len(ids), N.OVERLAP_DTYPE)
        rows["cur_id"] = ids
        rows["cur_len"] = np.diff(offs)
        return rows, np.arange(len(ids) + 1, dtype=np.int64)

    def stats(self):
        return {"ms_total": 10.0}


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
    # capacity = reads per second of library time
    assert abs(rep["capacity_reads_per_s"] - 72 / (srv.iterations * 0.010)) < 1e-6
    assert rep["p50_ms"] <= rep["p99_ms"] <= rep["max_ms"]
