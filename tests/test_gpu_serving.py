"""Serving loop: continuous batching over a resident genome index; each
request's overlaps are identical alone and mixed with concurrent requests,
and equal to the oracle's."""
import threading
import time

import numpy as np
import pytest

import cuflye_amd.synth as S
from cuflye_amd.overlap import OverlapConfig
from cuflye_amd.serve import OverlapServer, Request, request_rows

pytestmark = pytest.mark.gpu


def _contigs(genome, n):
    step = len(genome) // n
    offs = np.array([i * step for i in range(n)] + [len(genome)], np.int64)
    return offs


def test_serving_alone_equals_mixed(gpu, oracle_mod):
    genome = S.genome_with_repeats(2_000_000, 0.42, 0.08, 2000, 8000, 2, 50, 0.005, 0.05, seed=51)
    offs = _contigs(genome, 4)
    gpu.store_reads(genome, offs)
    gpu.build_index(15, solid_min=1, rep_max=64)
    cfg = OverlapConfig(kmer_size=15, minimum_overlap=1000)
    reqs_data = []
    for r in range(5):
        b, o = S.reads_from(genome, 40 + 30 * r, seed=100 + r)
        ids = np.arange(1_000_001 + 1000 * r, 1_000_001 + 1000 * r + len(o) - 1, dtype=np.int64)
        reqs_data.append((b, o, ids))
    # alone: one request at a time
    alone = []
    for b, o, ids in reqs_data:
        srv = OverlapServer(gpu, cfg, max_iteration_reads=64)
        srv.start()
        req = srv.submit(Request(0, b, o, ids))
        req.done_event.wait(120)
        srv.stop()
        alone.append(request_rows(req))
    # mixed: all requests in flight together, small iterations so they interleave
    srv = OverlapServer(gpu, cfg, max_iteration_reads=48)
    srv.start()
    reqs = []
    for k, (b, o, ids) in enumerate(reqs_data):
        reqs.append(srv.submit(Request(k, b, o, ids)))
        time.sleep(0.01)
    for r in reqs:
        assert r.done_event.wait(300)
    srv.stop()
    rep = srv.latency_report()
    assert rep["requests"] == len(reqs)
    for a, r in zip(alone, reqs):
        assert a.tobytes() == request_rows(r).tobytes()
    # and the oracle agrees with the request results
    orc = oracle_mod.Oracle(genome, offs)
    orc.build_index(15, solid_min=1, rep_max=64)
    b, o, ids = reqs_data[0]
    erows, _ = orc.overlaps(ids, cfg, threads=8, seqs=(b, o))
    assert erows.tobytes() == alone[0].tobytes()
    assert len(erows) > 0
