"""flyehip_overlaps_submit / _collect / _release through the C ABI: jobs
queued from one thread, collected from several host threads in random order
(as Flye's thread pool would ask for them), equal to flyehip_overlaps on the
same list - rows and divStats values - across several batches per job and
two jobs in the queue at once."""
import threading

import numpy as np
import pytest

import cuflye_amd.synth as S
from cuflye_amd import native as N
from cuflye_amd.overlap import FlyeHipError, OverlapConfig

pytestmark = pytest.mark.gpu


def test_submit_collect_from_threads(gpu):
    r = S.config1(scale=0.05, seed=31)
    print(f"[submit] reads={r.n_reads}", flush=True)
    gpu.store_reads(r.bases, r.offsets)
    gpu.build_index(15)
    ids = np.empty(2 * r.n_reads, np.int64)
    ids[0::2] = np.arange(1, r.n_reads + 1)
    ids[1::2] = -np.arange(1, r.n_reads + 1)
    cfg = OverlapConfig(kmer_size=15, minimum_overlap=3000)
    erows, eroff, edivs, edoff = gpu.get_seq_overlaps(ids, cfg, with_div_stats=True)
    half = ids[: len(ids) // 2]
    hrows, hroff = gpu.get_seq_overlaps(half, cfg)
    try:
        gpu.set_tuning(queries_per_batch=32, batches_in_flight=2)  # 64 query records per chunk
        t1 = gpu.submit_overlaps(ids, cfg, with_div_stats=True)
        t2 = gpu.submit_overlaps(half, cfg)
        got1, got2 = {}, {}
        errs = []
        order = np.random.default_rng(3).permutation(len(ids))

        def worker(w):
            try:
                for i in order[w::8]:
                    got1[int(i)] = gpu.collect_overlaps(t1, int(i), with_div_stats=True)
                    if i < len(half):
                        got2[int(i)] = gpu.collect_overlaps(t2, int(i))
            except Exception as e:  # noqa: BLE001
                errs.append(e)

        th = [threading.Thread(target=worker, args=(w,)) for w in range(8)]
        for t in th:
            t.start()
        for t in th:
            t.join(timeout=240)
        assert not errs, errs
        assert len(got1) == len(ids) and len(got2) == len(half)
        for i in range(len(ids)):
            rows, divs = got1[i]
            assert rows.tobytes() == erows[eroff[i]:eroff[i + 1]].tobytes(), f"query {i}"
            assert divs.tobytes() == edivs[edoff[i]:edoff[i + 1]].tobytes(), f"query {i} divStats"
        for i in range(len(half)):
            assert got2[i].tobytes() == hrows[hroff[i]:hroff[i + 1]].tobytes(), f"job 2 query {i}"
        gpu.release_overlaps(t1)
        gpu.release_overlaps(t2)
        with pytest.raises(FlyeHipError) as e:
            gpu.collect_overlaps(t1, 0)
        assert e.value.code == N.FLYEHIP_ERR_ARG
        with pytest.raises(FlyeHipError) as e:
            gpu.submit_overlaps(np.array([r.n_reads + 1], np.int64), cfg)
        assert e.value.code == N.FLYEHIP_ERR_ARG
    finally:
        gpu.set_tuning(queries_per_batch=2048, batches_in_flight=2)
