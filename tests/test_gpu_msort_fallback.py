"""The stable match sort's two paths: the per-query LDS counting sort holds a
histogram of every target id (2 x reads), so a read set of 20,000 reads
(40,000 ids, more than the LDS holds) takes the segmented radix sort
instead.  Both must give Flye's order: sorted candidates (std::sort's tie
order included) and overlap rows equal to the oracle's."""
import numpy as np
import pytest

import cuflye_amd.synth as S
from cuflye_amd.overlap import OverlapConfig

pytestmark = pytest.mark.gpu


def test_many_reads_take_the_radix_path(gpu, oracle_mod):
    genome = S.genome_with_repeats(2_000_000, 0.5, 0.02, 2000, 5000, 2, 4, 0.01, 0.01, seed=61)
    b, o = S.reads_from(genome, 20_000, seed=62, mean=700.0, n50=800.0, err=0.02)
    n = len(o) - 1
    print(f"[msort fallback] reads={n} bases={len(b)}", flush=True)
    # more target ids than the histogram's LDS holds: 160 KiB of 4-byte
    # words less the 4,112 words of the scatter's tile state
    assert 2 * n > 160 * 1024 // 4 - 4_112
    gpu.store_reads(b, o)
    gpu.build_index(15)
    orc = oracle_mod.Oracle(b, o)
    orc.build_index(15)
    ids = np.array([1, -1, 2, -2, 17, -17, 5000, -5000, 19_999, -19_999], np.int64)
    m, moff = gpu.collect_sorted_candidates(ids)
    assert int(moff[-1]) > 0
    for i, q in enumerate(ids):
        em, _ = orc.candidates(int(q), sorted_=True)
        assert m[moff[i]:moff[i + 1]].tobytes() == em.tobytes(), f"sorted q={q}"
    q2 = np.empty(400, np.int64)
    q2[0::2] = np.arange(1, 201)
    q2[1::2] = -np.arange(1, 201)
    cfg = OverlapConfig(kmer_size=15, minimum_overlap=500)
    rows, roff = gpu.get_seq_overlaps(q2, cfg)
    erows, eroff = orc.overlaps(q2, cfg, threads=8)
    assert np.array_equal(roff, eroff)
    assert rows.tobytes() == erows.tobytes()
    assert len(rows) > 0
