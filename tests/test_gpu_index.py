"""VertexIndex buckets far above the per-thread size: homopolymer runs and
tandem repeats give k-mers with thousands of occurrences, which the library
puts in (read, pos) order with one block per bucket.  The emission order of
the candidate lists reads the bucket order, so it is compared with the
oracle's."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _low_complexity_reads(seed: int):
    rng = np.random.default_rng(seed)
    unit = "".join(rng.choice(list("ACGT"), 37))
    reads = []
    for i in range(24):
        body = "".join(rng.choice(list("ACGT"), int(rng.integers(3000, 6000))))
        run = "A" * int(rng.integers(150, 500))           # homopolymer run
        tandem = unit * int(rng.integers(10, 30))           # tandem repeat
        cut = int(rng.integers(100, len(body) - 100))
        reads.append(body[:cut] + run + body[cut:] + tandem + "".join(rng.choice(list("ACGT"), 200)))
    bases = np.frombuffer("".join(reads).encode(), np.uint8).copy()
    off = np.zeros(len(reads) + 1, np.int64)
    off[1:] = np.cumsum([len(r) for r in reads])
    return bases, off


def test_large_buckets_match_oracle(gpu, oracle_mod):
    bases, off = _low_complexity_reads(17)
    print(f"[large buckets] reads={len(off) - 1} bases={len(bases)}", flush=True)
    orc = oracle_mod.Oracle(bases, off)
    oinfo = orc.build_index(15, 1, 1 << 20)
    t = orc.index_tables()
    big = int(np.max(np.diff(t["start"])))
    assert big > 2048  # the homopolymer k-mer alone: several 1,024-thread strides
    gpu.store_reads(bases, off)
    ginfo = gpu.build_index(15, solid_min=1, rep_max=1 << 20)
    assert ginfo.solid_kmers == oinfo["solid"]
    assert ginfo.indexed_positions == oinfo["positions"]
    ids = np.array([1, -1, 7, -12, 24], np.int64)
    m, moff, f, foff = gpu.collect_candidates(ids)
    for i, q in enumerate(ids):
        em, ef = orc.candidates(int(q), sorted_=False)
        got = m[moff[i]:moff[i + 1]]
        assert len(got) == len(em), f"q={q}"
        assert got.tobytes() == em.tobytes(), f"q={q}: emission order differs"
