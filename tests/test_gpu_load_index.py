"""flyehip_load_index through the C ABI: the reference's k=4 fixtures through
the HIP join reproduce the SHA-256 the reference records, and an index whose
buckets are in an order other than (read, position) - as Flye's may be -
drives the emission order, the std::sort tie order and the overlaps exactly
as in the oracle over the same tables."""
import hashlib

import numpy as np
import pytest

from fixture_index import canonical_text, fixture

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["read-window-smoke-v0", "kmer-encode-smoke-v0"])
def test_gpu_fixture_hash(gpu, name):
    f = fixture(name)
    print(f"[load_index] {name}", flush=True)
    gpu.store_reads(f["bases"], f["offsets"])
    info = gpu.load_index(4, f["kmers"], f["bucket_offsets"], f["ids"], f["pos"], f["repetitive"])
    assert info.indexed_positions == len(f["ids"])
    m, moff, filt, foff = gpu.collect_candidates(f["queries"])
    text = canonical_text(m)
    assert hashlib.sha256(text.encode()).hexdigest() == f["sha"], text


def test_gpu_loaded_order_matches_oracle(gpu, oracle_mod):
    import cuflye_amd.synth as S
    from cuflye_amd.overlap import OverlapConfig
    r = S.config1(scale=0.05, seed=12)
    print(f"[load_index permuted] reads={r.n_reads}", flush=True)
    orc = oracle_mod.Oracle(r.bases, r.offsets)
    orc.build_index(15)
    t = orc.index_tables()
    rng = np.random.default_rng(2)
    iid = t["ids"].astype(np.int64)
    ids = np.where(iid & 1, -((iid + 1) // 2), iid // 2 + 1)  # internal -> signed id
    pos = t["pos"].copy()
    for b in range(len(t["kmers"])):
        lo, hi = t["start"][b], t["start"][b + 1]
        p = lo + rng.permutation(hi - lo)
        ids[lo:hi], pos[lo:hi] = ids[p].copy(), pos[p].copy()
    orc.load_index(15, t["kmers"], t["start"], ids, pos, t["repetitive"])
    gpu.store_reads(r.bases, r.offsets)
    # k-mers in a shuffled order: the library keys buckets by k-mer
    perm = rng.permutation(len(t["kmers"]))
    sizes = np.diff(t["start"])[perm]
    off = np.zeros(len(perm) + 1, np.int64)
    off[1:] = np.cumsum(sizes)
    gid = np.concatenate([ids[t["start"][b]:t["start"][b + 1]] for b in perm])
    gpos = np.concatenate([pos[t["start"][b]:t["start"][b + 1]] for b in perm])
    gpu.load_index(15, t["kmers"][perm], off, gid, gpos, t["repetitive"])
    q = np.array([1, -1, 3, -4, 9, r.n_reads, -r.n_reads], np.int64)
    m, moff, filt, foff = gpu.collect_candidates(q)
    for i, qq in enumerate(q):
        em, ef = orc.candidates(int(qq))
        assert m[moff[i]:moff[i + 1]].tobytes() == em.tobytes(), f"emission q={qq}"
        assert np.array_equal(filt[foff[i]:foff[i + 1]], ef)
    sm, smoff = gpu.collect_sorted_candidates(q)
    for i, qq in enumerate(q):
        em, _ = orc.candidates(int(qq), sorted_=True)
        assert sm[smoff[i]:smoff[i + 1]].tobytes() == em.tobytes(), f"sorted q={qq}"
    ids_all = np.empty(2 * r.n_reads, np.int64)
    ids_all[0::2] = np.arange(1, r.n_reads + 1)
    ids_all[1::2] = -np.arange(1, r.n_reads + 1)
    cfg = OverlapConfig(kmer_size=15, minimum_overlap=3000)
    rows, roff = gpu.get_seq_overlaps(ids_all, cfg)
    erows, eroff = orc.overlaps(ids_all, cfg, threads=8)
    assert np.array_equal(roff, eroff)
    assert rows.tobytes() == erows.tobytes()
    assert len(rows) > 0


def test_gpu_load_index_rejects_bad_tables(gpu):
    from cuflye_amd import native as N
    from cuflye_amd.overlap import FlyeHipError
    f = fixture("read-window-smoke-v0")
    gpu.store_reads(f["bases"], f["offsets"])
    bad = f["kmers"].copy()
    bad[0] = 255  # TTTT: not in standard form (its reverse complement AAAA is smaller)
    with pytest.raises(FlyeHipError) as e:
        gpu.load_index(4, bad, f["bucket_offsets"], f["ids"], f["pos"], f["repetitive"])
    assert e.value.code == N.FLYEHIP_ERR_ARG
    dup = f["kmers"].copy()
    dup[1] = dup[0]
    with pytest.raises(FlyeHipError) as e:
        gpu.load_index(4, dup, f["bucket_offsets"], f["ids"], f["pos"], f["repetitive"])
    assert e.value.code == N.FLYEHIP_ERR_ARG
    far = f["pos"].copy()
    far[0] = 10_000
    with pytest.raises(FlyeHipError) as e:
        gpu.load_index(4, f["kmers"], f["bucket_offsets"], f["ids"], far, f["repetitive"])
    assert e.value.code == N.FLYEHIP_ERR_ARG
