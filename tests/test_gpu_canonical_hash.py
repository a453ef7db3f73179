"""flyehip_candidates_canonical_sha256 (streamed through the HIP join): equal
to the oracle's canonical hash on a read set, to the hash of the canonical
text built from the library's own candidate rows, and to the reference's
recorded fixture hashes over a loaded index."""
import hashlib

import numpy as np
import pytest

from fixture_index import canonical_text, fixture

pytestmark = pytest.mark.gpu


def test_gpu_canonical_hash_matches_oracle(gpu, oracle_mod):
    import cuflye_amd.synth as S
    r = S.config1(scale=0.05, seed=41)
    print(f"[canonical hash] reads={r.n_reads}", flush=True)
    gpu.store_reads(r.bases, r.offsets)
    gpu.build_index(15)
    orc = oracle_mod.Oracle(r.bases, r.offsets)
    orc.build_index(15)
    ids = np.empty(2 * r.n_reads, np.int64)
    ids[0::2] = np.arange(1, r.n_reads + 1)
    ids[1::2] = -np.arange(1, r.n_reads + 1)
    rng = np.random.default_rng(4)
    ids = ids[rng.permutation(len(ids))]  # any order: the hash sorts the ids
    g, gn = gpu.candidates_canonical_sha256(ids)
    o, on = orc.candidates_canonical_sha256(ids, threads=8)
    assert (g, gn) == (o, on) and gn > 0
    few = ids[:9]
    m, moff, _, _ = gpu.collect_candidates(few)
    g2, n2 = gpu.candidates_canonical_sha256(few)
    assert n2 == len(m)
    assert g2 == hashlib.sha256(canonical_text(m).encode()).hexdigest()


@pytest.mark.parametrize("name", ["read-window-smoke-v0", "kmer-encode-smoke-v0"])
def test_gpu_canonical_hash_fixture(gpu, name):
    f = fixture(name)
    gpu.store_reads(f["bases"], f["offsets"])
    gpu.load_index(4, f["kmers"], f["bucket_offsets"], f["ids"], f["pos"], f["repetitive"])
    h, n = gpu.candidates_canonical_sha256(f["queries"])
    assert h == f["sha"] and n == 6
