"""The oracle's streaming candidate-record-v1 canonical SHA-256 against the
reference's recorded fixture hashes and against the canonical text written
out in Python (tools/validate_candidate_dump.py:76-134 semantics; no GPU)."""
import hashlib

import numpy as np
import pytest

from fixture_index import canonical_text, fixture


@pytest.mark.parametrize("name", ["read-window-smoke-v0", "kmer-encode-smoke-v0"])
def test_oracle_canonical_hash_fixture(oracle_mod, name):
    f = fixture(name)
    orc = oracle_mod.Oracle(f["bases"], f["offsets"])
    orc.load_index(4, f["kmers"], f["bucket_offsets"], f["ids"], f["pos"], f["repetitive"])
    h, n = orc.candidates_canonical_sha256(f["queries"])
    assert h == f["sha"] and n == 6


def test_oracle_canonical_hash_equals_text(oracle_mod):
    import cuflye_amd.synth as S
    r = S.config1(scale=0.02, seed=5)
    orc = oracle_mod.Oracle(r.bases, r.offsets)
    orc.build_index(15)
    ids = np.array([3, -1, 2, -2, 1, 7], np.int64)
    h, n = orc.candidates_canonical_sha256(ids, threads=4)
    recs = []
    for q in ids:
        m, _ = orc.candidates(int(q))
        recs.extend(m)
    assert n == len(recs) > 0
    assert h == hashlib.sha256(canonical_text(recs).encode()).hexdigest()
