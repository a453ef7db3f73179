"""The oracle's join over an index given as tables (flyehip_load_index's
arguments) against the reference's k=4 fixtures: the canonical candidate text
hashes to the SHA-256 the reference records (no GPU)."""
import hashlib

import numpy as np
import pytest

from fixture_index import canonical_text, fixture


@pytest.mark.parametrize("name", ["read-window-smoke-v0", "kmer-encode-smoke-v0"])
def test_loaded_fixture_index_reproduces_reference_hash(oracle_mod, name):
    f = fixture(name)
    orc = oracle_mod.Oracle(f["bases"], f["offsets"])
    orc.load_index(4, f["kmers"], f["bucket_offsets"], f["ids"], f["pos"], f["repetitive"])
    recs = []
    for q in f["queries"]:
        m, _ = orc.candidates(int(q))
        recs.extend(m)
    text = canonical_text(recs)
    assert hashlib.sha256(text.encode()).hexdigest() == f["sha"], text


def test_loaded_index_keeps_the_given_bucket_order(oracle_mod):
    """The emission order inside a window follows the bucket order as given,
    not (read, position) order."""
    import cuflye_amd.synth as S
    r = S.config1(scale=0.02, seed=5)
    orc = oracle_mod.Oracle(r.bases, r.offsets)
    orc.build_index(15)
    t = orc.index_tables()
    rng = np.random.default_rng(1)
    iid = t["ids"].astype(np.int64)
    ids = np.where(iid & 1, -((iid + 1) // 2), iid // 2 + 1)  # internal -> signed id
    pos = t["pos"].copy()
    for b in range(len(t["kmers"])):
        lo, hi = t["start"][b], t["start"][b + 1]
        p = lo + rng.permutation(hi - lo)
        ids[lo:hi], pos[lo:hi] = ids[p].copy(), pos[p].copy()
    built, _ = orc.candidates(1)
    orc.load_index(15, t["kmers"], t["start"], ids, pos, t["repetitive"])
    loaded, _ = orc.candidates(1)
    key = lambda a: sorted(a.tolist())  # noqa: E731
    assert key(built) == key(loaded)
    assert built.tobytes() != loaded.tobytes()
