"""CPU tests of the oracle: pinned against the reference's committed fixtures,
closed forms and brute force (no GPU)."""
import hashlib
from pathlib import Path

import numpy as np
import pytest

from cuflye_amd.overlap import OverlapConfig

FIX = Path(__file__).resolve().parent / "golden" / "reference_fixtures"
JOIN_SHA = "c07761a6cf2dca8d2c3d511b938b454d8e272e0f035a147a59674f1c1c2c67ad"
WINDOW_SHA = "f0ef59dafc1a8efa5f007443d4c11191e3f03b2500c87874b40fa89f2803010d"


def canonical_candidates(records):
    rows = sorted(records)
    return "".join(f"{q}\t{qp}\t{k}\t{t}\t{tp}\t{s}\n" for q, qp, k, t, tp, s in rows)


def test_fixture_files_match_reference_hashes():
    for name, sha in (("kmer-encode-smoke-v0", JOIN_SHA), ("kmer-join-smoke-v0", JOIN_SHA),
                      ("read-window-smoke-v0", WINDOW_SHA)):
        data = (FIX / name / "expected.candidates.tsv").read_bytes()
        assert hashlib.sha256(data).hexdigest() == sha


def _toy_join(O, queries, index, repetitive, k):
    """The fixtures' join: lookup = standard form; repetitive lookups skipped;
    self hits (same id and position) dropped (read_window_smoke.cu:249-302 semantics)."""
    out = []
    for qid, qpos, kmer, lookup in queries:
        if lookup in repetitive:
            continue
        for tlookup, tid, tpos, strand in index:
            if tlookup != lookup:
                continue
            if tid == qid and tpos == qpos:
                continue
            out.append((qid, qpos, kmer, tid, tpos, strand))
    return out


def test_kmer_encode_fixture(oracle_mod):
    O = oracle_mod
    d = FIX / "kmer-encode-smoke-v0"
    k = 4
    q = [l.split("\t") for l in (d / "queries.tsv").read_text().splitlines()]
    idx = [l.split("\t") for l in (d / "index.tsv").read_text().splitlines()]
    rep = {O.standard_form(O.encode(l.strip()), k)[0] for l in (d / "repetitive-kmers.tsv").read_text().splitlines()}
    queries = [(int(a), int(b), O.encode(s), O.standard_form(O.encode(s), k)[0]) for a, b, s in q]
    index = [(O.standard_form(O.encode(s), k)[0], int(a), int(b), st) for a, b, st, s in idx]
    got = canonical_candidates(_toy_join(O, queries, index, rep, k))
    assert got == (d / "expected.candidates.tsv").read_text()
    assert hashlib.sha256(got.encode()).hexdigest() == JOIN_SHA


def test_kmer_join_fixture(oracle_mod):
    d = FIX / "kmer-join-smoke-v0"
    q = [tuple(map(int, l.split("\t"))) for l in (d / "queries.tsv").read_text().splitlines()]
    idx = [l.split("\t") for l in (d / "index.tsv").read_text().splitlines()]
    rep = {int(l) for l in (d / "repetitive-kmers.tsv").read_text().splitlines()}
    index = [(int(a), int(b), int(c), s) for a, b, c, s in idx]
    got = canonical_candidates(_toy_join(oracle_mod, q, index, rep, 4))
    assert hashlib.sha256(got.encode()).hexdigest() == JOIN_SHA


def test_read_window_fixture(oracle_mod):
    O = oracle_mod
    d = FIX / "read-window-smoke-v0"
    k = 4
    reads = [l.split("\t") for l in (d / "reads.tsv").read_text().splitlines()]
    idx = [l.split("\t") for l in (d / "index.tsv").read_text().splitlines()]
    rep = {O.standard_form(O.encode(l.strip()), k)[0] for l in (d / "repetitive-kmers.tsv").read_text().splitlines()}
    queries = []
    for rid, seq in reads:
        for pos in range(len(seq) - k + 1):
            x = O.encode(seq[pos:pos + k])
            queries.append((int(rid), pos, x, O.standard_form(x, k)[0]))
    index = [(O.standard_form(O.encode(s), k)[0], int(a), int(b), st) for a, b, st, s in idx]
    got = canonical_candidates(_toy_join(O, queries, index, rep, k))
    assert hashlib.sha256(got.encode()).hexdigest() == WINDOW_SHA


def test_encoding_closed_forms(oracle_mod):
    O = oracle_mod
    assert O.encode("CATC") == 77 and O.encode("AGGG") == 42 and O.encode("CTGT") == 123
    assert O.reverse_complement(O.encode("CATC"), 4) == O.encode("GATG")
    assert O.reverse_complement(O.encode("AAAAAAAAAAAAAAA"), 15) == O.encode("T" * 15)
    # canonical = min(kmer, rc); flipped unless kmer < rc; a palindrome flips
    assert O.standard_form(O.encode("GATG"), 4) == (O.encode("CATC"), True)
    assert O.standard_form(O.encode("CATC"), 4) == (O.encode("CATC"), False)
    assert O.standard_form(O.encode("ACGT"), 4) == (O.encode("ACGT"), True)
    for sid, iid in ((1, 0), (-1, 1), (2, 2), (-2, 3), (71, 140), (-71, 141)):
        assert O.signed_to_internal(sid) == iid
        assert O.internal_to_signed(iid) == sid


def _revcomp(s: str) -> str:
    return s[::-1].translate(str.maketrans("ACGT", "TGCA"))


def test_candidates_brute_force(oracle_mod):
    """Every emitted hit spells the query k-mer on the returned strand; every
    such pair (except the self window) is emitted; order is window-major."""
    O = oracle_mod
    rng = np.random.default_rng(7)
    k = 5
    base = "".join(rng.choice(list("ACGT"), 300))
    reads = [base[0:120], _revcomp(base[60:200]), base[100:260], base[150:300] + base[0:20]]
    bases = np.frombuffer("".join(reads).encode(), np.uint8)
    offs = np.cumsum([0] + [len(r) for r in reads]).astype(np.int64)
    orc = O.Oracle(bases, offs)
    info = orc.build_index(k, solid_min=1, rep_max=1000)
    strands = {}
    for i, r in enumerate(reads):
        strands[i + 1] = r
        strands[-(i + 1)] = _revcomp(r)
    for qid in (1, -1, 2, -3, 4):
        q = strands[qid]
        got, filt = orc.candidates(qid)
        assert len(filt) == 0
        exp = set()
        for cp in range(len(q) - k + 1):
            for tid, t in strands.items():
                for ep in range(len(t) - k + 1):
                    if t[ep:ep + k] == q[cp:cp + k] and not (tid == qid and ep == cp):
                        exp.add((cp, tid, ep))
        gotset = {(int(r["query_pos"]), int(r["target_id"]), int(r["target_pos"])) for r in got}
        assert gotset == exp
        assert len(gotset) == len(got)
        assert np.all(np.diff(got["query_pos"]) >= 0)
        for r in got:
            assert O.encode(q[r["query_pos"]:r["query_pos"] + k]) == r["kmer"]
    assert info["positions"] == sum(len(r) - k + 1 for r in reads)


def _params(**kw):
    return OverlapConfig(**kw)


def test_chain_closed_form_single_diagonal(oracle_mod):
    """A perfect diagonal of hits every 10 bp.  Flye starts score[0] at 0 (not
    k), so hit 1's best gain 0 + 10 does not beat k and hit 0 never joins:
    score[i] = 15 + 10 * (i - 1) for i >= 1 and the chain starts at hit 1."""
    O = oracle_mod
    n = 700
    cur = np.arange(100, 100 + 10 * n, 10, dtype=np.int32)
    ext = cur + 50
    p = _params(kmer_size=15, minimum_overlap=1000, check_overhang=False)
    rows, off = O.chain_lists(cur, ext, [0, n], [1], [-5], [8000], [8000], p)
    assert len(rows) == 1
    r = rows[0]
    assert (r["cur_begin"], r["cur_end"]) == (110, int(cur[-1]) + 14)
    assert (r["ext_begin"], r["ext_end"]) == (160, int(ext[-1]) + 14)
    # score = s[last] - s[first] + k - 1
    assert r["score"] == (15 + 10 * (n - 2)) - 15 + 14
    assert r["chain_length"] == n - 1
    rate = np.float32(n - 1) / np.float32(max(r["cur_end"] - r["cur_begin"], r["ext_end"] - r["ext_begin"]))
    div = np.float32(O.logf(float(np.float32(1.0) / rate))) / np.float32(15)
    assert r["seq_divergence"] == div


def test_chain_unit_spacing_diagonal(oracle_mod):
    """Hits at 1 bp spacing on one exact diagonal: every step gains
    min(1, 1, 15) = 1 and the scan stops right after its first predecessor
    (jumpDiv == 0 and dc < k).  The gap tiers, the max_jump_gap reject and the
    early break's effect are pinned in test_oracle_closed_form.py and by the
    reference vectors."""
    O = oracle_mod
    cur = np.array(list(range(0, 2000)) , np.int32)
    ext = cur.copy()
    p = _params(kmer_size=15, minimum_overlap=1000, check_overhang=False)
    rows, _ = O.chain_lists(cur, ext, [0, len(cur)], [1], [3], [5000], [5000], p)
    assert len(rows) == 1
    # each step gains min(1, 1, 15) = 1: score[i] = 15 + (i - 1) for i >= 1
    assert rows[0]["score"] == (15 + 1998) - 15 + 14
    assert rows[0]["chain_length"] == 1999
    assert rows[0]["cur_begin"] == 1


def test_chain_prefilters(oracle_mod):
    O = oracle_mod
    cur = np.arange(0, 900, 10, dtype=np.int32)
    p = _params(kmer_size=15, minimum_overlap=1000, check_overhang=False)
    rows, _ = O.chain_lists(cur, cur, [0, len(cur)], [1], [3], [5000], [5000], p)
    assert len(rows) == 0  # span 890 < minimum overlap


def test_read_alignment_fixture(oracle_mod):
    """tests/fixtures/read-alignment-replay-fixture-v0/simple: 2 alignments, one
    accepted chain of both, rows as oracle.read-alignment.tsv."""
    import json
    from cuflye_amd import native as N
    O = oracle_mod
    d = FIX / "read-alignment-replay-fixture-v0" / "simple"
    man = json.loads((d / "manifest.json").read_text())
    rows = [l.split("\t") for l in (d / "edge-overlaps.tsv").read_text().splitlines()]
    a = np.zeros(len(rows), N.EDGE_ALIGNMENT_DTYPE)
    for i, f in enumerate(rows):
        a[i] = (int(f[1]), int(f[5]), int(f[8]), int(f[2]), int(f[3]), int(f[4]), int(f[6]), int(f[7]),
                int(f[9]), int(f[10]), int(f[11]), int(f[12]), float(f[13]))
    params = N.ChainParams(man["maximum_jump"], man["max_read_overlap"], man["minimum_overlap"],
                           man["max_separation"])
    # the fixture's alignments are all chain inputs (long_edge above every edge)
    si, ko, segs, so = O.read_align_chains(a, [0, len(a)], 1 << 30, 0, params)
    assert list(si) == [0, 1]
    # chainReadAlignments yields the fixture's 2 candidate chains; the chain
    # divergence acceptance (Flye's base alignment, not replayable) comes from
    # chain-divergence.tsv as the reference replay takes it
    assert len(set(segs["chain_id"])) == man["candidate_chains"]
    flags = [l.split("\t")[2] == "1" for l in (d / "chain-divergence.tsv").read_text().splitlines()]
    out = []
    for s in segs:
        if not flags[s["chain_id"]]:
            continue
        item = a[si[s["input_index"]]]
        out.append("\t".join(str(x) for x in (s["chain_id"], s["segment_id"], item["read_id"], item["read_begin"],
                                               item["read_end"], item["read_len"], item["edge_id"],
                                               item["edge_seq_id"], item["edge_begin"], item["edge_end"],
                                               item["edge_len"], item["score"])) +
                   "\t" + format(float(item["seq_divergence"]), ".9g").rstrip())
    expected = (d / "oracle.read-alignment.tsv").read_text().splitlines()
    assert [l.rsplit("\t", 1)[0] for l in out] == [l.rsplit("\t", 1)[0] for l in expected]
    assert [float(l.rsplit("\t", 1)[1]) for l in out] == pytest.approx([float(l.rsplit("\t", 1)[1]) for l in expected])


def test_std_sort_perm_is_libstdcxx(oracle_mod):
    """orc_std_sort_perm is real std::sort: sorted, a permutation, and not
    stable on ties (introsort moves equal keys), which is what the GPU must match."""
    O = oracle_mod
    rng = np.random.default_rng(3)
    keys = rng.integers(0, 4, 200).astype(np.uint64)
    perm = O.std_sort_perm(keys)
    assert sorted(perm) == list(range(200))
    assert np.all(np.diff(keys[perm].astype(np.int64)) >= 0)
    stable = np.argsort(keys, kind="stable")
    assert not np.array_equal(perm, stable)
