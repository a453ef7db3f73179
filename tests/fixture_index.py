"""The reference's k=4 candidate fixtures as a read store plus VertexIndex
tables (the arguments of flyehip_load_index), for the CPU oracle and the HIP
join alike.

The fixtures (tests/fixtures/kmer-encode-smoke-v0 and read-window-smoke-v0 in
the reference) list, per query window, the positions iterKmerPos returns.
VertexIndex stores each bucket's positions for the canonical k-mer and flips
them (reverse strand, position len - p - k) when the query window is not in
standard form, so the tables stored here hold, for a bucket the queries reach
through a reverse-complement window, the reverse-strand equivalents of the
listed positions.  The fixtures carry no target sequences: target reads get
seeded filler of 200 bases (only their lengths matter to the join), and the
query reads are those of read-window's reads.tsv.  kmer-encode lists query
windows, not reads; its read 1 is built to hold exactly the windows that
produce hits (AGGG at 0, CTGT at 7, the repetitive TTCA at 11; its CGAT window
at 5 only ever hits itself, so it is left out).
"""
from pathlib import Path

import numpy as np

FIX = Path(__file__).resolve().parent / "golden" / "reference_fixtures"
K = 4
SHA = {"read-window-smoke-v0": "f0ef59dafc1a8efa5f007443d4c11191e3f03b2500c87874b40fa89f2803010d",
       "kmer-encode-smoke-v0": "c07761a6cf2dca8d2c3d511b938b454d8e272e0f035a147a59674f1c1c2c67ad"}
CODE = {"A": 0, "C": 1, "G": 2, "T": 3}


def _num(s: str) -> int:
    v = 0
    for ch in s:
        v = v * 4 + CODE[ch]
    return v


def _rc(s: str) -> str:
    return s[::-1].translate(str.maketrans("ACGT", "TGCA"))


def _canon(s: str) -> int:
    return min(_num(s), _num(_rc(s)))


def fixture(name: str):
    d = FIX / name
    rows = [l.split("\t") for l in (d / "index.tsv").read_text().splitlines()]
    rep = sorted({_canon(l.strip()) for l in (d / "repetitive-kmers.tsv").read_text().splitlines()})
    # query strands: -2 = AAACATCAAA extended by an A run (no index k-mer), so
    # that read 2 also holds the listed +2 position 10
    minus2 = "AAACATCAAA" + "A" * 10
    if name == "read-window-smoke-v0":
        read1 = dict(l.split("\t") for l in (d / "reads.tsv").read_text().splitlines())["1"]
    else:
        read1 = "AGGGTTTCTGTTTCA"
    rng = np.random.default_rng(4)
    reads = {1: read1, 2: _rc(minus2)}
    for r in range(3, 9):
        reads[r] = "".join(rng.choice(list("ACGT"), 200))
    strands = {}
    for r, s in reads.items():
        strands[r], strands[-r] = s, _rc(s)
    queries = [-2, 1]
    # canonical k-mers the queries reach through a reverse-complement window
    flipped = set()
    for q in queries:
        s = strands[q]
        for p in range(len(s) - K + 1):
            w = s[p:p + K]
            if _num(_rc(w)) < _num(w):
                flipped.add(_num(_rc(w)))
    buckets = {}
    for tid, tpos, _strand, kmer in rows:
        c = _canon(kmer)
        if c in rep:
            continue
        tid, tpos = int(tid), int(tpos)
        if c in flipped:
            tid, tpos = -tid, len(reads[abs(tid)]) - tpos - K
        buckets.setdefault(c, []).append((tid, tpos))
    kmers = np.array(list(buckets), np.uint64)
    off = np.zeros(len(kmers) + 1, np.int64)
    off[1:] = np.cumsum([len(buckets[int(c)]) for c in kmers])
    ids = np.array([e[0] for c in kmers for e in buckets[int(c)]], np.int64)
    pos = np.array([e[1] for c in kmers for e in buckets[int(c)]], np.int32)
    seqs = [reads[r] for r in range(1, 9)]
    bases = np.frombuffer("".join(seqs).encode(), np.uint8).copy()
    boff = np.zeros(len(seqs) + 1, np.int64)
    boff[1:] = np.cumsum([len(s) for s in seqs])
    return dict(bases=bases, offsets=boff, queries=np.array(queries, np.int64), kmers=kmers, bucket_offsets=off,
                ids=ids, pos=pos, repetitive=np.array(rep, np.uint64), sha=SHA[name])


def canonical_text(records) -> str:
    """candidate-record-v1 canonical text: the six fields, rows sorted."""
    rows = sorted((int(r["query_id"]), int(r["query_pos"]), int(r["kmer"]), int(r["target_id"]),
                   int(r["target_pos"]), "+" if int(r["target_id"]) > 0 else "-") for r in records)
    return "".join(f"{q}\t{qp}\t{k}\t{t}\t{tp}\t{s}\n" for q, qp, k, t, tp, s in rows)
