"""Dump formats and canonical hashes of the reference's parity gates.

candidate-record-v1 (docs/abi/candidate-record-v1.md:31-68): six TSV fields
    query_id, query_pos, kmer, target_id, target_pos, target_strand;
    canonical = records sorted as tuples, SHA-256 of the joined text
    (tools/validate_candidate_dump.py:80-134, diff_candidate_dumps.py:50-80).
overlap-range-v1 (docs/abi/overlap-range-v1.md:21-67): ten TSV fields, the
    divergence printed like C++ ``setprecision(9)`` and canonicalised with
    ``"0" or format(x, ".9g")`` (tools/validate_overlap_dump.py:24-27, 130-207).
read-alignment-v1 (docs/abi/read-alignment-v1.md:27-67): thirteen fields.

The canonical orderings are computed with numpy sorts on the record fields
(the same order as sorting the parsed tuples), so hashing a billion-record
run does not go through Python tuples.
"""
from __future__ import annotations

import hashlib

import numpy as np


def _fmt_div(x: float) -> str:
    x = float(x)
    return "0" if x == 0 else format(x, ".9g")


def candidate_lines(records: np.ndarray) -> list[str]:
    return [f"{r['query_id']}\t{r['query_pos']}\t{r['kmer']}\t{r['target_id']}\t{r['target_pos']}\t"
            f"{'+' if r['target_id'] > 0 else '-'}\n" for r in records]


def canonical_candidates(records: np.ndarray) -> np.ndarray:
    """Records in candidate-record-v1 canonical order (strand follows target_id)."""
    order = np.lexsort((records["target_pos"], records["target_id"], records["kmer"],
                        records["query_pos"], records["query_id"]))
    return records[order]


def candidate_sha256(records: np.ndarray, canonical: bool = True) -> str:
    recs = canonical_candidates(records) if canonical else records
    h = hashlib.sha256()
    chunk = 1 << 16
    for s in range(0, len(recs), chunk):
        h.update("".join(candidate_lines(recs[s:s + chunk])).encode())
    return h.hexdigest()


def overlap_lines(rows: np.ndarray) -> list[str]:
    return ["\t".join(str(int(r[f])) for f in ("cur_id", "cur_begin", "cur_end", "cur_len", "ext_id",
                                                "ext_begin", "ext_end", "ext_len", "score"))
            + "\t" + _fmt_div(r["seq_divergence"]) + "\n" for r in rows]


def canonical_overlaps(rows: np.ndarray) -> np.ndarray:
    order = np.lexsort(tuple(rows[f] for f in ("seq_divergence", "score", "ext_len", "ext_end", "ext_begin",
                                               "ext_id", "cur_len", "cur_end", "cur_begin", "cur_id")))
    return rows[order]


def overlap_sha256(rows: np.ndarray, canonical: bool = True) -> str:
    r = canonical_overlaps(rows) if canonical else rows
    return hashlib.sha256("".join(overlap_lines(r)).encode()).hexdigest()


def read_alignment_sha256(lines: list[str]) -> str:
    """Canonical SHA-256 of read-alignment-v1 rows (validate_read_alignment_dump.py:143-149)."""
    def key(line):
        f = line.rstrip("\n").split("\t")
        return tuple(int(x) for x in f[:12]) + (float(f[12]),)
    rows = sorted((l.rstrip("\n") for l in lines), key=key)
    return hashlib.sha256(("\n".join(rows) + "\n").encode()).hexdigest()
