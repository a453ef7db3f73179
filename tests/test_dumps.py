"""Dump writers and canonical hashes against the reference's fixtures."""
import hashlib
from pathlib import Path

import numpy as np

from cuflye_amd import dumps as D
from cuflye_amd import native as N

FIX = Path(__file__).resolve().parent / "golden" / "reference_fixtures"


def _records(path):
    rows = [l.split("\t") for l in path.read_text().splitlines()]
    a = np.zeros(len(rows), N.CANDIDATE_DTYPE)
    for i, f in enumerate(rows):
        a[i] = (int(f[0]), int(f[3]), int(f[2]), int(f[1]), int(f[4]))
        assert (f[5] == "+") == (int(f[3]) > 0)
    return a


def test_candidate_canonical_hash_matches_reference():
    for name, sha in (("kmer-join-smoke-v0", "c07761a6cf2dca8d2c3d511b938b454d8e272e0f035a147a59674f1c1c2c67ad"),
                      ("read-window-smoke-v0", "f0ef59dafc1a8efa5f007443d4c11191e3f03b2500c87874b40fa89f2803010d")):
        recs = _records(FIX / name / "expected.candidates.tsv")
        rng = np.random.default_rng(0)
        shuffled = recs[rng.permutation(len(recs))]
        assert D.candidate_sha256(shuffled) == sha
        assert "".join(D.candidate_lines(recs)) == (FIX / name / "expected.candidates.tsv").read_text()


def test_overlap_format_matches_cpp_setprecision9():
    rows = np.zeros(3, N.OVERLAP_DTYPE)
    rows["cur_id"] = [5, -3, 5]
    rows["ext_id"] = [2, 7, 1]
    rows["cur_len"] = rows["ext_len"] = 100
    rows["seq_divergence"] = [np.float32(0.0), np.float32(0.1), np.float32(1e-5)]
    lines = D.overlap_lines(rows)
    assert lines[0].endswith("\t0\n")
    assert lines[1].endswith("\t0.100000001\n")   # float 0.1 printed with 9 significant digits
    assert lines[2].endswith("\t9.99999975e-06\n")
    canon = D.canonical_overlaps(rows)
    assert list(canon["cur_id"]) == [-3, 5, 5] and list(canon["ext_id"]) == [7, 1, 2]
    assert len(D.overlap_sha256(rows)) == 64


def test_read_alignment_canonical_hash():
    lines = (FIX / "read-alignment-v1" / "simple.tsv").read_text().splitlines()
    expect = hashlib.sha256(("\n".join(sorted(lines, key=lambda l: tuple(
        int(x) for x in l.split("\t")[:12]) + (float(l.split("\t")[12]),))) + "\n").encode()).hexdigest()
    assert D.read_alignment_sha256(lines[::-1]) == expect
