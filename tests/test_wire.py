"""Binary wire formats (cuflye_amd/wire.py; no GPU): the CUFRALB0 layout of
docs/abi/read-alignment-compact-binary-v0.md:20-66, round trips, and every
check the reference's payload validator makes."""
import hashlib
import struct
from pathlib import Path

import numpy as np
import pytest

from cuflye_amd import native as N
from cuflye_amd import wire as W

GOLD = Path(__file__).resolve().parent / "golden"


def test_ralb_record_layout_matches_reference_table():
    # (offset, width) per field, docs/abi/read-alignment-compact-binary-v0.md:43-62
    table = [(0, 8), (8, 4), (12, 4), (16, 4), (20, 8), (28, 8), (36, 4), (40, 4), (44, 4), (48, 8), (56, 4),
             (60, 4), (64, 8), (72, 4), (76, 4), (80, 4), (84, 4), (88, 4)]
    got = [(W.RALB_RECORD.fields[n][1], W.RALB_RECORD.fields[n][0].itemsize) for n in W.RALB_RECORD.names]
    assert got == table and W.RALB_RECORD.itemsize == 92
    assert W.OVLB_RECORD.itemsize == 52 and W.CANB_RECORD.itemsize == 32


def _chains(oracle_mod):
    d = np.load(GOLD / "read_align_v1.npz")
    al, ao = d["alignments"], d["aln_offsets"]
    a = np.zeros(len(al), N.EDGE_ALIGNMENT_DTYPE)
    for j, f in enumerate(("read_begin", "read_end", "read_len", "edge_left_node", "edge_right_node", "edge_begin",
                           "edge_end", "edge_len", "score", "edge_id", "edge_seq_id")):
        a[f] = al[:, j]
    a["read_id"] = np.repeat(np.arange(1, len(ao)), np.diff(ao))
    a["seq_divergence"] = np.linspace(0.01, 0.2, len(a), dtype=np.float32)
    pv = d["params"]
    params = N.ChainParams(int(pv[0]), int(pv[1]), int(pv[2]), int(pv[3]))
    si, ko, segs, so = oracle_mod.read_align_chains(a, ao, 1 << 30, 0, params)
    return a, ao, si, ko, segs, so


def test_ralb_round_trip(oracle_mod):
    a, ao, si, ko, segs, so = _chains(oracle_mod)
    payload = W.ralb_encode(segs, so, si, ko, a, ao)
    info = W.validate_ralb(payload, expected_fixtures=len(ao) - 1, expected_records=len(segs),
                           expected_inputs=len(a), expected_sha256=hashlib.sha256(payload).hexdigest())
    assert info["records"] == len(segs) > 0
    assert len(payload) == 64 + 92 * len(segs)
    rec = W.ralb_decode(payload)
    slot = segs["read_slot"].astype(np.int64)
    orig = si[ko[slot] + segs["input_index"]]
    src = a[ao[slot] + orig]
    assert np.array_equal(rec["query_id"], src["read_id"])
    assert np.array_equal(rec["chain_id"], segs["chain_id"]) and np.array_equal(rec["segment_id"], segs["segment_id"])
    assert np.array_equal(rec["candidate_id"], orig)
    for f in ("read_begin", "read_end", "edge_id", "edge_left_node", "edge_right_node", "edge_end", "score"):
        assert np.array_equal(rec[f], src[f]), f
    assert rec["seq_divergence"].tobytes() == src["seq_divergence"].tobytes()
    # header fields, little-endian at the documented offsets
    assert payload[:8] == b"CUFRALB0"
    assert struct.unpack_from("<IIII", payload, 8) == (64, 92, 0, 0)
    assert struct.unpack_from("<QQQ", payload, 24) == (len(ao) - 1, len(segs), len(a))
    assert struct.unpack_from("<II", payload, 56) == (1, 0)


@pytest.mark.parametrize("mutate,msg", [
    (lambda b: b"CUFRALB1" + b[8:], "magic"),
    (lambda b: b[:8] + struct.pack("<I", 72) + b[12:], "header size"),
    (lambda b: b[:12] + struct.pack("<I", 96) + b[16:], "record size"),
    (lambda b: b[:16] + struct.pack("<I", 1) + b[20:], "version"),
    (lambda b: b[:56] + struct.pack("<I", 2) + b[60:], "output mode"),
    (lambda b: b[:60] + struct.pack("<I", 7) + b[64:], "reserved"),
    (lambda b: b[:-1], "payload size"),
    (lambda b: b[:48] + struct.pack("<II", 9, 1) + b[56:], "minimum"),
    (lambda b: b[:40], "shorter"),
])
def test_ralb_validation_fails_closed(oracle_mod, mutate, msg):
    a, ao, si, ko, segs, so = _chains(oracle_mod)
    payload = W.ralb_encode(segs, so, si, ko, a, ao)
    with pytest.raises(W.WireError, match=msg):
        W.validate_ralb(mutate(payload))


def test_ralb_expectations_and_divergence(oracle_mod):
    a, ao, si, ko, segs, so = _chains(oracle_mod)
    payload = W.ralb_encode(segs, so, si, ko, a, ao)
    with pytest.raises(W.WireError, match="fixture count"):
        W.validate_ralb(payload, expected_fixtures=1)
    with pytest.raises(W.WireError, match="SHA-256"):
        W.validate_ralb(payload, expected_sha256="0" * 64)
    bad = bytearray(payload)
    bad[64 + 88:64 + 92] = struct.pack("<f", float("nan"))
    with pytest.raises(W.WireError, match="non-finite"):
        W.validate_ralb(bytes(bad))


def test_ovlb_round_trip_and_canonical_hash(oracle_mod):
    from cuflye_amd.overlap import OverlapConfig
    import cuflye_amd.synth as S
    r = S.config1(scale=0.02, seed=9)
    orc = oracle_mod.Oracle(r.bases, r.offsets)
    orc.build_index(15)
    ids = np.array([1, -1, 2, -2, 3, -3, 4, -4], np.int64)
    rows, roff = orc.overlaps(ids, OverlapConfig(kmer_size=15, minimum_overlap=2000))
    assert len(rows) > 0
    rec = W.ovlb_decode(W.ovlb_encode(rows, roff))
    for f in W.OVLB_RECORD.names:
        assert rec[f].tobytes() == rows[f].tobytes(), f
    perm = np.random.default_rng(0).permutation(len(rec))
    assert W.overlap_canonical_sha256(rec) == W.overlap_canonical_sha256(rec[perm])


def test_canb_streaming_writer(tmp_path, oracle_mod):
    import cuflye_amd.synth as S
    r = S.config1(scale=0.02, seed=9)
    orc = oracle_mod.Oracle(r.bases, r.offsets)
    orc.build_index(15)
    w = W.CandidateWriter(tmp_path / "c.bin")
    allr = []
    for batch in ([1, -1], [2, -2, 3]):
        recs = [orc.candidates(q)[0] for q in batch]
        off = np.cumsum([0] + [len(x) for x in recs])
        cat = np.concatenate(recs)
        w.add(cat, off)
        allr.append(cat)
    digest = w.close()
    back = W.canb_read(tmp_path / "c.bin")
    cat = np.concatenate(allr)
    assert len(back) == len(cat)
    for f in W.CANB_RECORD.names:
        assert np.array_equal(back[f], cat[f]), f
    raw = (tmp_path / "c.bin").read_bytes()
    assert digest == hashlib.sha256(raw[64:]).hexdigest()
    assert struct.unpack_from("<QQ", raw, 24) == (5, len(cat))
