"""Compact binary wire formats: fixed-width little-endian records behind a
64-byte header, written field by field (not struct dumps), for the three
outputs of the path.

CUFRALB0  read-alignment chains, the reference's compact-binary-v0 layout
          (docs/abi/read-alignment-compact-binary-v0.md:20-66): header of
          magic, header/record bytes (64/92), version 0.0, fixture count,
          output records, total input records, min/max input records per
          fixture, output mode 1 (pre-divergence chains), reserved 0; one
          92-byte record per chain segment, fixtures (reads) in list order,
          segments in emission order.  validate_ralb() applies the checks of
          tools/validate_read_alignment_binary_payload.py.
CUFOVLB0  overlap rows (overlap-range-v1 fields plus the chain length), 52
          bytes a row, query records in list order.
CUFCANB0  candidate records (candidate-record-v1 fields), 32 bytes a record,
          written in batches so a dump of billions of records never sits in
          memory.

The canonical hash of each format sorts records as the text format's
canonical order does and hashes the binary records (overlap_canonical_sha256);
the candidate-record-v1 text hash itself is streamed by the library
(FlyeHip.candidates_canonical_sha256).
"""
from __future__ import annotations

import hashlib
import struct
from pathlib import Path

import numpy as np

HEADER_BYTES = 64

RALB_MAGIC = b"CUFRALB0"
RALB_RECORD = np.dtype({
    "names": ["query_id", "chain_id", "segment_id", "overlap_index", "candidate_id", "read_id", "read_begin",
              "read_end", "read_len", "edge_id", "edge_left_node", "edge_right_node", "edge_seq_id", "edge_begin",
              "edge_end", "edge_len", "score", "seq_divergence"],
    "formats": ["<i8", "<i4", "<i4", "<i4", "<i8", "<i8", "<i4", "<i4", "<i4", "<i8", "<i4", "<i4", "<i8", "<i4",
                "<i4", "<i4", "<i4", "<f4"],
    "offsets": [0, 8, 12, 16, 20, 28, 36, 40, 44, 48, 56, 60, 64, 72, 76, 80, 84, 88],
    "itemsize": 92,
})

OVLB_MAGIC = b"CUFOVLB0"
OVLB_RECORD = np.dtype({
    "names": ["cur_id", "ext_id", "cur_begin", "cur_end", "cur_len", "ext_begin", "ext_end", "ext_len", "score",
              "seq_divergence", "chain_length"],
    "formats": ["<i8", "<i8", "<i4", "<i4", "<i4", "<i4", "<i4", "<i4", "<i4", "<f4", "<i4"],
    "offsets": [0, 8, 16, 20, 24, 28, 32, 36, 40, 44, 48],
    "itemsize": 52,
})

CANB_MAGIC = b"CUFCANB0"
CANB_RECORD = np.dtype({
    "names": ["query_id", "query_pos", "kmer", "target_id", "target_pos"],
    "formats": ["<i8", "<i4", "<u8", "<i8", "<i4"],
    "offsets": [0, 8, 12, 20, 28],
    "itemsize": 32,
})


class WireError(ValueError):
    pass


def _header(magic: bytes, record_bytes: int, counts: tuple[int, int, int], minmax: tuple[int, int],
            mode: int) -> bytes:
    return (magic + struct.pack("<IIII", HEADER_BYTES, record_bytes, 0, 0) + struct.pack("<QQQ", *counts)
            + struct.pack("<IIII", minmax[0], minmax[1], mode, 0))


def _parse_header(data: bytes, magic: bytes, record_bytes: int, mode: int) -> dict:
    if len(data) < HEADER_BYTES:
        raise WireError(f"payload is shorter than the {HEADER_BYTES}-byte header")
    if data[:8] != magic:
        raise WireError(f"invalid magic {data[:8]!r}")
    hb, rb, vmaj, vmin = struct.unpack_from("<IIII", data, 8)
    groups, records, inputs = struct.unpack_from("<QQQ", data, 24)
    mn, mx, md, reserved = struct.unpack_from("<IIII", data, 48)
    if hb != HEADER_BYTES:
        raise WireError(f"unsupported header size {hb}")
    if rb != record_bytes:
        raise WireError(f"unsupported record size {rb}")
    if (vmaj, vmin) != (0, 0):
        raise WireError(f"unsupported version {vmaj}.{vmin}")
    if md != mode:
        raise WireError(f"unsupported output mode {md}")
    if reserved != 0:
        raise WireError("reserved header field is nonzero")
    if mn > mx:
        raise WireError("minimum input records exceed the maximum")
    if len(data) != hb + records * rb:
        raise WireError(f"payload size {len(data)} does not match expected {hb + records * rb}")
    return {"groups": groups, "records": records, "inputs": inputs, "min_inputs": mn, "max_inputs": mx}


# ---- CUFRALB0 ------------------------------------------------------------------

def ralb_encode(segments: np.ndarray, seg_offsets: np.ndarray, sorted_index: np.ndarray, kept_offsets: np.ndarray,
                alignments: np.ndarray, aln_offsets: np.ndarray) -> bytes:
    """Chain segments of flyehip_read_align_chains as a CUFRALB0 payload: one
    fixture per read (query id = its read id), overlap index = the segment's
    position in the read's sorted chain input, candidate id = the
    alignment's position in the read's input list."""
    n_reads = len(aln_offsets) - 1
    rec = np.zeros(len(segments), RALB_RECORD)
    if len(segments):
        slot = segments["read_slot"].astype(np.int64)
        inp = segments["input_index"].astype(np.int64)
        orig = sorted_index[kept_offsets[slot] + inp].astype(np.int64)
        a = alignments[aln_offsets[slot] + orig]
        rec["query_id"] = a["read_id"]
        rec["chain_id"] = segments["chain_id"]
        rec["segment_id"] = segments["segment_id"]
        rec["overlap_index"] = inp
        rec["candidate_id"] = orig
        for f in ("read_id", "read_begin", "read_end", "read_len", "edge_id", "edge_left_node", "edge_right_node",
                  "edge_seq_id", "edge_begin", "edge_end", "edge_len", "score", "seq_divergence"):
            rec[f] = a[f]
    per = np.diff(aln_offsets)
    mn = int(per.min()) if n_reads else 0
    mx = int(per.max()) if n_reads else 0
    head = _header(RALB_MAGIC, 92, (n_reads, len(rec), int(aln_offsets[-1] - aln_offsets[0])), (mn, mx), 1)
    return head + rec.tobytes()


def validate_ralb(data: bytes, expected_fixtures: int | None = None, expected_records: int | None = None,
                  expected_inputs: int | None = None, expected_sha256: str | None = None) -> dict:
    """The checks of tools/validate_read_alignment_binary_payload.py; also
    rejects non-finite divergences (the Flye-side rehydration rule of
    docs/abi/read-alignment-compact-binary-v0.md:94-99)."""
    h = _parse_header(data, RALB_MAGIC, 92, 1)
    digest = hashlib.sha256(data).hexdigest()
    for name, want, got in (("fixture count", expected_fixtures, h["groups"]),
                            ("output records", expected_records, h["records"]),
                            ("input records", expected_inputs, h["inputs"])):
        if want is not None and want != got:
            raise WireError(f"{name} {got} differs from the expected {want}")
    if expected_sha256 and digest != expected_sha256:
        raise WireError("SHA-256 mismatch")
    rec = np.frombuffer(data, RALB_RECORD, offset=HEADER_BYTES)
    if not np.all(np.isfinite(rec["seq_divergence"])):
        raise WireError("non-finite sequence divergence")
    h["sha256"] = digest
    return h


def ralb_decode(data: bytes) -> np.ndarray:
    validate_ralb(data)
    return np.frombuffer(data, RALB_RECORD, offset=HEADER_BYTES).copy()


# ---- CUFOVLB0 ------------------------------------------------------------------

def ovlb_encode(rows: np.ndarray, row_offsets: np.ndarray) -> bytes:
    rec = np.zeros(len(rows), OVLB_RECORD)
    for f in OVLB_RECORD.names:
        rec[f] = rows[f]
    per = np.diff(row_offsets)
    n = len(row_offsets) - 1
    head = _header(OVLB_MAGIC, 52, (n, len(rec), 0), (int(per.min()) if n else 0, int(per.max()) if n else 0), 2)
    return head + rec.tobytes()


def ovlb_decode(data: bytes) -> np.ndarray:
    _parse_header(data, OVLB_MAGIC, 52, 2)
    rec = np.frombuffer(data, OVLB_RECORD, offset=HEADER_BYTES)
    if not np.all(np.isfinite(rec["seq_divergence"])):
        raise WireError("non-finite sequence divergence")
    return rec.copy()


def overlap_canonical_sha256(rec: np.ndarray) -> str:
    """Binary records in overlap-range-v1 canonical order (numeric tuple of the
    ten dump fields, tools/validate_overlap_dump.py:130-207), hashed."""
    order = np.lexsort(tuple(rec[f] for f in ("chain_length", "seq_divergence", "score", "ext_len", "ext_end",
                                              "ext_begin", "ext_id", "cur_len", "cur_end", "cur_begin", "cur_id")))
    out = np.zeros(len(rec), OVLB_RECORD)
    for f in OVLB_RECORD.names:
        out[f] = rec[f][order]
    return hashlib.sha256(out.tobytes()).hexdigest()


# ---- CUFCANB0 ------------------------------------------------------------------

class CandidateWriter:
    """Streams candidate records to a CUFCANB0 file batch by batch; the header
    counts are patched in when the writer closes."""

    def __init__(self, path: str | Path):
        self.path = Path(path)
        self.f = open(self.path, "wb")
        self.f.write(b"\0" * HEADER_BYTES)
        self.records = 0
        self.queries = 0
        self.mn = None
        self.mx = 0
        self.hash = hashlib.sha256()

    def add(self, records: np.ndarray, offsets: np.ndarray) -> None:
        rec = np.zeros(len(records), CANB_RECORD)
        for f in CANB_RECORD.names:
            rec[f] = records[f]
        b = rec.tobytes()
        self.f.write(b)
        self.hash.update(b)
        per = np.diff(offsets)
        if len(per):
            self.mn = int(per.min()) if self.mn is None else min(self.mn, int(per.min()))
            self.mx = max(self.mx, int(per.max()))
        self.records += len(rec)
        self.queries += len(offsets) - 1

    def close(self) -> str:
        """Writes the header; returns the SHA-256 of the records in file order."""
        self.f.seek(0)
        self.f.write(_header(CANB_MAGIC, 32, (self.queries, self.records, 0), (self.mn or 0, self.mx), 3))
        self.f.close()
        return self.hash.hexdigest()


def canb_read(path: str | Path) -> np.ndarray:
    data = Path(path).read_bytes()
    _parse_header(data, CANB_MAGIC, 32, 3)
    return np.frombuffer(data, CANB_RECORD, offset=HEADER_BYTES).copy()
