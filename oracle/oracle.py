"""ctypes wrapper of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker that tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg compare the GPU library against.  It is a
restatement of Flye 2.9.6's overlap stage (see flye_oracle.cpp for the
reference lines each part follows) and never part of the product path.

Parity pins (what each oracle function is checked against):
  encode / reverse complement / standard form / id mapping
      - the reference fixtures kmer-encode-smoke-v0, kmer-join-smoke-v0,
        read-window-smoke-v0: expected candidates and their recorded SHA-256
        (c07761a6..., f0ef59da...); closed forms (tests/test_oracle.py).
  candidate loop + VertexIndex orientation
      - brute force over every (query window, target strand position) pair on
        tiny inputs (test_candidates_brute_force); the k=4 fixtures through
        Flye-ordered buckets reproduce their hashes (tests/test_load_index.py).
  candidate-record-v1 canonical hash
      - the fixtures' recorded SHA-256 (tests/test_canonical_hash.py).
  VertexIndex thresholds and buckets
      - a dictionary count of every k-mer on tiny reads, thresholds and bucket
        order by hand (test_index_brute_force).  Parity unpinned against Flye:
        the reference does not ship vertex_index.*.
  minimizer index
      - a brute force written from the definition, the 2 / (w + 1) density and
        Flye's recorded read-to-graph minimizer density, the full index
        filtered by the flags (tests/test_minimizers.py).  Selection parity
        unpinned against Flye: yieldMinimizers is not in the reference.
  std::sort points
      - real libstdc++ std::sort; against the reference's cpp_sort_in_place on
        tie-heavy keys (tests/golden/cpp_sort_v1.npz).
  prefilters, chaining DP (gap tiers, max_jump_gap, maxJump, early break),
  backtrack, overlapTest, divergence, filtered positions, selection, divStats
      - hand-worked closed forms (tests/test_oracle_closed_form.py); golden
        vectors of the reference's replay_group, chain_target and overlap_test
        (tests/golden/*.npz, tests/test_reference_vectors.py).
  chainReadAlignments
      - the reference fixture read-alignment-replay-fixture-v0/simple; golden
        vectors of chain_read_alignments (read_align_v1.npz).
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"

_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        build()
    L = C.CDLL(str(LIB_PATH))
    P, i32, i64, u64, f32 = C.c_void_p, C.c_int32, C.c_int64, C.c_uint64, C.c_float
    L.orc_encode_kmer.argtypes = [C.c_char_p, i32]
    L.orc_encode_kmer.restype = u64
    L.orc_reverse_complement.argtypes = [u64, i32]
    L.orc_reverse_complement.restype = u64
    L.orc_standard_form.argtypes = [u64, i32, P]
    L.orc_standard_form.restype = u64
    L.orc_signed_to_internal.argtypes = [i64]
    L.orc_signed_to_internal.restype = C.c_uint32
    L.orc_internal_to_signed.argtypes = [C.c_uint32]
    L.orc_internal_to_signed.restype = i64
    L.orc_create.argtypes = [P, P, i32]
    L.orc_create.restype = P
    L.orc_destroy.argtypes = [P]
    L.orc_build_index.argtypes = [P, i32, i32, i32, f32, P]
    L.orc_build_index.restype = C.c_int
    L.orc_build_index_minimizers.argtypes = [P, i32, i32, i32, i32, f32, P]
    L.orc_build_index_minimizers.restype = C.c_int
    L.orc_minimizer_flags.argtypes = [P, i64, i32, i32, P]
    L.orc_minimizer_flags.restype = i64
    L.orc_histogram.argtypes = [P, P, i64]
    L.orc_histogram.restype = i64
    L.orc_candidates.argtypes = [P, i64, P, i32, i32, P, i64, P, i64, P]
    L.orc_candidates.restype = i64
    L.orc_overlaps.argtypes = [P, P, i32, P, P, P, i32, P, i64, P, P, i64, P]
    L.orc_overlaps.restype = i64
    L.orc_chain_lists.argtypes = [P, P, P, i32, P, P, P, P, P, i32, P, i64, P]
    L.orc_chain_lists.restype = i64
    L.orc_read_align_chains.argtypes = [P, P, i32, i32, i32, P, P, P, P, i64, P]
    L.orc_read_align_chains.restype = i64
    L.orc_std_sort_perm.argtypes = [P, i64, P]
    L.orc_logf.argtypes = [f32]
    L.orc_logf.restype = f32
    L.orc_logf_array.argtypes = [P, i64, P]
    L.orc_logf_array.restype = None
    L.orc_chain_dp.argtypes = [P, P, i32, P, i32, i32, P, P]
    L.orc_chain_dp.restype = None
    L.orc_chain_group.argtypes = [P, P, i32, i64, i32, i64, i32, P, i32, P, P, i64]
    L.orc_chain_group.restype = i64
    L.orc_overlap_test.argtypes = [P, P]
    L.orc_overlap_test.restype = i32
    L.orc_div_stats.argtypes = [P, i64, i32, P, i64]
    L.orc_div_stats.restype = i64
    L.orc_index_tables.argtypes = [P, P, P, P, P, P, P]
    L.orc_index_tables.restype = None
    L.orc_load_index.argtypes = [P, i32, P, i64, P, P, P, P, i64]
    L.orc_load_index.restype = C.c_int
    L.orc_candidates_canonical_sha256.argtypes = [P, P, i32, i32, P, P]
    L.orc_candidates_canonical_sha256.restype = C.c_int
    _lib = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _dt():
    from cuflye_amd import native as N
    return N


def encode(kmer: str) -> int:
    return lib().orc_encode_kmer(kmer.encode(), len(kmer))


def reverse_complement(x: int, k: int) -> int:
    return lib().orc_reverse_complement(x, k)


def standard_form(x: int, k: int) -> tuple[int, bool]:
    f = C.c_int32(0)
    c = lib().orc_standard_form(x, k, C.byref(f))
    return c, bool(f.value)


def signed_to_internal(sid: int) -> int:
    return lib().orc_signed_to_internal(sid)


def internal_to_signed(iid: int) -> int:
    return lib().orc_internal_to_signed(iid)


def std_sort_perm(keys: np.ndarray) -> np.ndarray:
    k = np.ascontiguousarray(keys, np.uint64)
    out = np.empty(len(k), np.int32)
    lib().orc_std_sort_perm(_p(k), len(k), _p(out))
    return out


def logf(x: float) -> float:
    return lib().orc_logf(x)


def logf_array(x: np.ndarray) -> np.ndarray:
    """glibc logf of every element (float32)."""
    a = np.ascontiguousarray(x, np.float32)
    out = np.empty_like(a)
    lib().orc_logf_array(_p(a), len(a), _p(out))
    return out


def _params(params):
    return params.to_c() if hasattr(params, "to_c") else params


def minimizer_flags(seq, k: int, window: int) -> np.ndarray:
    """One byte per k-mer position of an ASCII sequence: 1 at its minimizers."""
    b = np.ascontiguousarray(np.frombuffer(seq, np.uint8) if isinstance(seq, (bytes, bytearray)) else seq, np.uint8)
    out = np.zeros(max(0, len(b) - k + 1), np.uint8)
    n = lib().orc_minimizer_flags(_p(b), len(b), k, window, _p(out))
    return out[:n]


def chain_dp(cur, ext, params, ext_sorted: bool = False, early_break: bool = True):
    """Score and backtrack tables of one group's DP, hits in their final order."""
    c = np.ascontiguousarray(cur, np.int32)
    e = np.ascontiguousarray(ext, np.int32)
    score = np.zeros(len(c), np.int32)
    bt = np.zeros(len(c), np.int32)
    p = _params(params)
    lib().orc_chain_dp(_p(c), _p(e), len(c), C.byref(p), int(ext_sorted), int(early_break), _p(score), _p(bt))
    return score, bt


def chain_group(cur, ext, cur_id: int, cur_len: int, ext_id: int, ext_len: int, filtered, params):
    """detectedOverlaps of one (query, target) group, hits in (extId, curPos) order."""
    N = _dt()
    c = np.ascontiguousarray(cur, np.int32)
    e = np.ascontiguousarray(ext, np.int32)
    f = np.ascontiguousarray(filtered if filtered is not None else [], np.int32)
    p = _params(params)
    cap = 16
    while True:
        rows = np.zeros(cap, N.OVERLAP_DTYPE)
        n = lib().orc_chain_group(_p(c), _p(e), len(c), cur_id, cur_len, ext_id, ext_len, _p(f), len(f),
                                  C.byref(p), _p(rows), cap)
        if n <= cap:
            return rows[:n]
        cap = n


def overlap_test(fields: dict, params) -> bool:
    """Flye's overlapTest of one proposal; `fields` names OVERLAP_DTYPE fields
    (cur_id, cur_begin, cur_end, cur_len, ext_id, ext_begin, ext_end, ext_len)."""
    N = _dt()
    r = np.zeros(1, N.OVERLAP_DTYPE)
    for name, value in fields.items():
        r[0][name] = value
    p = _params(params)
    return bool(lib().orc_overlap_test(_p(r), C.byref(p)))


def div_stats(rows, cur_len: int) -> np.ndarray:
    """OvlpDivStats values of one query's detected overlaps (rows in order)."""
    N = _dt()
    r = np.ascontiguousarray(rows, N.OVERLAP_DTYPE)
    out = np.zeros(cur_len // 10000 + 2, np.float32)
    n = lib().orc_div_stats(_p(r), len(r), cur_len, _p(out), len(out))
    return out[:n]


class Oracle:
    """Read store + VertexIndex + getSeqOverlaps restated on the CPU."""

    def __init__(self, bases: np.ndarray, offsets: np.ndarray):
        self.bases = np.ascontiguousarray(bases, np.uint8)
        self.offsets = np.ascontiguousarray(offsets, np.int64)
        self.h = lib().orc_create(_p(self.bases), _p(self.offsets), len(self.offsets) - 1)
        self.info = None

    def __del__(self):
        try:
            if self.h:
                lib().orc_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def build_index(self, k: int, solid_min: int = 0, rep_max: int = 0, rate: float = 10.0,
                    minimizer_window: int = 0) -> dict:
        """countKmers + buildIndex, or buildIndexMinimizers with a window
        (buckets of minimizer positions only)."""
        out = np.zeros(8, np.int64)
        lib().orc_build_index_minimizers(self.h, k, minimizer_window, solid_min, rep_max, rate, _p(out))
        self.info = dict(solid_min=int(out[0]), rep_max=int(out[1]), distinct=int(out[2]),
                         solid=int(out[3]), repetitive=int(out[4]), positions=int(out[5]))
        return self.info

    def load_index(self, k: int, kmers, offsets, ids, pos, repetitive=()):
        """An index given as tables (the arguments of flyehip_load_index)."""
        km = np.ascontiguousarray(kmers, np.uint64)
        off = np.ascontiguousarray(offsets, np.int64)
        ei = np.ascontiguousarray(ids, np.int64)
        ep = np.ascontiguousarray(pos, np.int32)
        rp = np.ascontiguousarray(repetitive, np.uint64)
        lib().orc_load_index(self.h, k, _p(km), len(km), _p(off), _p(ei), _p(ep), _p(rp), len(rp))

    def candidates_canonical_sha256(self, query_ids, threads: int = 1) -> tuple[str, int]:
        """candidate-record-v1 canonical SHA-256 (hex) and record count."""
        q = np.ascontiguousarray(query_ids, np.int64)
        out = (C.c_uint8 * 32)()
        n = C.c_int64(0)
        if lib().orc_candidates_canonical_sha256(self.h, _p(q), len(q), threads, out, C.byref(n)) != 0:
            raise RuntimeError("SHA-256 failed")
        return bytes(out).hex(), n.value

    def index_tables(self) -> dict:
        """kmers (ascending), bucket starts, entry ids / positions in bucket
        order and the repetitive k-mers of the built index."""
        counts = np.zeros(3, np.int64)
        lib().orc_index_tables(self.h, _p(counts), None, None, None, None, None)
        kmers = np.zeros(counts[0], np.uint64)
        start = np.zeros(counts[0] + 1, np.int64)
        ids = np.zeros(counts[1], np.uint32)
        pos = np.zeros(counts[1], np.int32)
        rep = np.zeros(counts[2], np.uint64)
        lib().orc_index_tables(self.h, _p(counts), _p(kmers), _p(start), _p(ids), _p(pos), _p(rep))
        return dict(kmers=kmers, start=start, ids=ids, pos=pos, repetitive=rep)

    def histogram(self) -> np.ndarray:
        n = lib().orc_histogram(self.h, None, 0)
        h = np.zeros(n, np.int64)
        lib().orc_histogram(self.h, _p(h), n)
        return h

    def candidates(self, query_id: int, sorted_: bool = False, seq: bytes | None = None):
        N = _dt()
        sbuf = None if seq is None else np.frombuffer(seq, np.uint8).copy()
        slen = 0 if seq is None else len(seq)
        nf = C.c_int64(0)
        n = lib().orc_candidates(self.h, query_id, _p(sbuf), slen, int(sorted_), None, 0, None, 0, C.byref(nf))
        out = np.zeros(n, N.CANDIDATE_DTYPE)
        filt = np.zeros(nf.value, np.int32)
        lib().orc_candidates(self.h, query_id, _p(sbuf), slen, int(sorted_), _p(out), n, _p(filt), len(filt),
                             C.byref(nf))
        return out, filt

    def overlaps(self, query_ids, params, threads: int = 1, seqs: tuple | None = None,
                 with_div_stats: bool = False):
        N = _dt()
        q = np.ascontiguousarray(query_ids, np.int64)
        roff = np.zeros(len(q) + 1, np.int64)
        doff = np.zeros(len(q) + 1, np.int64)
        sb, so = (None, None) if seqs is None else (np.ascontiguousarray(seqs[0], np.uint8),
                                                    np.ascontiguousarray(seqs[1], np.int64))
        p = params.to_c() if hasattr(params, "to_c") else params
        cap = 64 * len(q) + 1024
        while True:
            rows = np.zeros(cap, N.OVERLAP_DTYPE)
            divs = np.zeros(cap, np.float32)
            n = lib().orc_overlaps(self.h, _p(q), len(q), _p(sb), _p(so), C.byref(p), threads, _p(rows), cap,
                                   _p(roff), _p(divs), cap, _p(doff))
            if n <= cap and doff[-1] <= cap:
                break
            cap = max(n, int(doff[-1])) + 16
        if with_div_stats:
            return rows[:n], roff, divs[:doff[-1]], doff
        return rows[:n], roff


def chain_lists(cur, ext, offsets, cur_ids, ext_ids, cur_lens, ext_lens, params, threads: int = 1):
    N = _dt()
    args = [np.ascontiguousarray(a, t) for a, t in ((cur, np.int32), (ext, np.int32), (offsets, np.int64),
                                                     (cur_ids, np.int64), (ext_ids, np.int64),
                                                     (cur_lens, np.int32), (ext_lens, np.int32))]
    n = len(args[2]) - 1
    roff = np.zeros(n + 1, np.int64)
    p = params.to_c() if hasattr(params, "to_c") else params
    cap = n + 1024
    while True:
        rows = np.zeros(cap, N.OVERLAP_DTYPE)
        got = lib().orc_chain_lists(_p(args[0]), _p(args[1]), _p(args[2]), n, _p(args[3]), _p(args[4]),
                                    _p(args[5]), _p(args[6]), C.byref(p), threads, _p(rows), cap, _p(roff))
        if got <= cap:
            return rows[:got], roff
        cap = got + 16


def read_align_chains(alignments, aln_offsets, long_edge: int, big_aln: int, params):
    N = _dt()
    a = np.ascontiguousarray(alignments, N.EDGE_ALIGNMENT_DTYPE)
    ao = np.ascontiguousarray(aln_offsets, np.int64)
    n = len(ao) - 1
    si = np.zeros(max(1, len(a)), np.int32)
    ko = np.zeros(n + 1, np.int64)
    so = np.zeros(n + 1, np.int64)
    cap = len(a) * 4 + 64
    while True:
        segs = np.zeros(cap, N.SEGMENT_DTYPE)
        got = lib().orc_read_align_chains(_p(a), _p(ao), n, long_edge, big_aln, C.byref(params), _p(si), _p(ko),
                                          _p(segs), cap, _p(so))
        if got <= cap:
            return si[:ko[-1]], ko, segs[:got], so
        cap = got + 16
