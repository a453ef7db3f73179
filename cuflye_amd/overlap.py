"""Host-side mirror of the reference's overlap-detection seam, over the C ABI.

The reference selects the candidate backend from CUFLYE_CANDIDATE_BACKEND
(patches/flye/2.9.6/0002-cuflye-candidate-backend-seam.patch file lines 29-43,
0003-cuflye-cuda-backend-stub.patch file lines 7-26): unset or "cpu" is
Flye's own loop, "cuda" the CUDA adapter, anything else throws
"Unsupported CUFLYE_CANDIDATE_BACKEND: <value>".  This build adds "hip",
served by libflyehip.so; it has no CPU path of its own, so "cpu" here is an
error too (the CPU path is Flye itself).  Device and budget come from
CUFLYE_CUDA_DEVICE and CUFLYE_CUDA_MEMORY_BUDGET_BYTES as in patch 0003.

FlyeHip wraps one flyehip_ctx: a resident read store and VertexIndex in HBM,
and the getSeqOverlaps stages (candidate collection, match sort, chaining).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import native as N


class FlyeHipError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"flyehip error {code}: {message}")
        self.code = code


def candidate_backend(env: dict | None = None) -> str:
    """Returns "hip" or raises like Flye's candidateBackend()."""
    env = os.environ if env is None else env
    value = env.get("CUFLYE_CANDIDATE_BACKEND", "")
    if value == "hip":
        return "hip"
    if value in ("", "cpu"):
        raise RuntimeError("CUFLYE_CANDIDATE_BACKEND=cpu selects Flye's own CPU loop; "
                           "this library serves CUFLYE_CANDIDATE_BACKEND=hip")
    raise RuntimeError(f"Unsupported CUFLYE_CANDIDATE_BACKEND: {value}")


@dataclass
class OverlapConfig:
    """getSeqOverlaps parameters (patch 0008:151-171; defaults of Flye 2.9.6's
    config as recorded in tests/golden/cuflye-m4b-overlap-chain-replay-dgx-aarch64.json:54-74)."""
    kmer_size: int = 15
    minimum_overlap: int = 5000
    maximum_jump: int = 1500
    max_jump_gap: int = 500
    gap_jump_threshold: int = 100
    large_gap_penalty: float = 2.0
    small_gap_penalty: float = 0.5
    min_kmer_survival_rate: float = 0.01
    maximum_overhang: int = 1500
    check_overhang: bool = True
    force_local: bool = False
    only_max_ext: bool = True
    max_overlaps: int = 0
    max_divergence: float = 1.0
    sample_rate: float = 1.0
    keep_alignment: bool = False
    nucl_alignment: bool = False
    partition_bad_mappings: bool = False

    def to_c(self) -> N.OverlapParams:
        p = N.OverlapParams()
        for name, _ in N.OverlapParams._fields_:
            setattr(p, name, getattr(self, name))
        return p

    @staticmethod
    def read_to_graph(kmer_size: int, small_alignment_threshold: int) -> "OverlapConfig":
        """The detector ReadAligner builds (tools/replay_read_to_graph_source_pack.py:
        25-31, 324-349): maxOverhang 0 (no overhang check), min overlap =
        small_alignment_threshold, onlyMaxExt false, maxDivergence 1."""
        return OverlapConfig(kmer_size=kmer_size, minimum_overlap=small_alignment_threshold,
                             maximum_overhang=0, check_overhang=False, only_max_ext=False,
                             max_divergence=1.0)


def pack_sequences(seqs: list[bytes] | list[str]) -> tuple[np.ndarray, np.ndarray]:
    """Concatenates sequences into (bases uint8, offsets int64[n+1])."""
    bs = [s.encode() if isinstance(s, str) else s for s in seqs]
    offsets = np.zeros(len(bs) + 1, dtype=np.int64)
    offsets[1:] = np.cumsum([len(b) for b in bs])
    bases = np.frombuffer(b"".join(bs), dtype=np.uint8).copy() if bs else np.zeros(0, np.uint8)
    return bases, offsets


class FlyeHip:
    """One context of libflyehip.so on one GPU."""

    def __init__(self, device: int | None = None, memory_budget: int | None = None):
        self.lib = N.load_flyehip()
        if device is None:
            device = int(os.environ.get("CUFLYE_CUDA_DEVICE", "0") or 0)
        if memory_budget is None:
            memory_budget = int(os.environ.get("CUFLYE_CUDA_MEMORY_BUDGET_BYTES", "0") or 0)
        ctx = C.c_void_p()
        rc = self.lib.flyehip_open(device, memory_budget, C.byref(ctx))
        self.ctx = ctx
        if rc != N.FLYEHIP_OK:
            msg = self.lib.flyehip_last_error(ctx).decode() if ctx else "open failed"
            if ctx:
                self.lib.flyehip_close(ctx)
            self.ctx = None
            raise FlyeHipError(rc, msg)
        self.index_info: N.IndexInfo | None = None
        self.n_reads = 0

    # -- lifecycle ----------------------------------------------------------
    def close(self) -> None:
        if self.ctx:
            self.lib.flyehip_close(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int) -> None:
        if rc != N.FLYEHIP_OK:
            raise FlyeHipError(rc, self.lib.flyehip_last_error(self.ctx).decode())

    def stats(self) -> dict:
        s = N.Stats()
        self._check(self.lib.flyehip_get_stats(self.ctx, C.byref(s)))
        return s.as_dict()

    def set_tuning(self, dp_lanes_per_group: int = 0, queries_per_batch: int = 0,
                   batches_in_flight: int = 0) -> None:
        self._check(self.lib.flyehip_set_tuning(self.ctx, dp_lanes_per_group, queries_per_batch,
                                                batches_in_flight))

    # -- read store + VertexIndex --------------------------------------------
    def store_reads(self, bases: np.ndarray, offsets: np.ndarray) -> None:
        bases = np.ascontiguousarray(bases, dtype=np.uint8)
        offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        n = len(offsets) - 1
        self._check(self.lib.flyehip_store_reads(self.ctx, N.ptr(bases), N.ptr(offsets), n))
        self.n_reads = n

    def build_index(self, kmer_size: int, solid_min: int = 0, rep_max: int = 0,
                    repeat_kmer_rate: float = 10.0) -> N.IndexInfo:
        p = N.IndexParams(kmer_size, solid_min, rep_max, repeat_kmer_rate)
        info = N.IndexInfo()
        self._check(self.lib.flyehip_build_index(self.ctx, C.byref(p), C.byref(info)))
        self.index_info = info
        return info

    def build_index_minimizers(self, kmer_size: int, window: int, solid_min: int = 0, rep_max: int = 0,
                               repeat_kmer_rate: float = 10.0) -> N.IndexInfo:
        """VertexIndex::buildIndexMinimizers: the thresholds of build_index,
        buckets holding the minimizer positions (window `window` k-mers) only."""
        p = N.IndexParams(kmer_size, solid_min, rep_max, repeat_kmer_rate)
        info = N.IndexInfo()
        self._check(self.lib.flyehip_build_index_minimizers(self.ctx, C.byref(p), window, C.byref(info)))
        self.index_info = info
        return info

    def load_index(self, kmer_size: int, kmers, bucket_offsets, entry_ids, entry_pos,
                   repetitive=()) -> N.IndexInfo:
        """Flye's own VertexIndex (flyehip_load_index): canonical k-mers,
        bucket offsets, entries (signed id, position) in Flye's bucket order,
        repetitive canonical k-mers."""
        km = np.ascontiguousarray(kmers, np.uint64)
        bo = np.ascontiguousarray(bucket_offsets, np.int64)
        ei = np.ascontiguousarray(entry_ids, np.int64)
        ep = np.ascontiguousarray(entry_pos, np.int32)
        rp = np.ascontiguousarray(repetitive, np.uint64)
        info = N.IndexInfo()
        self._check(self.lib.flyehip_load_index(self.ctx, kmer_size, N.ptr(km), len(km), N.ptr(bo), N.ptr(ei),
                                                N.ptr(ep), N.ptr(rp), len(rp), C.byref(info)))
        self.index_info = info
        return info

    # -- candidate collection ---------------------------------------------------
    def collect_candidates(self, query_ids) -> tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
        """Emission-order matches + filtered positions (collectCandidateMatches)."""
        q = np.ascontiguousarray(query_ids, dtype=np.int64)
        moff = np.zeros(len(q) + 1, np.int64)
        foff = np.zeros(len(q) + 1, np.int64)
        need_m = C.c_int64(0)
        need_f = C.c_int64(0)
        cap_m, cap_f = 1 << 16, 1 << 12
        while True:
            m = np.zeros(cap_m, N.CANDIDATE_DTYPE)
            f = np.zeros(cap_f, np.int32)
            rc = self.lib.flyehip_collect_candidates(self.ctx, N.ptr(q), len(q), N.ptr(m), cap_m,
                                                     N.ptr(moff), C.byref(need_m), N.ptr(f), cap_f,
                                                     N.ptr(foff), C.byref(need_f))
            if rc == N.FLYEHIP_ERR_CAPACITY:
                cap_m, cap_f = max(cap_m, need_m.value), max(cap_f, need_f.value)
                continue
            self._check(rc)
            return m[:need_m.value], moff, f[:need_f.value], foff

    def candidates_canonical_sha256(self, query_ids) -> tuple[str, int]:
        """candidate-record-v1 canonical SHA-256 (hex) and record count of the
        query records' matches, streamed (no dump held)."""
        q = np.ascontiguousarray(query_ids, dtype=np.int64)
        out = (C.c_uint8 * 32)()
        n = C.c_int64(0)
        self._check(self.lib.flyehip_candidates_canonical_sha256(self.ctx, N.ptr(q), len(q), out, C.byref(n)))
        return bytes(out).hex(), n.value

    def collect_sorted_candidates(self, query_ids) -> tuple[np.ndarray, np.ndarray]:
        """Matches after Flye's std::sort by (extId, curPos) (candidate-record-v1 order)."""
        q = np.ascontiguousarray(query_ids, dtype=np.int64)
        moff = np.zeros(len(q) + 1, np.int64)
        need = C.c_int64(0)
        cap = 1 << 16
        while True:
            m = np.zeros(cap, N.CANDIDATE_DTYPE)
            rc = self.lib.flyehip_collect_sorted_candidates(self.ctx, N.ptr(q), len(q), N.ptr(m), cap,
                                                            N.ptr(moff), C.byref(need))
            if rc == N.FLYEHIP_ERR_CAPACITY:
                cap = need.value
                continue
            self._check(rc)
            return m[:need.value], moff

    # -- overlaps ---------------------------------------------------------------
    def get_seq_overlaps(self, query_ids, config: OverlapConfig, with_div_stats: bool = False,
                         per_query: bool = False, row_capacity: int = 0, div_capacity: int = 0):
        """getSeqOverlaps for each query record: (rows, row_offsets[, divs, div_offsets])."""
        q = np.ascontiguousarray(query_ids, dtype=np.int64)
        p = config.to_c()
        cap = row_capacity or max(1024, 64 * len(q))
        dcap = div_capacity or 64 * len(q)
        need = C.c_int64(0)
        roff = np.zeros(len(q) + 1, np.int64)
        while True:
            rows = np.zeros(cap, N.OVERLAP_DTYPE)
            if per_query:
                rc = self.lib.flyehip_overlaps_per_query(self.ctx, N.ptr(q), len(q), C.byref(p),
                                                         N.ptr(rows), cap, N.ptr(roff), C.byref(need))
                divs = doff = None
            else:
                divs = np.zeros(dcap, np.float32) if with_div_stats else None
                doff = np.zeros(len(q) + 1, np.int64) if with_div_stats else None
                rc = self.lib.flyehip_overlaps(self.ctx, N.ptr(q), len(q), C.byref(p), N.ptr(rows), cap,
                                               N.ptr(roff), C.byref(need), N.ptr(divs), dcap if with_div_stats else 0,
                                               N.ptr(doff))
            if rc == N.FLYEHIP_ERR_CAPACITY:
                # rows_needed and div_offsets[n] hold what the call needs
                cap = max(cap, need.value)
                if doff is not None:
                    dcap = max(dcap, int(doff[-1]))
                continue
            self._check(rc)
            rows = rows[:need.value]
            if with_div_stats and not per_query:
                return rows, roff, divs[:doff[-1]], doff
            return rows, roff

    # -- submitted jobs (flyehip_overlaps_submit / _collect / _release) ---------
    def submit_overlaps(self, query_ids, config: OverlapConfig, with_div_stats: bool = False) -> int:
        """Queues getSeqOverlaps for the query records; returns a ticket."""
        q = np.ascontiguousarray(query_ids, dtype=np.int64)
        self._pending_params = config.to_c()
        t = C.c_uint64(0)
        self._check(self.lib.flyehip_overlaps_submit(self.ctx, N.ptr(q), len(q), C.byref(self._pending_params),
                                                     int(with_div_stats), C.byref(t)))
        return t.value

    def collect_overlaps(self, ticket: int, query_index: int, with_div_stats: bool = False):
        """Rows (and divStats values) of one submitted query record; blocks
        until its batch is done.  Safe to call from several threads."""
        cap, dcap = 64, 16
        need, dneed = C.c_int64(0), C.c_int64(0)
        while True:
            rows = np.zeros(cap, N.OVERLAP_DTYPE)
            divs = np.zeros(dcap, np.float32) if with_div_stats else None
            rc = self.lib.flyehip_overlaps_collect(self.ctx, ticket, query_index, N.ptr(rows), cap, C.byref(need),
                                                   N.ptr(divs), dcap, C.byref(dneed))
            if rc == N.FLYEHIP_ERR_CAPACITY:
                cap, dcap = max(cap, need.value), max(dcap, dneed.value)
                continue
            self._check(rc)
            if with_div_stats:
                return rows[:need.value], divs[:dneed.value]
            return rows[:need.value]

    def release_overlaps(self, ticket: int) -> None:
        self._check(self.lib.flyehip_overlaps_release(self.ctx, ticket))

    def get_seq_overlaps_external(self, seqs_bases: np.ndarray, seqs_offsets: np.ndarray, query_ids,
                                  config: OverlapConfig):
        q = np.ascontiguousarray(query_ids, dtype=np.int64)
        b = np.ascontiguousarray(seqs_bases, np.uint8)
        o = np.ascontiguousarray(seqs_offsets, np.int64)
        p = config.to_c()
        cap = max(1024, 64 * len(q))
        need = C.c_int64(0)
        roff = np.zeros(len(q) + 1, np.int64)
        while True:
            rows = np.zeros(cap, N.OVERLAP_DTYPE)
            rc = self.lib.flyehip_overlaps_external(self.ctx, N.ptr(b), N.ptr(o), N.ptr(q), len(q), C.byref(p),
                                                    N.ptr(rows), cap, N.ptr(roff), C.byref(need))
            if rc == N.FLYEHIP_ERR_CAPACITY:
                cap = need.value
                continue
            self._check(rc)
            return rows[:need.value], roff

    def chain_lists(self, cur, ext, list_offsets, cur_ids, ext_ids, cur_lens, ext_lens,
                    config: OverlapConfig):
        cur = np.ascontiguousarray(cur, np.int32)
        ext = np.ascontiguousarray(ext, np.int32)
        lo = np.ascontiguousarray(list_offsets, np.int64)
        ci = np.ascontiguousarray(cur_ids, np.int64)
        ei = np.ascontiguousarray(ext_ids, np.int64)
        cl = np.ascontiguousarray(cur_lens, np.int32)
        el = np.ascontiguousarray(ext_lens, np.int32)
        n = len(lo) - 1
        p = config.to_c()
        cap = max(1024, n)
        need = C.c_int64(0)
        roff = np.zeros(n + 1, np.int64)
        while True:
            rows = np.zeros(cap, N.OVERLAP_DTYPE)
            rc = self.lib.flyehip_chain_lists(self.ctx, N.ptr(cur), N.ptr(ext), N.ptr(lo), n, N.ptr(ci),
                                              N.ptr(ei), N.ptr(cl), N.ptr(el), C.byref(p), N.ptr(rows), cap,
                                              N.ptr(roff), C.byref(need))
            if rc == N.FLYEHIP_ERR_CAPACITY:
                cap = need.value
                continue
            self._check(rc)
            return rows[:need.value], roff

    def divergence_logf(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        out = np.empty_like(x)
        self._check(self.lib.flyehip_divergence_logf(self.ctx, N.ptr(x), len(x), N.ptr(out)))
        return out

    def read_align_chains(self, alignments: np.ndarray, aln_offsets, long_edge: int, big_aln: int,
                          params: N.ChainParams):
        a = np.ascontiguousarray(alignments, N.EDGE_ALIGNMENT_DTYPE)
        ao = np.ascontiguousarray(aln_offsets, np.int64)
        n = len(ao) - 1
        sorted_index = np.zeros(max(1, len(a)), np.int32)
        kept = np.zeros(n + 1, np.int64)
        soff = np.zeros(n + 1, np.int64)
        need = C.c_int64(0)
        cap = max(1024, len(a))
        while True:
            segs = np.zeros(cap, N.SEGMENT_DTYPE)
            rc = self.lib.flyehip_read_align_chains(self.ctx, N.ptr(a), N.ptr(ao), n, long_edge, big_aln,
                                                    C.byref(params), N.ptr(sorted_index), N.ptr(kept),
                                                    N.ptr(segs), cap, N.ptr(soff), C.byref(need))
            if rc == N.FLYEHIP_ERR_CAPACITY:
                cap = need.value
                continue
            self._check(rc)
            return sorted_index[:kept[-1]], kept, segs[:need.value], soff
