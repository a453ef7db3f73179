"""ctypes bindings of include/flyehip.h (libflyehip.so) and of the synthetic
workload generators (libflyesynth.so).

The record layouts below are the C structs of include/flyehip.h, as numpy
structured dtypes so results come back as arrays without copies per row.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_DIR = Path(__file__).resolve().parent / "lib"

FLYEHIP_OK = 0
FLYEHIP_ERR_ARG = 1
FLYEHIP_ERR_HIP = 2
FLYEHIP_ERR_BUDGET = 3
FLYEHIP_ERR_CAPACITY = 4
FLYEHIP_ERR_STATE = 5
FLYEHIP_ERR_UNSUPPORTED = 6

# Every entry point include/flyehip.h declares.
EXPORTED = (
    "flyehip_open", "flyehip_close", "flyehip_last_error", "flyehip_abi_version",
    "flyehip_get_stats", "flyehip_set_tuning", "flyehip_store_reads",
    "flyehip_build_index", "flyehip_build_index_minimizers", "flyehip_collect_candidates",
    "flyehip_collect_sorted_candidates", "flyehip_overlaps",
    "flyehip_overlaps_external", "flyehip_overlaps_per_query",
    "flyehip_chain_lists", "flyehip_read_align_chains", "flyehip_divergence_logf",
    "flyehip_read_align_chains_noalloc", "flyehip_load_index", "flyehip_overlaps_submit",
    "flyehip_overlaps_collect", "flyehip_overlaps_release", "flyehip_candidates_canonical_sha256",
)


class IndexParams(C.Structure):
    _fields_ = [("kmer_size", C.c_int32), ("solid_min", C.c_int32),
                ("rep_max", C.c_int32), ("repeat_kmer_rate", C.c_float)]


class IndexInfo(C.Structure):
    _fields_ = [("kmer_size", C.c_int32), ("solid_min", C.c_int32), ("rep_max", C.c_int32),
                ("minimizer_window", C.c_int32), ("distinct_kmers", C.c_int64),
                ("solid_kmers", C.c_int64), ("repetitive_kmers", C.c_int64),
                ("indexed_positions", C.c_int64), ("hash_slots", C.c_int64),
                ("device_bytes", C.c_int64)]


class OverlapParams(C.Structure):
    _fields_ = [("kmer_size", C.c_int32), ("minimum_overlap", C.c_int32),
                ("maximum_jump", C.c_int32), ("max_jump_gap", C.c_int32),
                ("gap_jump_threshold", C.c_int32), ("large_gap_penalty", C.c_float),
                ("small_gap_penalty", C.c_float), ("min_kmer_survival_rate", C.c_float),
                ("maximum_overhang", C.c_int32), ("check_overhang", C.c_int32),
                ("force_local", C.c_int32), ("only_max_ext", C.c_int32),
                ("max_overlaps", C.c_int32), ("max_divergence", C.c_float),
                ("sample_rate", C.c_float), ("keep_alignment", C.c_int32),
                ("nucl_alignment", C.c_int32), ("partition_bad_mappings", C.c_int32)]


class ChainParams(C.Structure):
    _fields_ = [("maximum_jump", C.c_int32), ("max_read_overlap", C.c_int32),
                ("minimum_overlap", C.c_int32), ("max_separation", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("ms_total", C.c_double), ("ms_join", C.c_double), ("ms_sort", C.c_double),
                ("ms_group", C.c_double), ("ms_chain", C.c_double), ("ms_extract", C.c_double),
                ("queries", C.c_int64), ("windows", C.c_int64), ("matches", C.c_int64),
                ("groups", C.c_int64), ("overlaps", C.c_int64), ("batches", C.c_int64),
                ("device_bytes_peak", C.c_int64), ("tie_queries", C.c_int64),
                ("kernel_ms", C.c_double * 16), ("kernel_launches", C.c_int64 * 16),
                ("kernel_bytes", C.c_double * 16), ("dp_cells", C.c_int64), ("exact_walk_groups", C.c_int64)]

    def as_dict(self) -> dict:
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if name.startswith("kernel_") else v
        return out


# Kernels of flyehip_stats_t.kernel_* (FLYEHIP_K_* order in include/flyehip.h).
KERNEL_FAMILIES = ("join_count", "join_fill", "radix", "rank_keys", "leaf", "heap", "prefilter", "dp", "walk",
                   "round_count", "round_rank", "round_swap", "round_other", "gather", "score_keys", "claim")
KERNEL_NAMES = {
    "join_count": "joinCountKernel",
    "join_fill": "joinFillKernel",
    "radix": "matchCountSortKernel (hipcub DeviceSegmentedRadixSort above its target-id limit)",
    "rank_keys": "tieFlagsKernel + rankKeysKernel",
    "leaf": "segLeafKernel",
    "heap": "segHeapLdsKernel (+ segHeapGlobalKernel)",
    "prefilter": "chainRelevanceKernel (with the tie replay) + chainPrefilterKernel",
    "dp": "chainDpKernel<G>",
    "walk": "chainWalkKernel<size class, onlyMaxExt>",
    "round_count": "segCountKernel",
    "round_rank": "segRankKernel",
    "round_swap": "segSwapKernel",
    "round_other": "segMedian/Map/SplitKernel + chunk scans",
    "gather": "gatherHitsKernel",
    "score_keys": "chainScoreKeysKernel + chainScorePushKernel",
    "claim": "chainClaimKernel<onlyMaxExt>",
}
# Kernels that are one HIP kernel (a roofline names one of these).
SINGLE_KERNELS = ("join_count", "join_fill", "leaf", "heap", "dp", "walk", "round_count", "round_rank", "round_swap",
                  "gather", "claim")


CANDIDATE_DTYPE = np.dtype([("query_id", "<i8"), ("target_id", "<i8"), ("kmer", "<u8"),
                            ("query_pos", "<i4"), ("target_pos", "<i4")])
OVERLAP_DTYPE = np.dtype([("cur_id", "<i8"), ("ext_id", "<i8"), ("cur_begin", "<i4"),
                          ("cur_end", "<i4"), ("cur_len", "<i4"), ("ext_begin", "<i4"),
                          ("ext_end", "<i4"), ("ext_len", "<i4"), ("score", "<i4"),
                          ("seq_divergence", "<f4"), ("chain_length", "<i4"),
                          ("reserved", "<i4")])
EDGE_ALIGNMENT_DTYPE = np.dtype([("read_id", "<i8"), ("edge_id", "<i8"), ("edge_seq_id", "<i8"),
                                 ("read_begin", "<i4"), ("read_end", "<i4"), ("read_len", "<i4"),
                                 ("edge_left_node", "<i4"), ("edge_right_node", "<i4"),
                                 ("edge_begin", "<i4"), ("edge_end", "<i4"), ("edge_len", "<i4"),
                                 ("score", "<i4"), ("seq_divergence", "<f4")])
SEGMENT_DTYPE = np.dtype([("read_slot", "<i4"), ("chain_id", "<i4"), ("segment_id", "<i4"),
                          ("input_index", "<i4"), ("chain_score", "<i4"), ("reserved", "<i4")])

assert CANDIDATE_DTYPE.itemsize == 32
assert OVERLAP_DTYPE.itemsize == 56
assert EDGE_ALIGNMENT_DTYPE.itemsize == 64
assert SEGMENT_DTYPE.itemsize == 24

_P = C.c_void_p


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def load_flyehip(path: Path | None = None) -> C.CDLL:
    """Loads libflyehip.so (or the build named by FLYEHIP_LIBRARY); raises when
    it is not built (no CPU fallback)."""
    path = path or Path(os.environ.get("FLYEHIP_LIBRARY") or (LIB_DIR / "libflyehip.so"))
    if not path.exists():
        raise RuntimeError(f"{path} is not built: run __graft_entry__.build() "
                           "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(str(path))
    i32, i64, u64 = C.c_int32, C.c_int64, C.c_uint64
    sig = {
        "flyehip_open": ([C.c_int, u64, C.POINTER(_P)], C.c_int),
        "flyehip_close": ([_P], C.c_int),
        "flyehip_last_error": ([_P], C.c_char_p),
        "flyehip_abi_version": ([], C.c_int),
        "flyehip_get_stats": ([_P, C.POINTER(Stats)], C.c_int),
        "flyehip_set_tuning": ([_P, i32, i32, i32], C.c_int),
        "flyehip_store_reads": ([_P, _P, _P, i32], C.c_int),
        "flyehip_build_index": ([_P, C.POINTER(IndexParams), C.POINTER(IndexInfo)], C.c_int),
        "flyehip_build_index_minimizers": ([_P, C.POINTER(IndexParams), i32, C.POINTER(IndexInfo)], C.c_int),
        "flyehip_collect_candidates": ([_P, _P, i32, _P, i64, _P, _P, _P, i64, _P, _P], C.c_int),
        "flyehip_collect_sorted_candidates": ([_P, _P, i32, _P, i64, _P, _P], C.c_int),
        "flyehip_overlaps": ([_P, _P, i32, C.POINTER(OverlapParams), _P, i64, _P, _P, _P, i64, _P],
                             C.c_int),
        "flyehip_overlaps_external": ([_P, _P, _P, _P, i32, C.POINTER(OverlapParams), _P, i64, _P, _P],
                                      C.c_int),
        "flyehip_overlaps_per_query": ([_P, _P, i32, C.POINTER(OverlapParams), _P, i64, _P, _P],
                                       C.c_int),
        "flyehip_chain_lists": ([_P, _P, _P, _P, i32, _P, _P, _P, _P, C.POINTER(OverlapParams), _P,
                                 i64, _P, _P], C.c_int),
        "flyehip_read_align_chains": ([_P, _P, _P, i32, i32, i32, C.POINTER(ChainParams), _P, _P, _P,
                                       i64, _P, _P], C.c_int),
        "flyehip_divergence_logf": ([_P, _P, i64, _P], C.c_int),
        "flyehip_read_align_chains_noalloc": ([_P, _P, _P, i32, i32, i32, C.POINTER(ChainParams), _P, _P,
                                               _P, i64, _P, _P], C.c_int),
        "flyehip_load_index": ([_P, i32, _P, i64, _P, _P, _P, _P, i64, C.POINTER(IndexInfo)], C.c_int),
        "flyehip_overlaps_submit": ([_P, _P, i32, C.POINTER(OverlapParams), i32, C.POINTER(u64)], C.c_int),
        "flyehip_overlaps_collect": ([_P, u64, i32, _P, i64, _P, _P, i64, _P], C.c_int),
        "flyehip_overlaps_release": ([_P, u64], C.c_int),
        "flyehip_candidates_canonical_sha256": ([_P, _P, i32, _P, C.POINTER(i64)], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    return lib


def load_synth(path: Path | None = None) -> C.CDLL:
    path = path or (LIB_DIR / "libflyesynth.so")
    if not path.exists():
        raise RuntimeError(f"{path} is not built: run __graft_entry__.build()")
    lib = C.CDLL(str(path))
    d, i32, i64, u64 = C.c_double, C.c_int32, C.c_int64, C.c_uint64
    lib.syn_read_set.argtypes = [i64, d, i32, i32, d, i32, i32, i32, i32, d, d, d, d, d, i32, i32,
                                 d, d, d, d, d, u64]
    lib.syn_read_set.restype = _P
    lib.syn_reads_from.argtypes = [_P, i64, i32, d, d, i32, i32, d, d, d, d, d, u64]
    lib.syn_reads_from.restype = _P
    lib.syn_hit_lists.argtypes = [i32, d, i32, i32, i32, d, d, u64]
    lib.syn_hit_lists.restype = _P
    lib.syn_repeat_graph.argtypes = [i64, d, d, i32, i32, i32, i32, d, d, d, d, d, i32, i32, d, d, d,
                                     d, d, i32, i32, d, u64]
    lib.syn_repeat_graph.restype = _P
    lib.syn_destroy.argtypes = [_P]
    lib.syn_bases_size.argtypes = [_P, i32]
    lib.syn_bases_size.restype = i64
    lib.syn_count.argtypes = [_P, i32]
    lib.syn_count.restype = i32
    lib.syn_copy_bases.argtypes = [_P, i32, _P, _P]
    lib.syn_copy_read_truth.argtypes = [_P, _P, _P]
    lib.syn_copy_edges.argtypes = [_P, _P, _P, _P]
    lib.syn_hits_size.argtypes = [_P]
    lib.syn_hits_size.restype = i64
    lib.syn_copy_lists.argtypes = [_P, _P, _P, _P, _P, _P, _P, _P]
    return lib


ptr = _ptr
