"""Read-to-graph alignment (subsystem c): the host-side loop of
ReadAligner::alignReads over the C ABI.

The edge sequences are indexed at their minimizer positions
(flyehip_build_index_minimizers, window MINIMIZER_WINDOW).  For each read:
quickSeqOverlaps against that index - every read k-mer joined, as Flye's
IterKmers loop does (docs/abi/read-to-graph-minimizer-source-pack-v0.md, "Full
Query-Hit TSV") - in the read-to-graph detector shape of getSeqOverlaps (no
overhang check, min overlap = small alignment threshold, containment primary
selection;
tools/replay_read_to_graph_source_pack.py:25-31, 324-349), the chain-input
filter and std::sort by curBegin (patch 0039:72-77, 241-251), then
chainReadAlignments (tools/replay_read_alignment_chains.py:64-139).  The
per-chain base-alignment divergence check (Flye's C4 step) is not modelled;
accepted chains are reported before it.

Edge model: edge sequence i (0-based) has signed id i + 1; an overlap with
the reverse strand (ext_id < 0) is an alignment to the complement edge, whose
left/right nodes are the complements of the forward edge's right/left nodes
(node v's complement is v + n_nodes).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import native as N
from .overlap import OverlapConfig

LONG_EDGE = 900   # patch 0025 context: static const int LONG_EDGE = 900
BIG_ALN = 500     # patch 0039:215
SMALL_ALN = 100   # patch 0039:214 (ReadAligner::alignReads: static const int SMALL_ALN = 100)
# Minimizer window of the edge-sequence index: Flye recorded 7,725 minimizers
# over the 8 toy-hifi queries of its read-to-graph source pack (density 0.332,
# tests/golden/cuflye-m6d-read-to-graph-minimizer-source-pack-dgx-aarch64.json),
# the 2 / (w + 1) density of window 5.
MINIMIZER_WINDOW = 5


@dataclass
class ReadToGraphParams:
    kmer_size: int = 15
    small_alignment_threshold: int = SMALL_ALN
    minimizer_window: int = MINIMIZER_WINDOW
    maximum_jump: int = 1500
    max_read_overlap: int = 50
    minimum_overlap: int = 1000
    max_separation: int = 500

    def detector(self) -> OverlapConfig:
        return OverlapConfig.read_to_graph(self.kmer_size, self.small_alignment_threshold)

    def chain(self) -> N.ChainParams:
        return N.ChainParams(self.maximum_jump, self.max_read_overlap, self.minimum_overlap,
                             self.max_separation)


def edge_alignments(rows: np.ndarray, row_offsets: np.ndarray, left: np.ndarray, right: np.ndarray):
    """OverlapRange rows (read vs edge sequence) -> EdgeAlignment records."""
    n_nodes = int(max(left.max(initial=0), right.max(initial=0))) + 1
    out = np.zeros(len(rows), N.EDGE_ALIGNMENT_DTYPE)
    eid = rows["ext_id"]
    e = np.abs(eid) - 1
    fwd = eid > 0
    out["read_id"] = rows["cur_id"]
    out["edge_id"] = eid
    out["edge_seq_id"] = eid
    out["read_begin"] = rows["cur_begin"]
    out["read_end"] = rows["cur_end"]
    out["read_len"] = rows["cur_len"]
    out["edge_left_node"] = np.where(fwd, left[e], right[e] + n_nodes)
    out["edge_right_node"] = np.where(fwd, right[e], left[e] + n_nodes)
    out["edge_begin"] = rows["ext_begin"]
    out["edge_end"] = rows["ext_end"]
    out["edge_len"] = rows["ext_len"]
    out["score"] = rows["score"]
    out["seq_divergence"] = rows["seq_divergence"]
    return out, row_offsets


def read_alignment_rows(segments: np.ndarray, seg_offsets: np.ndarray, sorted_index: np.ndarray,
                        kept_offsets: np.ndarray, alignments: np.ndarray, aln_offsets: np.ndarray) -> list[str]:
    """read-alignment-v1 rows (docs/abi/read-alignment-v1.md:27-47), chain ids numbered per read."""
    out = []
    for s in segments:
        r = int(s["read_slot"])
        orig = int(sorted_index[kept_offsets[r] + s["input_index"]])
        a = alignments[aln_offsets[r] + orig]
        div = float(a["seq_divergence"])
        out.append("\t".join(str(int(x)) for x in (s["chain_id"], s["segment_id"], a["read_id"], a["read_begin"],
                                                   a["read_end"], a["read_len"], a["edge_id"], a["edge_seq_id"],
                                                   a["edge_begin"], a["edge_end"], a["edge_len"], a["score"]))
                   + "\t" + ("0" if div == 0 else format(div, ".9g")))
    return out
