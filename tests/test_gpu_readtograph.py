"""Read-to-graph chaining end to end on a synthetic repeat graph (scaled
config 4): edge-sequence minimizer index, quickSeqOverlaps (every read k-mer
joined, read-to-graph detector shape), chain-input filter + sort,
chainReadAlignments - GPU vs oracle, row for row; and the minimizer index's
bucket contents through the candidate lists of the edges themselves."""
import numpy as np
import pytest

import cuflye_amd.synth as S
from cuflye_amd import readtograph as R

pytestmark = pytest.mark.gpu


def test_read_to_graph_config4_small(gpu, oracle_mod):
    g = S.config4(scale=0.02, seed=41, coverage=6.0)
    reads = g.reads
    print(f"[c4] edges={len(g.left)} reads={reads.n_reads}", flush=True)
    prm = R.ReadToGraphParams()
    # edge index: all k-mers solid, none repetitive, minimizer positions
    gpu.store_reads(g.edges_bases, g.edges_offsets)
    ginfo = gpu.build_index_minimizers(prm.kmer_size, prm.minimizer_window, solid_min=1, rep_max=1 << 20)
    ids = np.arange(1, reads.n_reads + 1, dtype=np.int64)
    cfg = prm.detector()
    rows, roff = gpu.get_seq_overlaps_external(reads.bases, reads.offsets, ids, cfg)
    orc = oracle_mod.Oracle(g.edges_bases, g.edges_offsets)
    oinfo = orc.build_index(prm.kmer_size, solid_min=1, rep_max=1 << 20, minimizer_window=prm.minimizer_window)
    assert ginfo.indexed_positions == oinfo["positions"] and ginfo.minimizer_window == prm.minimizer_window
    assert ginfo.repetitive_kmers == oinfo["repetitive"]
    erows, eroff = orc.overlaps(ids, cfg, threads=8, seqs=(reads.bases, reads.offsets))
    assert np.array_equal(roff, eroff)
    assert rows.tobytes() == erows.tobytes()
    assert len(rows) > 0
    aln, aoff = R.edge_alignments(rows, roff, g.left, g.right)
    gi, gk, gs, go = gpu.read_align_chains(aln, aoff, R.LONG_EDGE, R.BIG_ALN, prm.chain())
    ei, ek, es, eo = oracle_mod.read_align_chains(aln, aoff, R.LONG_EDGE, R.BIG_ALN, prm.chain())
    assert np.array_equal(gk, ek) and np.array_equal(gi, ei) and np.array_equal(go, eo)
    assert gs.tobytes() == es.tobytes()
    lines = R.read_alignment_rows(gs, go, gi, gk, aln, aoff)
    assert len(lines) == len(gs) > 0
    # chains through repeat edges join unique edges across shared nodes
    assert (np.diff(go) > 1).any() or len(gs) > 0


@pytest.mark.parametrize("window", [1, 5, 12])
def test_minimizer_index_buckets_match_oracle(gpu, oracle_mod, window):
    g = S.config4(scale=0.01, seed=43, coverage=1.0)
    n = len(g.left)
    print(f"[minimizer index] edges={n} window={window}", flush=True)
    gpu.store_reads(g.edges_bases, g.edges_offsets)
    ginfo = gpu.build_index_minimizers(15, window, solid_min=1, rep_max=1 << 20)
    orc = oracle_mod.Oracle(g.edges_bases, g.edges_offsets)
    oinfo = orc.build_index(15, solid_min=1, rep_max=1 << 20, minimizer_window=window)
    assert ginfo.indexed_positions == oinfo["positions"]
    # candidate lists of the edges (both strands) read every bucket they hit
    # in bucket order: equal to the oracle's in content and emission order
    ids = np.concatenate([np.arange(1, n + 1), -np.arange(1, n + 1)]).astype(np.int64)
    m, moff, f, foff = gpu.collect_candidates(ids)
    for i, q in enumerate(ids):
        em, ef = orc.candidates(int(q), sorted_=False)
        got = m[moff[i]:moff[i + 1]]
        assert got.tobytes() == em.tobytes(), f"q={q}"
