"""The oracle against vectors written by the reference's own Python restatements
(tests/golden/gen_reference_vectors.py; no GPU).

  replay_group_v1   tools/replay_read_to_graph_source_pack.py:352-491 on tie-heavy
                    groups (libstdc++ sort order restated by the reference):
                    row order and integer fields exact; the reference computes
                    the divergence in double (math.log), the oracle in float
                    with glibc logf as Flye does, so those agree to 1e-6.
  chain_target_v1   tools/replay_overlap_chains.py:228-380 in eight parameter
                    families, float32 divergence through glibc logf: bit-exact.
  overlap_test_v1   tools/replay_overlap_chains.py:196-225: every case.
  read_align_v1     tools/replay_read_alignment_chains.py:64-139: accepted chains
                    (indices, scores) exact.
  cpp_sort_v1       tools/replay_read_to_graph_source_pack.py:95-181: the oracle's
                    real std::sort yields the same permutation of equal keys.
"""
from pathlib import Path

import numpy as np
import pytest

from cuflye_amd.overlap import OverlapConfig

GOLD = Path(__file__).resolve().parent / "golden"
DIV_TOL = 1e-6  # |float32 glibc logf path - float64 math.log path| of the divergence (max seen 3.3e-8)


def _load(name):
    return np.load(GOLD / name)


def read_to_graph_config(min_overlap: int) -> OverlapConfig:
    """The replay_group detector: maxOverhang 0, onlyMaxExt false, maxDiv 1."""
    return OverlapConfig.read_to_graph(15, int(min_overlap))


def test_replay_group_vectors(oracle_mod):
    d = _load("replay_group_v1.npz")
    ho, ro = d["hit_offsets"], d["row_offsets"]
    n = len(ho) - 1
    assert n >= 1000 and ro[-1] >= 5000
    for g in range(n):
        cur = d["cur"][ho[g]:ho[g + 1]]
        ext = d["ext"][ho[g]:ho[g + 1]]
        rows = oracle_mod.chain_group(cur, ext, int(d["cur_id"][g]), int(d["cur_len"][g]), int(d["ext_id"][g]),
                                      int(d["ext_len"][g]), None, read_to_graph_config(d["min_overlap"][g]))
        exp = d["rows_int"][ro[g]:ro[g + 1]]
        got = np.stack([rows["cur_begin"], rows["cur_end"], rows["ext_begin"], rows["ext_end"], rows["score"],
                        rows["chain_length"]], axis=1) if len(rows) else np.zeros((0, 6), np.int32)
        assert got.tolist() == exp.tolist(), f"group {g}"
        assert np.allclose(rows["seq_divergence"].astype(np.float64), d["rows_div_f64"][ro[g]:ro[g + 1]],
                           rtol=0, atol=DIV_TOL), f"group {g}"


def _ct_config(pi, pf) -> OverlapConfig:
    return OverlapConfig(kmer_size=int(pi[0]), minimum_overlap=int(pi[1]), maximum_jump=int(pi[2]),
                         max_jump_gap=int(pi[3]), gap_jump_threshold=int(pi[4]), maximum_overhang=int(pi[5]),
                         check_overhang=bool(pi[6]), force_local=bool(pi[7]), only_max_ext=bool(pi[8]),
                         large_gap_penalty=float(pf[0]), small_gap_penalty=float(pf[1]),
                         min_kmer_survival_rate=float(pf[2]), max_divergence=float(pf[3]),
                         sample_rate=float(pf[4]))


def chain_target_families():
    d = _load("chain_target_v1.npz")
    return [str(f) for f in d["families"]]


@pytest.mark.parametrize("family", chain_target_families())
def test_chain_target_vectors(oracle_mod, family):
    d = _load("chain_target_v1.npz")
    g_ = {k.split("__", 1)[1]: d[k] for k in d.files if k.startswith(family + "__")}
    cfg = _ct_config(g_["params_i"], g_["params_f"])
    ho, ro, fo = g_["hit_offsets"], g_["row_offsets"], g_["filtered_offsets"]
    for g in range(len(ho) - 1):
        rows = oracle_mod.chain_group(g_["cur"][ho[g]:ho[g + 1]], g_["ext"][ho[g]:ho[g + 1]], int(g_["cur_id"][g]),
                                      int(g_["cur_len"][g]), int(g_["ext_id"][g]), int(g_["ext_len"][g]),
                                      g_["filtered"][fo[g]:fo[g + 1]], cfg)
        exp = g_["rows"][ro[g]:ro[g + 1]]
        got = np.stack([rows["cur_begin"], rows["cur_end"], rows["ext_begin"], rows["ext_end"], rows["score"],
                        rows["seq_divergence"].view(np.int32)], axis=1) if len(rows) else np.zeros((0, 6), np.int32)
        assert got.tolist() == exp.tolist(), f"{family} group {g}"


def test_chain_target_vectors_cover_rows():
    """Every family yields rows (so each parameter shape is exercised)."""
    d = _load("chain_target_v1.npz")
    for f in chain_target_families():
        assert d[f + "__row_offsets"][-1] >= 20, f


def test_overlap_test_vectors(oracle_mod):
    d = _load("overlap_test_v1.npz")
    recs, params, expect = d["records"], d["params"], d["expect"]
    names = ("cur_id", "cur_begin", "cur_end", "cur_len", "ext_id", "ext_begin", "ext_end", "ext_len")
    for r, p, e in zip(recs, params, expect):
        cfg = OverlapConfig(minimum_overlap=int(p[0]), maximum_overhang=int(p[1]), check_overhang=bool(p[2]),
                            force_local=bool(p[3]))
        assert oracle_mod.overlap_test(dict(zip(names, (int(x) for x in r))), cfg) == bool(e), r.tolist()
    assert 0 < expect.sum() < len(expect)


def test_read_align_vectors(oracle_mod):
    from cuflye_amd import native as N
    d = _load("read_align_v1.npz")
    al, ao = d["alignments"], d["aln_offsets"]
    a = np.zeros(len(al), N.EDGE_ALIGNMENT_DTYPE)
    for j, f in enumerate(("read_begin", "read_end", "read_len", "edge_left_node", "edge_right_node", "edge_begin",
                           "edge_end", "edge_len", "score", "edge_id", "edge_seq_id")):
        a[f] = al[:, j]
    pv = d["params"]
    params = N.ChainParams(int(pv[0]), int(pv[1]), int(pv[2]), int(pv[3]))
    # every alignment is a chain input (long_edge above every edge length) and
    # read_begin values are distinct, so the sorted order is the given order
    si, ko, segs, so = oracle_mod.read_align_chains(a, ao, 1 << 30, 0, params)
    assert np.array_equal(ko, ao - ao[0])
    co, sidx, soff = d["chain_offsets"], d["segment_index"], d["segment_offsets"]
    for r in range(len(ao) - 1):
        mine = segs[so[r]:so[r + 1]]
        chains = []
        for cid in sorted(set(mine["chain_id"].tolist())):
            part = mine[mine["chain_id"] == cid]
            chains.append((int(part["chain_score"][0]), [int(x) for x in si[ko[r] + part["input_index"]]]))
        exp = [(int(d["chain_scores"][c]), sidx[soff[c]:soff[c + 1]].tolist()) for c in range(co[r], co[r + 1])]
        assert chains == exp, f"read {r}"


def test_cpp_sort_vectors(oracle_mod):
    d = _load("cpp_sort_v1.npz")
    off = d["offsets"]
    for i in range(len(off) - 1):
        keys = d["keys"][off[i]:off[i + 1]]
        assert oracle_mod.std_sort_perm(keys).tolist() == d["perm"][off[i]:off[i + 1]].tolist(), f"array {i}"


def test_vectors_exercise_the_early_break(oracle_mod):
    """The DP's break after an improvement on an exact diagonal step shorter
    than k changes the score table of some vector groups, so the vectors pin
    that rule (a DP without it fails them)."""
    d = _load("replay_group_v1.npz")
    ho = d["hit_offsets"]
    cfg = read_to_graph_config(500)
    differ = 0
    for g in range(len(ho) - 1):
        cur, ext = d["cur"][ho[g]:ho[g + 1]], d["ext"][ho[g]:ho[g + 1]]
        s1, b1 = oracle_mod.chain_dp(cur, ext, cfg)
        s0, b0 = oracle_mod.chain_dp(cur, ext, cfg, early_break=False)
        differ += int(not (np.array_equal(s1, s0) and np.array_equal(b1, b0)))
    assert differ > 0
