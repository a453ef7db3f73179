"""The claim walk (chainClaimKernel): Flye's score-ordered backtrack walks
computed without the score sort, and the groups it leaves to the exact
score sort (two surviving chain ends of equal score), checked against the
oracle, which walks in real std::sort order.

Both paths must run in each case: exact_walk_groups counts the groups that
took the exact sort, and must lie strictly between 0 and the group count.
"""
import numpy as np
import pytest

from cuflye_amd.overlap import OverlapConfig

pytestmark = pytest.mark.gpu


def _tie_lists():
    """Hit lists (sorted by curPos) of three kinds: a clean diagonal; a
    diagonal ending in two leaves of equal score off the same last row (the
    walk order of equal scores decides which leaf's chain holds the diagonal,
    and so curEnd); two parallel diagonals, one shifted by 2, with equal
    scores at their ends."""
    lists = []
    diag = [(100 * i, 100 * i) for i in range(12)]
    lists.append(diag)
    last = diag[-1]
    # (+100, +102) and (+102, +100): jumpDiv 2 each, the same gain
    lists.append(diag + [(last[0] + 100, last[1] + 102), (last[0] + 102, last[1] + 100)])
    par = sorted(diag + [(c + 2, e) for c, e in diag[1:]])
    lists.append(par)
    for shift in (1, 3, 7):
        lists.append(diag + [(last[0] + 100 + shift, last[1] + 100 + shift + 2),
                             (last[0] + 102 + shift, last[1] + 100 + shift)])
    return lists


def _pack(lists):
    cur = np.concatenate([np.array([c for c, _ in l], np.int32) for l in lists])
    ext = np.concatenate([np.array([e for _, e in l], np.int32) for l in lists])
    off = np.zeros(len(lists) + 1, np.int64)
    off[1:] = np.cumsum([len(l) for l in lists])
    n = len(lists)
    ids = np.arange(1, n + 1, dtype=np.int64)
    return cur, ext, off, ids, -(ids + 1000), np.full(n, 5000, np.int32), np.full(n, 5000, np.int32)


@pytest.mark.parametrize("only_max_ext", [True, False])
def test_claim_walk_equal_score_leaves(gpu, oracle_mod, only_max_ext):
    cfg = OverlapConfig(kmer_size=15, minimum_overlap=500, check_overhang=False, only_max_ext=only_max_ext)
    args = _pack(_tie_lists())
    print(f"[claim walk] hand-made lists, only_max_ext={only_max_ext}", flush=True)
    rows, roff = gpu.chain_lists(*args, cfg)
    st = gpu.stats()
    erows, eroff = oracle_mod.chain_lists(*args, cfg)
    assert np.array_equal(roff, eroff)
    assert rows.tobytes() == erows.tobytes()
    assert 0 < st["exact_walk_groups"] < len(args[2]) - 1


@pytest.mark.parametrize("only_max_ext", [True, False])
def test_claim_walk_dp_stress_sample(gpu, oracle_mod, only_max_ext):
    import cuflye_amd.synth as S
    h = S.config3(n_lists=20000, seed=31, max_hits=20000)
    cfg = OverlapConfig(kmer_size=15, minimum_overlap=200, check_overhang=False, only_max_ext=only_max_ext)
    args = (h.cur, h.ext, h.offsets, h.cur_id, h.ext_id, h.cur_len, h.ext_len)
    print(f"[claim walk] 20000 DP-stress lists, only_max_ext={only_max_ext}", flush=True)
    rows, roff = gpu.chain_lists(*args, cfg)
    st = gpu.stats()
    erows, eroff = oracle_mod.chain_lists(*args, cfg, threads=8)
    assert np.array_equal(roff, eroff)
    assert rows.tobytes() == erows.tobytes()
    assert 0 < st["exact_walk_groups"] < len(h.offsets) - 1
