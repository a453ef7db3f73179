"""Closed-form pins of the oracle's chaining pieces (no GPU).

Every expected value below is worked out by hand in the docstring or comment
next to it from the rules of tools/replay_overlap_chains.py:196-380 (the DP
recurrence, gap tiers, backtrack walk, overlapTest, divergence) and
patches/flye/2.9.6/0023-cuflye-gpu-first-overlap-substitution-perf-gate.patch
:245-269 (divStats); none is recomputed from a formula in the test.
"""
from collections import defaultdict

import numpy as np
import pytest

from cuflye_amd.overlap import OverlapConfig

K = 15


def _p(**kw):
    base = dict(kmer_size=K, minimum_overlap=20, check_overhang=False, only_max_ext=False)
    base.update(kw)
    return OverlapConfig(**base)


def _diag(n, step=15, start=0):
    c = np.arange(start, start + step * n, step, dtype=np.int32)
    return c, c.copy()


def test_dp_small_gap_tier(oracle_mod):
    """Diagonal hits 10 bp apart, then an off-diagonal hit with jumpDiv 50.
    s0 = 0; s1: 0 + min(10,10,15) = 10, break (jumpDiv 0, dc 10 < k), floored
    to k = 15 without a predecessor (10 <= 15); s2..s4 = 25, 35, 45.
    (140,190) from (40,40): dc 100, de 150, jumpDiv 50 <= 100 -> cost
    int(0.5 * 50) = 25, 45 + 15 - 25 = 35 (every other predecessor is lower)."""
    cur = np.array([0, 10, 20, 30, 40, 140], np.int32)
    ext = np.array([0, 10, 20, 30, 40, 190], np.int32)
    s, bt = oracle_mod.chain_dp(cur, ext, _p())
    assert s.tolist() == [0, 15, 25, 35, 45, 35]
    assert bt.tolist() == [-1, -1, 1, 2, 3, 4]


@pytest.mark.parametrize("ext_last,score,pred", [
    (500, 265, 20),   # jumpDiv 100 = threshold: small tier, int(0.5*100) = 50; 300 + 15 - 50
    (501, 113, 20),   # jumpDiv 101: large tier, int(2.0*101) = 202; 300 + 15 - 202
    (520, 75, 20),    # jumpDiv 120: int(2.0*120) = 240; 300 + 15 - 240
])
def test_dp_gap_tier_threshold(oracle_mod, ext_last, score, pred):
    """21 diagonal hits 15 apart score s_i = 15 i (s20 = 300); the last hit sits
    at cur 400 (dc 100 from hit 20) and ext `ext_last`."""
    c, e = _diag(21)
    cur = np.append(c, 400).astype(np.int32)
    ext = np.append(e, ext_last).astype(np.int32)
    s, bt = oracle_mod.chain_dp(cur, ext, _p())
    assert s[:21].tolist() == [0] + [15 * i for i in range(1, 21)]
    assert (int(s[21]), int(bt[21])) == (score, pred)


@pytest.mark.parametrize("ext_last,score,pred", [
    (1650, 65, 70),   # jumpDiv 500 = max_jump_gap: allowed, 1050 + 15 - int(2.0*500)
    (1651, 15, -1),   # jumpDiv 501: every predecessor rejected, floor k, no pointer
])
def test_dp_max_jump_gap(oracle_mod, ext_last, score, pred):
    """71 diagonal hits 15 apart (s70 = 1050), then a hit at cur 1150."""
    c, e = _diag(71)
    cur = np.append(c, 1150).astype(np.int32)
    ext = np.append(e, ext_last).astype(np.int32)
    s, bt = oracle_mod.chain_dp(cur, ext, _p())
    assert int(s[70]) == 1050
    assert (int(s[71]), int(bt[71])) == (score, pred)


@pytest.mark.parametrize("gap,score,pred", [(1499, 30, 1), (1500, 15, -1)])
def test_dp_max_jump_strict(oracle_mod, gap, score, pred):
    """0 < dc < maxJump is strict: (0,0), (15,15), then a diagonal step of
    `gap`: 1499 joins (15 + 15 = 30 > k), 1500 does not."""
    cur = np.array([0, 15, 15 + gap], np.int32)
    s, bt = oracle_mod.chain_dp(cur, cur.copy(), _p())
    assert (int(s[2]), int(bt[2])) == (score, pred)


def test_dp_ext_sorted_axis(oracle_mod):
    """With extLen > curLen the group is scanned in extPos order: hits given
    by ext ascending with cur out of order.  (0,0) -> (30,15) -> (45,30):
    (30,15) from (0,0): dc 30, de 15, jumpDiv 15 -> 0 + 15 - 7 = 8, floor 15,
    no pointer; (45,30) from (30,15): 15 + 15 = 30, pointer 1; (10,45) has
    dc <= 0 to (30,15) and (45,30) and jumpDiv 35 to (0,0): 0 + 10 - 17 < 0."""
    cur = np.array([0, 30, 45, 10], np.int32)
    ext = np.array([0, 15, 30, 45], np.int32)
    s, bt = oracle_mod.chain_dp(cur, ext, _p(), ext_sorted=True)
    assert s.tolist() == [0, 15, 30, 15]
    assert bt.tolist() == [-1, -1, 1, -1]


def _group(oracle_mod, cur, ext, cur_len=1000, ext_len=1000, filtered=(), **kw):
    return oracle_mod.chain_group(np.array(cur, np.int32), np.array(ext, np.int32), 1, cur_len, 2, ext_len,
                                  np.array(filtered, np.int32), _p(**kw))


def test_walk_marks_visited(oracle_mod):
    """Diagonal (0..60 step 15) then two hits at cur 75: (75,75) continues the
    diagonal (s = 75, pointer 4), (75,85) is a jumpDiv-10 step from (60,60)
    (60 + 15 - 5 = 70, pointer 4); neither reaches the other (dc = 0).
    Walk order by score: (75,75) walks 5-4-3-2-1 (s1 = 15 has no pointer):
    cur/ext [15, 89], score 75 - 15 + 14 = 74, 5 hits.  (75,85) then stops at
    hit 4, already visited: cur [60, 89], ext [60, 99], score 70 - 60 + 14 =
    24, 2 hits (without the marking it would walk down to hit 1)."""
    cur = [0, 15, 30, 45, 60, 75, 75]
    ext = [0, 15, 30, 45, 60, 75, 85]
    rows = _group(oracle_mod, cur, ext)
    got = [(r["cur_begin"], r["cur_end"], r["ext_begin"], r["ext_end"], r["score"], r["chain_length"])
           for r in rows]
    assert got == [(15, 89, 15, 89, 74, 5), (60, 89, 60, 99, 24, 2)]
    # divergence: logf(1 / (5 / 74)) / 15 and logf(1 / (2 / 39)) / 15 in float
    d0 = np.float32(oracle_mod.logf(float(np.float32(1) / (np.float32(5) / np.float32(74))))) / np.float32(15)
    assert rows[0]["seq_divergence"] == d0


def test_containment_and_only_max_ext(oracle_mod):
    """As above with the second hit at (75,70): 60 + 10 - 2 = 68, proposal cur
    [60, 89], ext [60, 84] lies inside the first (cur [15, 89], ext [15, 89])
    with a lower score, so the containment filter drops it; onlyMaxExt keeps
    the first maximum either way."""
    cur = [0, 15, 30, 45, 60, 75, 75]
    ext = [0, 15, 30, 45, 60, 75, 70]
    for only in (False, True):
        rows = _group(oracle_mod, cur, ext, only_max_ext=only)
        assert [(r["cur_begin"], r["ext_end"], r["score"]) for r in rows] == [(15, 89, 74)]
    # not contained (ext end 99 > 89): both survive containment, onlyMaxExt keeps one
    ext2 = [0, 15, 30, 45, 60, 75, 85]
    assert len(_group(oracle_mod, cur, ext2, only_max_ext=False)) == 2
    assert len(_group(oracle_mod, cur, ext2, only_max_ext=True)) == 1


def test_max_divergence_filter(oracle_mod):
    """Proposal 1 of test_walk_marks_visited has divergence logf(14.8)/15 =
    0.17966; a limit of 0.18 keeps it, 0.179 drops it (strict <)."""
    cur = [0, 15, 30, 45, 60, 75, 75]
    ext = [0, 15, 30, 45, 60, 75, 70]
    assert len(_group(oracle_mod, cur, ext, max_divergence=0.18)) == 1
    assert len(_group(oracle_mod, cur, ext, max_divergence=0.179)) == 0


def test_filtered_positions_inclusive(oracle_mod):
    """A diagonal at 1 bp spacing, hits 0..99: s_i = 15 + (i - 1) from hit 1,
    the chain is hits 1..99 (99 hits), cur [1, 113], range 112.  normLen =
    112 - f, rate = 99 / (112 - f): f = 13 gives rate 1 and divergence 0
    exactly; f = 12 gives rate 99/100 < 1.  Positions 1 and 113 (the range
    ends) count, 0 and 114 do not."""
    c = np.arange(0, 100, dtype=np.int32)
    inside = [1, 113] + list(range(20, 31))          # 13 positions in [1, 113]
    outside = [0, 114, 500]
    rows = _group(oracle_mod, c, c, cur_len=1000, ext_len=1000, filtered=sorted(inside + outside))
    assert (rows[0]["cur_begin"], rows[0]["cur_end"], rows[0]["chain_length"]) == (1, 113, 99)
    assert rows[0]["seq_divergence"] == 0.0
    twelve = sorted([1] + list(range(20, 31)) + outside)
    rows = _group(oracle_mod, c, c, cur_len=1000, ext_len=1000, filtered=twelve)
    assert rows[0]["seq_divergence"] > 0.0


def test_prefilter_unique_count_from_zero(oracle_mod):
    """The unique-position counter starts from prevPos = 0, so hits at cur 0
    do not count.  min_kmer_survival_rate 0.01 x minimum_overlap 300 = 3:
    positions {0, 0, 100, 400} count 2 unique (< 3, dropped); {0, 50, 100,
    400} count 3 (kept: a proposal exists)."""
    p = dict(minimum_overlap=300)
    assert len(_group(oracle_mod, [0, 0, 100, 400], [0, 0, 100, 400], **p)) == 0
    ok = _group(oracle_mod, [0, 50, 100, 400], [0, 50, 100, 400], **p)
    # the chain 1-2-3: cur [50, 414], score 45 - 15 + 14 = 44
    assert [(r["cur_begin"], r["cur_end"], r["score"]) for r in ok] == [(50, 414, 44)]
    s, _ = oracle_mod.chain_dp(np.array([0, 50, 100, 400], np.int32), np.array([0, 50, 100, 400], np.int32),
                               _p(**p))
    assert s.tolist() == [0, 15, 30, 45]


def test_prefilter_overhang(oracle_mod):
    """With the overhang check on, a group whose hits start further than
    maxOverhang into both reads (min(minCur, minExt) = 600 > 500) is dropped
    before the DP; 400 passes."""
    c = np.arange(600, 2600, 10, dtype=np.int32)
    kw = dict(check_overhang=True, maximum_overhang=500, minimum_overlap=1000)
    assert len(_group(oracle_mod, c, c, cur_len=2700, ext_len=2700, **kw)) == 0
    c2 = c - 200
    assert len(_group(oracle_mod, c2, c2, cur_len=2500, ext_len=2500, **kw)) == 1


BASE_ROW = dict(cur_id=1, cur_begin=0, cur_end=1000, cur_len=5000, ext_id=2, ext_begin=0, ext_end=1000,
                ext_len=5000)


@pytest.mark.parametrize("change,params,expect", [
    ({}, {}, True),
    ({"cur_end": 999}, {}, False),                               # cur range 999 < 1000
    ({"ext_end": 999}, {}, False),                               # ext range 999 < 1000
    ({"ext_end": 1500, "cur_end": 1000}, {}, True),              # |1500 - 1000| = 500 <= 0.5 * 1000
    ({"ext_end": 1501, "cur_end": 1000}, {}, False),             # 501 > 500
    ({"ext_id": 1, "ext_begin": 500, "ext_end": 1500}, {}, True),   # self: intersect 500 <= 500
    ({"ext_id": 1, "ext_begin": 499, "ext_end": 1499}, {}, False),  # self: intersect 501 > 500
    ({"ext_id": -1, "ext_begin": 4000, "ext_end": 5000}, {}, False),  # rc self: min(1000, 1000) - max(0, 0)
    ({"ext_id": -1, "ext_begin": 3500, "ext_end": 4500}, {}, True),   # rc: min(1000, 1500) - max(0, 500) = 500
    ({"ext_id": -1, "ext_begin": 3501, "ext_end": 4501}, {}, False),  # rc: min(1000, 1499) - max(0, 499) = 501
    ({"cur_begin": 600, "cur_end": 1600, "ext_begin": 600, "ext_end": 1600},
     {"check_overhang": True, "maximum_overhang": 599}, False),  # overhang max(600, 3400) > 599
    ({"cur_begin": 600, "cur_end": 1600, "ext_begin": 600, "ext_end": 1600, "cur_len": 1600},
     {"check_overhang": True, "maximum_overhang": 600}, True),   # max(min(600,600), min(0,3400)) = 600
    ({"cur_begin": 600, "cur_end": 1600, "ext_begin": 600, "ext_end": 1600},
     {"check_overhang": True, "maximum_overhang": 10, "force_local": True}, True),
    ({"cur_begin": 600, "cur_end": 1600, "ext_begin": 600, "ext_end": 1600},
     {"check_overhang": False, "maximum_overhang": 10}, True),
])
def test_overlap_test_branches(oracle_mod, change, params, expect):
    row = dict(BASE_ROW)
    row.update(change)
    p = _p(minimum_overlap=1000, **params)
    assert oracle_mod.overlap_test(row, p) is expect


def test_div_stats_windows(oracle_mod):
    """curLen 35000 -> windows 0..3 by curBegin / 10000; the longest curRange
    of a window wins (strictly longer: the first of equal ranges stays);
    windows without an overlap add nothing."""
    from cuflye_amd import native as N
    spec = [(500, 3000, 0.1), (9999, 5000, 0.2), (10000, 100, 0.3), (25000, 2000, 0.4), (29999, 2000, 0.5)]
    rows = np.zeros(len(spec), N.OVERLAP_DTYPE)
    for i, (cb, rng_, d) in enumerate(spec):
        rows[i]["cur_begin"], rows[i]["cur_end"], rows[i]["seq_divergence"] = cb, cb + rng_, d
    got = oracle_mod.div_stats(rows, 35000)
    assert got.tolist() == pytest.approx([0.2, 0.3, 0.4])


def _brute_index(reads, k, solid_min, rep_max):
    """Python-dict restatement of the index definition in include/flyehip.h,
    written without the oracle: canonical = min(kmer, rc) as strings mapped to
    2-bit numbers; counts over forward strands; buckets in (read, pos) order,
    the stored strand oriented to spell the canonical k-mer."""
    code = {"A": 0, "C": 1, "G": 2, "T": 3}
    comp = {"A": "T", "C": "G", "G": "C", "T": "A"}

    def num(s):
        v = 0
        for ch in s:
            v = v * 4 + code[ch]
        return v

    counts = defaultdict(int)
    occ = defaultdict(list)
    for n, r in enumerate(reads):
        for pos in range(len(r) - k + 1):
            w = r[pos:pos + k]
            rc = "".join(comp[ch] for ch in reversed(w))
            x, y = num(w), num(rc)
            canon = x if x < y else y
            counts[canon] += 1
            if x < y:
                occ[canon].append((2 * n, pos))
            else:
                occ[canon].append((2 * n + 1, len(r) - pos - k))
    hist = defaultdict(int)
    for c in counts.values():
        hist[c] += 1
    kmers = sorted(c for c, v in counts.items() if solid_min <= v <= rep_max)
    rep = sorted(c for c, v in counts.items() if v > rep_max and v >= solid_min)
    return counts, hist, kmers, rep, occ


def test_index_brute_force(oracle_mod):
    """Tiny reads with repeats: the oracle's solid set, repetitive set, bucket
    contents and bucket order equal the dict restatement; the automatic
    thresholds follow the stated rule on the brute-force histogram
    (first c >= 2 with hist[c] <= hist[c+1]; floor(rate x mean solid count))."""
    rng = np.random.default_rng(5)
    unit = "".join(rng.choice(list("ACGT"), 40))
    reads = []
    for i in range(12):
        body = "".join(rng.choice(list("ACGT"), int(rng.integers(60, 120))))
        reads.append(body[:30] + unit * int(rng.integers(1, 4)) + body[30:] + ("ACGTTG" * 3 if i % 3 == 0 else ""))
    bases = np.frombuffer("".join(reads).encode(), np.uint8).copy()
    off = np.zeros(len(reads) + 1, np.int64)
    off[1:] = np.cumsum([len(r) for r in reads])
    k = 7
    orc = oracle_mod.Oracle(bases, off)
    for solid_min, rep_max in ((2, 5), (1, 3), (3, 1000)):
        info = orc.build_index(k, solid_min, rep_max)
        counts, hist, kmers, rep, occ = _brute_index(reads, k, solid_min, rep_max)
        t = orc.index_tables()
        assert t["kmers"].tolist() == kmers
        assert t["repetitive"].tolist() == rep
        assert info["distinct"] == len(counts)
        for b, km in enumerate(kmers):
            entries = list(zip(t["ids"][t["start"][b]:t["start"][b + 1]].tolist(),
                               t["pos"][t["start"][b]:t["start"][b + 1]].tolist()))
            assert entries == occ[km]
    # automatic thresholds on the brute-force histogram
    info = orc.build_index(k, 0, 0, 10.0)
    counts, hist, _, _, _ = _brute_index(reads, k, 1, 1 << 30)
    top = max(hist)
    h = [hist.get(c, 0) for c in range(top + 2)]
    sm = next((c for c in range(2, len(h) - 1) if h[c] <= h[c + 1]), 2)
    assert info["solid_min"] == sm
    tot = sum(c * h[c] for c in range(sm, len(h)))
    dis = sum(h[c] for c in range(sm, len(h)))
    assert info["rep_max"] == max(int(np.floor(10.0 * tot / (dis + 1.0))), sm)
