"""Minimizer index of read-to-graph alignment (flyehip_build_index_minimizers'
definition, restated by the oracle): the oracle's minimizer flags against a
brute force written here from the definition, the expected density
2 / (w + 1) and the densities Flye recorded for its own read-to-graph
minimizers, and the oracle's minimizer index against its full index
filtered by the flags."""
import numpy as np
import pytest

M64 = (1 << 64) - 1
CODE = {ord("A"): 0, ord("C"): 1, ord("G"): 2, ord("T"): 3}

# query_sequence_length and query_minimizers of the eight toy-hifi queries
# whose minimizers Flye wrote with yieldMinimizers(querySeq, minWnd)
# (tests/golden/cuflye-m6d-read-to-graph-minimizer-source-pack-dgx-aarch64.json,
# source_pack.queries[*]; patch 0040 file line 133)
FLYE_QUERY_MINIMIZERS = [(4697, 1544), (4297, 1408), (2080, 706), (2799, 925), (1069, 353), (1612, 541),
                         (2479, 807), (4391, 1441)]


def _mix64(x: int) -> int:
    x ^= x >> 30
    x = (x * 0xbf58476d1ce4e5b9) & M64
    x ^= x >> 27
    x = (x * 0x94d049bb133111eb) & M64
    x ^= x >> 31
    return x


def _brute_flags(seq: bytes, k: int, w: int) -> np.ndarray:
    nk = max(0, len(seq) - k + 1)
    order = []
    for p in range(nk):
        x = 0
        r = 0
        for j in range(k):
            b = CODE[seq[p + j]]
            x = (x << 2) | b
            r |= (3 - b) << (2 * j)
        order.append(_mix64(min(x, r)))
    flags = np.zeros(nk, np.uint8)
    if nk == 0:
        return flags
    starts = range(0, nk - w + 1) if nk >= w else [0]
    for s in starts:
        window = order[s:s + w]
        flags[s + window.index(min(window))] = 1  # leftmost minimum
    return flags


def _random_seq(rng, n: int) -> bytes:
    return bytes(rng.choice(np.frombuffer(b"ACGT", np.uint8), n))


@pytest.mark.parametrize("k,w", [(15, 5), (5, 3), (4, 1), (7, 10)])
def test_minimizer_flags_match_brute_force(oracle_mod, k, w):
    rng = np.random.default_rng(100 + k * w)
    unit = _random_seq(rng, 9)
    cases = [_random_seq(rng, n) for n in (0, k - 1, k, k + 1, k + w - 2, k + w - 1, 300)]
    cases.append(unit * 30)                      # tandem repeat: equal k-mers inside windows
    cases.append(b"A" * 80 + _random_seq(rng, 60))  # homopolymer: all-equal windows
    for seq in cases:
        got = oracle_mod.minimizer_flags(seq, k, w)
        assert np.array_equal(got, _brute_flags(seq, k, w)), (k, w, len(seq))


def test_window_one_selects_every_position(oracle_mod):
    seq = _random_seq(np.random.default_rng(5), 500)
    assert oracle_mod.minimizer_flags(seq, 15, 1).all()


def test_density_matches_random_minimizers_and_flye_records(oracle_mod):
    rng = np.random.default_rng(9)
    # random (w, k) minimizers select about 2 / (w + 1) of the positions
    seq = _random_seq(rng, 200_000)
    for w in (3, 5, 10):
        d = oracle_mod.minimizer_flags(seq, 15, w).mean()
        assert abs(d - 2.0 / (w + 1)) < 0.02 * 2.0 / (w + 1), (w, d)
    # Flye's recorded read-to-graph minimizer density is that of window 5,
    # the window the read-to-graph index uses here (k = 17 for the toy-hifi
    # run; 15 changes the position count by 0.2%)
    total_pos = sum(n - 17 + 1 for n, _ in FLYE_QUERY_MINIMIZERS)
    flye_density = sum(m for _, m in FLYE_QUERY_MINIMIZERS) / total_pos
    ours = {w: oracle_mod.minimizer_flags(seq, 15, w).mean() for w in (4, 5, 6)}
    assert abs(ours[5] - flye_density) < 0.01
    assert abs(ours[4] - flye_density) > 0.05 and abs(ours[6] - flye_density) > 0.03


def test_minimizer_index_is_the_full_index_at_minimizer_positions(oracle_mod):
    import cuflye_amd.synth as S
    r = S.config1(scale=0.01, seed=21)
    k, w = 15, 5
    full = oracle_mod.Oracle(r.bases, r.offsets)
    finfo = full.build_index(k, 1, 1 << 20)
    mini = oracle_mod.Oracle(r.bases, r.offsets)
    minfo = mini.build_index(k, 1, 1 << 20, minimizer_window=w)
    assert minfo["solid_min"] == finfo["solid_min"] and minfo["rep_max"] == finfo["rep_max"]
    assert minfo["repetitive"] == finfo["repetitive"]
    ft, mt = full.index_tables(), mini.index_tables()
    # expected: every full-index entry whose stored-strand position is a
    # minimizer of its read (entries on the reverse strand point at the
    # mirrored position)
    flags = [oracle_mod.minimizer_flags(r.bases[r.offsets[i]:r.offsets[i + 1]].tobytes(), k, w)
             for i in range(r.n_reads)]
    lens = np.diff(r.offsets)
    keep = np.zeros(len(ft["ids"]), bool)
    for e, (iid, pos) in enumerate(zip(ft["ids"], ft["pos"])):
        n = int(iid) // 2
        fwd_pos = int(pos) if int(iid) % 2 == 0 else int(lens[n]) - int(pos) - k
        keep[e] = flags[n][fwd_pos] == 1
    exp_kmers, exp_start, exp_ids, exp_pos = [], [0], [], []
    for b, km in enumerate(ft["kmers"]):
        sel = np.flatnonzero(keep[ft["start"][b]:ft["start"][b + 1]]) + ft["start"][b]
        if len(sel) == 0:
            continue
        exp_kmers.append(km)
        exp_ids.extend(ft["ids"][sel])
        exp_pos.extend(ft["pos"][sel])
        exp_start.append(exp_start[-1] + len(sel))
    assert np.array_equal(mt["kmers"], np.array(exp_kmers, np.uint64))
    assert np.array_equal(mt["start"], np.array(exp_start, np.int64))
    assert np.array_equal(mt["ids"], np.array(exp_ids, np.uint32))
    assert np.array_equal(mt["pos"], np.array(exp_pos, np.int32))
    assert 0.25 < len(mt["ids"]) / len(ft["ids"]) < 0.42
