"""GPU parity: libflyehip.so (through the C ABI) against the CPU oracle on the
same seeded inputs.  Integer fields and the float divergence must match bit
for bit; candidate lists must match in order (Flye's emission order, and
std::sort's tie order after the match sort)."""
import numpy as np
import pytest

import cuflye_amd.synth as S
from cuflye_amd import native as N
from cuflye_amd.overlap import OverlapConfig

pytestmark = pytest.mark.gpu


def _bytes(a):
    return np.ascontiguousarray(a).view(np.uint8)


def assert_rows_equal(got, exp, what):
    assert len(got) == len(exp), f"{what}: {len(got)} rows vs {len(exp)}"
    if len(got) and not np.array_equal(_bytes(got), _bytes(exp)):
        bad = np.nonzero(got != exp)[0]
        i = int(bad[0]) if len(bad) else 0
        raise AssertionError(f"{what}: first mismatch at row {i}: {got[i]} vs {exp[i]} ({len(bad)} rows differ)")


def both_strands(n):
    ids = np.empty(2 * n, np.int64)
    ids[0::2] = np.arange(1, n + 1)
    ids[1::2] = -np.arange(1, n + 1)
    return ids


@pytest.fixture(scope="module")
def c1(gpu, oracle_mod):
    r = S.config1(scale=0.1, seed=21)
    print(f"[c1] reads={r.n_reads} bases={len(r.bases)}", flush=True)
    orc = oracle_mod.Oracle(r.bases, r.offsets)
    oinfo = orc.build_index(15)
    gpu.store_reads(r.bases, r.offsets)
    ginfo = gpu.build_index(15)
    return r, orc, oinfo, ginfo


def test_index_matches_oracle(c1):
    r, orc, oinfo, ginfo = c1
    assert ginfo.solid_min == oinfo["solid_min"]
    assert ginfo.rep_max == oinfo["rep_max"]
    assert ginfo.distinct_kmers == oinfo["distinct"]
    assert ginfo.solid_kmers == oinfo["solid"]
    assert ginfo.repetitive_kmers == oinfo["repetitive"]
    assert ginfo.indexed_positions == oinfo["positions"]


def test_candidates_emission_order(gpu, c1):
    r, orc, _, _ = c1
    ids = np.array([1, -1, 5, -7, r.n_reads, -r.n_reads], np.int64)
    m, moff, f, foff = gpu.collect_candidates(ids)
    for i, q in enumerate(ids):
        em, ef = orc.candidates(int(q), sorted_=False)
        assert_rows_equal(m[moff[i]:moff[i + 1]], em, f"emission q={q}")
        assert np.array_equal(f[foff[i]:foff[i + 1]], ef)


def test_candidates_sorted_tie_order(gpu, c1):
    r, orc, _, _ = c1
    ids = both_strands(min(r.n_reads, 40))
    m, moff = gpu.collect_sorted_candidates(ids)
    st = gpu.stats()
    for i, q in enumerate(ids):
        em, _ = orc.candidates(int(q), sorted_=True)
        assert_rows_equal(m[moff[i]:moff[i + 1]], em, f"sorted q={q}")
    print("tie queries", st["tie_queries"], "of", len(ids))


def test_overlaps_config1(gpu, c1):
    r, orc, _, _ = c1
    ids = both_strands(r.n_reads)
    cfg = OverlapConfig(kmer_size=15, minimum_overlap=5000)
    rows, roff, divs, doff = gpu.get_seq_overlaps(ids, cfg, with_div_stats=True)
    erows, eroff, edivs, edoff = orc.overlaps(ids, cfg, threads=8, with_div_stats=True)
    assert np.array_equal(roff, eroff)
    assert_rows_equal(rows, erows, "config1 overlaps")
    assert np.array_equal(doff, edoff)
    assert np.array_equal(divs.view(np.uint32), edivs.view(np.uint32))
    assert len(rows) > 0


def test_overlaps_per_query_path_equal(gpu, c1):
    r, orc, _, _ = c1
    ids = both_strands(min(r.n_reads, 12))
    cfg = OverlapConfig(kmer_size=15, minimum_overlap=5000)
    a, ao = gpu.get_seq_overlaps(ids, cfg)
    b, bo = gpu.get_seq_overlaps(ids, cfg, per_query=True)
    assert np.array_equal(ao, bo)
    assert_rows_equal(a, b, "batched vs per-query")


def test_overlaps_config2_small(gpu, oracle_mod):
    r = S.config2(scale=0.03, seed=22)
    print(f"[c2] reads={r.n_reads} bases={len(r.bases)}", flush=True)
    orc = oracle_mod.Oracle(r.bases, r.offsets)
    orc.build_index(15)
    gpu.store_reads(r.bases, r.offsets)
    gpu.build_index(15)
    ids = both_strands(r.n_reads)
    cfg = OverlapConfig(kmer_size=15, minimum_overlap=3000)
    rows, roff = gpu.get_seq_overlaps(ids, cfg)
    erows, eroff = orc.overlaps(ids, cfg, threads=8)
    assert np.array_equal(roff, eroff)
    assert_rows_equal(rows, erows, "config2 overlaps")
    # batches dealt to 1 .. 3 concurrent streams: scheduling only
    try:
        for fl in (1, 3):
            gpu.set_tuning(queries_per_batch=64, batches_in_flight=fl)
            frows, froff = gpu.get_seq_overlaps(ids, cfg)
            assert np.array_equal(froff, eroff)
            assert_rows_equal(frows, erows, f"config2 overlaps, {fl} batches in flight")
        # every DP lookback width (1 .. 64 lanes per group): scheduling only
        for g in (1, 4, 16, 32, 64):
            print(f"[c2] dp lanes {g}", flush=True)
            gpu.set_tuning(dp_lanes_per_group=g, queries_per_batch=2048, batches_in_flight=2)
            frows, froff = gpu.get_seq_overlaps(ids, cfg)
            assert np.array_equal(froff, eroff)
            assert_rows_equal(frows, erows, f"config2 overlaps, {g} DP lanes per group")
    finally:
        gpu.set_tuning(dp_lanes_per_group=8, queries_per_batch=2048, batches_in_flight=2)
    # read-to-graph detector shape on the same data: containment filter path
    cfg2 = OverlapConfig.read_to_graph(15, 1000)
    rows, roff = gpu.get_seq_overlaps(ids[:60], cfg2)
    erows, eroff = orc.overlaps(ids[:60], cfg2, threads=8)
    assert_rows_equal(rows, erows, "config2 read-to-graph shape")


def test_candidates_sorted_tie_order_ont(gpu, oracle_mod):
    # ONT-like errors and repeat families: long match lists (tens of
    # thousands per query, partitioned across many waves) with equal
    # (extId, curPos) keys throughout
    r = S.config2(scale=0.03, seed=24)
    orc = oracle_mod.Oracle(r.bases, r.offsets)
    orc.build_index(15)
    gpu.store_reads(r.bases, r.offsets)
    gpu.build_index(15)
    ids = both_strands(min(r.n_reads, 60))
    m, moff = gpu.collect_sorted_candidates(ids)
    st = gpu.stats()
    assert st["tie_queries"] > 0
    assert int(np.max(np.diff(moff))) > 4096
    for i, q in enumerate(ids):
        em, _ = orc.candidates(int(q), sorted_=True)
        assert_rows_equal(m[moff[i]:moff[i + 1]], em, f"sorted q={q}")


def test_chain_lists_config3_small(gpu, oracle_mod):
    h = S.config3(n_lists=4000, seed=23, max_hits=20000)
    print(f"[c3] lists={len(h.offsets) - 1} hits={len(h.cur)}", flush=True)
    cfg = OverlapConfig(kmer_size=15, minimum_overlap=200, check_overhang=False)
    rows, roff = gpu.chain_lists(h.cur, h.ext, h.offsets, h.cur_id, h.ext_id, h.cur_len, h.ext_len, cfg)
    erows, eroff = oracle_mod.chain_lists(h.cur, h.ext, h.offsets, h.cur_id, h.ext_id, h.cur_len, h.ext_len,
                                          cfg, threads=8)
    assert np.array_equal(roff, eroff)
    assert_rows_equal(rows, erows, "config3 lists")
    assert len(rows) > 100
    cfg2 = OverlapConfig(kmer_size=15, minimum_overlap=200, check_overhang=False, only_max_ext=False)
    rows, roff = gpu.chain_lists(h.cur, h.ext, h.offsets, h.cur_id, h.ext_id, h.cur_len, h.ext_len, cfg2)
    erows, eroff = oracle_mod.chain_lists(h.cur, h.ext, h.offsets, h.cur_id, h.ext_id, h.cur_len, h.ext_len,
                                          cfg2, threads=8)
    assert_rows_equal(rows, erows, "config3 lists, containment selection")


def _diagonal_lists(rng, sizes, twin):
    """Long clean diagonals: DP scores grow along the list, so the score-order
    keys are nearly reverse sorted, the input on which libstdc++'s introsort
    exhausts its depth budget and heap-sorts.  With `twin` every list carries
    a second, identically spaced diagonal, so every score occurs twice and
    the heap sort's tie order decides the walk order."""
    cur, ext, offs = [], [], [0]
    for n in sizes:
        m = n // 2 if twin else n
        step = rng.integers(5, 9, size=m).cumsum().astype(np.int32)
        c = 100 + step
        e = 50 + step
        if twin:
            c = np.concatenate([c, c])
            e = np.concatenate([e, e + 3 * int(step[-1]) // 2])
            o = np.argsort(c, kind="stable")
            c, e = c[o], e[o]
        cur.append(c)
        ext.append(e)
        offs.append(offs[-1] + len(c))
    return np.concatenate(cur), np.concatenate(ext), np.array(offs, np.int64)


@pytest.mark.parametrize("twin", [False, True])
def test_chain_lists_heap_fallback(gpu, oracle_mod, twin):
    rng = np.random.default_rng(31 + int(twin))
    sizes = [int(x) for x in rng.integers(2500, 24000, size=24)]
    cur, ext, offs = _diagonal_lists(rng, sizes, twin)
    n = len(offs) - 1
    ci = np.arange(1, n + 1, dtype=np.int64)
    ei = ci + 1000
    cl = np.full(n, int(cur.max()) + 500, np.int32)
    el = cl + 7  # extLen > curLen: the extPos sort runs too
    for only_max in (True, False):
        cfg = OverlapConfig(kmer_size=15, minimum_overlap=200, check_overhang=False, only_max_ext=only_max)
        rows, roff = gpu.chain_lists(cur, ext, offs, ci, ei, cl, el, cfg)
        erows, eroff = oracle_mod.chain_lists(cur, ext, offs, ci, ei, cl, el, cfg, threads=8)
        assert np.array_equal(roff, eroff)
        assert_rows_equal(rows, erows, f"diagonal lists twin={twin} only_max={only_max}")
        assert len(rows) >= n


def _random_alignments(rng, n_reads):
    rows = []
    offs = [0]
    for r in range(n_reads):
        m = int(rng.integers(0, 40))
        pos = 0
        for _ in range(m):
            pos += int(rng.integers(-200, 3000))
            pos = max(pos, 0)
            ln = int(rng.integers(100, 5000))
            elen = int(rng.integers(500, 20000))
            eb = int(rng.integers(0, 2000))
            ee = min(elen - 1, eb + ln + int(rng.integers(-50, 50)))
            left = int(rng.integers(0, 12))
            right = int(rng.integers(0, 12))
            rows.append((r + 1, int(rng.integers(1, 30)), int(rng.integers(1, 30)), pos, pos + ln, 100000, left,
                         right, eb, max(eb, ee), elen, int(rng.integers(1, 5)) * 100, 0.01))
        offs.append(len(rows))
    a = np.zeros(len(rows), N.EDGE_ALIGNMENT_DTYPE)
    for i, t in enumerate(rows):
        a[i] = t
    return a, np.array(offs, np.int64)


def test_read_align_chains(gpu, oracle_mod):
    rng = np.random.default_rng(5)
    a, off = _random_alignments(rng, 500)
    params = N.ChainParams(1500, 50, 1000, 500)
    gi, gk, gs, go = gpu.read_align_chains(a, off, 900, 500, params)
    ei, ek, es, eo = oracle_mod.read_align_chains(a, off, 900, 500, params)
    assert np.array_equal(gk, ek)
    assert np.array_equal(gi, ei)
    assert np.array_equal(go, eo)
    assert_rows_equal(gs, es, "read-align segments")
    assert len(gs) > 0
