"""2-bit packing of sequences on the device (the read store and the external
query path): a byte other than A/C/G/T fails with the sequence and position,
and a store whose offsets do not start at 0 packs exactly like the sliced
arrays.  Index facts and overlap rows are compared against the oracle."""
import numpy as np
import pytest

import cuflye_amd.synth as S
from cuflye_amd import native as N
from cuflye_amd.overlap import FlyeHip, FlyeHipError, OverlapConfig

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = FlyeHip(device=0)
    yield c
    c.close()


def test_non_acgt_byte_names_sequence_and_position(ctx):
    bases = np.frombuffer(b"ACGTACGTACGTACGTACGT" + b"ACGTNCGT", np.uint8).copy()
    offs = np.array([0, 20, 28], np.int64)
    with pytest.raises(FlyeHipError, match="sequence 1 has a non-ACGT byte 0x4e at 4") as e:
        ctx.store_reads(bases, offs)
    assert e.value.code == N.FLYEHIP_ERR_ARG
    lower = np.frombuffer(b"ACGTacgt", np.uint8).copy()
    with pytest.raises(FlyeHipError, match="sequence 0 has a non-ACGT byte 0x61 at 4"):
        ctx.store_reads(lower, np.array([0, 8], np.int64))


def test_offset_store_matches_sliced_store_and_oracle(ctx, oracle_mod):
    r = S.config1(scale=0.02, seed=61)
    print(f"[pack] reads={r.n_reads} bases={len(r.bases)}", flush=True)
    # the same reads behind a 37-byte prefix the offsets skip
    pad = np.frombuffer(b"N" * 37, np.uint8)
    bases = np.concatenate([pad, r.bases])
    offs = r.offsets + 37
    ctx.store_reads(bases, offs)
    a = ctx.build_index(15)
    orc = oracle_mod.Oracle(r.bases, r.offsets)
    o = orc.build_index(15)
    assert a.solid_min == o["solid_min"] and a.indexed_positions == o["positions"]
    ids = np.arange(1, min(r.n_reads, 120) + 1, dtype=np.int64)
    cfg = OverlapConfig(kmer_size=15, minimum_overlap=1000)
    rows, roff = ctx.get_seq_overlaps(ids, cfg)
    erows, eroff = orc.overlaps(ids, cfg, threads=8)
    assert np.array_equal(roff, eroff) and rows.tobytes() == erows.tobytes()
    # the external path packs the query sequences the same way: the reads
    # handed over again (from an offset slice) give the same rows
    xrows, xroff = ctx.get_seq_overlaps_external(bases, offs[: len(ids) + 1], ids, cfg)
    assert np.array_equal(xroff, roff) and xrows.tobytes() == rows.tobytes()


def test_external_non_acgt_fails(ctx):
    r = S.config1(scale=0.01, seed=62)
    ctx.store_reads(r.bases, r.offsets)
    ctx.build_index(15)
    q = np.frombuffer(b"ACGT" * 300 + b"ACGX", np.uint8).copy()
    with pytest.raises(FlyeHipError, match="sequence 0 has a non-ACGT byte 0x58 at 1203"):
        ctx.get_seq_overlaps_external(q, np.array([0, len(q)], np.int64), np.array([10**9], np.int64),
                                      OverlapConfig(kmer_size=15, minimum_overlap=1000))
