"""The HIP chain path (flyehip_chain_lists through the C ABI) against vectors
written by the reference's own Python restatements
(tests/golden/gen_reference_vectors.py), and the fail-closed shapes.

  replay_group_v1   tools/replay_read_to_graph_source_pack.py:352-491, tie-heavy
                    groups: row order and integer fields exact, divergence
                    within 1e-6 (the reference computes it in double).
  chain_target_v1   tools/replay_overlap_chains.py:228-380, the families without
                    filtered positions (flyehip_chain_lists has none): bit-exact
                    rows including the float32 divergence.
"""
from pathlib import Path

import numpy as np
import pytest

from cuflye_amd import native as N
from cuflye_amd.overlap import FlyeHipError, OverlapConfig

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden"
DIV_TOL = 1e-6


def _run_lists(gpu, cur, ext, ho, sel, cur_id, cur_len, ext_id, ext_len, cfg):
    """chain_lists over the selected groups; returns per-group row arrays."""
    offs = [0]
    cs, es = [], []
    for g in sel:
        cs.append(cur[ho[g]:ho[g + 1]])
        es.append(ext[ho[g]:ho[g + 1]])
        offs.append(offs[-1] + ho[g + 1] - ho[g])
    rows, roff = gpu.chain_lists(np.concatenate(cs), np.concatenate(es), np.array(offs, np.int64),
                                 cur_id[sel], ext_id[sel], cur_len[sel], ext_len[sel], cfg)
    return [rows[roff[i]:roff[i + 1]] for i in range(len(sel))]


def test_gpu_replay_group_vectors(gpu):
    d = np.load(GOLD / "replay_group_v1.npz")
    ho, ro = d["hit_offsets"], d["row_offsets"]
    print(f"[replay_group] {len(ho) - 1} groups", flush=True)
    for mo in sorted(set(d["min_overlap"].tolist())):
        sel = np.nonzero(d["min_overlap"] == mo)[0]
        per = _run_lists(gpu, d["cur"], d["ext"], ho, sel, d["cur_id"], d["cur_len"], d["ext_id"], d["ext_len"],
                         OverlapConfig.read_to_graph(15, int(mo)))
        for g, rows in zip(sel, per):
            exp = d["rows_int"][ro[g]:ro[g + 1]]
            got = np.stack([rows["cur_begin"], rows["cur_end"], rows["ext_begin"], rows["ext_end"],
                            rows["score"], rows["chain_length"]], axis=1) if len(rows) else np.zeros((0, 6))
            assert got.tolist() == exp.tolist(), f"group {g} (min overlap {mo})"
            assert np.allclose(rows["seq_divergence"].astype(np.float64), d["rows_div_f64"][ro[g]:ro[g + 1]],
                               rtol=0, atol=DIV_TOL), f"group {g}"


def test_gpu_chain_target_vectors(gpu):
    d = np.load(GOLD / "chain_target_v1.npz")
    done = 0
    for fam in [str(f) for f in d["families"]]:
        g_ = {k.split("__", 1)[1]: d[k] for k in d.files if k.startswith(fam + "__")}
        if g_["filtered_offsets"][-1] > 0:
            continue
        pi, pf = g_["params_i"], g_["params_f"]
        cfg = OverlapConfig(kmer_size=int(pi[0]), minimum_overlap=int(pi[1]), maximum_jump=int(pi[2]),
                            max_jump_gap=int(pi[3]), gap_jump_threshold=int(pi[4]), maximum_overhang=int(pi[5]),
                            check_overhang=bool(pi[6]), force_local=bool(pi[7]), only_max_ext=bool(pi[8]),
                            large_gap_penalty=float(pf[0]), small_gap_penalty=float(pf[1]),
                            min_kmer_survival_rate=float(pf[2]), max_divergence=float(pf[3]),
                            sample_rate=float(pf[4]))
        ho, ro = g_["hit_offsets"], g_["row_offsets"]
        print(f"[chain_target {fam}] {len(ho) - 1} groups", flush=True)
        sel = np.arange(len(ho) - 1)
        per = _run_lists(gpu, g_["cur"], g_["ext"], ho, sel, g_["cur_id"], g_["cur_len"], g_["ext_id"],
                         g_["ext_len"], cfg)
        for g, rows in zip(sel, per):
            exp = g_["rows"][ro[g]:ro[g + 1]]
            got = np.stack([rows["cur_begin"], rows["cur_end"], rows["ext_begin"], rows["ext_end"],
                            rows["score"], rows["seq_divergence"].view(np.int32)], axis=1) if len(rows) \
                else np.zeros((0, 6))
            assert got.tolist() == exp.tolist(), f"{fam} group {g}"
        done += 1
    assert done >= 5


@pytest.mark.parametrize("field", ["keep_alignment", "nucl_alignment", "partition_bad_mappings", "max_overlaps"])
def test_unsupported_shapes_fail_closed(gpu, field):
    """Shapes the reference replay refuses (tools/replay_overlap_chains.py:127-137;
    max_overlaps: cuda/cuflye_cuda_overlap_chain_replay.cu:727-734) return
    FLYEHIP_ERR_UNSUPPORTED and write no rows."""
    import cuflye_amd.synth as S
    r = S.config1(scale=0.01, seed=3)
    gpu.store_reads(r.bases, r.offsets)
    gpu.build_index(15)
    cfg = OverlapConfig(kmer_size=15, minimum_overlap=1000)
    setattr(cfg, field, 5 if field == "max_overlaps" else True)
    ids = np.array([1, -1, 2], np.int64)
    with pytest.raises(FlyeHipError) as e:
        gpu.get_seq_overlaps(ids, cfg, row_capacity=64)
    assert e.value.code == N.FLYEHIP_ERR_UNSUPPORTED
    # the raw entry point leaves the caller's rows untouched
    import ctypes as C
    rows = np.zeros(8, N.OVERLAP_DTYPE)
    rows["score"] = -777
    roff = np.zeros(len(ids) + 1, np.int64)
    need = C.c_int64(0)
    p = cfg.to_c()
    rc = gpu.lib.flyehip_overlaps(gpu.ctx, N.ptr(ids), len(ids), C.byref(p), N.ptr(rows), len(rows),
                                  N.ptr(roff), C.byref(need), None, 0, None)
    assert rc == N.FLYEHIP_ERR_UNSUPPORTED
    assert np.all(rows["score"] == -777)


def test_div_stats_capacity_retry(gpu, oracle_mod):
    """A div_stats buffer smaller than the windows of one long query returns
    FLYEHIP_ERR_CAPACITY with div_offsets[n] = the values needed; a second
    call with that capacity returns the oracle's values."""
    import ctypes as C
    import cuflye_amd.synth as S
    r = S.config1(scale=0.05, seed=8)
    gpu.store_reads(r.bases, r.offsets)
    gpu.build_index(15)
    orc = oracle_mod.Oracle(r.bases, r.offsets)
    orc.build_index(15)
    lens = np.diff(r.offsets)
    q = int(np.argmax(lens)) + 1
    ids = np.array([q, -q], np.int64)
    cfg = OverlapConfig(kmer_size=15, minimum_overlap=3000)
    p = cfg.to_c()
    rows = np.zeros(4096, N.OVERLAP_DTYPE)
    roff = np.zeros(3, np.int64)
    doff = np.zeros(3, np.int64)
    divs = np.zeros(1, np.float32)
    need = C.c_int64(0)
    rc = gpu.lib.flyehip_overlaps(gpu.ctx, N.ptr(ids), 2, C.byref(p), N.ptr(rows), len(rows), N.ptr(roff),
                                  C.byref(need), N.ptr(divs), 1, N.ptr(doff))
    erows, eroff, edivs, edoff = orc.overlaps(ids, cfg, with_div_stats=True)
    if edoff[-1] <= 1:
        pytest.skip("the longest read has at most one divStats window")
    assert rc == N.FLYEHIP_ERR_CAPACITY
    assert doff[-1] == edoff[-1]
    divs = np.zeros(int(doff[-1]), np.float32)
    rc = gpu.lib.flyehip_overlaps(gpu.ctx, N.ptr(ids), 2, C.byref(p), N.ptr(rows), len(rows), N.ptr(roff),
                                  C.byref(need), N.ptr(divs), len(divs), N.ptr(doff))
    assert rc == N.FLYEHIP_OK
    assert np.array_equal(divs.view(np.uint32), edivs.view(np.uint32))
    # the wrapper retries by itself
    grows, groff, gdivs, gdoff = gpu.get_seq_overlaps(ids, cfg, with_div_stats=True, div_capacity=1)
    assert np.array_equal(gdivs.view(np.uint32), edivs.view(np.uint32))
