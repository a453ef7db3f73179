#!/usr/bin/env python3
"""Writes golden vectors from the reference's own Python restatements.

Run in the build container only (it reads the reference tools from
/root/reference/tools; the GPU box has no reference).  The outputs are data
(seeded inputs and the reference functions' outputs) kept beside this script:

  replay_group_v1.npz     tools/replay_read_to_graph_source_pack.py:352-491
                          replay_group on tie-heavy groups (its libstdc++
                          std::sort restatement, cpp_sort_in_place :95-181, keeps
                          equal keys in Flye's order); groups whose sort reached
                          the depth limit are left out, because that restatement
                          falls back to a stable sort there instead of heapsort.
  chain_target_v1.npz     tools/replay_overlap_chains.py:228-380 chain_target in
                          several parameter families (overhang check, force
                          local, onlyMaxExt, filtered positions, max divergence),
                          float32 divergence through glibc logf; only groups
                          whose result does not depend on the order of equal
                          keys are kept (chain_target sorts stably where Flye's
                          std::sort is unstable).
  overlap_test_v1.npz     tools/replay_overlap_chains.py:196-225 overlap_test.
  read_align_v1.npz       tools/replay_read_alignment_chains.py:64-139
                          chain_read_alignments on reads whose chain scores are
                          distinct (the reference sorts the chains stably).
  cpp_sort_v1.npz         cpp_sort_in_place on tie-heavy integer keys.

Usage: python tests/golden/gen_reference_vectors.py [--reference /root/reference]
"""
from __future__ import annotations

import argparse
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
SEED = 20261204


class _Capture:
    """Keeps the locals of one function at its return (sys.setprofile)."""

    def __init__(self, code):
        self.code = code
        self.locals = None

    def __enter__(self):
        self.locals = None

        def prof(frame, event, arg):
            if event == "return" and frame.f_code is self.code:
                self.locals = dict(frame.f_locals)
        sys.setprofile(prof)
        return self

    def __exit__(self, *exc):
        sys.setprofile(None)
        return False


def _hit_list(rng: np.random.Generator, n: int, tie_heavy: bool):
    """One group's hits in (curPos ascending, given order for equal curPos):
    a drifting true diagonal with random spacing plus uniform spurious hits."""
    n_true = max(2, int(n * rng.uniform(0.4, 0.9)))
    step = rng.integers(1, 45, n_true)
    cur = np.cumsum(step) + int(rng.integers(0, 2000))
    drift = np.cumsum(rng.normal(0.0, rng.uniform(0.0, 3.0), n_true)).astype(np.int64)
    ext = cur + int(rng.integers(-1500, 3000)) + drift
    n_spur = n - n_true
    span_c = (int(cur.min()) - 200, int(cur.max()) + 200)
    span_e = (int(ext.min()) - 200, int(ext.max()) + 200)
    sc = rng.integers(span_c[0], span_c[1], n_spur)
    se = rng.integers(span_e[0], span_e[1], n_spur)
    cur = np.concatenate([cur, sc])
    ext = np.concatenate([ext, se])
    if tie_heavy and n > 4:
        # repeated query positions (a k-mer seen twice in one target) and
        # repeated target positions
        m = int(rng.integers(1, max(2, n // 5)))
        src = rng.integers(0, n, m)
        dst = rng.integers(0, n, m)
        cur[dst] = cur[src]
        m = int(rng.integers(1, max(2, n // 8)))
        src = rng.integers(0, n, m)
        dst = rng.integers(0, n, m)
        ext[dst] = ext[src]
    lo = min(int(cur.min()), int(ext.min()))
    if lo < 0:
        cur = cur - lo
        ext = ext - lo
    # group order: curPos ascending, equal curPos in a random (emission) order
    perm = rng.permutation(len(cur))
    cur, ext = cur[perm], ext[perm]
    order = np.argsort(cur, kind="stable")
    return cur[order].astype(np.int32), ext[order].astype(np.int32)


def _overlap_list(rng: np.random.Generator, n: int):
    """A read-read overlap: hits along one diagonal (with indel drift) inside
    the region two reads of lengths ql, el share, trimmed by a random overhang
    at either end, plus uniform spurious hits; group order as _hit_list."""
    ql = int(rng.integers(2500, 14000))
    el = int(rng.integers(2500, 14000))
    d = int(rng.integers(-(ql - 1200), el - 1200))
    lo, hi = max(0, -d), min(ql, el - d) - 15
    lo += int(rng.integers(0, 400)) if rng.random() < 0.5 else 0
    hi -= int(rng.integers(0, 400)) if rng.random() < 0.5 else 0
    if hi - lo < 50:
        lo, hi = max(0, -d), min(ql, el - d) - 15
    n_true = max(3, int(n * rng.uniform(0.5, 0.95)))
    cur = np.sort(rng.integers(lo, max(lo + 1, hi), n_true))
    drift = np.cumsum(rng.normal(0.0, rng.uniform(0.0, 2.5), n_true)).astype(np.int64)
    ext = np.clip(cur + d + drift, 0, el - 15)
    n_spur = n - n_true
    cur = np.concatenate([cur, rng.integers(0, ql - 15, n_spur)])
    ext = np.concatenate([ext, rng.integers(0, el - 15, n_spur)])
    perm = rng.permutation(len(cur))
    cur, ext = cur[perm], ext[perm]
    order = np.argsort(cur, kind="stable")
    return cur[order].astype(np.int32), ext[order].astype(np.int32), ql, el


def _ids(rng):
    qid = int(rng.integers(1, 5000)) * (1 if rng.random() < 0.5 else -1)
    u = rng.random()
    if u < 0.1:
        eid = qid                     # self (overlapTest self intersection)
    elif u < 0.2:
        eid = -qid                    # reverse-complement self
    else:
        eid = int(rng.integers(1, 5000)) * (1 if rng.random() < 0.5 else -1)
    return qid, eid


def _lengths(rng, cur, ext):
    span_c = int(cur.max()) + 15
    span_e = int(ext.max()) + 15
    ql = span_c + int(rng.integers(0, 3000))
    el = span_e + int(rng.integers(0, 3000))
    return ql, el


def _pack(lists, **extra):
    off = np.zeros(len(lists) + 1, np.int64)
    off[1:] = np.cumsum([len(c) for c, _ in lists])
    cur = np.concatenate([c for c, _ in lists]).astype(np.int32) if lists else np.zeros(0, np.int32)
    ext = np.concatenate([e for _, e in lists]).astype(np.int32) if lists else np.zeros(0, np.int32)
    return dict(hit_offsets=off, cur=cur, ext=ext, **extra)


def gen_replay_group(R, rng, n_groups=400):
    flag = {"hit": False}
    base_key = R.CmpToKey

    class FlagKey(base_key):
        def __call__(self, obj):
            flag["hit"] = True
            return base_key.__call__(self, obj)

    R.CmpToKey = FlagKey
    lists, meta, rows = [], [], []
    row_off = [0]
    dropped = 0
    try:
        while len(lists) < n_groups:
            n = int(rng.integers(3, 500)) if rng.random() < 0.85 else int(rng.integers(500, 1200))
            if rng.random() < 0.5:
                cur, ext = _hit_list(rng, n, tie_heavy=True)
                qid, eid = _ids(rng)
                ql, el = _lengths(rng, cur, ext)
            else:
                cur, ext, ql, el = _overlap_list(rng, n)
                qid, eid = _ids(rng)
                if abs(qid) == abs(eid):
                    el = ql = max(ql, el)
                # repeated positions (a k-mer seen twice in one target)
                m = int(rng.integers(1, max(2, n // 5)))
                cur[rng.integers(0, n, m)] = cur[rng.integers(0, n, m)]
                ext[rng.integers(0, n, m)] = ext[rng.integers(0, n, m)]
                order = np.argsort(cur, kind="stable")
                cur, ext = cur[order], ext[order]
            min_ovl = int(rng.choice([100, 300, 500, 1000]))
            group = [R.Match(int(c), int(e), eid, 0, i) for i, (c, e) in enumerate(zip(cur, ext))]
            flag["hit"] = False
            det, _ = R.replay_group(qid, ql, eid, el, group, kmer_size=15, max_jump=1500,
                                    min_overlap=min_ovl, force_local=False)
            if flag["hit"]:
                dropped += 1
                continue
            lists.append((cur, ext))
            meta.append((qid, ql, eid, el, min_ovl))
            for o in det:
                rows.append((o.cur_begin, o.cur_end, o.ext_begin, o.ext_end, o.score, o.chain_length,
                             o.seq_divergence))
            row_off.append(len(rows))
    finally:
        R.CmpToKey = base_key
    m = np.array(meta, np.int64)
    r = np.array(rows, np.float64).reshape(-1, 7)
    out = _pack(lists, cur_id=m[:, 0], cur_len=m[:, 1].astype(np.int32), ext_id=m[:, 2],
                ext_len=m[:, 3].astype(np.int32), min_overlap=m[:, 4].astype(np.int32),
                row_offsets=np.array(row_off, np.int64), rows_int=r[:, :6].astype(np.int32),
                rows_div_f64=r[:, 6])
    print(f"replay_group: {len(lists)} groups, {len(rows)} rows, {dropped} dropped (depth limit)")
    return out


CT_FAMILIES = [
    # (name, parameter overrides, with filtered positions)
    ("default_overhang", dict(maximum_overhang=1500, check_overhang=True, force_local=False, only_max_ext=True,
                              max_divergence=1.0, minimum_overlap=1000), False),
    ("tight_overhang", dict(maximum_overhang=300, check_overhang=True, force_local=False, only_max_ext=True,
                            max_divergence=1.0, minimum_overlap=500), False),
    ("force_local", dict(maximum_overhang=100, check_overhang=True, force_local=True, only_max_ext=True,
                         max_divergence=1.0, minimum_overlap=500), False),
    ("containment", dict(maximum_overhang=0, check_overhang=False, force_local=False, only_max_ext=False,
                         max_divergence=1.0, minimum_overlap=300), False),
    ("containment_div", dict(maximum_overhang=0, check_overhang=False, force_local=False, only_max_ext=False,
                             max_divergence=0.3, minimum_overlap=200), False),
    ("filtered_maxext", dict(maximum_overhang=1500, check_overhang=True, force_local=False, only_max_ext=True,
                             max_divergence=1.0, minimum_overlap=500), True),
    ("filtered_containment", dict(maximum_overhang=0, check_overhang=False, force_local=False,
                                  only_max_ext=False, max_divergence=0.35, minimum_overlap=200), True),
    ("sample_rate_half", dict(maximum_overhang=1500, check_overhang=True, force_local=False, only_max_ext=False,
                              max_divergence=1.0, minimum_overlap=300, sample_rate=0.5), False),
]

CT_BASE = dict(kmer_size=15, maximum_jump=1500, max_jump_gap=500, chain_gap_jump_threshold=100,
               chain_large_gap_penalty=2.0, chain_small_gap_penalty=0.5, min_kmer_survival_rate=0.01,
               sample_rate=1.0, max_overlaps=0)


def _tie_free_ct(loc) -> bool:
    if loc is None or "score_table" not in loc:
        return True  # returned at a prefilter: no sort ran
    k = int(loc["kmer_size"])
    scores = [s for s in loc["score_table"] if s > k]
    if len(set(scores)) != len(scores):
        return False
    if loc.get("ext_sorted"):
        tp = [m.target_pos for m in loc["work_matches"]]
        if len(set(tp)) != len(tp):
            return False
    ov = [o.score for o in loc.get("ext_overlaps", [])]
    return len(set(ov)) == len(ov)


def gen_chain_target_family(O, rng, fi, per_family=220):
    fam_out = {}
    for name, over, with_filt in [CT_FAMILIES[fi]]:
        params = dict(CT_BASE)
        params.update(over)
        lists, meta, rows, filt_all = [], [], [], []
        row_off, filt_off = [0], [0]
        tried = 0
        with _Capture(O.chain_target.__code__) as cap:
            while len(lists) < per_family:
                tried += 1
                n = int(rng.integers(3, 500))
                if rng.random() < 0.8:
                    cur, ext, ql, el = _overlap_list(rng, n)
                    qid, eid = _ids(rng)
                    if abs(qid) == abs(eid):
                        el = ql = max(ql, el)
                else:
                    cur, ext = _hit_list(rng, n, tie_heavy=False)
                    qid, eid = _ids(rng)
                    ql, el = _lengths(rng, cur, ext)
                if with_filt:
                    nf = int(rng.integers(0, 60))
                    filt = np.sort(rng.integers(0, ql, nf)).astype(np.int32)
                else:
                    filt = np.zeros(0, np.int32)
                manifest = {"query_id": qid, "query_length": ql, "parameters": params}
                cands = [O.Candidate(qid, int(c), 0, eid, int(e), "+" if eid > 0 else "-")
                         for c, e in zip(cur, ext)]
                det = O.chain_target(cands, manifest, {eid: el}, [int(x) for x in filt])
                if not _tie_free_ct(cap.locals):
                    continue
                lists.append((cur, ext))
                meta.append((qid, ql, eid, el))
                filt_all.append(filt)
                filt_off.append(filt_off[-1] + len(filt))
                for o in det:
                    rows.append((o.cur_begin, o.cur_end, o.ext_begin, o.ext_end, o.score,
                                 np.float32(o.seq_divergence).view(np.int32)))
                row_off.append(len(rows))
        m = np.array(meta, np.int64)
        pv = np.array([params["kmer_size"], params["minimum_overlap"], params["maximum_jump"],
                       params["max_jump_gap"], params["chain_gap_jump_threshold"],
                       params["maximum_overhang"], int(params["check_overhang"]), int(params["force_local"]),
                       int(params["only_max_ext"])], np.int64)
        pf = np.array([params["chain_large_gap_penalty"], params["chain_small_gap_penalty"],
                       params["min_kmer_survival_rate"], params["max_divergence"], params["sample_rate"]],
                      np.float32)
        d = _pack(lists, cur_id=m[:, 0], cur_len=m[:, 1].astype(np.int32), ext_id=m[:, 2],
                  ext_len=m[:, 3].astype(np.int32), row_offsets=np.array(row_off, np.int64),
                  rows=np.array(rows, np.int32).reshape(-1, 6),
                  filtered=np.concatenate(filt_all).astype(np.int32),
                  filtered_offsets=np.array(filt_off, np.int64), params_i=pv, params_f=pf)
        for key, v in d.items():
            fam_out[f"{name}__{key}"] = v
        print(f"chain_target[{name}]: {len(lists)} groups ({tried} tried), {len(rows)} rows", flush=True)
    return fam_out


def gen_overlap_test(O, rng, n=6000):
    recs, params_rows, expect = [], [], []
    for _ in range(n):
        min_ovl = int(rng.choice([100, 500, 1000, 5000]))
        cur_len = int(rng.integers(1000, 40000))
        ext_len = int(rng.integers(1000, 40000))
        cb = int(rng.integers(0, cur_len))
        ce = min(cur_len, cb + int(rng.integers(0, 3 * min_ovl + 200)))
        if rng.random() < 0.5:
            # length difference near the half-of-min boundary
            er = int((ce - cb) * rng.uniform(0.3, 1.7))
        else:
            er = int(rng.integers(0, 3 * min_ovl + 200))
        eb = int(rng.integers(0, max(1, ext_len)))
        ee = eb + er
        qid = int(rng.integers(1, 100)) * (1 if rng.random() < 0.5 else -1)
        u = rng.random()
        if u < 0.25:
            # self overlap near the diagonal (the intersection branch)
            eid, ext_len = qid, cur_len
            eb = max(0, cb + int(rng.integers(-min_ovl, min_ovl)))
            ee = eb + er
        elif u < 0.5:
            eid, ext_len = -qid, cur_len
        else:
            eid = int(rng.integers(1, 100))
        max_oh = int(rng.choice([0, 100, 1500, 5000]))
        check = bool(rng.random() < 0.7)
        force = bool(rng.random() < 0.3)
        params = {"minimum_overlap": min_ovl, "force_local": force, "check_overhang": check,
                  "maximum_overhang": max_oh}
        o = O.Overlap(cur_id=qid, cur_begin=cb, cur_end=ce, cur_len=cur_len, ext_id=eid, ext_begin=eb,
                      ext_end=ee, ext_len=ext_len, score=0, seq_divergence=0.0)
        recs.append((qid, cb, ce, cur_len, eid, eb, ee, ext_len))
        params_rows.append((min_ovl, max_oh, int(check), int(force)))
        expect.append(int(O.overlap_test(o, params)))
    e = np.array(expect, np.int8)
    print(f"overlap_test: {n} cases, {int(e.sum())} pass")
    return dict(records=np.array(recs, np.int64), params=np.array(params_rows, np.int32), expect=e)


def gen_read_align(A, rng, n_reads=1500):
    alns, off, chains, chain_off, seg_idx, seg_off = [], [0], [], [0], [], [0]
    params = {"maximum_jump": 1500, "max_read_overlap": 400, "minimum_overlap": 1000, "max_separation": 500}
    kept = 0
    with _Capture(A.chain_read_alignments.__code__) as cap:
        while kept < n_reads:
            m = int(rng.integers(1, 40))
            read_len = int(rng.integers(5000, 60000))
            begins = np.sort(rng.choice(read_len, size=m, replace=False))
            nodes = int(rng.integers(2, 8))
            recs = []
            for i, b in enumerate(begins):
                ln = int(rng.integers(200, 8000))
                e = min(read_len, int(b) + ln)
                edge_len = int(rng.integers(500, 20000))
                eb = int(rng.integers(0, 3000)) if rng.random() < 0.7 else int(rng.integers(0, edge_len))
                ee = min(edge_len, eb + (e - int(b)) + int(rng.integers(-100, 100)))
                if rng.random() < 0.6:
                    ee = edge_len - int(rng.integers(0, 2000))
                ee = max(eb + 1, min(edge_len, ee))
                left = int(rng.integers(0, nodes))
                right = int(rng.integers(0, nodes))
                recs.append(A.EdgeOverlap(i, 7, int(b), e, read_len, int(rng.integers(1, 50)), left, right,
                                          int(rng.integers(1, 50)), eb, ee, edge_len,
                                          int(rng.integers(100, 5000)), 0.0))
            acc = A.chain_read_alignments(recs, params)
            act = cap.locals["active"] if cap.locals else []
            sc = [c.score for c in act]
            if len(set(sc)) != len(sc):
                continue
            kept += 1
            for r in recs:
                alns.append((r.read_begin, r.read_end, r.read_len, r.edge_left_node, r.edge_right_node,
                             r.edge_begin, r.edge_end, r.edge_len, r.score, r.edge_id, r.edge_seq_id))
            off.append(len(alns))
            for c in acc:
                chains.append(c.score)
                seg_idx.extend(c.indices)
                seg_off.append(len(seg_idx))
            chain_off.append(len(chains))
    pv = np.array([params["maximum_jump"], params["max_read_overlap"], params["minimum_overlap"],
                   params["max_separation"]], np.int32)
    print(f"read_align: {n_reads} reads, {len(alns)} alignments, {len(chains)} accepted chains")
    return dict(alignments=np.array(alns, np.int32), aln_offsets=np.array(off, np.int64),
                chain_scores=np.array(chains, np.int32), chain_offsets=np.array(chain_off, np.int64),
                segment_index=np.array(seg_idx, np.int32), segment_offsets=np.array(seg_off, np.int64),
                params=pv)


def gen_cpp_sort(R, rng, n_arrays=400):
    flag = {"hit": False}
    base_key = R.CmpToKey

    class FlagKey(base_key):
        def __call__(self, obj):
            flag["hit"] = True
            return base_key.__call__(self, obj)

    R.CmpToKey = FlagKey
    keys_all, perm_all, off = [], [], [0]
    try:
        while len(off) - 1 < n_arrays:
            n = int(rng.integers(0, 3000))
            kind = rng.integers(0, 4)
            if kind == 0:
                keys = rng.integers(0, max(1, n // 4), n)
            elif kind == 1:
                keys = np.sort(rng.integers(0, max(1, n // 3), n))
            elif kind == 2:
                keys = np.sort(rng.integers(0, max(1, n // 3), n))[::-1].copy()
            else:
                keys = rng.integers(0, 8, n)
            items = list(range(n))
            flag["hit"] = False
            kl = [int(x) for x in keys]
            R.cpp_sort_in_place(items, lambda a, b: kl[a] < kl[b])
            if flag["hit"]:
                continue
            keys_all.append(keys.astype(np.uint64))
            perm_all.append(np.array(items, np.int32))
            off.append(off[-1] + n)
    finally:
        R.CmpToKey = base_key
    print(f"cpp_sort: {len(off) - 1} arrays, {off[-1]} keys")
    return dict(keys=np.concatenate(keys_all), perm=np.concatenate(perm_all), offsets=np.array(off, np.int64))


def _modules(reference: str):
    tools = Path(reference) / "tools"
    if not tools.is_dir():
        raise SystemExit(f"no reference tools under {tools}")
    if str(tools) not in sys.path:
        sys.path.insert(0, str(tools))
    import replay_overlap_chains as O  # noqa: E402
    import replay_read_alignment_chains as A  # noqa: E402
    import replay_read_to_graph_source_pack as R  # noqa: E402
    if O.LOGF is None:
        raise SystemExit("glibc logf not found through ctypes")
    return O, A, R


def _part(job):
    """One seeded part; runs in a worker process."""
    reference, kind, idx = job
    O, A, R = _modules(reference)
    kinds = ("replay_group", "chain_target", "overlap_test", "read_align", "cpp_sort")
    rng = np.random.default_rng([SEED, kinds.index(kind), idx])
    if kind == "replay_group":
        return kind, idx, gen_replay_group(R, rng)
    if kind == "chain_target":
        return kind, idx, gen_chain_target_family(O, rng, idx)
    if kind == "overlap_test":
        return kind, idx, gen_overlap_test(O, rng)
    if kind == "read_align":
        return kind, idx, gen_read_align(A, rng)
    return kind, idx, gen_cpp_sort(R, rng)


def _concat_groups(parts):
    """Joins replay_group parts: offsets shift, per-group arrays concatenate."""
    out = {}
    hit_base, row_base = 0, 0
    for pi, p in enumerate(parts):
        skip = 1 if pi else 0  # later parts drop their leading 0 offset
        for key, v in p.items():
            if key == "hit_offsets":
                v = v[skip:] + hit_base
            elif key == "row_offsets":
                v = v[skip:] + row_base
            out.setdefault(key, []).append(v)
        hit_base += int(p["hit_offsets"][-1])
        row_base += int(p["row_offsets"][-1])
    return {k: np.concatenate(v) for k, v in out.items()}


def main() -> int:
    from concurrent.futures import ProcessPoolExecutor

    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--reference", default="/root/reference")
    ap.add_argument("--jobs", type=int, default=6)
    args = ap.parse_args()
    _modules(args.reference)
    jobs = [(args.reference, "replay_group", i) for i in range(4)]
    jobs += [(args.reference, "chain_target", i) for i in range(len(CT_FAMILIES))]
    jobs += [(args.reference, k, 0) for k in ("overlap_test", "read_align", "cpp_sort")]
    results = {}
    with ProcessPoolExecutor(max_workers=args.jobs) as pool:
        for kind, idx, data in pool.map(_part, jobs):
            results.setdefault(kind, {})[idx] = data
    rg = _concat_groups([results["replay_group"][i] for i in sorted(results["replay_group"])])
    np.savez_compressed(HERE / "replay_group_v1.npz", **rg)
    ct = {}
    for i in sorted(results["chain_target"]):
        ct.update(results["chain_target"][i])
    ct["families"] = np.array([f[0] for f in CT_FAMILIES])
    np.savez_compressed(HERE / "chain_target_v1.npz", **ct)
    np.savez_compressed(HERE / "overlap_test_v1.npz", **results["overlap_test"][0])
    np.savez_compressed(HERE / "read_align_v1.npz", **results["read_align"][0])
    np.savez_compressed(HERE / "cpp_sort_v1.npz", **results["cpp_sort"][0])
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
