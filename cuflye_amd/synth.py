"""Seeded synthetic workloads of the benchmark configurations.

Each function returns plain numpy arrays; the generators themselves are C++
(csrc/synth.cpp, libflyesynth.so) so a 200 Mb genome is generated in seconds.
Sizes default to the configurations' full sizes; `scale` shrinks genome and
read counts for tests (the read-length and error models stay the same).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import native as N

_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        _LIB = N.load_synth()
    return _LIB


@dataclass
class ReadSet:
    bases: np.ndarray        # uint8 ACGT
    offsets: np.ndarray      # int64, n + 1
    genome_len: int
    start: np.ndarray        # template start per read
    strand: np.ndarray       # 1 forward, 0 reverse

    @property
    def n_reads(self) -> int:
        return len(self.offsets) - 1

    def read(self, i: int) -> bytes:
        return self.bases[self.offsets[i]:self.offsets[i + 1]].tobytes()


def _take_reads(h) -> ReadSet:
    lib = _lib()
    nb = lib.syn_bases_size(h, 0)
    n = lib.syn_count(h, 0)
    bases = np.empty(nb, np.uint8)
    offs = np.empty(n + 1, np.int64)
    lib.syn_copy_bases(h, 0, N.ptr(bases), N.ptr(offs))
    start = np.empty(n, np.int64)
    strand = np.empty(n, np.int32)
    lib.syn_copy_read_truth(h, N.ptr(start), N.ptr(strand))
    glen = lib.syn_bases_size(h, 2)
    return ReadSet(bases, offs, int(glen), start, strand)


def lognormal_mean_sd(mean: float, sd: float) -> tuple[float, float]:
    s2 = math.log(1.0 + (sd / mean) ** 2)
    return math.log(mean) - s2 / 2.0, math.sqrt(s2)


def lognormal_mean_n50(mean: float, n50: float) -> tuple[float, float]:
    # lognormal: mean = exp(mu + s2/2), N50 = exp(mu + s2)
    s2 = 2.0 * math.log(n50 / mean)
    return math.log(mean) - s2 / 2.0, math.sqrt(s2)


def config1(seed: int = 1, scale: float = 1.0) -> ReadSet:
    """500 kb uniform genome + 20 inserted exact 6 kb copies; 30x HiFi-like
    reads, log-normal mean 12 kb sd 3 kb (min 2 kb), 0.1% substitutions."""
    mu, sigma = lognormal_mean_sd(12000.0, 3000.0)
    h = _lib().syn_read_set(int(500_000 * scale), 0.5, max(1, int(20 * scale)), 6000, 0.0, 0, 0, 0, 0, 0.0,
                            0.0, 30.0, mu, sigma, 2000, 1_000_000, 0.001, 0.0, 0.0, 1.0, 1.0, seed)
    try:
        return _take_reads(h)
    finally:
        _lib().syn_destroy(h)


def config2(seed: int = 2, scale: float = 1.0, coverage: float = 40.0) -> ReadSet:
    """5 Mb genome, GC 50%, 2% of bases in 5-10 kb repeat families of 2-6
    copies at 1% divergence; 40x ONT-like reads (log-normal mean 20 kb, N50
    ~30 kb, max 150 kb), 7% errors ins:del:sub = 2:2:1 with deletions three
    times likelier inside homopolymers."""
    mu, sigma = lognormal_mean_n50(20000.0, 30000.0)
    h = _lib().syn_read_set(int(5_000_000 * scale), 0.5, 0, 0, 0.02, 5000, 10000, 2, 6, 0.01, 0.01,
                            coverage, mu, sigma, 500, 150_000, 0.07, 2.0, 2.0, 1.0, 3.0, seed)
    try:
        return _take_reads(h)
    finally:
        _lib().syn_destroy(h)


@dataclass
class HitLists:
    cur: np.ndarray
    ext: np.ndarray
    offsets: np.ndarray
    cur_id: np.ndarray
    ext_id: np.ndarray
    cur_len: np.ndarray
    ext_len: np.ndarray


def config3(n_lists: int = 2_000_000, seed: int = 3, max_hits: int = 200_000) -> HitLists:
    """Pareto(alpha 1.5, min 8) list sizes capped at 200,000: a drifting
    diagonal (random walk sd 0.05 per sqrt(bp), gaps up to 50 bp) plus 30%
    uniform spurious hits per list."""
    lib = _lib()
    h = lib.syn_hit_lists(n_lists, 1.5, 8, max_hits, 25, 0.05, 0.3, seed)
    try:
        nh = lib.syn_hits_size(h)
        n = lib.syn_count(h, 3)
        out = HitLists(np.empty(nh, np.int32), np.empty(nh, np.int32), np.empty(n + 1, np.int64),
                       np.empty(n, np.int64), np.empty(n, np.int64), np.empty(n, np.int32),
                       np.empty(n, np.int32))
        lib.syn_copy_lists(h, N.ptr(out.cur), N.ptr(out.ext), N.ptr(out.offsets), N.ptr(out.cur_id),
                           N.ptr(out.ext_id), N.ptr(out.cur_len), N.ptr(out.ext_len))
        return out
    finally:
        lib.syn_destroy(h)


@dataclass
class RepeatGraph:
    edges_bases: np.ndarray
    edges_offsets: np.ndarray
    left: np.ndarray
    right: np.ndarray
    is_repeat: np.ndarray
    reads: ReadSet


def config4(seed: int = 4, scale: float = 1.0, coverage: float = 30.0) -> RepeatGraph:
    """20 Mb genome with 3% repeats (10-25 kb copies, 2-6 per family, 0.5-2%
    divergence) cut into a repeat graph whose unique stretches are paths of
    edges with Pareto lengths (minimum 1 kb, shape 1.02, at most 2 Mb): about
    3,000 edges of 1 kb - 2 Mb at scale 1; 30x ONT-like reads as in config2."""
    lib = _lib()
    mu, sigma = lognormal_mean_n50(20000.0, 30000.0)
    h = lib.syn_repeat_graph(int(20_000_000 * scale), 0.5, 0.03, 10000, 25000, 2, 6, 0.005, 0.02, coverage,
                             mu, sigma, 500, 150_000, 0.07, 2.0, 2.0, 1.0, 3.0, 1000, 2_000_000, 1.02, seed)
    try:
        reads = _take_reads(h)
        ne = lib.syn_count(h, 1)
        nb = lib.syn_bases_size(h, 1)
        eb = np.empty(nb, np.uint8)
        eo = np.empty(ne + 1, np.int64)
        lib.syn_copy_bases(h, 1, N.ptr(eb), N.ptr(eo))
        left = np.empty(ne, np.int32)
        right = np.empty(ne, np.int32)
        rep = np.empty(ne, np.int32)
        lib.syn_copy_edges(h, N.ptr(left), N.ptr(right), N.ptr(rep))
        return RepeatGraph(eb, eo, left, right, rep, reads)
    finally:
        lib.syn_destroy(h)


def genome_with_repeats(length: int, gc: float, rep_frac: float, rep_min: int, rep_max: int,
                        copies_min: int, copies_max: int, div_min: float, div_max: float,
                        seed: int) -> np.ndarray:
    """A genome alone (config 5's 200 Mb reference: GC 42%, 8% repeats in
    families of 2-50 copies at 0.5-5% divergence)."""
    lib = _lib()
    h = lib.syn_read_set(length, gc, 0, 0, rep_frac, rep_min, rep_max, copies_min, copies_max, div_min,
                         div_max, 0.0, 0.0, 0.0, 1, 1, 0.0, 0.0, 0.0, 1.0, 1.0, seed)
    try:
        glen = lib.syn_bases_size(h, 2)
        g = np.empty(glen, np.uint8)
        lib.syn_copy_bases(h, 2, N.ptr(g), None)
        return g
    finally:
        lib.syn_destroy(h)


def reads_from(genome: np.ndarray, n_reads: int, seed: int, mean: float = 20000.0,
               n50: float = 30000.0, err: float = 0.07) -> tuple[np.ndarray, np.ndarray]:
    lib = _lib()
    mu, sigma = lognormal_mean_n50(mean, n50)
    h = lib.syn_reads_from(N.ptr(genome), len(genome), n_reads, mu, sigma, 500, 150_000, err, 2.0, 2.0, 1.0,
                           3.0, seed)
    try:
        nb = lib.syn_bases_size(h, 0)
        n = lib.syn_count(h, 0)
        b = np.empty(nb, np.uint8)
        o = np.empty(n + 1, np.int64)
        lib.syn_copy_bases(h, 0, N.ptr(b), N.ptr(o))
        return b, o
    finally:
        lib.syn_destroy(h)
