"""Serving loop: a resident k-mer index in HBM and continuous batching of
incoming read batches (overlap requests).

Requests arrive at any time (a thread calls `submit`); the loop thread keeps
one GPU iteration in flight at a time and fills each iteration with query
records taken in arrival order from every admitted request, up to the
iteration's query budget and the device-memory admission budget.  A request
that arrives while an iteration runs is admitted into the next one as soon as
slots and memory free up, beside the unfinished parts of earlier requests, so
a short request never waits for a long one to drain.  Every query record is
processed independently on the device, so a request's overlaps are the same
whether it runs alone or mixed with others (tests/test_gpu_serving.py).

Latency of a request = completion time - arrival time; the loop reports
p50/p99 over requests and the sustained reads/s.
"""
from __future__ import annotations

import queue
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import native as N
from .overlap import FlyeHip, OverlapConfig


@dataclass
class Request:
    rid: int
    bases: np.ndarray          # uint8 ACGT, concatenated reads
    offsets: np.ndarray        # int64, n + 1
    query_ids: np.ndarray      # int64 signed ids, one per read
    arrival: float = 0.0
    done_event: threading.Event = field(default_factory=threading.Event)
    rows: list = field(default_factory=list)       # per read: overlap rows
    next_read: int = 0
    finished: float = 0.0

    @property
    def n_reads(self) -> int:
        return len(self.offsets) - 1

    def bytes_per_read(self) -> float:
        return float(self.offsets[-1]) / max(1, self.n_reads)


class OverlapServer:
    """Continuous-batching front end over one FlyeHip context."""

    # device bytes a query base costs in one iteration (windows, matches and
    # chaining scratch at ~0.35 matches per base for a genome index; measured
    # peaks stay below this estimate)
    BYTES_PER_BASE = 400

    def __init__(self, ctx: FlyeHip, config: OverlapConfig, admission_bytes: int = 64 << 30,
                 max_iteration_reads: int = 4096):
        self.ctx = ctx
        self.config = config
        self.admission_bytes = admission_bytes
        self.max_iteration_reads = max_iteration_reads
        self.inbox: "queue.Queue[Request]" = queue.Queue()
        self.active: list[Request] = []
        self.completed: list[Request] = []
        self.iterations = 0
        self.iteration_ms: list[float] = []
        self.iteration_reads: list[int] = []
        self._stop = threading.Event()
        self._thread: threading.Thread | None = None
        self.error: BaseException | None = None

    # -- client side ----------------------------------------------------------
    def submit(self, req: Request) -> Request:
        req.arrival = time.perf_counter()
        req.rows = [None] * req.n_reads
        self.inbox.put(req)
        return req

    def start(self) -> None:
        self._thread = threading.Thread(target=self._run, name="flyehip-serve", daemon=True)
        self._thread.start()

    def stop(self) -> None:
        self._stop.set()
        if self._thread:
            self._thread.join()
        if self.error:
            raise self.error

    # -- loop -------------------------------------------------------------------
    def _admit(self) -> None:
        while True:
            try:
                self.active.append(self.inbox.get_nowait())
            except queue.Empty:
                return

    def _form_iteration(self):
        """Query records for one GPU iteration, taken in arrival order."""
        picks = []
        reads = 0
        budget = self.admission_bytes
        for req in self.active:
            while req.next_read < req.n_reads and reads < self.max_iteration_reads:
                i = req.next_read
                nb = int(req.offsets[i + 1] - req.offsets[i])
                cost = nb * self.BYTES_PER_BASE
                if picks and cost > budget:
                    return picks
                budget -= cost
                picks.append((req, i))
                req.next_read += 1
                reads += 1
            if reads >= self.max_iteration_reads:
                break
        return picks

    def _run(self) -> None:
        try:
            while not self._stop.is_set() or not self.inbox.empty() or self.active:
                self._admit()
                picks = self._form_iteration()
                if not picks:
                    if self._stop.is_set() and self.inbox.empty():
                        break
                    time.sleep(0.0005)
                    continue
                seqs = [req.bases[req.offsets[i]:req.offsets[i + 1]] for req, i in picks]
                offs = np.zeros(len(seqs) + 1, np.int64)
                offs[1:] = np.cumsum([len(s) for s in seqs])
                bases = np.concatenate(seqs) if seqs else np.zeros(0, np.uint8)
                ids = np.array([req.query_ids[i] for req, i in picks], np.int64)
                rows, roff = self.ctx.get_seq_overlaps_external(bases, offs, ids, self.config)
                self.iterations += 1
                self.iteration_ms.append(float(self.ctx.stats()["ms_total"]))
                self.iteration_reads.append(len(picks))
                now = time.perf_counter()
                for j, (req, i) in enumerate(picks):
                    req.rows[i] = rows[roff[j]:roff[j + 1]].copy()
                for req in list(self.active):
                    if req.next_read >= req.n_reads and all(r is not None for r in req.rows):
                        req.finished = now
                        self.active.remove(req)
                        self.completed.append(req)
                        req.done_event.set()
        except BaseException as exc:  # surfaced by stop()
            self.error = exc
            for req in self.active:
                req.done_event.set()

    # -- report -------------------------------------------------------------------
    def latency_report(self) -> dict:
        lat = np.array([r.finished - r.arrival for r in self.completed])
        reads = sum(r.n_reads for r in self.completed)
        if not len(lat):
            return {"requests": 0}
        t0 = min(r.arrival for r in self.completed)
        t1 = max(r.finished for r in self.completed)
        return {
            "requests": int(len(lat)),
            "reads": int(reads),
            "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "max_ms": float(lat.max() * 1e3),
            # offered load while the server keeps up (first arrival to last finish)
            "reads_per_s": float(reads / (t1 - t0)) if t1 > t0 else 0.0,
            "iterations": self.iterations,
            # the library's time per iteration (one flyehip_overlaps_external call)
            "iteration_ms_mean": float(np.mean(self.iteration_ms)) if self.iteration_ms else 0.0,
            "iteration_reads_mean": float(np.mean(self.iteration_reads)) if self.iteration_reads else 0.0,
            "capacity_reads_per_s": (float(sum(self.iteration_reads) / (sum(self.iteration_ms) / 1e3))
                                     if self.iteration_ms and sum(self.iteration_ms) > 0 else 0.0),
        }


def request_rows(req: Request) -> np.ndarray:
    parts = [r for r in req.rows if r is not None and len(r)]
    return np.concatenate(parts) if parts else np.zeros(0, N.OVERLAP_DTYPE)
