"""Serving loop: a resident k-mer index in HBM and continuous batching of
incoming read batches (overlap requests).

Requests arrive at any time (a thread calls `submit`).  Each context given to
the server drives one loop thread, so as many GPU iterations are in flight as
there are contexts (each context holds the index).  A loop thread fills its iteration with query records taken in
arrival order from every admitted request (oldest request first, a large
request split across iterations), up to the iteration's read budget and the
device-memory admission budget.  A request that arrives while iterations run
is admitted into the next one as soon as slots and memory free up, beside the
unfinished parts of earlier requests, so a short request never waits for a
long one to drain.  Every query record is processed independently on the
device, so a request's overlaps are the same whether it runs alone or mixed
with others (tests/test_gpu_serving.py).

Admission: the device bytes an iteration needs per query base are measured
from the library's own accounting.  A context's work buffers only grow, so
flyehip_stats_t.device_bytes_peak less the resident index is what the
largest iteration so far needed; whenever an iteration raises it, that
figure over the iteration's bases becomes the estimate (until the first
one, `INITIAL_BYTES_PER_BASE`; later raises come from larger iterations,
whose figure carries less of the fixed per-call buffers).  The iterations in
flight may together claim at most `admission_bytes`.

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
    rows_done: int = 0
    finished: float = 0.0

    @property
    def n_reads(self) -> int:
        return len(self.offsets) - 1


class OverlapServer:
    """Continuous-batching front end over one or more FlyeHip contexts (each
    with the same index; one iteration in flight per context).  On one GPU a
    single context is fastest: a second context's iterations run beside the
    first's and slow both, and admission then splits the reads into smaller
    iterations."""

    # device bytes per query base before the first measurement (above the
    # measured figure for a genome index)
    INITIAL_BYTES_PER_BASE = 400.0

    def __init__(self, ctx: FlyeHip | list, config: OverlapConfig, admission_bytes: int = 64 << 30,
                 max_iteration_reads: int = 4096):
        self.contexts = list(ctx) if isinstance(ctx, (list, tuple)) else [ctx]
        self.ctx = self.contexts[0]
        self.config = config
        self.admission_bytes = admission_bytes
        self.max_iteration_reads = max_iteration_reads
        self.inbox: "queue.Queue[Request]" = queue.Queue()
        self.active: list[Request] = []
        self.completed: list[Request] = []
        self.iterations = 0
        self.iteration_ms: list[float] = []
        self.iteration_reads: list[int] = []
        self.iteration_spans: list[tuple[float, float]] = []   # wall clock per iteration
        self.bytes_per_base = self.INITIAL_BYTES_PER_BASE
        self.bytes_measured = False
        self.bytes_in_flight = 0.0
        self._lock = threading.Lock()
        self._stop = threading.Event()
        self._threads: list[threading.Thread] = []
        self.error: BaseException | None = None

    # -- client side ----------------------------------------------------------
    def submit(self, req: Request) -> Request:
        req.arrival = time.perf_counter()
        req.rows = [None] * req.n_reads
        self.inbox.put(req)
        return req

    def start(self) -> None:
        for k, c in enumerate(self.contexts):
            th = threading.Thread(target=self._run, args=(c,), name=f"flyehip-serve-{k}", daemon=True)
            th.start()
            self._threads.append(th)

    def stop(self) -> None:
        self._stop.set()
        for th in self._threads:
            th.join()
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
        """Query records for one GPU iteration, taken in arrival order within
        the read budget and what the admission budget leaves."""
        picks = []
        reads = 0
        room = self.admission_bytes - self.bytes_in_flight
        cost_total = 0.0
        for req in self.active:
            while req.next_read < req.n_reads and reads < self.max_iteration_reads:
                i = req.next_read
                cost = float(req.offsets[i + 1] - req.offsets[i]) * self.bytes_per_base
                if cost_total + cost > room and (picks or self.bytes_in_flight > 0):
                    return picks, cost_total
                cost_total += cost
                picks.append((req, i))
                req.next_read += 1
                reads += 1
            if reads >= self.max_iteration_reads:
                break
        return picks, cost_total

    def _run(self, ctx: FlyeHip) -> None:
        info = getattr(ctx, "index_info", None)
        index_bytes = float(info.device_bytes) if info is not None else 0.0
        last_peak = 0.0
        try:
            while True:
                with self._lock:
                    self._admit()
                    picks, cost = self._form_iteration()
                    self.bytes_in_flight += cost
                    # nothing left to take and nothing more will come (the
                    # iterations other threads run finish their own reads)
                    drained = (not picks and self._stop.is_set() and self.inbox.empty()
                               and all(r.next_read >= r.n_reads for r in self.active))
                if drained:
                    break
                if not picks:
                    time.sleep(0.0005)
                    continue
                t_begin = time.perf_counter()
                seqs = [req.bases[req.offsets[i]:req.offsets[i + 1]] for req, i in picks]
                offs = np.zeros(len(seqs) + 1, np.int64)
                offs[1:] = np.cumsum([len(s) for s in seqs])
                bases = np.concatenate(seqs) if seqs else np.zeros(0, np.uint8)
                ids = np.array([req.query_ids[i] for req, i in picks], np.int64)
                try:
                    rows, roff = ctx.get_seq_overlaps_external(bases, offs, ids, self.config)
                    st = ctx.stats()
                finally:
                    with self._lock:
                        self.bytes_in_flight -= cost
                with self._lock:
                    self.iterations += 1
                    self.iteration_ms.append(float(st["ms_total"]))
                    self.iteration_reads.append(len(picks))
                    self.iteration_spans.append((t_begin, time.perf_counter()))
                    peak = float(st.get("device_bytes_peak", 0))
                    if offs[-1] > 0 and peak > last_peak and peak > index_bytes:
                        self.bytes_per_base = (peak - index_bytes) / float(offs[-1])
                        self.bytes_measured = True
                    last_peak = max(last_peak, peak)
                    now = time.perf_counter()
                    for j, (req, i) in enumerate(picks):
                        req.rows[i] = rows[roff[j]:roff[j + 1]].copy()
                        req.rows_done += 1
                    for req in list(self.active):
                        if req.rows_done >= req.n_reads:
                            req.finished = now
                            self.active.remove(req)
                            self.completed.append(req)
                            req.done_event.set()
        except BaseException as exc:  # surfaced by stop()
            self.error = exc
            with self._lock:
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
        busy = _union_seconds(self.iteration_spans)
        return {
            "requests": int(len(lat)),
            "reads": int(reads),
            "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "max_ms": float(lat.max() * 1e3),
            # offered load while the server keeps up (first arrival to last finish)
            "reads_per_s": float(reads / (t1 - t0)) if t1 > t0 else 0.0,
            "iterations": self.iterations,
            "iterations_in_flight": len(self.contexts),
            # the library's time per iteration (one flyehip_overlaps_external call)
            "iteration_ms_mean": float(np.mean(self.iteration_ms)) if self.iteration_ms else 0.0,
            "iteration_reads_mean": float(np.mean(self.iteration_reads)) if self.iteration_reads else 0.0,
            # reads per second of wall clock while at least one iteration ran
            "capacity_reads_per_s": (float(sum(self.iteration_reads) / busy) if busy > 0 else 0.0),
            "busy_s": busy,
            "admission_bytes_per_base": self.bytes_per_base,
        }


def _union_seconds(spans) -> float:
    total = 0.0
    end = None
    for a, b in sorted(spans):
        if end is None or a > end:
            total += b - a
            end = b
        elif b > end:
            total += b - end
            end = b
    return total


def request_rows(req: Request) -> np.ndarray:
    parts = [r for r in req.rows if r is not None and len(r)]
    return np.concatenate(parts) if parts else np.zeros(0, N.OVERLAP_DTYPE)
