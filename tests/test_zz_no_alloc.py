"""flyehip_read_align_chains_noalloc (device buffers in, device buffers out,
intermediate buffers in the library's fixed per-device array) against
flyehip_read_align_chains on the same input, through the C ABI: two calls in
a row, the second with a smaller input, so the second reads the array after
the first has filled it."""
import ctypes as C

import numpy as np
import pytest

from cuflye_amd import native as N

pytestmark = pytest.mark.gpu


def _alignments(rng, n_reads):
    rows, offs = [], [0]
    for r in range(n_reads):
        m = int(rng.integers(0, 40))
        pos = 0
        for _ in range(m):
            pos = max(0, pos + int(rng.integers(-200, 3000)))
            ln = int(rng.integers(100, 5000))
            elen = int(rng.integers(500, 20000))
            eb = int(rng.integers(0, 2000))
            ee = min(elen - 1, eb + ln + int(rng.integers(-50, 50)))
            rows.append((r + 1, int(rng.integers(1, 30)), int(rng.integers(1, 30)), pos, pos + ln, 100000,
                         int(rng.integers(0, 12)), int(rng.integers(0, 12)), eb, max(eb, ee), elen,
                         int(rng.integers(1, 5)) * 100, 0.01))
        offs.append(len(rows))
    a = np.zeros(len(rows), N.EDGE_ALIGNMENT_DTYPE)
    for i, t in enumerate(rows):
        a[i] = t
    return a, np.array(offs, np.int64)


class _Hip:
    """The caller's own device memory, through the HIP runtime libflyehip.so
    links (torch bundles a second HIP runtime, which cannot share the device
    with this one inside one process)."""

    def __init__(self):
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.held = []

    def put(self, arr: np.ndarray) -> C.c_void_p:
        arr = np.ascontiguousarray(arr)
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), max(8, arr.nbytes)) == 0
        self.held.append(p)
        assert self.rt.hipMemcpy(p, arr.ctypes.data_as(C.c_void_p), arr.nbytes, 1) == 0  # host to device
        return p

    def get(self, p: C.c_void_p, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(like)
        assert self.rt.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.held:
            self.rt.hipFree(p)
        self.held = []


def _noalloc(gpu, a, off, long_edge, big_aln, params, seg_cap=None):
    n = len(off) - 1
    hip = _Hip()
    try:
        cap = max(1, 4 * len(a)) if seg_cap is None else seg_cap
        si0 = np.full(max(1, len(a)), -1, np.int32)
        ko0 = np.full(n + 1, -1, np.int64)
        seg0 = np.zeros(max(1, cap), N.SEGMENT_DTYPE)
        d_a, d_off = hip.put(a), hip.put(off)
        d_si, d_ko, d_seg, d_so = hip.put(si0), hip.put(ko0), hip.put(seg0), hip.put(ko0)
        need = C.c_int64(-1)
        rc = gpu.lib.flyehip_read_align_chains_noalloc(gpu.ctx, d_a, d_off, n, long_edge, big_aln, C.byref(params),
                                                       d_si, d_ko, d_seg, cap, d_so, C.byref(need))
        ko = hip.get(d_ko, ko0)
        si = hip.get(d_si, si0)
        return (rc, si[:ko[-1]] if rc == 0 else None, ko, hip.get(d_seg, seg0), hip.get(d_so, ko0), need.value)
    finally:
        hip.free()


def test_noalloc_equals_allocating_entry(gpu):
    params = N.ChainParams(1500, 50, 1000, 500)
    rng = np.random.default_rng(41)
    for n_reads in (3000, 400):  # the second call is smaller
        a, off = _alignments(rng, n_reads)
        print(f"[noalloc] reads={n_reads} alignments={len(a)}", flush=True)
        ei, ek, es, eo = gpu.read_align_chains(a, off, 900, 500, params)
        rc, gi, gk, gs, go, need = _noalloc(gpu, a, off, 900, 500, params)
        assert rc == N.FLYEHIP_OK, gpu.lib.flyehip_last_error(gpu.ctx).decode()
        assert need == len(es) > 0
        assert np.array_equal(gk, ek)
        assert np.array_equal(gi, ei)
        assert np.array_equal(go, eo)
        assert gs[:need].tobytes() == es.tobytes()


def test_noalloc_limits_and_capacity(gpu):
    params = N.ChainParams(1500, 50, 1000, 500)
    a, off = _alignments(np.random.default_rng(42), 200)
    _, _, es, _ = gpu.read_align_chains(a, off, 900, 500, params)
    rc, _, _, _, _, need = _noalloc(gpu, a, off, 900, 500, params, seg_cap=1)
    assert rc == N.FLYEHIP_ERR_CAPACITY and need == len(es)
    # more reads than the fixed array serves: refused before any device work
    need = C.c_int64(0)
    rc = gpu.lib.flyehip_read_align_chains_noalloc(gpu.ctx, None, None, (1 << 17) + 1, 900, 500, C.byref(params),
                                                   None, None, None, 0, None, C.byref(need))
    assert rc == N.FLYEHIP_ERR_ARG
    assert "no-allocation limit" in gpu.lib.flyehip_last_error(gpu.ctx).decode()
