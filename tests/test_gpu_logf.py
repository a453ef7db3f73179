"""The device's divergence logf equals glibc's logf on every float in
[1, 2^25) (the values 1 / matchRate takes), bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_logf_exhaustive(gpu, oracle_mod):
    """All 209,715,200 floats of [1, 2^25), in blocks of 2^24: the device
    result against glibc's logf (called through the oracle library, which
    links the host libm) bit for bit."""
    lo = int(np.float32(1.0).view(np.uint32))
    hi = int(np.float32(2.0 ** 25).view(np.uint32))
    step = 1 << 24
    checked = 0
    for start in range(lo, hi, step):
        u = np.arange(start, min(start + step, hi), dtype=np.uint32)
        x = u.view(np.float32)
        got = gpu.divergence_logf(x)
        ref = oracle_mod.logf_array(x)
        bad = np.flatnonzero(got.view(np.uint32) != ref.view(np.uint32))
        assert len(bad) == 0, f"{len(bad)} arguments differ, first {x[bad[0]]!r}: {got[bad[0]]!r} vs {ref[bad[0]]!r}"
        checked += len(x)
    assert checked == hi - lo == 25 * (1 << 23)


def test_logf_table_arguments(gpu):
    """Every argument where glibc's logf differs from the correctly rounded
    value (the table the build generates by calling glibc) comes out as glibc's."""
    from pathlib import Path
    import ctypes
    table = Path(__file__).resolve().parent.parent / "cuflye_amd" / "lib" / "logf_table.bin"
    t = np.fromfile(table, dtype=np.uint32).reshape(-1, 2)
    x = t[:, 0].view(np.float32)
    got = gpu.divergence_logf(x)
    assert np.array_equal(got.view(np.uint32), t[:, 1])
    libm = ctypes.CDLL("libm.so.6")
    logf = libm.logf
    logf.restype = ctypes.c_float
    logf.argtypes = [ctypes.c_float]
    sample = np.random.default_rng(1).choice(len(x), 2000, replace=False)
    for i in sample:
        assert np.float32(logf(float(x[i]))).view(np.uint32) == t[i, 1]
