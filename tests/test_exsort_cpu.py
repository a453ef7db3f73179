"""The std::sort emulation (cuflye_amd/csrc/exsort.hpp) against the host
libstdc++ std::sort, compiled for the CPU (tests/native/test_exsort.cpp)."""
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def test_exsort_matches_libstdcxx(tmp_path):
    exe = tmp_path / "test_exsort"
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", str(exe), str(ROOT / "tests/native/test_exsort.cpp")],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "failures=0" in out.stdout
