"""C-ABI surface checks that need no GPU: the library loads and exports every
entry point include/flyehip.h declares; the backend selector mirrors Flye's."""
import ctypes
import subprocess
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared_entry_points():
    text = (ROOT / "include" / "flyehip.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(flyehip_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    from cuflye_amd import native as N
    lib = N.load_flyehip()
    names = declared_entry_points()
    assert len(names) >= 15
    assert set(names) == set(N.EXPORTED)
    for name in names:
        assert hasattr(lib, name), name
    assert lib.flyehip_abi_version() == 4


def test_open_without_gpu_fails_closed():
    """No silent CPU path: without a device, open returns an error."""
    import torch  # noqa: F401  (only to know whether a GPU exists here)
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from cuflye_amd.overlap import FlyeHip, FlyeHipError
    with pytest.raises(FlyeHipError):
        FlyeHip(device=0)


def test_backend_selector():
    from cuflye_amd.overlap import candidate_backend
    assert candidate_backend({"CUFLYE_CANDIDATE_BACKEND": "hip"}) == "hip"
    with pytest.raises(RuntimeError, match="Unsupported CUFLYE_CANDIDATE_BACKEND: cuda2"):
        candidate_backend({"CUFLYE_CANDIDATE_BACKEND": "cuda2"})
    with pytest.raises(RuntimeError):
        candidate_backend({})


def test_record_layouts_match_header():
    from cuflye_amd import native as N
    assert ctypes.sizeof(N.OverlapParams) == 18 * 4
    assert ctypes.sizeof(N.IndexInfo) == 16 + 6 * 8
    assert N.OVERLAP_DTYPE.itemsize == 56 and N.CANDIDATE_DTYPE.itemsize == 32


def test_struct_sizes_match_c_compiler(tmp_path):
    """ctypes mirrors against the C compiler's layout of include/flyehip.h."""
    from cuflye_amd import native as N
    src = tmp_path / "sizes.c"
    src.write_text('#include <stdio.h>\n#include "flyehip.h"\n'
                   'int main(void){printf("%zu %zu %zu %zu\\n", sizeof(flyehip_stats_t), '
                   'sizeof(flyehip_overlap_params_t), sizeof(flyehip_index_info_t), '
                   'sizeof(flyehip_overlap_t)); return 0;}\n')
    exe = tmp_path / "sizes"
    root = Path(__file__).resolve().parents[1]
    subprocess.run(["gcc", "-I", str(root / "include"), "-o", str(exe), str(src)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    assert got == [ctypes.sizeof(N.Stats), ctypes.sizeof(N.OverlapParams), ctypes.sizeof(N.IndexInfo),
                   N.OVERLAP_DTYPE.itemsize]
