"""Builds the native parts of cuflye_amd in-tree.

Products (under cuflye_amd/lib/, not tracked by git):
  libflyehip.so    - the gfx950 HIP library behind include/flyehip.h
  libflyesynth.so  - seeded synthetic workload generators (host only)
  logf_table.bin   - glibc logf agreement table embedded into libflyehip.so
The oracle (oracle/Makefile) is built by __graft_entry__.build() separately.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "lib"
OBJ = LIB / "obj"

HIP_SOURCES = ["pipeline.hip", "msort.hip", "segsort.hip", "chain.hip", "readalign.hip", "runtime.hip"]
HOST_SOURCES = ["wire.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}")
    out = "\n".join(line for line in proc.stdout.splitlines()
                    if "packed-fp32-ops" not in line and line.strip())
    if out:
        sys.stderr.write(out + "\n")


def build_logf_table() -> Path:
    table = LIB / "logf_table.bin"
    gen = OBJ / "gen_logf_table"
    src = CSRC / "gen_logf_table.cpp"
    if not gen.exists() or gen.stat().st_mtime < src.stat().st_mtime:
        _run(["g++", "-O2", "-std=c++17", "-o", str(gen), str(src)])
    if not table.exists() or table.stat().st_mtime < gen.stat().st_mtime:
        _run([str(gen), str(table)])
    return table


def build(verbose: bool = False) -> None:
    LIB.mkdir(exist_ok=True)
    OBJ.mkdir(exist_ok=True)
    table = build_logf_table()
    blob = OBJ / "logf_blob.o"
    _run(["g++", "-c", "-fPIC", "-O2", f"-DFH_LOGF_TABLE_PATH={table}",
          "-o", str(blob), str(CSRC / "logf_blob.cpp")])

    def compile_one(name: str) -> Path:
        obj = OBJ / (name + ".o")
        _run([HIPCC, *HIP_FLAGS, "-c", "-o", str(obj), str(CSRC / name)])
        return obj

    with ThreadPoolExecutor(max_workers=len(HIP_SOURCES)) as pool:
        objs = list(pool.map(compile_one, HIP_SOURCES))
    # host-only parts (the canonical dump hash) with the host compiler
    host_objs = []
    for name in HOST_SOURCES:
        obj = OBJ / (name + ".o")
        _run(["g++", "-O3", "-std=c++17", "-fPIC", "-Wall", "-c", "-o", str(obj), str(CSRC / name)])
        host_objs.append(str(obj))
    tmp = LIB / "libflyehip.so.tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp),
          *[str(o) for o in objs], *host_objs, str(blob), "-lcrypto"])
    os.replace(tmp, LIB / "libflyehip.so")
    stmp = LIB / "libflyesynth.so.tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-o", str(stmp),
          str(CSRC / "synth.cpp")])
    os.replace(stmp, LIB / "libflyesynth.so")


if __name__ == "__main__":
    build()
    print("built", LIB)
