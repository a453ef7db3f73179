"""cuflye_amd - Flye's overlap-detection hot path on AMD MI355X (gfx950).

The product is libflyehip.so (C ABI in include/flyehip.h); this package
holds its Python bindings (native), the host-side mirror of the reference's
backend seam (overlap), and the seeded workload generators (synth).
"""
from .overlap import FlyeHip, FlyeHipError, OverlapConfig, candidate_backend, pack_sequences

__all__ = ["FlyeHip", "FlyeHipError", "OverlapConfig", "candidate_backend", "pack_sequences"]
