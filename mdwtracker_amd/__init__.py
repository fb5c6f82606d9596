"""mdwtracker_amd — MI355X-native multi-density Woodcock tracking.

The compute path is libmdwt.so (hand-written HIP kernels for gfx950 behind the
C ABI in include/mdwt.h). ``MdwtRenderer`` binds that ABI one call per entry
point; ``Renderer`` mirrors the reference's dtracker::Renderer method names.
Both load libmdwt.so on first use and raise ImportError if it is not built:
there is no CPU fallback. ``mdwtracker_amd.scenes`` (numpy) generates the
seeded workloads and ``mdwtracker_amd.build`` compiles the library.
"""
from __future__ import annotations

import importlib

__all__ = ["MdwtRenderer", "MdwtError", "Renderer"]


def __getattr__(name):
    if name in ("MdwtRenderer", "MdwtError"):
        return getattr(importlib.import_module(".capi", __name__), name)
    if name == "Renderer":
        return importlib.import_module(".renderer", __name__).Renderer
    raise AttributeError(name)
