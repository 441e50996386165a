"""MI355X-native coarse diff2 search (RELION), host-side Python binding.

The compute path is ``librelion_coarse_amd.so`` (gfx950 HIP kernels behind a
C ABI, declared in ``include/relion_coarse.h``); this package loads it with
ctypes and adds the seeded synthetic-input generator used by the tests and
the benchmark.
"""
from . import synth  # noqa: F401

__all__ = ["synth"]
