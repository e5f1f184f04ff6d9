"""MI355X (gfx950) FCFB + Welch hot path of spektri behind a C ABI.

The compute lives in ``libspektri_gpu.so`` (HIP); this package holds its
build script, a ctypes binding, a host-side mirror of the reference's DSP
interface (``dsp``), and the seeded synthetic workloads (``workloads``).
"""
from . import _lib  # noqa: F401

__all__ = ["_lib"]
