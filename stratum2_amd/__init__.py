"""MI355X (gfx950) path-tracing hot path of Stratum2's TestRenderer.

The product is the HIP library stratum2_amd/_build/libstratum2.so behind the
C ABI of include/stratum2.h; this package is its thin host-side mirror
(ctypes structures, seeded scene generators, a TestRenderer-like wrapper that
hands torch device tensors to the ABI).
"""
from . import abi, scenes  # noqa: F401
from .renderer import Context, Scene, TestRenderer  # noqa: F401
