"""quetzalpic_amd: MI355X-native quasistatic xi-slice loop (QuetzalPIC hot path).

The product is the HIP library behind include/quetzal_hip.h; `capi` is its
ctypes binding (no CPU fallback), `build` compiles it for gfx950.
"""
from . import capi  # noqa: F401

__all__ = ["capi"]
