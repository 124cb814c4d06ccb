"""zap_amd — gfx950 (MI355X) texture block encoders and mip generation behind zap_tex.h's C entry points.

The product is libzaptex.so (zap_amd/csrc, C ABI in include/zaptex.h, drop-in header include/zap_tex.h);
this package builds it (zap_amd.build) and binds it for Python callers (zap_amd.binding).
"""
from .binding import (ASTC, BC1, BC3, BC4, BC5, BC6H, BC7, FORMATS, Library, ZapTexError, bc_block_bytes, bc_size, load,
                      mip_levels, mip_sizes)

__all__ = ["ASTC", "BC1", "BC3", "BC4", "BC5", "BC6H", "BC7", "FORMATS", "Library", "ZapTexError", "bc_block_bytes", "bc_size",
           "load", "mip_levels", "mip_sizes"]
