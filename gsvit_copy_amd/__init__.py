"""MI355X-native engine for GSvit's 3-D FDTD time-stepping path.

The product is the C-ABI library lib/libgsvit_amd.so (drop-in sv_gpu_new /
sv_pool_gpu_* entry points plus a plain session API) and the gsvit_amd driver;
this package only locates and binds them.
"""
from .engine import Batch, Session, library, LIB_PATH  # noqa: F401

__all__ = ["Batch", "Session", "library", "LIB_PATH"]
