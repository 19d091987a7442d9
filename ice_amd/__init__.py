"""MI355X-native explicit residual-and-update path of the ICE dispersed-phase solver.

The compute path is the HIP library ``libice_gpu.so`` (include/ice_gpu.h); this
package holds its ctypes binding (`native`), the host-side mesh metrics
(`geometry`), case builders (`cases`) and the Fortran host (`fortran/`).
"""
from . import native  # noqa: F401

__all__ = ["native", "geometry", "cases"]
