"""MI355X (gfx950) implementation of DMOLP's two-phase partitioner path.

The product is libdmolp_hip.so with the C ABI of include/dmolp_hip.h; this
package is the Python view of it (ctypes) plus the build script.
"""
from . import _abi  # noqa: F401

__all__ = ["api", "build"]
