"""FyzzXYZZY step pipeline on MI355X (gfx950): Python side of the C-ABI.

The compute path is the HIP library libfyzz_hip.so in this directory; this
package only binds it (ctypes) and builds the seeded synthetic workloads.
"""
from ._lib import Context, FyzzError, LibraryMissing, lib  # noqa: F401
from . import scenes  # noqa: F401

__all__ = ["Context", "FyzzError", "LibraryMissing", "lib", "scenes"]
