"""MI355X (gfx950) occupancy discovery, inter-workgroup barriers and the
persistent graph kernels built on them.

The product is the shared library ``libgdb_amd.so`` (HIP kernels behind the C
ABI in ``include/gdb_amd.h``); this package is its Python binding.
"""
from .api import *  # noqa: F401,F403
from .api import lib  # noqa: F401
