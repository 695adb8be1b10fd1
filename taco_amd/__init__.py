"""taco_amd — MI355X (gfx950) TCL4 engine: F/C/R -> MIKX -> GW -> L4 -> RK4.

The compute path is hand-written HIP in ``libtaco_tcl4.so`` behind the C ABI
declared in ``include/taco_tcl4.h``; this package mirrors the reference's
host interface on top of it.  There is no CPU fallback.
"""
from .tcl4 import *  # noqa: F401,F403
from .tcl4 import __all__ as _tcl4_all

__all__ = list(_tcl4_all)
