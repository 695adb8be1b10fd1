"""TACO Python bindings under the reference's package name.

python/taco/__init__.py of jcarrots/taco loads its native extension and
exposes version(); this package does the same over the MI355X engine
(libtaco_tcl4.so through taco_amd) and re-exports the engine's host
interface: System, the bath producers, the TCL4 builders, RK4 and the
gradient tapes (taco.adjoint).  With the repository's python/ directory on
sys.path, ``import taco`` works as it does with the reference.
"""
import os as _os
import sys as _sys

_root = _os.path.dirname(_os.path.dirname(_os.path.dirname(_os.path.abspath(__file__))))
if _root not in _sys.path:
    _sys.path.insert(0, _root)

try:
    from taco_amd import _lib as _native_lib

    _native_lib.lib()
except Exception as exc:  # the reference raises ImportError when its extension is missing
    raise ImportError("taco native library not built: run python -m taco_amd.build") from exc

from taco_amd import *  # noqa: E402,F401,F403
from taco_amd import adjoint, workloads  # noqa: E402,F401
from taco_amd import __all__ as _engine_all  # noqa: E402
from taco_amd.tcl4 import version  # noqa: E402

_native = _native_lib

__all__ = ["_native", "version", "adjoint", "workloads"] + [n for n in _engine_all if n != "version"]
