"""Top-level ``prrtc`` module name for code written against the reference's
pybind module (python/bindings.cpp:907-1285), which its plugins and loader
import as ``import prrtc`` with the compiled module's directory on
PYTHONPATH. Putting this directory on PYTHONPATH (or sys.path) gives the same
import here: the names are those of ``prismplan_amd.prrtc``, backed by the
HIP library.

    export PYTHONPATH=/path/to/repo/prismplan_amd/pythonpath:$PYTHONPATH
"""
import sys as _sys
from pathlib import Path as _Path

try:
    import prismplan_amd as _pkg  # noqa: F401
except ImportError:   # the repository root is not on the path yet
    _sys.path.insert(0, str(_Path(__file__).resolve().parents[2]))

from prismplan_amd.prrtc import *  # noqa: E402,F401,F403
from prismplan_amd.prrtc import __all__  # noqa: E402,F401
