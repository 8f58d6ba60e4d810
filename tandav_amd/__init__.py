"""MI355X gravity engine for Tandav (jdonnert/Tandav) -- Python front end.

The engine is the C-ABI shared library ``libtandav_gpu.so`` built from
``csrc/*.hip`` (see ``include/tandav_gpu.h``).  This package loads it with
ctypes and mirrors the reference's force-path calls (Domain_Decomposition,
Gravity_Tree_Build, Gravity_Tree_Acceleration, Compute_Acceleration, ...).
There is no CPU fallback: if the library is missing, importing ``Engine``
fails with an error that says how to build it.
"""
from .engine import Engine, TDVError, load_library, NODE_DTYPE, TOP_DTYPE

__all__ = ["Engine", "TDVError", "load_library", "NODE_DTYPE", "TOP_DTYPE"]
