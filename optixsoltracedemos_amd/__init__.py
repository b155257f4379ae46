"""MI355X-native SolTrace-style ray tracer (HIP, gfx950).

The product is the C-ABI library ``_lib/libsoltrace_hip.so`` (include/soltrace_hip.h)
and the C++ drop-in classes in ``lib/``; this Python package is a thin ctypes
host layer (``tracer``) plus synthetic scene builders (``scenes``).
"""
from . import _abi, scenes  # noqa: F401

__all__ = ["_abi", "scenes"]
