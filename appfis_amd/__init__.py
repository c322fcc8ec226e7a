"""appfis_amd -- MI355X (gfx950) backend of the APPFIS iterative-stencil runtime.

The product is the C ABI shared library ``libappfis_hip.so`` (include/appfis_hip.h)
and the C++ drop-in headers under ``Appfis/`` that sit on it.  This Python package is
host plumbing for tests and benchmarks: ``appfis_amd.hip`` binds the C ABI with ctypes.
"""
from . import hip  # noqa: F401

__all__ = ["hip"]
