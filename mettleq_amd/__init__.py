"""MI355X (gfx950) statevector engine behind MettleQ's statevector API.

Importing the package is cheap and never touches the GPU; the HIP library
(libmettleq_sv.so, built by ``python -m mettleq_amd.build``) is loaded on the
first call that needs it and its absence raises NativeUnavailableError.
"""
from __future__ import annotations

__version__ = "0.1.0"

__all__ = ["__version__"]
