"""openpv_amd -- MI355X-native synaptic delivery + LCA hot path of OpenPV.

libpvhip.so (include/pvhip.h): device layer and HIP kernels for gfx950.
libpvhost.so (include/pvhost.h): C++ host geometry / tables.
The Python modules are thin ctypes faces over those libraries.
"""
from . import abi  # noqa: F401

__all__ = ["abi"]
