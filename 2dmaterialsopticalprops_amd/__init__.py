"""MI355X-native drop-in for the RPA response path of Optabs.x / Sloss.x.

librpa_hip.so (csrc/, C ABI in include/rpa_hip.h) holds the device path:
vertices, spectral-function accumulation, Kramers-Kronig transform, Coulomb,
Dyson/W and epilogues.  The Fortran drivers under fortran/ are the
file-compatible replacements of the reference executables; this Python package
exposes the same ABI through ctypes (_abi) and a host-side runner (host) that
performs the reference's setup steps for in-memory inputs.
"""

from . import _abi  # noqa: F401
from ._abi import Context, RpaError  # noqa: F401
