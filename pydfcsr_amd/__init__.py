"""MI355X (gfx950) implementation of pyDFCSR's 2D CSR wake step.

The hot path runs in libdfcsr.so (hand-written HIP, C ABI in
include/dfcsr.h); this package is the host side that mirrors the
reference's interface: CSR2D, DF_tracker, Beam, Lattice.
"""
from ._lib import DfcsrError, load as load_library  # noqa: F401

__version__ = "0.1.0"


def __getattr__(name):
    # import the torch-heavy driver lazily
    if name == "CSR2D":
        from .csr import CSR2D
        return CSR2D
    if name == "DF_tracker":
        from .deposit import DF_tracker
        return DF_tracker
    if name == "Beam":
        from .beams import Beam
        return Beam
    if name == "Lattice":
        from .lattice import Lattice
        return Lattice
    raise AttributeError(name)
