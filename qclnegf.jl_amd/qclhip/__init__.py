"""qclhip: MI355X (gfx950) operator backend for the SCBA inner loop of QCLNEGF.jl.

The compute runs in hand-written HIP kernels of libqclhip.so (plus rocBLAS
for the plain GEMMs of the kernel contractions); this package is the host
mirror of the reference's operator interface over that library's C-ABI.
"""
from . import _lib
from ._lib import (ArgumentError, DimensionMismatch, DomainError, QCLHipError, SingularException,
                   UnsupportedError)

__all__ = ["_lib", "ArgumentError", "DimensionMismatch", "DomainError", "QCLHipError",
           "SingularException", "UnsupportedError"]
