"""Host-side Python bindings of the MI355X LSMPQS level-set library.

The product is the C-ABI shared library built from ../csrc (one per
QSSLIB_REAL); this package mirrors its records with ctypes, loads it, and
builds synthetic porous-medium inputs in the reference's formats.
"""
from .abi import types_for, make_grid, default_options, Arrays  # noqa: F401
