"""MI355X-native RF -> volume pipeline of the ultrasound toolkit.

The product is ``libcuda_toolkit.so`` (hand-written HIP kernels for gfx950
behind the reference's C++/C ABI, see ``include/``).  This package builds it
in-tree and exposes it to Python through ctypes; ``api`` mirrors the
reference's entry points for Python callers and tests.
"""
from . import _lib  # noqa: F401
from ._lib import load  # noqa: F401

__all__ = ["load"]
