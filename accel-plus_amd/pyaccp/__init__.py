"""Python view of the accp C ABI (ctypes), for tests and the benchmark.

The product is libaccp.so (HIP kernels + C ABI, include/accp.h) and the C++
host around it; this module only loads that library and converts arrays.
"""
from . import elements, lattices  # noqa: F401
from .abi import AccpError, Library, load  # noqa: F401
