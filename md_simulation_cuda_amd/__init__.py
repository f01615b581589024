"""MI355X-native 2D binary Lennard-Jones MD step (gfx950).

The compute path is libmdx.so (HIP kernels behind the C ABI of include/mdx.h);
abi.py binds it with ctypes.  The header-compatible C++ classes live in
host/ and build into libmd_env.so.
"""
from . import abi  # noqa: F401

__all__ = ["abi"]
