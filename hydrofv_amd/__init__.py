"""hydrofv_amd: hydroFV's MUSCL-Hancock finite-volume update on AMD MI355X.

The product is ``libhydrofv_amd.so`` (C99 host API + HIP kernels for gfx950),
built in-tree by ``__graft_entry__.build()`` / ``make -C hydrofv_amd/csrc``.
This package locates it and attaches ctypes prototypes; there is no CPU
implementation of the integrator to fall back to, so a missing library is an
error.
"""
from __future__ import annotations

import ctypes as C
import os

from . import abi

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libhydrofv_amd.so")

_lib = None


class LibraryMissing(RuntimeError):
    pass


def load() -> C.CDLL:
    """Load libhydrofv_amd.so (once) and return it with prototypes attached."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryMissing(
                f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
                "or `make -C hydrofv_amd/csrc`")
        # PyTorch for ROCm ships its own HIP runtime under the soname
        # libamdhip64.so.7, which the library also links.  Imported first,
        # torch's copy satisfies the library's dependency and the process
        # holds one runtime; loaded the other way round, torch loads a second
        # runtime that finds no GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = C.CDLL(LIB_PATH)
        abi.bind(lib)
        from . import ensemble, gpu
        gpu.bind(lib)
        ensemble.bind(lib)
        _lib = lib
    return _lib


__all__ = ["load", "LIB_PATH", "LibraryMissing", "abi"]
