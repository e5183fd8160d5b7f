"""MI355X-native OVERTURNS-2D iteration: gfx950 HIP kernels behind a C-ABI.

The product is ``libot2d.so`` (built in-tree by ``build.py`` from ``csrc/``)
whose entry points are declared in ``include/ot2d.h``.  This package holds
the host-side pieces around it: the ctypes mirror of the ABI (``abi``), the
grid and deck writer (``grid``), the request loop (``serve``) and ``Case``, a thin
Python driver equivalent to the Fortran one in ``fortran/``.

There is no CPU fallback: if the HIP library is missing or no GPU is
visible, ``load_library`` raises.
"""
from __future__ import annotations

import ctypes as _C
import os as _os

from . import abi, grid  # noqa: F401

_HERE = _os.path.dirname(_os.path.abspath(__file__))
LIBPATH = _os.path.join(_HERE, "libot2d.so")
_lib = None


def load_library(path: str = LIBPATH) -> _C.CDLL:
    """Load libot2d.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not _os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m turns2d_urans_git_amd.build` "
                           "(the HIP path has no CPU fallback)")
    lib = _C.CDLL(path)
    lib.ot2d_last_error.restype = _C.c_char_p
    lib.ot2d_last_error.argtypes = []
    _lib = lib
    return lib


def last_error() -> str:
    lib = load_library()
    return lib.ot2d_last_error().decode()


def default_params(**overrides) -> abi.Params:
    return abi.default_params(load_library(), "ot2d_", **overrides)


class Case(abi.CaseAPI):
    """A case resident on the current HIP device (see include/ot2d.h)."""

    def __init__(self, params: abi.Params, meshes):
        super().__init__(load_library(), "ot2d_", params, meshes, errfn=last_error)


class Batch(abi.CaseAPI):
    """Independent cases advanced together (ot2d_batch_create): one Params
    and one mesh description per case."""

    def __init__(self, params_list, meshes):
        super().__init__(load_library(), "ot2d_", params_list, meshes, errfn=last_error, batch=True)
