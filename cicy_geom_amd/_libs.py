"""Loaders for the native libraries built in-tree by cicy_geom_amd.build.

libcicy_dea.so  -- the HIP solver behind the C-ABI of include/cicy_dea.h (gfx950).
libcicy_net.so  -- host C++ network generator and DEA system assembly (netgen.h).

Both are loaded from the package directory only.  A missing library raises
ImportError naming the build step; there is no fallback path.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
_cache: dict[str, ctypes.CDLL] = {}

DEA_LIB = os.path.join(_HERE, "libcicy_dea.so")
NET_LIB = os.path.join(_HERE, "libcicy_net.so")


def _load(path: str) -> ctypes.CDLL:
    if path in _cache:
        return _cache[path]
    if not os.path.exists(path):
        raise ImportError(
            f"{os.path.basename(path)} is not built; run "
            "`python -c 'import __graft_entry__ as g; g.build()'` from the repository root")
    lib = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
    _cache[path] = lib
    return lib


def dea_lib() -> ctypes.CDLL:
    return _load(DEA_LIB)


def net_lib() -> ctypes.CDLL:
    return _load(NET_LIB)
