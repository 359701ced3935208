"""Loader of the HIP library. There is no fallback: if the library is missing, importing fails loudly."""
from __future__ import annotations

import ctypes
from pathlib import Path

from . import abi

LIB_PATH = Path(__file__).resolve().parent / "_build" / "libnoa_hip.so"

_lib: ctypes.CDLL | None = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"noa_amd: the HIP library {LIB_PATH} is not built. Run `python -m noa_amd.build` "
                "(or __graft_entry__.build()). There is no CPU fallback.")
        _lib = abi.declare(ctypes.CDLL(str(LIB_PATH)))
        version = _lib.noa_hip_abi_version()
        if version != abi.ABI_VERSION:
            raise ImportError(f"noa_amd: ABI version mismatch (library {version}, bindings {abi.ABI_VERSION})")
    return _lib


def last_error() -> str:
    return lib().noa_hip_last_error().decode("utf-8", "replace")
