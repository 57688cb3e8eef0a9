"""Loader for libmphip.so, the gfx950 kernel library.

There is no fallback: if the library is missing or cannot initialise the device, every
entry raises.  `build()` in the repository root compiles it in place
(magnetophotonic.jl_amd/csrc/Makefile -> mphip/libmphip.so).
"""
from __future__ import annotations

import ctypes as C
import os

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmphip.so")

_lib = None
_initialised = False


class MPError(RuntimeError):
    """A libmphip entry point returned a non-zero status."""


def lib():
    """The loaded libmphip (ctypes.CDLL) with every signature bound."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MPError(f"libmphip.so not found at {LIB_PATH}; build it with __graft_entry__.build()")
        h = C.CDLL(LIB_PATH)
        abi.bind(h, "mp_", abi.SIGNATURES)
        abi.bind(h, "mp_", abi.DRIVER_SIGNATURES)
        abi.bind(h, "mp_", abi.NOALLOC_SIGNATURES)
        abi.bind(h, "mp_", abi.HOOK_SIGNATURES)
        abi.bind(h, "", abi.RUNTIME_SIGNATURES)
        abi.bind(h, "", abi.SURROGATE_SIGNATURES)
        _lib = h
    return _lib


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = lib().mp_last_error().decode(errors="replace")
        raise MPError(f"{what or 'libmphip'} failed (status {rc}): {msg}")


def init(device: int = 0) -> None:
    """mp_init on first use (selects the device, creates the context)."""
    global _initialised
    if not _initialised:
        check(lib().mp_init(int(device)), "mp_init")
        _initialised = True


def call(name: str, *args) -> None:
    """Call mp_<name>(*args) and raise MPError on a non-zero status."""
    check(getattr(lib(), "mp_" + name)(*args), "mp_" + name)
