"""The reference's text formats: the IC file C/init.c reads and the state
dumps C/check_value.c writes, through the library's native reader/writers
(include/mhd_hip.h).  Host (numpy) arrays in the state layout."""
from __future__ import annotations

import os

import numpy as np

from ._lib import check, lib


def _host(a, n, what):
    a = np.ascontiguousarray(a, dtype=np.float32).reshape(-1)
    if a.size != n:
        raise ValueError(f"{what} has {a.size} floats, expected {n}")
    return a


def read_source_init(path, nx: int, ny: int, nz: int):
    """(u, b) from a source_init file (C/init.c:34-52): 8 floats per cell,
    i outermost, k innermost."""
    n = nx * ny * nz
    u = np.empty(5 * n, np.float32)
    b = np.empty(3 * n, np.float32)
    check(lib().mhd_read_source_init(os.fsencode(path), u.ctypes.data, b.ctypes.data, nx, ny, nz),
          f"read_source_init {path}")
    return u, b


def write_source_init(path, u, b, nx: int, ny: int, nz: int):
    """The inverse of read_source_init (9 significant digits: exact round trip)."""
    n = nx * ny * nz
    u, b = _host(u, 5 * n, "u"), _host(b, 3 * n, "b")
    check(lib().mhd_write_source_init(os.fsencode(path), u.ctypes.data, b.ctypes.data, nx, ny, nz),
          f"write_source_init {path}")


def write_check_dumps(directory, u, b, nx: int, ny: int, nz: int):
    """C_version_check_u1..u5.dat and _b1..b3.dat in directory (C/check_value.c:53-69)."""
    n = nx * ny * nz
    u, b = _host(u, 5 * n, "u"), _host(b, 3 * n, "b")
    check(lib().mhd_write_check_dumps(os.fsencode(directory), u.ctypes.data, b.ctypes.data, nx, ny,
                                      nz), f"write_check_dumps {directory}")
