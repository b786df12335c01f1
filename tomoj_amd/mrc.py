"""MRC files through the library's reader and writer (tj_mrc_* in
include/tomoj_hip.h): InputOutput/MRC_Reader.java and MRC_Writer.java."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional, Tuple

import numpy as np

from ._lib import call


class _Info(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("mode", C.c_int),
                ("little_endian", C.c_int), ("extra_bytes", C.c_int), ("fei", C.c_int),
                ("amin", C.c_float), ("amax", C.c_float), ("amean", C.c_float)]


@dataclass
class MrcInfo:
    nx: int
    ny: int
    nz: int
    mode: int
    little_endian: bool
    extra_bytes: int
    fei: bool
    amin: float
    amax: float
    amean: float


def read_info(path: str) -> MrcInfo:
    """Header as MRC_Reader.getMRCFileInfo reads it (MRC_Reader.java:176-233)."""
    info = _Info()
    call("tj_mrc_read_info", str(path).encode(), C.byref(info))
    return MrcInfo(info.nx, info.ny, info.nz, info.mode, bool(info.little_endian), info.extra_bytes,
                   bool(info.fei), info.amin, info.amax, info.amean)


def read(path: str) -> Tuple[np.ndarray, MrcInfo, Optional[np.ndarray]]:
    """(data nz x ny x nx float32, header, FEI tilt angles or None)."""
    info = read_info(path)
    out = np.empty((info.nz, info.ny, info.nx), np.float32)
    tilts = np.empty(info.nz, np.float32)
    call("tj_mrc_read", str(path).encode(), out.ctypes.data_as(C.c_void_p), tilts.ctypes.data_as(C.c_void_p))
    return out, info, (tilts if info.fei else None)


def write(path: str, data: np.ndarray) -> None:
    """A float stack or volume (nz x ny x nx) as MRC_Writer writes a float image."""
    a = np.ascontiguousarray(data, np.float32)
    if a.ndim == 2:
        a = a[None]
    nz, ny, nx = a.shape
    call("tj_mrc_write", str(path).encode(), a.ctypes.data_as(C.c_void_p), nx, ny, nz)
