"""ResolutionEstimation's Fourier shell correlation on the GPU.

``ResolutionEstimation`` follows TomoJ's class of that name
(src/TomoJ_Base/.../tomography/ResolutionEstimation.java): ``fsc()`` runs the
two half reconstructions of generateFSCData (:115-165: the odd images, then the
even ones, each through the CPU-manager driver with the user's parameters) and
correlates them on the device (``tj_fsc``); ``fsc_volumes`` is the static
fsc(ImagePlus, ImagePlus) (:456-557) for host volumes.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import call
from .parameters import (ALL_PROJECTIONS, EVEN_PROJECTIONS, ODD_PROJECTIONS, ReconstructionParameters,
                         ReconstructionThreadManager, TiltSeriesData)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def fsc_volumes(even: np.ndarray, odd: np.ndarray, device: int = 0) -> np.ndarray:
    """fsc(recFSCeven, recFSCodd) of two (D, H, W) volumes: a 4 x (W/2+1) array
    of frequency, FSC, 2/sqrt(n) and mean |E - O| per shell."""
    e = np.ascontiguousarray(even, dtype=np.float32)
    o = np.ascontiguousarray(odd, dtype=np.float32)
    if e.shape != o.shape or e.ndim != 3:
        raise ValueError("fsc_volumes: two volumes of the same (D, H, W) shape")
    D, H, W = e.shape
    out = np.zeros((4, W // 2 + 1), np.float64)
    call("tj_fsc_volumes", int(device), _ptr(e), _ptr(o), W, H, D, _ptr(out))
    return out


class ResolutionEstimation:
    """ResolutionEstimation(tsSignal, parameters) (:70-74), FSC part."""

    def __init__(self, ts: TiltSeriesData, params: ReconstructionParameters, device: int = 0):
        self.ts = ts
        self.recParameters = params
        self.device = device
        self.recManFSCodd: ReconstructionThreadManager | None = None
        self.recManFSCeven: ReconstructionThreadManager | None = None

    def generateFSCData(self, first_index: int = 0) -> None:
        """:115-165: the odd half, then the even half, then back to all images."""
        p = self.recParameters
        self.recManFSCodd = self.recManFSCodd or ReconstructionThreadManager(self.ts, self.device)
        p.setFscType(ODD_PROJECTIONS)
        self.recManFSCodd.createRec2(p)
        self.recManFSCodd.doRec(p, first_index)
        self.recManFSCeven = self.recManFSCeven or ReconstructionThreadManager(self.ts, self.device)
        p.setFscType(EVEN_PROJECTIONS)
        self.recManFSCeven.createRec2(p)
        self.recManFSCeven.doRec(p, first_index)
        p.setFscType(ALL_PROJECTIONS)

    def fsc(self, first_index: int = 0) -> np.ndarray:
        """:559-568: both halves, then their correlation on the device."""
        self.generateFSCData(first_index)
        W = self.recParameters.width
        out = np.zeros((4, W // 2 + 1), np.float64)
        call("tj_fsc", self.recManFSCeven.projector._ctx, self.recManFSCodd.projector._ctx, _ptr(out))
        return out

    def close(self) -> None:
        for m in (self.recManFSCodd, self.recManFSCeven):
            if m is not None:
                m.close()
        self.recManFSCodd = self.recManFSCeven = None
