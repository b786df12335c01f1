"""ResolutionEstimation's Fourier shell correlation on the GPU.

``ResolutionEstimation`` follows TomoJ's class of that name
(src/TomoJ_Base/.../tomography/ResolutionEstimation.java): ``fsc()`` runs the
two half reconstructions of generateFSCData (:115-165: the odd images, then the
even ones, each through the CPU-manager driver with the user's parameters) and
correlates them on the device (``tj_fsc``); ``fsc_volumes`` is the static
fsc(ImagePlus, ImagePlus) (:456-557) for host volumes; ``SSNR()`` the
spectral SNR of a reconstruction against one of pure noise (:167-260).
"""
from __future__ import annotations

import ctypes as C
import dataclasses

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
    """ResolutionEstimation(tsSignal, parameters) (:70-74): FSC and SSNR."""

    def __init__(self, ts: TiltSeriesData, params: ReconstructionParameters, device: int = 0):
        self.ts = ts
        self.recParameters = params
        self.device = device
        self.recManFSCodd: ReconstructionThreadManager | None = None
        self.recManFSCeven: ReconstructionThreadManager | None = None
        self.recThManSignal: ReconstructionThreadManager | None = None
        self.recThManNoise: ReconstructionThreadManager | None = None
        self.ringWidth = 4          # :37
        self.pixelSize = 1.0        # :38

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

    def SSNR(self, computeVSSNR: bool = False, noise_seed: int = 1, first_index: int = 0) -> np.ndarray:
        """SSNR(computeVSSNR) (:167-260): the signal reconstruction, a tilt
        series of N(0, 1) noise images (generateSSNRNoiseData, :93-112; TomoJ
        draws them from ImageJ's unseeded generator, here from `noise_seed`)
        supplied and reconstructed the same way, and the nine rows of the ring
        statistics (``tj_ssnr``); returns a 9 x (W/2+1) array.  With
        computeVSSNR the per-pixel SSNR maps form a new tilt series (TomoJ's
        defaults: identity alignment, normalised) reconstructed with the same
        parameters into ``recVSSNR`` (:250-258)."""
        p = self.recParameters
        p.setFscType(ALL_PROJECTIONS)
        self.recThManSignal = self.recThManSignal or ReconstructionThreadManager(self.ts, self.device)
        self.recThManSignal.createRec2(p)
        self.recThManSignal.doRec(p, first_index)
        rng = np.random.default_rng(noise_seed)
        noise = rng.standard_normal(np.shape(self.ts.images)).astype(np.float32)
        tsn = dataclasses.replace(self.ts, images=noise)
        self.recThManNoise = self.recThManNoise or ReconstructionThreadManager(tsn, self.device)
        self.recThManNoise.ts = tsn
        self.recThManNoise.createRec2(p)
        self.recThManNoise.doRec(p, first_index)
        W = self.ts.width
        out = np.zeros((9, W // 2 + 1), np.float64)
        nf = C.c_int()
        maps = np.zeros(np.shape(self.ts.images), np.float32) if computeVSSNR else None
        call("tj_ssnr", self.recThManSignal.projector._ctx, self.recThManNoise.projector._ctx, int(self.ringWidth),
             float(self.pixelSize), _ptr(out), C.byref(nf), _ptr(maps) if computeVSSNR else None)
        self.ssnr = out[:, :nf.value]
        if computeVSSNR:
            self.tsVSSNR = TiltSeriesData(maps, list(self.ts.tilts), normalize=True)
            m = ReconstructionThreadManager(self.tsVSSNR, self.device)
            try:
                m.createRec2(p)
                m.doRec(p, first_index)
                self.recVSSNR = m.getRec2()
            finally:
                m.close()
        return self.ssnr

    def close(self) -> None:
        for m in (self.recThManSignal, self.recThManNoise):
            if m is not None:
                m.close()
        self.recThManSignal = self.recThManNoise = None
        for m in (self.recManFSCodd, self.recManFSCeven):
            if m is not None:
                m.close()
        self.recManFSCodd = self.recManFSCeven = None
