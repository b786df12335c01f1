"""The GPU manager's side of the boundary: ProjectorGPU's lifecycle and
ReconstructionThreadManagerGPU.doGPURec over the HIP library.

TomoJ's GPU path (src/TomoJ_GPU/.../tomography/ReconstructionThreadManagerGPU.
java:172-308) drives an array of ``ProjectorGPU`` objects (projectors/
ProjectorGPU.java): per device it takes Y-slabs of the reconstruction sized by
the device memory (``getDevice().getOptimumMemoryUse()`` or
``computeMaximumSizeForVolume()``), calls ``initForBP`` / ``initForIterative``
for the slab, runs TomoReconstruction2's WBP / OSSART loop through the Projector
interface with the slab's row range, copies the usable rows back with
``updateFromGPU(startY, endY, usableStart, usableEnd)`` and frees the device
memory with ``releaseCL_Memory``.

``VoxelProjector3DHIPGPU`` is that lifecycle over one library context per slab:
the slab's rows plus one halo row below (TomoJ's sampler takes a volume's last
row, and ImageJ's bilinear the image's, by the nearest value, so a slab's own
last row must not be its context's last), centres lowered by the slab's first
row, the full images supplied once (getPixels: alignment, normalisation,
threshold; and for WBP the weighting, a filter of whole images) and cropped
per slab.  The Projector calls take absolute rows, as TomoJ's do.  The volume
moves between host and device only at initForIterative (the slab's rows),
updateFromGPU and copyInGPUBuffer, never per iteration.  ``doGPURec`` is
doGPURec's loop for one device.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, List, Optional, Sequence

import numpy as np

from ._lib import call
from .parameters import (BAYESIAN, BP, DEFINED_PROJECTIONS, OSSART, TVM, WBP, ReconstructionParameters,
                         TiltSeriesData, supply)
from .projector import VoxelProjector3DHIP
from . import reconstruction as R


class VoxelProjector3DHIPGPU:
    """ProjectorGPU (ProjectorGPU.java:31-438) over the HIP library for one
    device.  `rec` is the reconstruction's host volume (D, H, W), TomoJ's
    rec2 ImageStack; `vol_center` its centres."""

    def __init__(self, ts: TiltSeriesData, rec: np.ndarray, vol_center: Sequence[float],
                 weighting_diameter: Optional[float] = None, device: int = 0):
        self.ts = ts
        self.rec = rec
        self.D, self.H, self.W = rec.shape
        self.vol_center = [float(v) for v in vol_center]
        self.weighting_diameter = weighting_diameter
        self.device = device
        self.scale = 1.0
        self.long_object = False
        self.positivity = False
        self.fista = False
        self.P: Optional[VoxelProjector3DHIP] = None
        self.y0 = self.y1 = 0
        self._images: Optional[np.ndarray] = None
        self.errorVolume: Optional[np.ndarray] = None       # the full-size ImagePlus of createErrorVolume
        self.errorVolumes: Dict[int, np.ndarray] = {}         # endOfIteration's saves, by iteration
        self.currentIteration = 0

    # --------------------------------------------------------------- settings
    def setScale(self, value: float) -> None:                          # :141-143
        if float(value) != 1.0:
            raise ValueError("the slab contexts support scale 1 (detector width = volume width)")
        self.scale = float(value)

    def setLongObjectCompensation(self, on: bool) -> None:             # :145-147
        if on:
            raise ValueError("long-object compensation takes a whole-image maximum: no slabs")
        self.long_object = False

    def setPositivityConstraint(self, on: bool) -> None:               # :149-151
        self.positivity = bool(on)
        if self.P is not None:
            self.P.setPositivityConstraint(self.positivity)

    def setFista(self, on: bool) -> None:
        """The FISTA wrapper (FistaProjector3DGPU.java) run inside the library:
        its extrapolation step needs no volume round trip to the host."""
        self.fista = bool(on)
        if self.P is not None:
            self.P.setFista(self.fista)

    def computeMaximumSizeForVolume(self) -> int:                      # :162-166
        b = C.c_size_t()
        call("tj_max_volume_bytes", int(self.device), C.byref(b))
        return int(b.value)

    def getOptimumMemoryUse(self) -> int:
        """GPUDevice.getOptimumMemoryUse as doGPURec reads it (:184-186): 0, so
        computeMaximumSizeForVolume sizes the slabs."""
        return 0

    # ------------------------------------------------------------- lifecycle
    def _supplied(self) -> np.ndarray:
        """The full images TiltSeries.getPixels gives (weighted for WBP), once."""
        if self._images is None:
            ts = self.ts
            with VoxelProjector3DHIP((ts.width, ts.height, 1), (ts.width, ts.height), device=self.device) as S:
                supply(S, ts)
                imgs = np.stack([S.get_loaded_projection(k) for k in range(len(ts.tilts))])
                if self.weighting_diameter:
                    S.setWeightingFilter(self.weighting_diameter, ts.tilts, ts.tilt_axes())
                    imgs = np.stack([S.weighting(im) for im in imgs])
            self._images = imgs
        return self._images

    def _open(self, startY: int, endY: int, copy: bool) -> None:
        """The slab context.  Iterative slabs get NaN measured pixels in the
        halo row: its differences are then 0 (difference skips NaN), so the
        errors cover the slab's own detector rows, as the slab's error sums of
        doGPURec do, and the backprojection's next-row reads stay finite."""
        if self.P is not None:
            self.releaseCL_Memory()
        ts = self.ts
        if not ts.axis_vertical or ts.height != self.H or ts.width != self.W:
            raise ValueError("slab contexts need the vertical tilt axis and a W x H detector")
        nc = min(endY - startY + 1, self.H - startY)      # the slab and one halo row
        imgs = np.ascontiguousarray(self._supplied()[:, startY:startY + nc, :])
        if copy and nc > endY - startY:
            imgs[:, nc - 1, :] = np.nan
        W, H = self.W, self.H
        pc = [(W - 1) / 2.0, (H - 1) / 2.0, (W - 1) / 2.0, (H - 1) / 2.0]
        vc = self.vol_center
        self.P = VoxelProjector3DHIP((W, nc, self.D), (W, nc), device=self.device, positivity=self.positivity,
                                     proj_center=(pc[0], pc[1] - startY, pc[2], pc[3] - startY),
                                     vol_center=(vc[0], vc[1] - startY, vc[2]))
        self.P.load_tilt_series(imgs, ts.eulers())
        if self.fista:
            self.P.setFista(True)
        self.y0, self.y1 = startY, endY
        if copy:
            self.P.set_volume(np.ascontiguousarray(self.rec[:, startY:startY + nc, :]))

    def initForBP(self, startY: int, endY: int, updateNb: int) -> None:       # :175-190
        """A zero slab (initVolMemBuffer without copy) for the weighted
        backprojection of rows [startY, endY)."""
        self._open(startY, endY, copy=False)

    def initForIterative(self, startY: int, endY: int, update: int, kernelType: int = 1,
                         errorVolume: bool = False) -> None:                   # :211-216
        """The slab's rows of the reconstruction on the device, the images
        resident, and the error volume when asked."""
        self._open(startY, endY, copy=True)
        self.currentIteration = 0
        if errorVolume:
            self.createErrorVolume()

    def createErrorVolume(self) -> None:                                       # :280-291
        if self.errorVolume is None:
            self.errorVolume = np.zeros((self.D, self.H, self.W), np.float32)
        self.P.createErrorVolume()

    def updateFromGPU(self, startY: int, endY: int, usableStartY: Optional[int] = None,
                      usableEndY: Optional[int] = None) -> None:               # :297-315
        """Rows [usableStartY, usableEndY) (default [startY, endY)) of the
        device slab into the host reconstruction."""
        a = startY if usableStartY is None else usableStartY
        b = endY if usableEndY is None else usableEndY
        vol = self.P.get_volume()
        self.rec[:, a:b, :] = vol[:, a - self.y0:b - self.y0, :]

    def copyInGPUBuffer(self, startY: int, endY: int) -> None:                 # :317-330
        """The host rows [startY, endY) (and the halo row) to the device slab."""
        nc = self.P.H
        self.P.set_volume(np.ascontiguousarray(self.rec[:, self.y0:self.y0 + nc, :]))

    def releaseCL_Memory(self) -> None:                                        # :376-403
        if self.P is not None:
            self.P.close()
            self.P = None

    # ------------------------------------------------- Projector (absolute rows)
    def _rows(self, startY: int, endY: int):
        return max(startY, self.y0) - self.y0, min(endY, self.y1) - self.y0

    def addProjection(self, index: int) -> None:
        self.P.addProjection(index)

    def clearAllProjections(self) -> None:
        self.P.clearAllProjections()

    def project(self, startY: int = 0, endY: Optional[int] = None) -> None:
        self.P.project(*self._rows(startY, self.H if endY is None else endY))

    def projectAndDifference(self, factor: float, startY: int = 0, endY: Optional[int] = None) -> np.ndarray:
        return self.P.projectAndDifference(factor, *self._rows(startY, self.H if endY is None else endY))

    def backproject(self, startY: int = 0, endY: Optional[int] = None) -> None:
        self.P.backproject(*self._rows(startY, self.H if endY is None else endY))

    def getProjection(self, index: int) -> np.ndarray:
        return self.P.getProjection(index)

    def startOfIteration(self) -> None:
        self.P.startOfIteration()

    def endOfIteration(self) -> None:
        """VoxelProjector3D.endOfIteration (:826-838): with an error volume,
        its slab rows go into the full error volume, which is saved (kept by
        iteration here) and cleared."""
        self.P.endOfIteration()
        self.currentIteration += 1
        if self.errorVolume is not None and self.P.has_error_volume:
            ev = self.P.getErrorVolume()
            self.errorVolume[:, self.y0:self.y1, :] = ev[:, :self.y1 - self.y0, :]
            self.errorVolumes[self.currentIteration] = self.errorVolume.copy()
            self.P.clearErrorVolume()


def doGPURec(rec2: np.ndarray, ts: TiltSeriesData, projectors: List[VoxelProjector3DHIPGPU],
             params: ReconstructionParameters, use: Optional[Sequence[bool]] = None, saveError: bool = False,
             saveForAll: bool = False, first_index: int = 0, max_bytes: Optional[int] = None,
             slab_errors: Optional[list] = None) -> Optional[List[float]]:
    """ReconstructionThreadManagerGPU.doGPURec (:172-308) for the devices of
    `projectors` taken one after the other: Y-slabs of at most
    memory / (W * D * 4) rows (``max_bytes`` overrides the device figure), per
    slab initForBP + WBP or initForIterative + OSSART through the Projector
    interface with the slab's rows, updateFromGPU over the usable rows (the
    whole slab: computeUsablePart gives {startY, endY, endY} for a vertical
    axis, :310-314) and releaseCL_Memory.  Returns the per-iteration error
    sums of the slabs (TomoJ adds err1[0] of the slabs), or None for
    WBP/BP; `slab_errors`, when given, receives each slab's err rows."""
    D, H, W = rec2.shape
    use = [True] * len(projectors) if use is None else list(use)
    t = params.reconstructionType
    if t in (TVM, BAYESIAN):
        raise ValueError("TVM and Bayesian regularisation couple neighbouring rows: no slabs")
    errors: Optional[List[float]] = None if params.nbIterations == 0 else []
    E = ts.eulers()
    offset = 0
    for d, P in enumerate(projectors):
        if not use[d]:
            continue
        mem = P.getOptimumMemoryUse() or P.computeMaximumSizeForVolume()
        if max_bytes is not None:
            mem = max_bytes
        ymax = max(1, int(mem // (W * D * 4)))
        while offset < H:
            endY = offset + min(ymax, H - offset)
            if t in (BP, WBP):
                P.initForBP(offset, endY, 1)
                R.wbp(P, len(ts.tilts), 0, offset, endY, available=params.getAvailableIndexes(ts))
            else:
                n = params.nbIterations
                if saveError and saveForAll:
                    P.initForIterative(offset, endY, params.updateNb, 1, True)
                    P.setPositivityConstraint(params.positivityConstraint)
                    err = R.ossart(P, E, n, params.relaxationCoefficient, params.updateNb, first_index, offset, endY)
                elif saveError:
                    P.initForIterative(offset, endY, params.updateNb, 1, False)
                    P.setPositivityConstraint(params.positivityConstraint)
                    err1 = R.ossart(P, E, n - 1, params.relaxationCoefficient, params.updateNb, first_index,
                                    offset, endY)
                    P.createErrorVolume()
                    err2 = R.ossart(P, E, 1, params.relaxationCoefficient, params.updateNb, first_index,
                                    offset, endY)
                    err = np.concatenate([err1, err2], axis=0)
                else:
                    P.initForIterative(offset, endY, params.updateNb, 1, False)
                    P.setPositivityConstraint(params.positivityConstraint)
                    err = R.ossart(P, E, n, params.relaxationCoefficient, params.updateNb, first_index, offset,
                                   endY)
                if slab_errors is not None:
                    slab_errors.append(err)
                for e in range(err.shape[0]):
                    if len(errors) <= e:
                        errors.append(float(err[e, 0]))
                    else:
                        errors[e] += float(err[e, 0])
            P.updateFromGPU(offset, endY, offset, endY)
            P.releaseCL_Memory()
            offset = endY
    return errors
