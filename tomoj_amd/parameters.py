"""Host mirror of TomoJ's reconstruction parameters, projector factory and
CPU-manager driver, over the HIP projector.

``ReconstructionParameters`` follows
src/TomoJ_Base/src/main/java/fr/curie/tomoj/tomography/ReconstructionParameters.java:12-411
(field names, defaults and the two factories), ``getProjector`` its
:313-333, and ``ReconstructionThreadManager`` the CPU path of
ReconstructionThreadManager.java (createRec2 :169-196, doCPUrec :218-255).
The projections are the tilt series' images: with ``normalize`` they are
normalised on the device as ``ts.getPixels(i)`` does for unaligned images
(identity alignment); otherwise they are used as given (TomoJ's
-noNormalisation).  The Euler matrices are ``getEulerMatrix(ts, i)``.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from typing import Callable, Dict, List, Optional, Sequence

import numpy as np

from ._lib import call
from .projector import (BayesianVoxelProjector3DHIP, BgARTVoxelProjector3DHIP, TVMProjector3DHIP,
                        VoxelProjector3DHIP)
from .reconstruction import tilt_eulers

BP = 0
WBP = 1
OSSART = 2
# projection subsets (ReconstructionParameters.java:17-20)
ALL_PROJECTIONS = 10
EVEN_PROJECTIONS = 11
ODD_PROJECTIONS = 12
DEFINED_PROJECTIONS = 13
TVM = 3        # AdvancedReconstructionParameters.java:10
BAYESIAN = 4   # AdvancedReconstructionParameters.java:11
BGART = 6      # AdvancedReconstructionParameters.java:13


@dataclass
class TiltSeriesData:
    """What the projector path reads from a TiltSeries: the projections
    (n x ph x pw, float32) and the tilt geometry."""
    images: np.ndarray
    tilts: Sequence[float]
    tilt_axis_deg: float = 0.0
    axis_vertical: bool = True
    normalize: bool = False            # TomoJ's default is True (TiltSeries.java:128)
    normalization_min_size: int = 0
    # affine alignment: n x 6 forward transforms (m00, m10, m01, m11, m02, m12) of
    # AffineAlignment.getTransform(i), None for the identity alignment
    transforms: Optional[np.ndarray] = None
    fill: int = 1                      # TiltSeries.FILL_AVG (TiltSeries.java:130)
    stats: Optional[np.ndarray] = None  # n x (mean, stdDev) of the ImageJ ROI, when rotated
    # setThreshold / setThresholdHisteresis (TiltSeries.java:407-416): fractions,
    # applied by getPixels after the alignment when apply_threshold is set
    threshold: float = 5e-324          # Double.MIN_VALUE (:101-103)
    threshold_hysteresis: float = 5e-324
    apply_threshold: bool = False

    @property
    def width(self) -> int:
        return int(self.images.shape[2])

    @property
    def height(self) -> int:
        return int(self.images.shape[1])

    def eulers(self) -> np.ndarray:
        return tilt_eulers(self.tilts, self.tilt_axis_deg, self.axis_vertical)

    @classmethod
    def from_mrc(cls, path: str, tilts: Optional[Sequence[float]] = None, **kw) -> "TiltSeriesData":
        """A tilt series opened from an MRC stack as TomoJ opens one
        (TomoJ.java:42-49 -> MRC_Reader, InputOutput/MRC_Reader.java:176-256):
        the images as floats, the tilt angles from the FEI extended header when
        the file has one, else `tilts`."""
        from . import mrc
        data, info, fei = mrc.read(path)
        if fei is not None and tilts is None:
            tilts = [float(a) for a in fei]
        if tilts is None:
            raise ValueError(f"{path}: no FEI tilt angles in the file; pass the tilts")
        if len(tilts) != data.shape[0]:
            raise ValueError(f"{path}: {data.shape[0]} images but {len(tilts)} tilt angles")
        return cls(data, list(tilts), **kw)

    def tilt_axes(self) -> np.ndarray:
        return np.full(len(self.tilts), 0.0 if self.axis_vertical else float(self.tilt_axis_deg))


def supply(P: VoxelProjector3DHIP, ts: TiltSeriesData) -> None:
    """Makes TiltSeries.getPixels(i) resident in P for every image
    (TiltSeries.java:678-687): the raw images, then the affine alignment or,
    for unaligned images, the normalisation, then the optional threshold."""
    P.load_tilt_series(ts.images, ts.eulers())
    if ts.transforms is not None:
        # getPixels(i) through the affine alignment (AffineAlignment.applyTransformOnImage)
        P.align_projections(ts.transforms, ts.tilts, ts.normalization_min_size, ts.normalize, ts.fill,
                            ts.stats)
    elif ts.normalize:
        # getPixels(i) with normalisation, for unaligned images (identity alignment)
        P.normalize_projections(ts.tilts, ts.normalization_min_size)
    if ts.apply_threshold:
        # then applyThreshold (TiltSeries.java:682-684)
        P.threshold_projections(ts.threshold, ts.threshold_hysteresis, ts.fill)


class ReconstructionParameters:
    """ReconstructionParameters.java:12-411 (reconstruction type, iterations,
    relaxation, update size, weighting radius, constraints)."""

    def __init__(self, width: int, height: int, depth: int):
        self.width, self.height, self.depth = int(width), int(height), int(depth)
        self.recNbRaysPerPixels = 1
        self.reconstructionType = BP
        self.nbIterations = 0
        self.relaxationCoefficient = 0.0
        self.updateNb = 0
        self.positivityConstraint = False
        self.fista = False
        self.weightingRadius = 0.0
        self.rescaleData = True            # field default (:35)
        self.longObjectCompensation = False
        self.reconstructionCenterModifiers: Optional[Sequence[float]] = None
        self._tempNbCall = 0
        self.projectionType = 0
        self.availableIndexes: Optional[List[int]] = None
        self.saveErrorVolume = False       # (:39-40)
        self.saveErrorVolumeAll = False

    # factories (:279-311)
    @staticmethod
    def createOSSARTParameters(width, height, depth, nbIterations, relaxationCoefficient, updateNb,
                               reconstructionCenterModifiers=None) -> "ReconstructionParameters":
        p = ReconstructionParameters(width, height, depth)
        p.reconstructionType = OSSART
        p.nbIterations = int(nbIterations)
        p.relaxationCoefficient = float(relaxationCoefficient)
        p.updateNb = int(updateNb)
        p.weightingRadius = float("nan")
        p.reconstructionCenterModifiers = reconstructionCenterModifiers
        return p

    @staticmethod
    def createWBPParameters(width, height, depth, weightingRadius) -> "ReconstructionParameters":
        p = ReconstructionParameters(width, height, depth)
        wr = float(weightingRadius)
        p.reconstructionType = BP if (math.isnan(wr) or wr <= 0) else WBP
        p.weightingRadius = wr
        p.updateNb = 0
        p.nbIterations = 0
        p.relaxationCoefficient = float("nan")
        return p

    def getRelaxationCoefficient(self) -> float:
        """A negative coefficient means the 1/k schedule (:88-91)."""
        if self.relaxationCoefficient < 0:
            self._tempNbCall += 1
            return 1.0 / self._tempNbCall
        return self.relaxationCoefficient

    def resetTmpNbCall(self) -> None:
        self._tempNbCall = 0

    def setSaveErrorVolume(self, saveErrorVolume: bool, saveErrorVolumeAll: Optional[bool] = None) -> None:
        """:196-202 (both overloads)."""
        self.saveErrorVolume = bool(saveErrorVolume)
        if saveErrorVolumeAll is not None:
            self.saveErrorVolumeAll = bool(saveErrorVolumeAll)

    def isSaveErrorVolume(self) -> bool:
        return self.saveErrorVolume

    def isSaveErrorVolumeAll(self) -> bool:
        return self.saveErrorVolumeAll

    def setFista(self, fista: bool) -> None:
        self.fista = bool(fista)

    def isFista(self) -> bool:
        return self.fista

    # projection subsets (:174-261)
    def getFscType(self) -> int:
        return self.projectionType

    def setFscType(self, fscType: int) -> None:
        self.projectionType = int(fscType)
        self.availableIndexes = None

    getProjectionType = getFscType
    setProjectionType = setFscType

    def getAvailableIndexes(self, ts: "TiltSeriesData") -> List[int]:
        """:212-256, computed once and kept: EVEN / ODD keep that parity (at
        most N/2 images), any other type all of them."""
        if self.availableIndexes is None:
            n = len(ts.tilts)
            out = (C.c_int * max(n, 1))()
            k = call("tj_available_indexes", n, int(self.projectionType), out)
            self.availableIndexes = [int(out[i]) for i in range(k)]
        return self.availableIndexes

    def setAvailableIndexes(self, availableIndexes: Sequence[int]) -> None:
        """:258-261.  TomoJ also sets reconstructionType (not projectionType) to
        DEFINED_PROJECTIONS, which sends doCPUrec down its default branch (the
        regularisation loop, OS-SART's with the projector's hooks)."""
        self.availableIndexes = [int(v) for v in availableIndexes]
        self.reconstructionType = DEFINED_PROJECTIONS

    def volume_centers(self) -> List[float]:
        """TomoReconstruction2 centres (n-1)/2 (:63-71) shifted by the centre
        modifiers (ReconstructionThreadManager.java:179-182)."""
        c = [(self.width - 1) / 2.0, (self.height - 1) / 2.0, (self.depth - 1) / 2.0]
        if self.reconstructionCenterModifiers is not None:
            c = [c[k] + float(self.reconstructionCenterModifiers[k]) for k in range(3)]
        return c

    def getProjector(self, ts: TiltSeriesData, device: int = 0,
                     vol_center: Optional[Sequence[float]] = None) -> VoxelProjector3DHIP:
        """:313-333: a voxel projector with the weighting filter for WBP, the
        positivity constraint for OS-SART, the scale ts.W / rec.W when
        rescaleData, long-object compensation, and the FISTA wrapper.
        `vol_center`: the centres of the reconstruction the projector works on
        (TomoJ's projector takes them from rec2; default: this object's)."""
        return self._projector(VoxelProjector3DHIP, ts, device,
                               self.reconstructionType == OSSART and self.positivityConstraint, vol_center)

    def _projector(self, cls, ts: TiltSeriesData, device: int, positivity: bool,
                   vol_center: Optional[Sequence[float]] = None):
        scale = ts.width / float(self.width) if self.rescaleData else 1.0
        P = cls((self.width, self.height, self.depth), (ts.width, ts.height), device=device,
                scale=scale, positivity=positivity, long_object_compensation=self.longObjectCompensation,
                vol_center=self.volume_centers() if vol_center is None else list(vol_center))
        supply(P, ts)
        if self.reconstructionType == WBP:
            P.setWeightingFilter(self.weightingRadius * ts.width, ts.tilts, ts.tilt_axes())
        if self.fista:
            P.setFista(True)
        return P


class _BgartFactory:
    """Builds BgARTVoxelProjector3DHIP with the parameters' (k, darkBG, median)."""

    def __init__(self, k: float, dark: bool, median: bool):
        self.k, self.dark, self.median = k, dark, median

    def __call__(self, *args, **kw):
        return BgARTVoxelProjector3DHIP(*args, k=self.k, darkBG=self.dark, doMedian=self.median, **kw)


class AdvancedReconstructionParameters(ReconstructionParameters):
    """AdvancedReconstructionParameters (TomoJ_AdvancedReconstruction/.../tomography/
    AdvancedReconstructionParameters.java:8-283) for the TVM, BAYESIAN and BGART
    types: their settings and defaults (:19-27), the createTVMParameters,
    createBayesianParameters and createBgARTParameters factories (:112-155) and
    getProjector's branches for them (:179-203, 214-218)."""

    def __init__(self, width: int, height: int, depth: int):
        super().__init__(width, height, depth)
        self.regularizationWeight = 0.0
        self.tvmTheta = 25.0
        self.tvmG = 1.0
        self.tvmDt = 0.1
        self.tvmIterations = 5
        self.bgartK = 1.0
        self.bgartMedian = True
        self.bgartDarkBG = True

    @staticmethod
    def createTVMParameters(width, height, depth, nbIterations, *args) -> "AdvancedReconstructionParameters":
        """(w, h, d, nbIterations, regularizationWeight) sets the relaxation to 1
        (:112-120); (w, h, d, nbIterations, relaxationCoefficient, tvmTheta, tvmG,
        tvmDt, tvmIterations) sets the TV settings (:122-133).  Both: update 0
        (every projection per subset, SIRT), no weighting."""
        p = AdvancedReconstructionParameters(width, height, depth)
        p.reconstructionType = TVM
        p.nbIterations = int(nbIterations)
        p.weightingRadius = float("nan")
        p.updateNb = 0
        if len(args) == 1:
            p.relaxationCoefficient = 1.0
            p.regularizationWeight = float(args[0])
        elif len(args) == 5:
            p.relaxationCoefficient = float(args[0])
            p.tvmTheta, p.tvmG, p.tvmDt = (float(v) for v in args[1:4])
            p.tvmIterations = int(args[4])
        else:
            raise TypeError("createTVMParameters takes 5 or 9 arguments")
        return p

    @staticmethod
    def createBayesianParameters(width, height, depth, nbIterations, regularizationAlpha,
                                 regularizationWeight) -> "AdvancedReconstructionParameters":
        """:135-144: the alpha is stored as the relaxation coefficient; update 0."""
        p = AdvancedReconstructionParameters(width, height, depth)
        p.reconstructionType = BAYESIAN
        p.nbIterations = int(nbIterations)
        p.weightingRadius = float("nan")
        p.relaxationCoefficient = float(regularizationAlpha)
        p.regularizationWeight = float(regularizationWeight)
        p.updateNb = 0
        return p

    @staticmethod
    def createBgARTParameters(width, height, depth, nbIterations, relaxationCoefficient, updateNb,
                              bgartK, bgartMedian, bgartDarkBG) -> "AdvancedReconstructionParameters":
        """:145-155: OS-SART settings plus the BgART k, median and dark-background flags."""
        p = AdvancedReconstructionParameters(width, height, depth)
        p.reconstructionType = BGART
        p.nbIterations = int(nbIterations)
        p.weightingRadius = float("nan")
        p.relaxationCoefficient = float(relaxationCoefficient)
        p.updateNb = int(updateNb)
        p.bgartK = float(bgartK)
        p.bgartMedian = bool(bgartMedian)
        p.bgartDarkBG = bool(bgartDarkBG)
        return p

    def getProjector(self, ts: TiltSeriesData, device: int = 0, vol_center: Optional[Sequence[float]] = None):
        """The TVM, BAYESIAN and BGART branches of getProjector (:179-203).  TVM:
        TVMProjector3D with the relaxation as regularisation alpha, the
        positivity and long-object settings, the TV settings, the scale when
        rescaleData, and the FISTA wrapper; TomoJ leaves its regularisation
        lambda at 0.1 (the call that would set it is commented out, :182).
        BAYESIAN: BayesianVoxelProjector3D with alpha = the relaxation
        coefficient and lambda = the regularisation weight, positivity,
        long-object setting and scale, never wrapped in FISTA.  BGART (:214-218):
        BgARTVoxelProjector3D(k, darkBG, median) with the scale when rescaleData
        and the FISTA wrapper, TomoJ's defaults otherwise (no positivity, no
        long-object compensation).  Other types go to the base class."""
        if self.reconstructionType == BGART:
            loc, self.longObjectCompensation = self.longObjectCompensation, False
            try:
                return self._projector(_BgartFactory(self.bgartK, self.bgartDarkBG, self.bgartMedian), ts,
                                       device, False, vol_center)
            finally:
                self.longObjectCompensation = loc
        if self.reconstructionType == BAYESIAN:
            fista, self.fista = self.fista, False
            try:
                P = self._projector(BayesianVoxelProjector3DHIP, ts, device, self.positivityConstraint, vol_center)
            finally:
                self.fista = fista
            P.setRegularizationAlpha(self.relaxationCoefficient)
            P.setRegularizationLambda(self.regularizationWeight)
            return P
        if self.reconstructionType != TVM:
            return super().getProjector(ts, device, vol_center)
        P = self._projector(TVMProjector3DHIP, ts, device, self.positivityConstraint, vol_center)
        P.setRegularizationAlpha(self.relaxationCoefficient)
        P.setTvmTheta(self.tvmTheta)
        P.setTvmG(self.tvmG)
        P.setTvmDt(self.tvmDt)
        P.setTvmIteration(self.tvmIterations)
        return P


class ReconstructionThreadManager:
    """The CPU-manager driver (ReconstructionThreadManager.java:79-255) with the
    projector on the GPU.  `rec2` is the reconstruction TomoJ keeps across
    calls; its device copy lives in the projector and is read back when the
    host asks for it."""

    def __init__(self, ts: TiltSeriesData, device: int = 0):
        self.ts = ts
        self.device = device
        self.projector: Optional[VoxelProjector3DHIP] = None
        self.params: Optional[ReconstructionParameters] = None
        self.rec2: Optional[np.ndarray] = None          # (D, H, W), ImageStack layout
        self.rec2_centers: Optional[List[float]] = None
        self._device_newer = False                      # the projector's volume is newer than rec2
        self._rec2_zero = True                          # rec2 holds only zeros (a fresh reconstruction)
        self._thprojs: Optional[Dict[int, np.ndarray]] = None   # getProjection's cache (thprojs)
        self.errorVolumes: Dict[int, np.ndarray] = {}   # iteration -> error volume (endOfIteration's saves)
        self.error_volume_sink: Optional[Callable[[int, np.ndarray], None]] = None

    def _sync_host(self) -> None:
        if self._device_newer and self.projector is not None:
            self.rec2 = self.projector.get_volume()
            self._device_newer = False
            self._rec2_zero = False

    def createRec2(self, params: ReconstructionParameters) -> None:
        """createRec2 (:169-196): a new zero reconstruction of the parameters'
        size, with centres (n-1)/2 shifted by reconstructionCenterModifiers;
        when one of that size already exists TomoJ keeps it ("reconstruction
        exists, add some iterations", :184) together with the centres it was
        created with, and the next reconstruction continues from it."""
        self._sync_host()
        self._thprojs = None
        shape = (params.depth, params.height, params.width)
        if self.rec2 is None or self.rec2.shape != shape:
            self.rec2 = np.zeros(shape, np.float32)
            self.rec2_centers = params.volume_centers()
            self._rec2_zero = True
        if self.projector is not None:
            self.projector.close()
            self.projector = None
        self.params = params

    def resetRec2(self) -> None:
        """resetRec2 / setRec2(null) (:199-207): the next createRec2 starts from zero."""
        if self.projector is not None:
            self.projector.close()
            self.projector = None
        self.rec2 = None
        self.rec2_centers = None
        self._device_newer = False
        self._rec2_zero = True
        self._thprojs = None

    def _projector_for(self, params: ReconstructionParameters) -> VoxelProjector3DHIP:
        """doCPUrec's params.getProjector(ts, rec2) (:223): a projector over rec2
        (created for these parameters if rec2 does not exist or has another
        size); a projector already made for the same parameters object keeps
        its device volume."""
        if self.rec2 is None or self.rec2.shape != (params.depth, params.height, params.width):
            self.createRec2(params)
        if self.projector is not None and params is self.params:
            return self.projector
        self._sync_host()
        if self.projector is not None:
            self.projector.close()
        self.params = params
        self.projector = params.getProjector(self.ts, self.device, vol_center=self.rec2_centers)
        if not self._rec2_zero:
            self.projector.set_volume(self.rec2)
        return self.projector

    def _save_error_volume(self, iteration: int, P: VoxelProjector3DHIP) -> None:
        """VoxelProjector3D.endOfIteration with an error volume (:828-838):
        save it (here: keep it, and hand it to error_volume_sink if set) and
        clear it."""
        ev = P.getErrorVolume()
        self.errorVolumes[iteration] = ev
        if self.error_volume_sink is not None:
            self.error_volume_sink(iteration, ev)
        P.clearErrorVolume()

    def _ossart_saving_errors(self, P, params, iterations: int, done: int, first_index: int) -> np.ndarray:
        rows = []
        for k in range(iterations):
            rows.append(P.run_ossart(1, params.relaxationCoefficient, params.updateNb, first_index=first_index))
            self._save_error_volume(done + k + 1, P)
        return np.concatenate(rows, axis=0) if rows else np.zeros((0, 5))

    def doRec(self, params: ReconstructionParameters, saveErrorVol: Optional[bool] = None,
              saveForAll: Optional[bool] = None, first_index: int = 0) -> Optional[np.ndarray]:
        """doCPUrec (:218-255): WBP/BP return None; OS-SART returns the
        per-iteration err1 rows.  OSSART reads the relaxation field directly
        (TomoReconstruction2.java:645), so the 1/k schedule never applies.
        TomoJ draws the first image of each iteration at random
        (TomoReconstruction2.java:755); here it is `first_index`.

        Error volume (OSSART with a plain VoxelProjector3D; TomoJ prints that a
        FISTA-wrapped projector is not a VoxelProjector3D and makes none):
        saveErrorVol and saveForAll create it before the loop and save it at
        every iteration's end; saveErrorVol alone runs nbIterations - 1
        iterations, then creates it and runs the last iteration, whose error
        volume is saved (:229-246).  Saved volumes go to `errorVolumes`
        keyed by the projector's iteration count (TomoJ's errorVolume_<n>.tif).
        The flags default to the parameters' isSaveErrorVolume /
        isSaveErrorVolumeAll, as performReconstruction passes them (:83-84)."""
        saveErrorVol = params.isSaveErrorVolume() if saveErrorVol is None else bool(saveErrorVol)
        saveForAll = params.isSaveErrorVolumeAll() if saveForAll is None else bool(saveForAll)
        P = self._projector_for(params)
        self._thprojs = None
        avail = params.getAvailableIndexes(self.ts)
        P.setAvailableIndexes(None if len(avail) == len(self.ts.tilts) and avail == list(range(len(avail)))
                              else avail)
        self._device_newer = True
        if params.reconstructionType in (BP, WBP):
            P.run_wbp()
            return None
        if params.reconstructionType == OSSART and saveErrorVol and not params.fista:
            n = params.nbIterations
            if saveForAll:
                P.createErrorVolume()
                return self._ossart_saving_errors(P, params, n, 0, first_index)
            first = P.run_ossart(max(n - 1, 0), params.relaxationCoefficient, params.updateNb,
                                 first_index=first_index)
            P.createErrorVolume()
            last = self._ossart_saving_errors(P, params, 1, max(n - 1, 0), first_index)
            return np.concatenate([first.reshape(-1, 5), last], axis=0)
        if params.reconstructionType in (OSSART, TVM, BAYESIAN, BGART, DEFINED_PROJECTIONS):
            # TVM, BAYESIAN, BGART and DEFINED_PROJECTIONS go through doCPUrec's default
            # branch, TomoReconstruction2.regularization (:783-878): the OS-SART
            # loop with the projector's own iteration hooks.
            return P.run_ossart(params.nbIterations, params.relaxationCoefficient, params.updateNb,
                                first_index=first_index)
        raise ValueError(f"reconstruction type {params.reconstructionType} is not on this path")

    def getProjection(self, index: int) -> np.ndarray:
        """getProjection(int) (:257-270): the forward projection of image
        `index` from the current reconstruction through the projector
        (addProjection, project, getProjection(0), clearAllProjections),
        cached until the reconstruction changes."""
        if self.projector is None:
            if self.params is None:
                raise RuntimeError("getProjection: no reconstruction yet")
            self._projector_for(self.params)
        if self._thprojs is None:
            self._thprojs = {}
        if index not in self._thprojs:
            P = self.projector
            P.clearAllProjections()
            P.addProjection(int(index))
            P.project()
            self._thprojs[index] = P.getProjection(0)
            P.clearAllProjections()
        return self._thprojs[index]

    def performReconstruction(self, params: ReconstructionParameters, first_index: int = 0):
        """createRec2 then doRec with the parameters' error-volume flags (:83-84, :70-71)."""
        self.createRec2(params)
        return self.doRec(params, first_index=first_index)

    def getRec2(self) -> np.ndarray:
        """The reconstruction in ImageStack layout (D, H, W)."""
        self._sync_host()
        return self.rec2

    def close(self) -> None:
        if self.projector is not None:
            self.projector.close()
            self.projector = None
