"""Host-side mirror of TomoJ's projector plugin interface over libtomoj_hip.

``VoxelProjector3DHIP`` exposes the method names and argument meaning of
``Projector`` (src/TomoJ_Base/.../tomography/projectors/Projector.java:14-159)
as implemented by ``VoxelProjector3D`` (VoxelProjector3D.java:32-885), plus the
GPU lifecycle calls ``ProjectorGPU`` offers its slab driver (initForIterative /
updateFromGPU / releaseCL_Memory).  Every call goes through the C ABI; errors
from the library raise ``TjError``.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import call, tj_geometry, tj_timing, tj_tuning


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


class VoxelProjector3DHIP:
    """Parallel-beam voxel projector on one MI355X.

    vol_shape = (W, H, D) of the reconstruction, proj_shape = (W, H) of the tilt
    series.  Centres default to TomoJ's (n-1)/2; ``proj_center`` takes the four
    ``projectionCenters`` and ``vol_center`` the three ``volumeCenters``.
    """

    def __init__(self, vol_shape: Sequence[int], proj_shape: Optional[Sequence[int]] = None, *,
                 device: int = 0, scale: float = 1.0, positivity: bool = False,
                 long_object_compensation: bool = False, proj_center=None, vol_center=None,
                 storage: int = _lib.TJ_STORAGE_F32):
        W, H, D = (int(v) for v in vol_shape)
        pw, ph = (W, H) if proj_shape is None else (int(proj_shape[0]), int(proj_shape[1]))
        g = tj_geometry()
        _lib.load().tj_geometry_default(C.byref(g), W, H, D, pw, ph)
        if proj_center is not None:
            for k in range(4):
                g.proj_center[k] = float(proj_center[k])
        if vol_center is not None:
            for k in range(3):
                g.vol_center[k] = float(vol_center[k])
        g.scale = float(scale)
        g.positivity = int(bool(positivity))
        g.long_object_comp = int(bool(long_object_compensation))
        g.storage = int(storage)
        self.geometry = g
        self.W, self.H, self.D, self.pw, self.ph = W, H, D, pw, ph
        self._ctx = C.c_void_p()
        call("tj_open", int(device), C.byref(g), C.byref(self._ctx))
        self.n_loaded = 0
        self.has_error_volume = False

    # ------------------------------------------------------------------ lifecycle
    def close(self) -> None:
        """ProjectorGPU.releaseCL_Memory (ProjectorGPU.java:376-403)."""
        if self._ctx:
            call("tj_close", self._ctx)
            self._ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------ data
    def load_tilt_series(self, stack: np.ndarray, eulers: np.ndarray) -> None:
        """Makes the supplied projections (TiltSeries.getPixels(i) for each i) and
        their Euler matrices resident on the device."""
        stack = _f32(stack)
        eulers = _f64(eulers).reshape(-1, 9)
        n = eulers.shape[0]
        if stack.shape != (n, self.ph, self.pw):
            raise ValueError(f"stack shape {stack.shape} != {(n, self.ph, self.pw)}")
        call("tj_load_projections", self._ctx, _ptr(stack), n, _ptr(eulers))
        self.n_loaded = n

    def synth_tilt_series(self, ellipsoids: np.ndarray, eulers: np.ndarray, noise_rel: float,
                          noise_seed: int) -> None:
        """Generates analytic projections of an ellipsoid phantom on the device."""
        ell = _f64(ellipsoids).reshape(-1, 16)
        eulers = _f64(eulers).reshape(-1, 9)
        call("tj_synth_projections", self._ctx, _ptr(ell), ell.shape[0], _ptr(eulers),
             eulers.shape[0], float(noise_rel), C.c_uint64(int(noise_seed)))
        self.n_loaded = eulers.shape[0]

    def normalize_projections(self, tilts_deg, min_size: int = 0) -> np.ndarray:
        """TiltSeries.getPixels with normalisation and the identity alignment,
        applied to the resident projections; returns n x (mean, stdDev)."""
        t = _f64(tilts_deg)
        if t.shape[0] != self.n_loaded:
            raise ValueError(f"{t.shape[0]} tilt angles for {self.n_loaded} projections")
        stats = np.zeros((self.n_loaded, 2), np.float64)
        call("tj_normalize_projections", self._ctx, _ptr(t), int(min_size), _ptr(stats))
        return stats

    def align_projections(self, transforms, tilts_deg=None, min_size: int = 0, normalize: bool = True,
                          fill: int = 1, stats=None) -> np.ndarray:
        """TiltSeries.getPixels through AffineAlignment.applyTransformOnImage
        (AffineAlignment.java:136-254) applied to the resident projections:
        `transforms` n x 6 forward transforms (m00, m10, m01, m11, m02, m12);
        `fill` TJ_FILL_NONE/AVG/NAN; `stats` n x (mean, stdDev) of the ImageJ
        ROI, or None to compute them for rotation-free transforms.  Returns the
        statistics used (zeros when neither normalisation nor FILL_AVG needs them)."""
        T = _f64(transforms).reshape(-1, 6)
        if T.shape[0] != self.n_loaded:
            raise ValueError(f"{T.shape[0]} transforms for {self.n_loaded} projections")
        t = None if tilts_deg is None else _f64(tilts_deg)
        sin = None if stats is None else _f64(stats).reshape(-1, 2)
        out = np.zeros((self.n_loaded, 2), np.float64)
        call("tj_align_projections", self._ctx, _ptr(T), _ptr(t) if t is not None else None, int(min_size),
             1 if normalize else 0, int(fill), _ptr(sin) if sin is not None else None, _ptr(out))
        return out

    def threshold_projections(self, threshold: float, hysteresis: float = 5e-324,
                              fill: int = _lib.TJ_FILL_AVG) -> None:
        """TiltSeries.applyThreshold (TiltSeries.java:720-806) on every resident
        projection, as getPixels applies it after the alignment."""
        call("tj_threshold_projections", self._ctx, float(threshold), float(hysteresis), int(fill))

    def get_loaded_projection(self, index: int) -> np.ndarray:
        out = np.empty((self.ph, self.pw), np.float32)
        call("tj_get_loaded_projection", self._ctx, int(index), _ptr(out))
        return out

    def set_eulers(self, eulers: np.ndarray) -> None:
        eulers = _f64(eulers).reshape(-1, 9)
        call("tj_set_eulers", self._ctx, _ptr(eulers), eulers.shape[0])

    def set_volume(self, vol_zyx: np.ndarray, startY: int = 0, endY: Optional[int] = None) -> None:
        """ProjectorGPU.initForIterative's slab upload (ProjectorGPU.java:195-254).
        ``vol_zyx`` is the full (D, H, W) host volume; rows [startY, endY) are sent."""
        vol = _f32(vol_zyx)
        if vol.shape != (self.D, self.H, self.W):
            raise ValueError(f"volume shape {vol.shape} != {(self.D, self.H, self.W)}")
        call("tj_set_volume", self._ctx, _ptr(vol), int(startY), self.H if endY is None else int(endY))

    def get_volume(self, startY: int = 0, endY: Optional[int] = None,
                   out: Optional[np.ndarray] = None) -> np.ndarray:
        """ProjectorGPU.updateFromGPU (ProjectorGPU.java:297-315)."""
        if out is None:
            out = np.zeros((self.D, self.H, self.W), np.float32)
        assert out.dtype == np.float32 and out.flags["C_CONTIGUOUS"] and out.shape == (self.D, self.H, self.W)
        call("tj_get_volume", self._ctx, _ptr(out), int(startY), self.H if endY is None else int(endY))
        return out

    def reset_volume(self) -> None:
        call("tj_reset_volume", self._ctx)

    # ------------------------------------------------------------------ Projector SPI
    def setPositivityConstraint(self, on: bool) -> None:
        """setPositivityConstraint (VoxelProjector3D.java:102-104)."""
        call("tj_set_positivity", self._ctx, int(bool(on)))
        self.geometry.positivity = int(bool(on))

    def createErrorVolume(self) -> None:
        """VoxelProjector3D.createErrorVolume (:858-864): from now on every
        backprojection of differences also backprojects their squares into a
        zero-initialised error volume on the device."""
        call("tj_set_error_volume", self._ctx, 1)
        self.has_error_volume = True

    def dropErrorVolume(self) -> None:
        call("tj_set_error_volume", self._ctx, 0)
        self.has_error_volume = False

    def getErrorVolume(self) -> np.ndarray:
        """The error volume in ImageStack layout (D, H, W)."""
        out = np.empty((self.D, self.H, self.W), np.float32)
        call("tj_get_error_volume", self._ctx, _ptr(out))
        return out

    def clearErrorVolume(self) -> None:
        """clearErrorVolume (:842-849)."""
        call("tj_clear_error_volume", self._ctx)

    def addProjection(self, index: int) -> None:
        call("tj_add_projection", self._ctx, int(index))

    def clearAllProjections(self) -> None:
        call("tj_clear_projections", self._ctx)

    def num_added(self) -> int:
        return _lib.load().tj_num_added(self._ctx)

    def project(self, startY: int = 0, endY: Optional[int] = None) -> None:
        call("tj_project", self._ctx, int(startY), self.H if endY is None else int(endY))

    def projectAndDifference(self, factor: float, startY: int = 0, endY: Optional[int] = None) -> np.ndarray:
        errors = np.zeros(max(self.num_added(), 1), np.float64)
        call("tj_project_and_difference", self._ctx, float(factor), int(startY),
             self.H if endY is None else int(endY), _ptr(errors))
        return errors[: self.num_added()]

    def backproject(self, startY: int = 0, endY: Optional[int] = None) -> None:
        call("tj_backproject", self._ctx, int(startY), self.H if endY is None else int(endY))

    def getProjection(self, index: int) -> np.ndarray:
        out = np.empty((self.ph, self.pw), np.float32)
        call("tj_get_projection", self._ctx, int(index), _ptr(out), None)
        return out

    def getNormImage(self, index: int) -> np.ndarray:
        out = np.empty((self.ph, self.pw), np.float32)
        call("tj_get_projection", self._ctx, int(index), None, _ptr(out))
        return out

    def getDifferenceImage(self, index: int) -> np.ndarray:
        out = np.empty((self.ph, self.pw), np.float32)
        call("tj_get_difference", self._ctx, int(index), _ptr(out))
        return out

    def startOfIteration(self) -> None:
        call("tj_start_iteration", self._ctx)

    def endOfIteration(self) -> None:
        call("tj_end_iteration", self._ctx)

    def setFista(self, on: bool = True) -> None:
        """Wraps the projector like FistaProjector3D (FistaProjector3D.java:5-72):
        the iteration hooks (and the fused loop) apply FISTA's extrapolation."""
        call("tj_set_fista", self._ctx, int(bool(on)))

    def setTvm(self, iterations: int, theta: float = 25.0, g: float = 1.0, lam: float = 0.1) -> None:
        """Turns the projector into TVMProjector3D (TVMProjector3D.java:24-157):
        every endOfIteration (and every iteration of the fused loop) ends with
        TotalVariationDenoising_3D.denoise3D(volume, theta, g, (float) lam,
        iterations).  iterations = 0 turns it off."""
        call("tj_set_tvm", self._ctx, int(iterations), float(theta), float(g), float(lam))

    def setBayesian(self, on: bool, alpha: float = 0.1, lam: float = 0.1) -> None:
        """Turns the projector into BayesianVoxelProjector3D
        (BayesianVoxelProjector3D.java:25-204): the iteration hooks keep a
        distance volume (-lam x the [1,-2,1]^3 convolution of the volume plus
        the backprojected differences) and add (float) alpha times it to the
        volume at the end of each iteration."""
        call("tj_set_bayesian", self._ctx, int(bool(on)), float(alpha), float(lam))

    def setBgart(self, on: bool, k: float = 1.0, darkBG: bool = True, doMedian: bool = True,
                 bgValue: float = float("nan")) -> None:
        """Turns the projector into BgARTVoxelProjector3D
        (BgARTVoxelProjector3D.java:13-152): the collapse filter runs once now
        (the constructor) and at the end of every iteration; with doMedian each
        iteration starts with the 3-D cross median.  bgValue NaN takes the
        background from the histogram."""
        call("tj_set_bgart", self._ctx, int(bool(on)), float(k), int(bool(darkBG)), int(bool(doMedian)),
             float(bgValue))

    def tv_denoise(self, theta: float, g: float, dt: float, iterations: int) -> None:
        """TotalVariationDenoising_3D.denoise3D (TotalVariationDenoising_3D.java:84-202)
        of the resident volume, once."""
        call("tj_tv_denoise", self._ctx, float(theta), float(g), float(dt), int(iterations))

    def getCompletion(self) -> float:
        return _lib.load().tj_get_completion(self._ctx)

    def isWeighting(self) -> bool:
        return getattr(self, "_weighting", False)

    def setWeightingFilter(self, diameter: float, tilts_deg, axis_deg=None) -> None:
        """setWeightingFilter(new FFTWeighting(ts, diameter)) (FFTWeighting.java:21-65)."""
        t = _f64(tilts_deg)
        a = None if axis_deg is None else _f64(axis_deg)
        call("tj_set_weighting", self._ctx, float(diameter), _ptr(t),
             None if a is None else _ptr(a), len(t))
        self._weighting = diameter > 0

    def setRampFilter(self, hamming: bool = True, cutoff: float = 1.0) -> None:
        call("tj_set_ramp_filter", self._ctx, int(bool(hamming)), float(cutoff))
        self._weighting = True

    def getWeightingMask(self) -> np.ndarray:
        out = np.empty((self.ph, self.pw), np.float32)
        call("tj_get_weighting_mask", self._ctx, _ptr(out))
        return out

    def weighting(self, image: np.ndarray) -> np.ndarray:
        """FFTWeighting.weighting on one image (returns a new array)."""
        img = _f32(image).copy()
        call("tj_apply_weighting", self._ctx, _ptr(img))
        return img

    def interrupt(self) -> None:
        call("tj_interrupt", self._ctx)

    # ------------------------------------------------------------------ fused loops
    def setAvailableIndexes(self, indexes=None) -> None:
        """The images run_ossart / run_wbp walk (ReconstructionParameters.getAvailableIndexes);
        None restores all."""
        if indexes is None:
            call("tj_set_available_indexes", self._ctx, None, 0)
            return
        a = np.ascontiguousarray(indexes, dtype=np.int32)
        call("tj_set_available_indexes", self._ctx, _ptr(a), int(a.size))

    def run_ossart(self, iterations: int, relaxation: float, update: int, first_index: int = 0,
                   startY: int = 0, endY: Optional[int] = None) -> np.ndarray:
        """TomoReconstruction2.OSSART over the resident projections (iterations x 5 err1 rows)."""
        err = np.zeros((max(iterations, 1), 5), np.float64)
        call("tj_run_ossart", self._ctx, int(iterations), float(relaxation), int(update), int(first_index),
             int(startY), self.H if endY is None else int(endY), _ptr(err))
        return err[:iterations]

    def run_wbp(self, startY: int = 0, endY: Optional[int] = None) -> None:
        """TomoReconstruction2.WBP over the resident projections (volume not cleared)."""
        call("tj_run_wbp", self._ctx, int(startY), self.H if endY is None else int(endY))

    def enable_timing(self, on: bool = True) -> None:
        call("tj_enable_timing", self._ctx, int(bool(on)))

    def get_tuning(self) -> dict:
        t = tj_tuning()
        call("tj_get_tuning", self._ctx, C.byref(t))
        return {k: getattr(t, k) for k, _ in tj_tuning._fields_}

    def set_tuning(self, **kw) -> None:
        cur = self.get_tuning()
        cur.update(kw)
        t = tj_tuning(**cur)
        call("tj_set_tuning", self._ctx, C.byref(t))

    def timing(self) -> dict:
        t = tj_timing()
        call("tj_get_timing", self._ctx, C.byref(t))
        return {k: getattr(t, k) for k, _ in tj_timing._fields_}


class TVMProjector3DHIP(VoxelProjector3DHIP):
    """TVMProjector3D (TomoJ_AdvancedReconstruction/.../projectors/TVMProjector3D.java:24-157)
    on the GPU: the voxel projector whose endOfIteration denoises the volume by
    total variation.  The setters keep TomoJ's names and defaults (:26-32);
    as in TomoJ, endOfIteration sets dt to the regularisation lambda (:152), so
    setTvmDt has no effect on the result, and regularisationAlpha is kept but
    not used."""

    def __init__(self, *args, **kw):
        super().__init__(*args, **kw)
        self.regularizationLambda = 0.1
        self.regularizationAlpha = 0.1
        self.theta = 25.0
        self.dt = 0.05
        self.g = 1.0
        self.tvmIteration = 5
        self._push()

    def _push(self) -> None:
        self.setTvm(self.tvmIteration, self.theta, self.g, self.regularizationLambda)

    def setRegularizationLambda(self, v: float) -> None:
        self.regularizationLambda = float(v)
        self._push()

    def setRegularizationAlpha(self, v: float) -> None:
        self.regularizationAlpha = float(v)

    def setTvmTheta(self, v: float) -> None:
        self.theta = float(np.float32(v))
        self._push()

    def setTvmDt(self, v: float) -> None:
        self.dt = float(np.float32(v))

    def setTvmG(self, v: float) -> None:
        self.g = float(np.float32(v))
        self._push()

    def setTvmIteration(self, n: int) -> None:
        self.tvmIteration = int(n)
        self._push()


class BayesianVoxelProjector3DHIP(VoxelProjector3DHIP):
    """BayesianVoxelProjector3D (TomoJ_AdvancedReconstruction/.../projectors/
    BayesianVoxelProjector3D.java:25-204) on the GPU, with TomoJ's setter names
    and defaults (:26-27)."""

    def __init__(self, *args, **kw):
        super().__init__(*args, **kw)
        self.regularizationLambda = 0.1
        self.regularizationAlpha = 0.1
        self._push()

    def _push(self) -> None:
        self.setBayesian(True, self.regularizationAlpha, self.regularizationLambda)

    def setRegularizationLambda(self, v: float) -> None:
        self.regularizationLambda = float(v)
        self._push()

    def setRegularizationAlpha(self, v: float) -> None:
        self.regularizationAlpha = float(v)
        self._push()


class BgARTVoxelProjector3DHIP(VoxelProjector3DHIP):
    """BgARTVoxelProjector3D (TomoJ_AdvancedReconstruction/.../projectors/
    BgARTVoxelProjector3D.java:13-152) on the GPU.  The constructor takes
    TomoJ's (k, darkBG, doMedian) after the projector arguments and, like
    TomoJ's (:20-27), runs the collapse filter once on the volume."""

    def __init__(self, *args, k: float = 1.0, darkBG: bool = True, doMedian: bool = True, **kw):
        super().__init__(*args, **kw)
        self.k = float(k)
        self.darkBG = bool(darkBG)
        self.doMedian = bool(doMedian)
        self.bgValue = float("nan")
        self._push()

    def _push(self) -> None:
        self.setBgart(True, self.k, self.darkBG, self.doMedian, self.bgValue)

    def collapseFilter(self) -> None:
        """collapseFilter (:30-78) of the resident volume, once."""
        self._push()


def phantom_ellipsoids(seed: int, k: int, W: int, H: int, D: int) -> np.ndarray:
    """Seeded ellipsoid parameters (k x 16), generated by the library's host code."""
    out = np.zeros((k, 16), np.float64)
    _lib.load().tj_phantom_ellipsoids(C.c_uint64(int(seed)), int(k), int(W), int(H), int(D), _ptr(out))
    return out
