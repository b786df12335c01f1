"""Update loops of TomoReconstruction2 driven through the projector interface.

``ossart`` and ``wbp`` restate TomoReconstruction2.OSSART / osartIteration /
orderShuffleART / WBP (src/TomoJ_Base/.../tomography/TomoReconstruction2.java
:499-781) on top of any object with the Projector methods (addProjection,
clearAllProjections, projectAndDifference, backproject, start/endOfIteration),
exactly as the Java loop calls them.  ``VoxelProjector3DHIP.run_ossart`` runs
the same loop inside the library without the per-call host round trips; both
must give the same volume.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import List, Optional, Sequence

import numpy as np

from . import _lib

JAVA_FLOAT_MAX = 3.4028234663852886e38


def tilt_euler(tilt_deg: float, tilt_axis_deg: float = 0.0, axis_vertical: bool = True) -> np.ndarray:
    """AffineAlignment.getEulerMatrix (align/AffineAlignment.java:382-395) =
    MatrixUtils.eulerAngles2Matrix(ta, tilt, -ta) (MatrixUtils.java:226-251), with
    Java 8's Math.toRadians (angdeg / 180.0 * PI)."""
    ta = 0.0 if axis_vertical else float(tilt_axis_deg)
    phi, theta, psi = ta / 180.0 * math.pi, float(tilt_deg) / 180.0 * math.pi, (-ta) / 180.0 * math.pi
    cphi, sphi = math.cos(phi), math.sin(phi)
    ct, st = math.cos(theta), math.sin(theta)
    cpsi, spsi = math.cos(psi), math.sin(psi)
    return np.array([
        [cpsi * ct * cphi - spsi * sphi, cpsi * ct * sphi + spsi * cphi, -cpsi * st],
        [-spsi * ct * cphi - cpsi * sphi, -spsi * ct * sphi + cpsi * cphi, spsi * st],
        [st * cphi, st * sphi, ct],
    ])


def tilt_eulers(tilts_deg: Sequence[float], tilt_axis_deg: float = 0.0, axis_vertical: bool = True) -> np.ndarray:
    return np.stack([tilt_euler(t, tilt_axis_deg, axis_vertical) for t in tilts_deg])


def order_shuffle(eulers: np.ndarray, first_index: int, available: Optional[Sequence[int]] = None) -> List[int]:
    """orderShuffleART (:735-781) with the first pick fixed: each next image is the
    remaining one with the smallest |row2(E_prev) . row2(E_o)|, computed by the
    library's tj_order_shuffle.  With a list of available images TomoJ runs the
    chain over positions 0..len-1 using the Euler matrices of images 0..len-1
    (:741) and emits the listed images."""
    e = np.ascontiguousarray(np.asarray(eulers, np.float64).reshape(-1, 9))
    n = e.shape[0] if available is None else len(available)
    if n == 0:
        return []
    e = np.ascontiguousarray(e[:n])
    order = np.zeros(n, np.int32)
    _lib.load().tj_order_shuffle(e.ctypes.data_as(C.c_void_p), n, int(first_index),
                                 order.ctypes.data_as(C.c_void_p))
    out = [int(v) for v in order]
    return out if available is None else [int(available[k]) for k in out]


def osart_iteration(projector, eulers: np.ndarray, vol_w: int, vol_h: int, relaxation: float,
                    update_nb: int, first_index: int, startY: int, endY: int,
                    available: Optional[Sequence[int]] = None) -> List[float]:
    """TomoReconstruction2.osartIteration (:636-733); returns err1 (5 values).
    ``available``: the images to walk (getAvailableIndexes); update, the flush
    at position n-1 and the error normalisation keep the full count n."""
    projector.clearAllProjections()
    n = len(eulers)
    indexes = order_shuffle(eulers, first_index, available)
    update = n if update_nb == 0 else update_nb
    factorimg = float(np.float32(relaxation / update))
    diff_index = 0
    som, errmin, errmax = 0.0, JAVA_FLOAT_MAX, 0.0
    imin = imax = 0
    for t in range(len(indexes)):
        projector.addProjection(indexes[t])
        diff_index += 1
        if diff_index == update or t == n - 1:
            error = projector.projectAndDifference(factorimg, startY, endY)
            for e in range(len(error)):
                som += error[e]
                if error[e] > errmax:
                    errmax = error[e]
                    imax = indexes[t - len(error) + e + 1]
                elif error[e] < errmin:
                    errmin = error[e]
                    imin = indexes[t - len(error) + e + 1]
            projector.backproject(startY, endY)
            projector.clearAllProjections()
            diff_index = 0
    wh = vol_w * vol_h
    return [math.sqrt(som / float(n * vol_w * vol_h)), math.sqrt(errmin / wh), imin,
            math.sqrt(errmax / wh), imax]


def ossart(projector, eulers: np.ndarray, iterations: int, relaxation: float, update_nb: int,
           first_index: int = 0, startY: int = 0, endY=None,
           available: Optional[Sequence[int]] = None) -> np.ndarray:
    """TomoReconstruction2.OSSART (:548-634) through the projector interface."""
    endY = projector.H if endY is None else endY
    out = []
    for _ in range(iterations):
        projector.startOfIteration()
        out.append(osart_iteration(projector, eulers, projector.W, projector.H, relaxation, update_nb,
                                   first_index, startY, endY, available))
        projector.endOfIteration()
    return np.array(out, np.float64).reshape(iterations, 5)


def wbp(projector, n: int, update_nb: int = 0, startY: int = 0, endY=None,
        available: Optional[Sequence[int]] = None) -> None:
    """TomoReconstruction2.WBP(ts, projector, params, startY, endY) (:499-538)
    over getAvailableIndexes (all n images by default)."""
    endY = projector.H if endY is None else endY
    images = list(range(n)) if available is None else [int(v) for v in available]
    index = 0
    for i, im in enumerate(images):
        projector.addProjection(im)
        index += 1
        if index >= update_nb or i == len(images) - 1:
            projector.backproject(startY, endY)
            projector.clearAllProjections()
            index = 0
