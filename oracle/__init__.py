"""CPU oracle for the TomoJ projector path -- TEST INFRASTRUCTURE ONLY.

Python bindings (ctypes + numpy) over ``oracle/libtomoj_oracle.so``, built from
``oracle/tomoj_oracle.c`` by ``oracle/Makefile``.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg import this
package, and only to check (or time beside) the HIP path; the product never
routes through it.

Each function names the TomoJ source it restates (paths relative to
``src/TomoJ_Base/src/main/java/fr/curie/tomoj/``).  Volumes use TomoJ's host
layout ``[z][y][x]`` (ImageStack: one slice per z, x fastest); projections are
``[n][ph][pw]``; Euler matrices ``[n][3][3]`` row-major float64.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libtomoj_oracle.so")
_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, no FMA contraction)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE],
                       check=True)
    return _LIB_PATH


class Geom(C.Structure):
    """``or_geom``: the projector state TomoJ keeps in ``Projector`` /
    ``VoxelProjector3D`` fields (Projector.java:15-21, VoxelProjector3D.java:33-47)."""
    _fields_ = [("vw", C.c_int), ("vh", C.c_int), ("vd", C.c_int),
                ("pw", C.c_int), ("ph", C.c_int),
                ("pc", C.c_double * 4), ("vc", C.c_double * 3),
                ("scale", C.c_double), ("positivity", C.c_int), ("loc", C.c_int),
                ("nthreads", C.c_int)]


def make_geom(vw, vh, vd, pw=None, ph=None, pc=None, vc=None, scale=1.0,
              positivity=False, loc=False, nthreads=1) -> Geom:
    """Geometry with TomoJ's default centres: projection centres ``(n-1)/2``
    (TiltSeries.java:161-170) and volume centres ``(n-1)/2``
    (TomoReconstruction2.java:63-71)."""
    pw = vw if pw is None else pw
    ph = vh if ph is None else ph
    if pc is None:
        pc = ((pw - 1.0) / 2, (ph - 1.0) / 2, (pw - 1) / 2.0, (ph - 1) / 2.0)
    if vc is None:
        vc = ((vw - 1) / 2.0, (vh - 1) / 2.0, (vd - 1) / 2.0)
    g = Geom()
    g.vw, g.vh, g.vd, g.pw, g.ph = vw, vh, vd, pw, ph
    for k in range(4):
        g.pc[k] = pc[k]
    for k in range(3):
        g.vc[k] = vc[k]
    g.scale = scale
    g.positivity = int(bool(positivity))
    g.loc = int(bool(loc))
    g.nthreads = int(nthreads)
    return g


def _L():
    global _lib
    if _lib is None:
        build()
        lib = C.CDLL(_LIB_PATH)
        P = C.c_void_p
        G = C.POINTER(Geom)
        d, i, l, u64 = C.c_double, C.c_int, C.c_long, C.c_uint64
        sigs = {
            "or_euler_matrix": (None, [d, d, d, P]),
            "or_tilt_euler": (None, [d, d, i, P]),
            "or_ij_interpolated_value": (d, [P, i, i, d, d]),
            "or_getpixel": (C.c_float, [P, i, i, i, d, d, i]),
            "or_fp_factors": (None, [P, P]),
            "or_project_one": (None, [G, P, P, P, P, i, i]),
            "or_project_one_f64": (None, [G, P, P, P, P, i, i]),
            "or_difference": (d, [P, P, P, P, l, d]),
            "or_difference_f64": (d, [P, P, P, P, l, d]),
            "or_backproject_one": (None, [G, P, P, P, i, i]),
            "or_backproject_one_f64": (None, [G, P, P, P, i, i]),
            "or_order_shuffle": (None, [P, i, i, P]),
            "or_ossart": (i, [G, P, P, P, i, i, d, i, i, i, i, P]),
            "or_ossart_errvol": (i, [G, P, P, P, i, i, d, i, i, i, i, P, P]),
            "or_convert_diff_to_error": (None, [P, l]),
            "or_ossart_f64": (i, [G, P, P, P, i, i, d, i, i, i, i, P]),
            "or_ossart_avail": (i, [G, P, P, P, i, P, i, i, d, i, i, i, i, P]),
            "or_order_shuffle_avail": (None, [P, i, P, i, P]),
            "or_available_indexes": (i, [i, i, P]),
            "or_weighting_mask": (None, [i, i, P, P, i, d, P]),
            "or_fft2": (None, [P, P, i, i, i]),
            "or_weighting": (None, [P, P, i, i]),
            "or_weighting_f64": (None, [P, P, i, i]),
            "or_wbp": (None, [G, P, P, P, i, P, i, i, i]),
            "or_wbp_f64": (None, [G, P, P, P, i, P, i, i]),
            "or_ellipsoids": (None, [u64, i, i, i, i, P]),
            "or_phantom_project": (None, [P, i, P, i, i, i, d, d, P]),
            "or_gaussian_noise": (None, [u64, l, P]),
            "or_phantom_volume": (None, [P, i, i, i, i, d, d, d, P]),
            "or_num_procs": (i, []),
            "or_fista_step": (None, [P, P, l, P]),
            "or_roi_rect": (None, [i, d, i, P, P]),
            "or_roi_stats": (None, [P, i, i, i, i, P, P]),
            "or_normalize": (None, [P, l, d, d]),
            "or_affine_inverse": (i, [P, P]),
            "or_affine_resample": (None, [P, i, i, P, i, i, d, d, P]),
            "or_roi_rect_affine": (None, [i, i, d, i, P, P]),
            "or_rect_stats": (None, [P, i, i, P, P, P]),
            "or_tv_denoise3d": (None, [P, i, i, i, C.c_float, C.c_float, C.c_float, i]),
            "or_tv_denoise2d": (None, [P, i, i, C.c_float, C.c_float, C.c_float, i]),
            "or_bayes_convolve": (None, [P, P, i, i, i, d]),
            "or_bayes_end": (None, [P, P, l, d, i]),
            "or_threshold": (None, [P, i, i, d, d, i]),
            "or_poly_stats": (None, [P, i, i, P, P, P]),
            "or_median_cross": (None, [P, i, i, i]),
            "or_collapse": (d, [P, l, d, i, d]),
            "or_roi_poly_affine": (None, [i, i, d, i, P, P]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _p(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"], "arrays passed to the oracle must be C-contiguous"
    return a.ctypes.data_as(C.c_void_p)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


# --------------------------------------------------------------------------
# geometry
# --------------------------------------------------------------------------

def euler_matrix(phi, theta, psi) -> np.ndarray:
    """MatrixUtils.eulerAngles2Matrix (Utility .../MatrixUtils.java:226-251), degrees."""
    E = np.zeros(9)
    _L().or_euler_matrix(phi, theta, psi, _p(E))
    return E.reshape(3, 3)


def tilt_eulers(tilts_deg, tilt_axis_deg=0.0, axis_vertical=True) -> np.ndarray:
    """AffineAlignment.getEulerMatrix for each tilt (align/AffineAlignment.java:382-395)."""
    out = np.zeros((len(tilts_deg), 9))
    for k, t in enumerate(tilts_deg):
        _L().or_tilt_euler(float(t), float(tilt_axis_deg), int(axis_vertical), _p(out[k]))
    return out.reshape(-1, 3, 3)


def fp_factors(E) -> np.ndarray:
    """Per-projection constants of projectJosephPartial (VoxelProjector3D.java:682-694):
    [icf, zcf, divy, yfx, zfx, divx, dirX, dirY, normIncrement]."""
    E = _c(E, np.float64)
    out = np.zeros(9)
    _L().or_fp_factors(_p(E), _p(out))
    return out


def getpixel(vol, x, y, z) -> float:
    """TomoReconstruction2.getPixel (TomoReconstruction2.java:154-190)."""
    vol = _c(vol, np.float32)
    D, H, W = vol.shape
    return _L().or_getpixel(_p(vol), W, H, D, float(x), float(y), int(z))


def ij_interpolated_value(img, x, y) -> float:
    """ImageJ ImageProcessor.getInterpolatedValue restatement (third-party, ij 1.53a)."""
    img = _c(img, np.float32)
    H, W = img.shape
    return _L().or_ij_interpolated_value(_p(img), W, H, float(x), float(y))


# --------------------------------------------------------------------------
# projector SPI pieces
# --------------------------------------------------------------------------

def project(g: Geom, vol, eulers, start_y=0, end_y=None, f64=False):
    """VoxelProjector3D.project (:333-376): one projection + norm image per Euler matrix."""
    end_y = g.vh if end_y is None else end_y
    dt = np.float64 if f64 else np.float32
    vol = _c(vol, dt)
    eulers = _c(eulers, np.float64).reshape(-1, 9)
    n = eulers.shape[0]
    proj = np.zeros((n, g.ph, g.pw), dt)
    norm = np.zeros((n, g.ph, g.pw), dt)
    fn = _L().or_project_one_f64 if f64 else _L().or_project_one
    for p in range(n):
        fn(C.byref(g), _p(vol), _p(eulers[p]), _p(proj[p]), _p(norm[p]), start_y, end_y)
    return proj, norm


def difference(exp, th, norm, factor, f64=False):
    """VoxelProjector3D.difference (:414-433) -> (diff, error)."""
    dt = np.float64 if f64 else np.float32
    exp, th, norm = _c(exp, dt), _c(th, dt), _c(norm, dt)
    diff = np.zeros_like(exp)
    fn = _L().or_difference_f64 if f64 else _L().or_difference
    err = fn(_p(exp), _p(th), _p(norm), _p(diff), exp.size, float(factor))
    return diff, err


def project_and_difference(g: Geom, vol, exps, eulers, factor, start_y=0, end_y=None, f64=False):
    """VoxelProjector3D.projectAndDifference (:387-401) -> (diffs, errors, projs, norms)."""
    projs, norms = project(g, vol, eulers, start_y, end_y, f64)
    dt = np.float64 if f64 else np.float32
    exps = _c(exps, dt)
    diffs = np.zeros_like(projs)
    errors = np.zeros(len(projs))
    for p in range(len(projs)):
        diffs[p], errors[p] = difference(exps[p], projs[p], norms[p], factor, f64)
    return diffs, errors, projs, norms


def backproject(g: Geom, vol, images, eulers, start_y=0, end_y=None, f64=False):
    """VoxelProjector3D.backproject (:155-208) without weighting: adds every image
    into ``vol`` (in place, returned)."""
    end_y = g.vh if end_y is None else end_y
    dt = np.float64 if f64 else np.float32
    assert vol.dtype == dt and vol.flags["C_CONTIGUOUS"]
    images = _c(images, dt)
    eulers = _c(eulers, np.float64).reshape(-1, 9)
    fn = _L().or_backproject_one_f64 if f64 else _L().or_backproject_one
    for p in range(images.shape[0]):
        fn(C.byref(g), _p(vol), _p(images[p]), _p(eulers[p]), start_y, end_y)
    return vol


def available_indexes(n, projection_type) -> np.ndarray:
    """ReconstructionParameters.getAvailableIndexes (:212-256): ALL 10, EVEN 11, ODD 12."""
    out = np.zeros(max(n, 1), np.int32)
    k = _L().or_available_indexes(int(n), int(projection_type), _p(out))
    return out[:k].copy()


def order_shuffle_available(eulers, available, first_index=0) -> np.ndarray:
    """orderShuffleART (:735-781) over a list of available images."""
    eulers = _c(eulers, np.float64).reshape(-1, 9)
    av = _c(available, np.int32)
    order = np.zeros(max(av.size, 1), np.int32)
    _L().or_order_shuffle_avail(_p(eulers), av.size, _p(av), int(first_index), _p(order))
    return order[:av.size].copy()


def order_shuffle(eulers, first_index=0) -> np.ndarray:
    """TomoReconstruction2.orderShuffleART (:735-781) with a fixed first index."""
    eulers = _c(eulers, np.float64).reshape(-1, 9)
    order = np.zeros(eulers.shape[0], np.int32)
    _L().or_order_shuffle(_p(eulers), eulers.shape[0], int(first_index), _p(order))
    return order


def convert_diff_to_error(diff):
    """VoxelProjector3D.convertDiffToErrorImage (:315-324) on a copy: squared in
    float, NaN kept."""
    out = _c(diff, np.float32).copy()
    _L().or_convert_diff_to_error(_p(out), out.size)
    return out


def ossart(g: Geom, vol, projs, eulers, iterations, lam, update, first_index=0,
           start_y=0, end_y=None, f64=False, available=None, errvol=None):
    """TomoReconstruction2.OSSART (:548-634).  Updates ``vol`` in place; returns the
    per-iteration ``err1`` rows [avg, min, imin, max, imax] (iterations x 5).
    ``available``: the list getAvailableIndexes gives (FP32 only).  ``errvol``
    (FP32, ImageStack layout): VoxelProjector3D's error volume, accumulating the
    backprojected squared differences of every subset (:181-195)."""
    end_y = g.vh if end_y is None else end_y
    dt = np.float64 if f64 else np.float32
    assert vol.dtype == dt and vol.flags["C_CONTIGUOUS"]
    projs = _c(projs, dt)
    eulers = _c(eulers, np.float64).reshape(-1, 9)
    err = np.zeros((iterations, 5))
    if errvol is not None:
        assert not f64 and available is None
        assert errvol.dtype == np.float32 and errvol.flags["C_CONTIGUOUS"] and errvol.shape == vol.shape
        rc = _L().or_ossart_errvol(C.byref(g), _p(vol), _p(projs), _p(eulers), eulers.shape[0], iterations,
                                   float(lam), int(update), int(first_index), start_y, end_y, _p(err), _p(errvol))
        if rc != 0:
            raise MemoryError("oracle OS-SART allocation failed")
        return err
    if available is not None:
        assert not f64
        av = _c(available, np.int32)
        rc = _L().or_ossart_avail(C.byref(g), _p(vol), _p(projs), _p(eulers), eulers.shape[0], _p(av),
                                  av.size, iterations, float(lam), int(update), int(first_index), start_y,
                                  end_y, _p(err))
        if rc != 0:
            raise MemoryError("oracle OS-SART allocation failed")
        return err
    fn = _L().or_ossart_f64 if f64 else _L().or_ossart
    rc = fn(C.byref(g), _p(vol), _p(projs), _p(eulers), eulers.shape[0], iterations,
            float(lam), int(update), int(first_index), start_y, end_y, _p(err))
    if rc != 0:
        raise MemoryError("oracle OS-SART allocation failed")
    return err


def fista_step(vol, prev, t: float) -> float:
    """FistaProjector3D.endOfIteration (FistaProjector3D.java:43-59) on ``vol``
    (in place) against ``prev`` (the volume at startOfIteration); returns the new t."""
    assert vol.dtype == np.float32 and vol.flags["C_CONTIGUOUS"]
    prev = _c(prev, np.float32)
    tt = np.array([t], np.float64)
    _L().or_fista_step(_p(vol), _p(prev), vol.size, _p(tt))
    return float(tt[0])


def roi_rect(width, tilt_deg, min_size=0):
    """Columns [x0, x1) of TiltSeries.getImageStatistics' ROI (TiltSeries.java:809-835)
    with the identity alignment."""
    a, b = C.c_int(), C.c_int()
    _L().or_roi_rect(int(width), float(tilt_deg), int(min_size), C.byref(a), C.byref(b))
    return a.value, b.value


def normalized_projections(images, tilts_deg, min_size=0):
    """TiltSeries.getPixels(i) with normalisation and the identity alignment
    (AffineAlignment.java:136-156): (float)((x - mean) / stdDev) with the ROI
    statistics; returns (images, means, stddevs)."""
    imgs = np.array(images, np.float32, copy=True)
    n, h, w = imgs.shape
    means, sds = np.zeros(n), np.zeros(n)
    for k in range(n):
        x0, x1 = roi_rect(w, tilts_deg[k], min_size)
        m, sd = C.c_double(), C.c_double()
        _L().or_roi_stats(_p(np.ascontiguousarray(imgs[k])), w, h, x0, x1, C.byref(m), C.byref(sd))
        means[k], sds[k] = m.value, sd.value
        row = np.ascontiguousarray(imgs[k])
        _L().or_normalize(_p(row), row.size, m.value, sd.value)
        imgs[k] = row
    return imgs, means, sds


def affine_inverse(m):
    """java.awt.geom.AffineTransform.createInverse of (m00, m10, m01, m11, m02, m12);
    None when not invertible."""
    a = _c(m, np.float64)
    r = np.zeros(6, np.float64)
    return r if _L().or_affine_inverse(_p(a), _p(r)) else None


def roi_rect_affine(width, height, tilt_deg, transform, min_size=0):
    """[x0, x1, y0, y1] of getImageStatistics' ROI mapped by a rotation-free transform."""
    a = _c(transform, np.float64)
    r = np.zeros(4, np.int32)
    _L().or_roi_rect_affine(int(width), int(height), float(tilt_deg), int(min_size), _p(a), _p(r))
    return [int(v) for v in r]


def rect_stats(img, rect):
    img = _c(img, np.float32)
    r = _c(rect, np.int32)
    m, sd = C.c_double(), C.c_double()
    _L().or_rect_stats(_p(img), img.shape[1], img.shape[0], _p(r), C.byref(m), C.byref(sd))
    return m.value, sd.value


def aligned_projections(images, transforms, tilts_deg=None, min_size=0, normalize=True, fill=1, stats=None):
    """TiltSeries.getPixels(i) through AffineAlignment.applyTransformOnImage
    (AffineAlignment.java:136-254) for every image; stats n x (mean, stdDev) or
    None (the mapped ROI: its rectangle for rotation-free transforms, else the
    pixels whose centres lie in its polygon).  Returns (images, stats)."""
    imgs = np.array(images, np.float32, copy=True)
    n, h, w = imgs.shape
    T = _c(transforms, np.float64).reshape(n, 6)
    st = np.zeros((n, 2))
    for k in range(n):
        src = np.ascontiguousarray(imgs[k])
        need = normalize or fill == 1
        if need:
            if stats is not None:
                st[k] = stats[k]
            elif T[k][1] == 0.0 and T[k][2] == 0.0:
                st[k] = rect_stats(src, roi_rect_affine(w, h, tilts_deg[k], T[k], min_size))
            else:
                st[k] = poly_stats(src, roi_poly_affine(w, h, tilts_deg[k], T[k], min_size))
        if np.array_equal(T[k], [1.0, 0.0, 0.0, 1.0, 0.0, 0.0]):
            if normalize:
                _L().or_normalize(_p(src), src.size, st[k][0], st[k][1])
            imgs[k] = src
            continue
        inv = affine_inverse(T[k])
        if inv is None:
            raise ValueError(f"transform {k} is not invertible")
        out = np.empty_like(src)
        _L().or_affine_resample(_p(src), w, h, _p(inv), int(bool(normalize)), int(fill), float(st[k][0]),
                                float(st[k][1]) if need else 1.0, _p(out))
        imgs[k] = out
    return imgs, st


def weighting_mask(sx, sy, tilts_deg, diameter, axis_deg=None) -> np.ndarray:
    """FFTWeighting.createWeightingMask (filters/FFTWeighting.java:31-65)."""
    tilts = _c(tilts_deg, np.float64)
    axis = np.zeros_like(tilts) if axis_deg is None else _c(axis_deg, np.float64)
    mask = np.zeros((sy, sx), np.float32)
    _L().or_weighting_mask(sx, sy, _p(tilts), _p(axis), len(tilts), float(diameter), _p(mask))
    return mask


def fft2(re, im, sign=-1):
    """2D complex DFT (double), in place on copies; returns (re, im)."""
    re = _c(re, np.float64).copy()
    im = _c(im, np.float64).copy()
    sy, sx = re.shape
    _L().or_fft2(_p(re), _p(im), sx, sy, int(sign))
    return re, im


def weighting(img, mask, f64=False):
    """FFTWeighting.weighting (:87-112): real(ifft2(mask * fft2(img)))."""
    dt = np.float64 if f64 else np.float32
    img = _c(img, dt).copy()
    mask = _c(mask, dt)
    sy, sx = img.shape
    (_L().or_weighting_f64 if f64 else _L().or_weighting)(_p(img), _p(mask), sx, sy)
    return img


def wbp(g: Geom, vol, projs, eulers, mask=None, update=0, start_y=0, end_y=None, f64=False):
    """TomoReconstruction2.WBP (:499-538); ``vol`` updated in place."""
    end_y = g.vh if end_y is None else end_y
    dt = np.float64 if f64 else np.float32
    assert vol.dtype == dt and vol.flags["C_CONTIGUOUS"]
    projs = _c(projs, dt)
    eulers = _c(eulers, np.float64).reshape(-1, 9)
    n = eulers.shape[0]
    if f64:
        m = None if mask is None else _c(mask, np.float64)
        _L().or_wbp_f64(C.byref(g), _p(vol), _p(projs), _p(eulers), n,
                        None if m is None else _p(m), start_y, end_y)
    else:
        m = None if mask is None else _c(mask, np.float32)
        _L().or_wbp(C.byref(g), _p(vol), _p(projs), _p(eulers), n,
                    None if m is None else _p(m), int(update), start_y, end_y)
    return vol


def tv_denoise3d(vol, theta, g, dt, iterations) -> np.ndarray:
    """TotalVariationDenoising_3D.denoise3D (TotalVariationDenoising_3D.java:84-202) of a
    (D, H, W) float32 volume; returns the denoised copy.  theta, g and dt are floats
    as in Java."""
    out = np.array(vol, np.float32, order="C", copy=True)
    D, H, W = out.shape
    _L().or_tv_denoise3d(_p(out), W, H, D, float(np.float32(theta)), float(np.float32(g)),
                         float(np.float32(dt)), int(iterations))
    return out


def tv_denoise2d(img, theta, g, dt, iterations) -> np.ndarray:
    """TotalVariationDenoising_3D.denoise2D (:206-260) of an (H, W) float32 image."""
    out = np.array(img, np.float32, order="C", copy=True)
    h, w = out.shape
    _L().or_tv_denoise2d(_p(out), w, h, float(np.float32(theta)), float(np.float32(g)),
                         float(np.float32(dt)), int(iterations))
    return out


def bayes_convolve(vol, factor) -> np.ndarray:
    """BayesianVoxelProjector3D.convolve(rec, kernel, dist, factor) (:175-198) of a
    (D, H, W) float32 volume; returns the distance volume."""
    vol = _c(vol, np.float32)
    D, H, W = vol.shape
    dist = np.empty_like(vol)
    _L().or_bayes_convolve(_p(vol), _p(dist), W, H, D, float(factor))
    return dist


def bayes_iteration(g: Geom, vol, projs, eulers, alpha, lam, first_index=0):
    """One iteration of TomoReconstruction2.regularization (:783-878) with
    BayesianVoxelProjector3D and update = N (the Bayesian factory's): the
    distance volume is -lam times the convolved volume, every projection's
    difference (factor (float)(alpha/N)) is backprojected into it, then
    vol += (float)alpha * dist and the positivity step of ``g``.  ``vol`` is
    updated in place; returns the err row of osartIteration."""
    assert vol.dtype == np.float32 and vol.flags["C_CONTIGUOUS"]
    eulers = _c(eulers, np.float64).reshape(-1, 9)
    n = eulers.shape[0]
    order = order_shuffle(eulers, first_index)
    dist = bayes_convolve(vol, -float(lam))
    factor = float(np.float32(alpha / n))
    diffs, errs, _, _ = project_and_difference(g, vol, _c(projs, np.float32)[order], eulers[order], factor)
    backproject(g, dist, diffs, eulers[order])
    _L().or_bayes_end(_p(vol), _p(dist), vol.size, float(alpha), int(g.positivity))
    wh = float(g.vw * g.vh)
    som, emin, emax, imin, imax = 0.0, 3.4028234663852886e38, 0.0, 0, 0
    for t in range(n):
        e = errs[t]
        som += e
        if e > emax:
            emax, imax = e, int(order[t])
        elif e < emin:
            emin, imin = e, int(order[t])
    return [math.sqrt(som / (n * wh)), math.sqrt(emin / wh), imin, math.sqrt(emax / wh), imax]


def median_cross(vol) -> np.ndarray:
    """BgARTVoxelProjector3D.medianFilter3Dcross (:123-151) of a (D, H, W) volume (a copy)."""
    out = np.array(vol, np.float32, order="C", copy=True)
    D, H, W = out.shape
    _L().or_median_cross(_p(out), W, H, D)
    return out


def collapse(vol, k, dark_bg=True, bg_value=float("nan")):
    """BgARTVoxelProjector3D.collapseFilter (:34-78); returns (volume copy, background)."""
    out = np.array(vol, np.float32, order="C", copy=True)
    mv = _L().or_collapse(_p(out), out.size, float(k), int(bool(dark_bg)), float(bg_value))
    return out, mv


def bgart_iteration(g: Geom, vol, projs, eulers, lam, update, k, dark_bg=True, median=True, first_index=0):
    """One iteration of TomoReconstruction2.regularization with BgARTVoxelProjector3D:
    the cross median (startOfIteration), osartIteration, the collapse filter
    (endOfIteration).  ``vol`` in place; returns the err row."""
    if median:
        vol[...] = median_cross(vol)
    e = ossart(g, vol, projs, eulers, 1, lam, update, first_index)[0]
    vol[...] = collapse(vol, k, dark_bg)[0]
    return e


def roi_poly_affine(width, height, tilt_deg, transform, min_size=0) -> np.ndarray:
    """The int corners (xs[4], ys[4]) of getImageStatistics' ROI polygon under
    the forward transform (m00, m10, m01, m11, m02, m12)."""
    m = _c(transform, np.float64)
    out = np.zeros(8, np.int32)
    _L().or_roi_poly_affine(int(width), int(height), float(tilt_deg), int(min_size), _p(m), _p(out))
    return out


def poly_stats(img, poly):
    """Mean and sample standard deviation over the pixels whose centres lie in
    the polygon poly = (xs[4], ys[4])."""
    img = _c(img, np.float32)
    h, w = img.shape
    pl = _c(poly, np.int32)
    m, sd = C.c_double(), C.c_double()
    _L().or_poly_stats(_p(img), w, h, _p(pl), C.byref(m), C.byref(sd))
    return m.value, sd.value


def threshold(img, threshold, hysteresis=5e-324, fill=1) -> np.ndarray:
    """TiltSeries.applyThreshold (TiltSeries.java:720-806) of one (H, W) float32 image;
    returns the thresholded copy.  fill: 0 none (0), 1 average, 2 NaN; the
    hysteresis fraction defaults to Double.MIN_VALUE as setThreshold leaves it."""
    out = np.array(img, np.float32, order="C", copy=True)
    h, w = out.shape
    _L().or_threshold(_p(out), w, h, float(threshold), float(hysteresis), int(fill))
    return out


def ssnr_sums(sig_exps, sig_th, noi_exps, noi_th, ring_width=4) -> np.ndarray:
    """ResolutionEstimation.addToVectors1D (:581-697) over every tilt, FP64
    transforms: NaN -> 0, the float differences measured - projected, the power
    spectra |fft2|^2 of difference and projection for signal and noise, and
    ring sums (a frequency at radius w*W adds to rings
    [max(0, ceil(w*W - ringWidth)), floor(w*W))).  Returns sums[5][W]:
    thSignal, signal - thSignal, thNoise, noise - thNoise, count."""
    se, st, ne, nt = (np.nan_to_num(_c(a, np.float32), nan=0.0, posinf=np.inf, neginf=-np.inf)
                      for a in (sig_exps, sig_th, noi_exps, noi_th))
    n, h, w = se.shape

    def freq(m):
        i = np.arange(m)
        return np.where(i <= (m >> 1), i, i - m) / m

    fy, fx = np.meshgrid(freq(h) ** 2, freq(w) ** 2, indexing="ij")
    wi = np.sqrt(fx + fy) * w
    l0 = np.maximum(0, np.ceil(wi - ring_width)).astype(np.int64)
    lf = np.floor(wi).astype(np.int64)
    sums = np.zeros((5, w))
    for t in range(n):
        pw = [np.abs(np.fft.fft2(a.astype(np.float64))) ** 2
              for a in (st[t], (se[t] - st[t]).astype(np.float32), nt[t], (ne[t] - nt[t]).astype(np.float32))]
        for k in range(ring_width + 1):
            l = l0 + k
            ok = l < lf
            for row in range(4):
                np.add.at(sums[row], l[ok], pw[row][ok])
            np.add.at(sums[4], l[ok], 1.0)
    return sums


def vssnr_maps(sig_exps, sig_th, noi_exps, noi_th) -> np.ndarray:
    """The per-pixel SSNR maps of SSNR(true) (:648-663) with FP64 transforms:
    10 log10(max(ISSNR/alpha - 1, 0) + 1) (NaN kept, as Java's Math.max keeps
    it), laid out by swapQuadrant2D (:720-733): row y to (y + H/2) mod H, its
    first W/2 entries to the right half, the next W/2 to the left, an odd
    width's last column left 0."""
    se, st, ne, nt = (np.nan_to_num(_c(a, np.float32), nan=0.0, posinf=np.inf, neginf=-np.inf)
                      for a in (sig_exps, sig_th, noi_exps, noi_th))
    n, h, w = se.shape
    cx, cy = w // 2, h // 2
    out = np.zeros((n, h, w), np.float32)
    for t in range(n):
        p = [np.abs(np.fft.fft2(a.astype(np.float64))) ** 2
             for a in ((se[t] - st[t]).astype(np.float32), st[t], (ne[t] - nt[t]).astype(np.float32), nt[t])]
        with np.errstate(divide="ignore", invalid="ignore"):
            aux = (p[1] / p[0]) / (p[3] / p[2]) - 1.0
            m = (10.0 * np.log10(np.where(np.isnan(aux), aux, np.maximum(aux, 0.0)) + 1.0)).astype(np.float32)
        for y in range(h):
            yc = (y + cy) % h
            out[t, yc, cx:2 * cx] = m[y, :cx]
            out[t, yc, :cx] = m[y, cx:2 * cx]
    return out


def ssnr_rows(sums, pixel_size=1.0) -> np.ndarray:
    """SSNR (:212-260) from the ring sums: nine rows over the non-negative
    frequencies; the dB rows divide by TomoJ's counter imgno = W + 1."""
    SS, SN, NS, NN, K = (np.array(r, np.float64) for r in sums)
    w = SS.size
    with np.errstate(divide="ignore", invalid="ignore"):
        s_snr, n_snr = SS / SN, NS / NN
        SS, SN, NS, NN = SS / K, SN / K, NS / K, NN / K
        imgno = w + 1
        r = w // 2 + 1
        i = np.arange(r)
        ssnr = s_snr[:r] / n_snr[:r]
        corr = np.where(ssnr > 1, 10 * np.log10(np.where(ssnr > 1, ssnr - 1, 1.0)), -1000.0)
        return np.array([i, (i / w) / pixel_size, corr, s_snr[:r], 10 * np.log10(SS[:r] / imgno),
                         10 * np.log10(SN[:r] / imgno), n_snr[:r], 10 * np.log10(NS[:r] / imgno),
                         10 * np.log10(NN[:r] / imgno)])


def fsc_mask(D, H, W) -> np.ndarray:
    """ResolutionEstimation.fsc's mask (ResolutionEstimation.java:470-489), (D, H, W)
    float32: 1 within radius D of (W/2, H/2, D/2), a raised cosine over the next
    20 voxels, 0 beyond; dist^2 = dx^2 + (dz^2 + dy^2)."""
    z, y, x = np.meshgrid(np.arange(D), np.arange(H), np.arange(W), indexing="ij")
    dz = (z - D / 2.0) ** 2
    dy = (y - H / 2.0) ** 2
    dist = np.sqrt((x - W / 2.0) ** 2 + (dz + dy))
    taper = (1 + np.cos((dist - D) * math.pi / 20)).astype(np.float32) * np.float32(0.5)
    return np.where(dist < D, np.float32(1), np.where(dist < D + 20, taper, np.float32(0))).astype(np.float32)


def fsc(even, odd) -> np.ndarray:
    """ResolutionEstimation.fsc(even, odd) (:456-557) in double precision: the masked
    volumes (float products) transformed with an FP64 FFT, the shell index
    round(|f| W) from convertToFrequency (:699-701), and per shell the sums of
    Re(E conj O), |E|^2, |O|^2, |E - O| and the count; returns fsc[4][W/2+1]."""
    even = _c(even, np.float32)
    odd = _c(odd, np.float32)
    D, H, W = even.shape
    m = fsc_mask(D, H, W)
    E = np.fft.fftn((even * m).astype(np.float64))
    O = np.fft.fftn((odd * m).astype(np.float64))

    def freq(n):
        i = np.arange(n)
        return np.where(i <= (n >> 1), i, i - n) / n

    fz, fy, fx = np.meshgrid(freq(D) ** 2, freq(H) ** 2, freq(W) ** 2, indexing="ij")
    wi = np.floor(np.sqrt(fx + fy + fz) * W + 0.5).astype(np.int64).ravel()
    R = W // 2 + 1
    keep = wi < R
    wi = wi[keep]
    e = E.ravel()[keep]
    o = O.ravel()[keep]
    num = np.bincount(wi, (e * np.conj(o)).real, R)
    de = np.bincount(wi, np.abs(e) ** 2, R)
    do = np.bincount(wi, np.abs(o) ** 2, R)
    err = np.bincount(wi, np.abs(e - o), R)
    cnt = np.bincount(wi, None, R).astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.array([np.arange(R) / W, num / np.sqrt(de * do), 2 / np.sqrt(cnt), err / cnt])


# --------------------------------------------------------------------------
# synthetic data
# --------------------------------------------------------------------------

def ellipsoids(seed, k, W, H, D) -> np.ndarray:
    """Seeded ellipsoid parameters, k x 16 (density, centre[3], semi-axes[3], R[9])."""
    out = np.zeros((k, 16))
    _L().or_ellipsoids(C.c_uint64(seed), k, W, H, D, _p(out))
    return out


def phantom_project(ell, eulers, pw, ph, pc0=None, pc1=None) -> np.ndarray:
    """Analytic FP64 line integrals of the ellipsoid phantom (n x ph x pw)."""
    ell = _c(ell, np.float64)
    eulers = _c(eulers, np.float64).reshape(-1, 9)
    pc0 = (pw - 1.0) / 2 if pc0 is None else pc0
    pc1 = (ph - 1.0) / 2 if pc1 is None else pc1
    out = np.zeros((eulers.shape[0], ph, pw))
    _L().or_phantom_project(_p(ell), ell.shape[0], _p(eulers), eulers.shape[0], pw, ph,
                            float(pc0), float(pc1), _p(out))
    return out


def gaussian_noise(seed, count) -> np.ndarray:
    out = np.zeros(count)
    _L().or_gaussian_noise(C.c_uint64(seed), int(count), _p(out))
    return out


def phantom_volume(ell, W, H, D) -> np.ndarray:
    ell = _c(ell, np.float64)
    vol = np.zeros((D, H, W), np.float32)
    _L().or_phantom_volume(_p(ell), ell.shape[0], W, H, D, (W - 1) / 2.0, (H - 1) / 2.0,
                           (D - 1) / 2.0, _p(vol))
    return vol


def num_procs() -> int:
    return _L().or_num_procs()
