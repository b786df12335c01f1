"""The allocation-free projectAndDifference (tj_project_and_difference_noalloc)
against the context entry point tj_project_and_difference on the same inputs:
difference images bit for bit, FP64 errors to 1e-12 relative (another
summation order), on two calls in a row, the second smaller.  Device buffers
are the caller's (hipMalloc from this test through ctypes)."""
import ctypes as C

import numpy as np
import pytest

from tomoj_amd import _lib


class _Hip:
    def __init__(self):
        self.h = C.CDLL("libamdhip64.so")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipDeviceSynchronize.argtypes = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), C.c_size_t(nbytes)) == 0
        return p

    def free(self, p):
        assert self.h.hipFree(p) == 0

    def h2d(self, dst, a):
        a = np.ascontiguousarray(a)
        assert self.h.hipMemcpy(dst, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0

    def d2h(self, a, src):
        assert self.h.hipMemcpy(a.ctypes.data_as(C.c_void_p), src, C.c_size_t(a.nbytes), 2) == 0


def _case(oracle, W, H, D, step, seed, nan_pixels):
    tilts = np.arange(-60, 61, step, dtype=np.float64)
    E = oracle.tilt_eulers(tilts)
    ell = oracle.ellipsoids(seed, 6, W, H, D)
    exps = oracle.phantom_project(ell, E, W, H).astype(np.float32)
    rng = np.random.default_rng(seed)
    if nan_pixels:
        k = rng.integers(0, exps.size, nan_pixels)
        exps.reshape(-1)[k] = np.nan
    vol = (rng.random((D, H, W), dtype=np.float32) * np.float32(0.05)).astype(np.float32)
    return E, exps, vol


def _context_result(tj, E, exps, vol, factor):
    D, H, W = vol.shape
    n = len(E)
    with tj.VoxelProjector3DHIP((W, H, D), device=0) as P:
        P.load_tilt_series(exps, E)
        P.set_volume(vol)
        for k in range(n):
            P.addProjection(k)
        err = P.projectAndDifference(factor)
        diffs = np.stack([P.getDifferenceImage(k) for k in range(n)])
    return diffs, np.asarray(err)


def _noalloc_result(hip, E, exps, vol, factor):
    D, H, W = vol.shape
    n = len(E)
    lib = _lib.load()
    g = _lib.tj_geometry()
    lib.tj_geometry_default(C.byref(g), W, H, D, W, H)
    vol_yzx = np.ascontiguousarray(vol.transpose(1, 0, 2))
    d_vol = hip.alloc(vol_yzx.nbytes)
    d_exp = hip.alloc(exps.nbytes)
    d_diff = hip.alloc(exps.nbytes)
    try:
        hip.h2d(d_vol, vol_yzx)
        hip.h2d(d_exp, exps)
        eul = np.ascontiguousarray(E.reshape(n, 9), dtype=np.float64)
        errors = np.zeros(n, np.float64)
        print(f"[noalloc] {W}x{H}x{D}, {n} images", flush=True)
        _lib.call("tj_project_and_difference_noalloc", 0, C.byref(g), d_vol, d_exp,
                  eul.ctypes.data_as(C.c_void_p), n, float(factor), 0, H, d_diff,
                  errors.ctypes.data_as(C.c_void_p), None)
        diffs = np.empty_like(exps)
        hip.d2h(diffs, d_diff)
    finally:
        hip.free(d_vol)
        hip.free(d_exp)
        hip.free(d_diff)
    return diffs, errors


def test_noalloc_rejects_before_any_device_work():
    """Sizes beyond the scratch array and geometries outside the single-axis LDS
    path are refused on the host, with a message, before any HIP call."""
    lib = _lib.load()
    g = _lib.tj_geometry()
    lib.tj_geometry_default(C.byref(g), 512, 512, 256, 512, 512)
    assert lib.tj_noalloc_scratch_bytes(C.byref(g), 121) <= _lib.TJ_NOALLOC_SCRATCH_BYTES
    n_big = 4096
    assert lib.tj_noalloc_scratch_bytes(C.byref(g), n_big) > _lib.TJ_NOALLOC_SCRATCH_BYTES
    dummy = np.zeros(16, np.float32)
    p = dummy.ctypes.data_as(C.c_void_p)
    eul = np.tile(np.eye(3).reshape(9), (n_big, 1))
    errors = np.zeros(n_big)
    rc = lib.tj_project_and_difference_noalloc(0, C.byref(g), p, p, eul.ctypes.data_as(C.c_void_p), n_big, 1.0,
                                               0, 512, p, errors.ctypes.data_as(C.c_void_p), None)
    assert rc == _lib.TJ_ENOMEM and "scratch" in _lib.last_error()
    # a tilted (non-vertical) axis: not the LDS path
    eul2 = np.array([[1, 0, 0, 0, np.cos(0.3), -np.sin(0.3), 0, np.sin(0.3), np.cos(0.3)]] * 2, np.float64)
    rc = lib.tj_project_and_difference_noalloc(0, C.byref(g), p, p, eul2.ctypes.data_as(C.c_void_p), 2, 1.0,
                                               0, 512, p, errors.ctypes.data_as(C.c_void_p), None)
    assert rc == _lib.TJ_EUNSUP, _lib.last_error()
    g.long_object_comp = 1
    rc = lib.tj_project_and_difference_noalloc(0, C.byref(g), p, p, eul.ctypes.data_as(C.c_void_p), 2, 1.0,
                                               0, 512, p, errors.ctypes.data_as(C.c_void_p), None)
    assert rc == _lib.TJ_EUNSUP and "long-object" in _lib.last_error()


@pytest.mark.gpu
def test_noalloc_matches_project_and_difference_twice(tj, oracle):
    hip = _Hip()
    factor = float(np.float32(1.0 / 7))
    cases = [(96, 20, 40, 4.0, 11, 9), (48, 12, 24, 15.0, 12, 0)]  # second call: smaller volume, fewer images
    for W, H, D, step, seed, nanpx in cases:
        E, exps, vol = _case(oracle, W, H, D, step, seed, nanpx)
        want_d, want_e = _context_result(tj, E, exps, vol, factor)
        got_d, got_e = _noalloc_result(hip, E, exps, vol, factor)
        assert np.array_equal(got_d, want_d, equal_nan=True), (W, H, D, int((got_d != want_d).sum()))
        np.testing.assert_allclose(got_e, want_e, rtol=1e-12, atol=0)
        assert np.all(want_e > 0)
