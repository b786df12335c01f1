"""CPU checks of the C ABI library and the host-side logic (no GPU compute)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "tomoj_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tj_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from tomoj_amd import _lib
    lib = _lib.load()
    names = declared_symbols()
    assert len(names) >= 30
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.EXPORTED), set(names) ^ set(_lib.EXPORTED)


def test_geometry_defaults_follow_tomoj():
    """Centres (n-1)/2 (TiltSeries.java:161-170, TomoReconstruction2.java:63-71), scale 1."""
    from tomoj_amd import _lib
    g = _lib.tj_geometry()
    _lib.load().tj_geometry_default(C.byref(g), 512, 400, 256, 600, 500)
    assert list(g.proj_center) == [299.5, 249.5, 299.5, 249.5]
    assert list(g.vol_center) == [255.5, 199.5, 127.5]
    assert g.scale == 1.0 and g.positivity == 0 and g.long_object_comp == 0


def test_ellipsoid_parameters_match_oracle(oracle):
    """The library's host-side phantom generator and the oracle's draw identical streams."""
    from tomoj_amd import phantom_ellipsoids
    for seed, k, dims in [(1, 3, (32, 24, 16)), (0x70A10002, 40, (512, 512, 256))]:
        a = phantom_ellipsoids(seed, k, *dims)
        b = oracle.ellipsoids(seed, k, *dims)
        assert np.array_equal(a, b)


def test_open_without_device_fails_cleanly():
    """Without a HIP device tj_open returns an error code and a message (no crash)."""
    from tomoj_amd import _lib
    from tomoj_amd import VoxelProjector3DHIP, TjError
    try:
        hip = C.CDLL("libamdhip64.so")
        n = C.c_int(0)
        if hip.hipGetDeviceCount(C.byref(n)) == 0 and n.value > 0:
            pytest.skip("a HIP device is present")
    except OSError:
        pass
    with pytest.raises(TjError) as ei:
        VoxelProjector3DHIP((8, 8, 8))
    assert ei.value.code in (_lib.TJ_EDEVICE, _lib.TJ_EINVAL)
    assert _lib.last_error()


def test_byte_model_numbers():
    """Per-iteration algorithmic bytes of the slab-streaming model for the named
    workloads (FP32): 2.162 GB (256x256x128, 61 tilts, SIRT), 38.23 GB
    (512x512x256, 121 tilts, 10 subsets), 66.42 GB (512^3, 121 tilts, SIRT)."""
    from tomoj_amd import workloads as W
    assert W.algorithmic_bytes_per_iteration(W.SIRT_256) / 1e9 == pytest.approx(2.162, abs=5e-4)
    assert W.algorithmic_bytes_per_iteration(W.OSSART_512) / 1e9 == pytest.approx(38.23, abs=5e-3)
    assert W.algorithmic_bytes_per_iteration(W.SIRT_512_CUBE) / 1e9 == pytest.approx(66.42, abs=5e-3)
    assert W.OSSART_512.subsets() == 10
    assert W.OSSART_512.voxel_ray_updates_per_iteration() == 2 * 121 * 512 * 512 * 256


def test_tilt_schemes():
    from tomoj_amd import workloads as W
    assert W.SIRT_256.n == 61 and W.SIRT_256.tilts[0] == -60 and W.SIRT_256.tilts[-1] == 60
    assert W.OSSART_512.n == 121 and W.OSSART_512.tilts[60] == 0
