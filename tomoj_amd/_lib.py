"""ctypes binding of libtomoj_hip.so (the C ABI in include/tomoj_hip.h).

The library must exist (built by ``tomoj_amd.build.build()``); loading fails
loudly otherwise -- there is no CPU fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtomoj_hip.so")

TJ_OK = 0
TJ_EINVAL = -1
TJ_ENOMEM = -2
TJ_EDEVICE = -3
TJ_ESTATE = -4
TJ_EUNSUP = -5
TJ_ECANCELLED = -6
TJ_EIO = -7

TJ_FILL_NONE = 0
TJ_FILL_AVG = 1
TJ_FILL_NAN = 2

TJ_STORAGE_F32 = 0
TJ_STORAGE_BF16 = 1
TJ_STORAGE_F16 = 2
TJ_STORAGE_F64 = 3
STORAGE_NAMES = {"f32": TJ_STORAGE_F32, "bf16": TJ_STORAGE_BF16, "f16": TJ_STORAGE_F16, "f64": TJ_STORAGE_F64}

TJ_RUN_GRAPH = 1
TJ_NOALLOC_SCRATCH_BYTES = 16 << 20   # include/tomoj_hip.h


class TjError(RuntimeError):
    """A libtomoj_hip call returned a negative status."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class TjCancelled(TjError):
    pass


class tj_geometry(C.Structure):
    _fields_ = [("vol_w", C.c_int), ("vol_h", C.c_int), ("vol_d", C.c_int),
                ("proj_w", C.c_int), ("proj_h", C.c_int),
                ("proj_center", C.c_double * 4), ("vol_center", C.c_double * 3),
                ("scale", C.c_double), ("positivity", C.c_int),
                ("long_object_comp", C.c_int), ("storage", C.c_int)]


class tj_timing(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("fp_ms", C.c_double), ("res_ms", C.c_double),
                ("bp_ms", C.c_double), ("other_ms", C.c_double), ("fp_launches", C.c_int),
                ("res_launches", C.c_int), ("bp_launches", C.c_int),
                ("bpprep_ms", C.c_double), ("bpprep_launches", C.c_int), ("fused_subsets", C.c_int),
                ("fast_bp_launches", C.c_int)]


class tj_tuning(C.Structure):
    _fields_ = [("fp_kernel", C.c_int), ("fp_rows", C.c_int), ("fp_rays", C.c_int), ("fp_zchunk", C.c_int),
                ("bp_kernel", C.c_int), ("bp_rows", C.c_int), ("bp_zchunk", C.c_int), ("bp_group", C.c_int),
                ("bp_fast", C.c_int), ("fp_groups", C.c_int), ("fp_prefetch", C.c_int)]


class tj_job_desc(C.Structure):
    _fields_ = [("vol_w", C.c_int), ("vol_h", C.c_int), ("vol_d", C.c_int),
                ("proj_w", C.c_int), ("proj_h", C.c_int), ("n_proj", C.c_int),
                ("storage", C.c_int), ("positivity", C.c_int), ("relaxation", C.c_double),
                ("max_iterations", C.c_int), ("fp32_master", C.c_int),
                ("fp32_projections", C.c_int), ("reserved", C.c_int)]


class tj_batch_tuning(C.Structure):
    _fields_ = [("fp_rays", C.c_int), ("fp_lds_kb", C.c_int), ("bp_rows", C.c_int),
                ("bp_zchunk", C.c_int), ("bp_group", C.c_int), ("blocks_per_cu", C.c_int),
                ("fp_packed", C.c_int), ("fp_cols", C.c_int), ("fp_shared", C.c_int)]


# name -> (restype, argtypes); every function returning int is status-checked by call()
_P = C.c_void_p
_SIGS = {
    "tj_geometry_default": (None, [C.POINTER(tj_geometry), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "tj_last_error": (C.c_char_p, []),
    "tj_version": (C.c_char_p, []),
    "tj_device_info": (C.c_int, [C.c_int, C.c_char_p, C.c_int]),
    "tj_copy_probe": (C.c_int, [C.c_int, C.c_size_t, C.c_int, C.POINTER(C.c_double)]),
    "tj_open": (C.c_int, [C.c_int, C.POINTER(tj_geometry), C.POINTER(_P)]),
    "tj_close": (C.c_int, [_P]),
    "tj_load_projections": (C.c_int, [_P, _P, C.c_int, _P]),
    "tj_set_eulers": (C.c_int, [_P, _P, C.c_int]),
    "tj_set_volume": (C.c_int, [_P, _P, C.c_int, C.c_int]),
    "tj_get_volume": (C.c_int, [_P, _P, C.c_int, C.c_int]),
    "tj_reset_volume": (C.c_int, [_P]),
    "tj_set_positivity": (C.c_int, [_P, C.c_int]),
    "tj_set_error_volume": (C.c_int, [_P, C.c_int]),
    "tj_get_error_volume": (C.c_int, [_P, _P]),
    "tj_clear_error_volume": (C.c_int, [_P]),
    "tj_add_projection": (C.c_int, [_P, C.c_int]),
    "tj_clear_projections": (C.c_int, [_P]),
    "tj_num_added": (C.c_int, [_P]),
    "tj_project": (C.c_int, [_P, C.c_int, C.c_int]),
    "tj_project_and_difference": (C.c_int, [_P, C.c_double, C.c_int, C.c_int, _P]),
    "tj_get_projection": (C.c_int, [_P, C.c_int, _P, _P]),
    "tj_get_difference": (C.c_int, [_P, C.c_int, _P]),
    "tj_noalloc_scratch_bytes": (C.c_size_t, [C.POINTER(tj_geometry), C.c_int]),
    "tj_project_and_difference_noalloc": (C.c_int, [C.c_int, C.POINTER(tj_geometry), _P, _P, _P, C.c_int,
                                                    C.c_double, C.c_int, C.c_int, _P, _P, _P]),
    "tj_backproject": (C.c_int, [_P, C.c_int, C.c_int]),
    "tj_set_weighting": (C.c_int, [_P, C.c_double, _P, _P, C.c_int]),
    "tj_set_ramp_filter": (C.c_int, [_P, C.c_int, C.c_double]),
    "tj_get_weighting_mask": (C.c_int, [_P, _P]),
    "tj_apply_weighting": (C.c_int, [_P, _P]),
    "tj_start_iteration": (C.c_int, [_P]),
    "tj_end_iteration": (C.c_int, [_P]),
    "tj_set_fista": (C.c_int, [_P, C.c_int]),
    "tj_set_tvm": (C.c_int, [_P, C.c_int, C.c_double, C.c_double, C.c_double]),
    "tj_set_bayesian": (C.c_int, [_P, C.c_int, C.c_double, C.c_double]),
    "tj_set_bgart": (C.c_int, [_P, C.c_int, C.c_double, C.c_int, C.c_int, C.c_double]),
    "tj_available_indexes": (C.c_int, [C.c_int, C.c_int, _P]),
    "tj_order_shuffle": (None, [_P, C.c_int, C.c_int, _P]),
    "tj_threshold_projections": (C.c_int, [_P, C.c_double, C.c_double, C.c_int]),
    "tj_max_volume_bytes": (C.c_int, [C.c_int, C.POINTER(C.c_size_t)]),
    "tj_fsc": (C.c_int, [_P, _P, _P]),
    "tj_ssnr": (C.c_int, [_P, _P, C.c_int, C.c_double, _P, _P, _P]),
    "tj_fsc_volumes": (C.c_int, [C.c_int, _P, _P, C.c_int, C.c_int, C.c_int, _P]),
    "tj_set_available_indexes": (C.c_int, [_P, _P, C.c_int]),
    "tj_tv_denoise": (C.c_int, [_P, C.c_float, C.c_float, C.c_float, C.c_int]),
    "tj_get_completion": (C.c_double, [_P]),
    "tj_interrupt": (C.c_int, [_P]),
    "tj_run_ossart": (C.c_int, [_P, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, _P]),
    "tj_run_wbp": (C.c_int, [_P, C.c_int, C.c_int]),
    "tj_last_error_sum": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "tj_phantom_ellipsoids": (None, [C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_int, _P]),
    "tj_synth_projections": (C.c_int, [_P, _P, C.c_int, _P, C.c_int, C.c_double, C.c_uint64]),
    "tj_get_loaded_projection": (C.c_int, [_P, C.c_int, _P]),
    "tj_normalize_projections": (C.c_int, [_P, _P, C.c_int, _P]),
    "tj_align_projections": (C.c_int, [_P, _P, _P, C.c_int, C.c_int, C.c_int, _P, _P]),
    "tj_mrc_read_info": (C.c_int, [C.c_char_p, _P]),
    "tj_mrc_read": (C.c_int, [C.c_char_p, _P, _P]),
    "tj_mrc_write": (C.c_int, [C.c_char_p, _P, C.c_int, C.c_int, C.c_int]),
    "tj_get_timing": (C.c_int, [_P, C.POINTER(tj_timing)]),
    "tj_enable_timing": (C.c_int, [_P, C.c_int]),
    "tj_set_tuning": (C.c_int, [_P, C.POINTER(tj_tuning)]),
    "tj_get_tuning": (C.c_int, [_P, C.POINTER(tj_tuning)]),
    "tj_job_bytes": (C.c_size_t, [C.POINTER(tj_job_desc)]),
    "tj_batch_open": (C.c_int, [C.c_int, C.c_size_t, C.POINTER(_P)]),
    "tj_batch_close": (C.c_int, [_P]),
    "tj_batch_used": (C.c_size_t, [_P]),
    "tj_batch_budget": (C.c_size_t, [_P]),
    "tj_batch_add_job": (C.c_int, [_P, C.POINTER(tj_job_desc), _P, C.POINTER(C.c_int)]),
    "tj_batch_remove_job": (C.c_int, [_P, C.c_int]),
    "tj_batch_set_projections": (C.c_int, [_P, C.c_int, _P]),
    "tj_batch_synth_job": (C.c_int, [_P, C.c_int, _P, C.c_int, C.c_double, C.c_uint64]),
    "tj_batch_get_projection": (C.c_int, [_P, C.c_int, C.c_int, _P]),
    "tj_batch_get_difference": (C.c_int, [_P, C.c_int, C.c_int, _P]),
    "tj_batch_set_volume": (C.c_int, [_P, C.c_int, _P]),
    "tj_batch_get_volume": (C.c_int, [_P, C.c_int, _P]),
    "tj_batch_get_volume_f64": (C.c_int, [_P, C.c_int, _P]),
    "tj_batch_reset_volume": (C.c_int, [_P, C.c_int]),
    "tj_batch_convert": (C.c_int, [_P, C.c_int, C.c_int]),
    "tj_batch_run": (C.c_int, [_P, _P, C.c_int, C.c_int, C.c_int]),
    "tj_batch_last_run_ms": (C.c_double, [_P]),
    "tj_batch_errors": (C.c_int, [_P, C.c_int, _P, C.c_int]),
    "tj_batch_set_tuning": (C.c_int, [_P, C.POINTER(tj_batch_tuning)]),
    "tj_batch_get_tuning": (C.c_int, [_P, C.POINTER(tj_batch_tuning)]),
    "tj_batch_autotune": (C.c_int, [_P, _P, C.c_int, C.POINTER(tj_batch_tuning), C.POINTER(C.c_double)]),
}

EXPORTED = tuple(_SIGS)

_lib = None


def load() -> C.CDLL:
    """Loads libtomoj_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m tomoj_amd.build` "
                "(hipcc, gfx950). There is no CPU fallback.")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def last_error() -> str:
    return load().tj_last_error().decode(errors="replace")


def device_info(device: int = 0) -> dict:
    import json
    buf = C.create_string_buffer(1024)
    call("tj_device_info", int(device), buf, len(buf))
    return json.loads(buf.value.decode())


def copy_probe(bytes_: int, repeats: int = 5, device: int = 0) -> float:
    """Milliseconds per device-to-device copy of `bytes_` with a plain kernel."""
    ms = C.c_double()
    call("tj_copy_probe", int(device), C.c_size_t(int(bytes_)), int(repeats), C.byref(ms))
    return ms.value


def call(name: str, *args) -> int:
    """Calls a status-returning entry point and raises TjError on failure."""
    rc = getattr(load(), name)(*args)
    if rc < 0:
        cls = TjCancelled if rc == TJ_ECANCELLED else TjError
        raise cls(rc, f"{name}: {last_error()}")
    return rc
