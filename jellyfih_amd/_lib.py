"""ctypes binding of the engine's C ABI (include/jellymx.h).

Loads ``libjellymx.so`` from this package directory. There is no fallback: if the
library is missing or fails to load, importing the engine raises, so a run can never
silently measure or check something other than the HIP path.
"""
from __future__ import annotations

import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libjellymx.so")

JM_ABI_VERSION = 4
JM_NOALLOC_SCRATCH_BYTES = 32 << 20
JM_DIFF_MIXED = 3264   # jm_diff_create: fp32 buffers, fp64 arithmetic

# jm_status
JM_OK, JM_ERR_ARG, JM_ERR_HIP, JM_ERR_STATE, JM_ERR_NOMEM = range(5)
# jm_path
JM_PATH_FUSED, JM_PATH_PLAIN = 0, 1
# jm_field
(JM_FIELD_X, JM_FIELD_V, JM_FIELD_C, JM_FIELD_F, JM_FIELD_MATERIAL, JM_FIELD_JP, JM_FIELD_FIBER,
 JM_FIELD_GLOW, JM_FIELD_GRID_V, JM_FIELD_GRID_M, JM_FIELD_FITNESS) = range(11)
# jm_param
(JM_PARAM_ACTUATION, JM_PARAM_CONTRACTION, JM_PARAM_FREQ, JM_PARAM_PAYLOAD_DENSITY,
 JM_PARAM_E_JELLY) = range(5)
# jm_stats_mode
JM_STATS_PAYLOAD, JM_STATS_BODY, JM_STATS_ANY = range(3)

# Every symbol include/jellymx.h declares.
EXPORTS = (
    "jm_default_config", "jm_create", "jm_destroy", "jm_set_tuning", "jm_get_tuning",
    "jm_load_particles", "jm_set_instance_param", "jm_get_instance_param", "jm_set_time",
    "jm_get_time", "jm_substep", "jm_compute_stats", "jm_run_batch_headless", "jm_read_field",
    "jm_write_field", "jm_live_count", "jm_read_device_order", "jm_sort_now", "jm_timed_block", "jm_sync",
    "jm_stream", "jm_last_error", "jm_water_mask", "jm_polygon_self_intersects",
    "jm_points_in_polygon_grid", "jm_generate_phenotype", "jm_fill_tank",
    "jm_diff_create", "jm_diff_destroy", "jm_diff_set_water_start",
    "jm_fwd_block", "jm_bwd_block", "jm_fwd_block_noalloc",
    "jm_render_clear", "jm_render_abyss", "jm_render_flat", "jm_render_tone_map",
    "jm_render_vorticity_grid", "jm_render_vorticity", "jm_read_frame", "jm_read_vorticity",
    "jm_read_frame_u8",
)


class JmConfig(ctypes.Structure):
    _fields_ = [
        ("abi_version", ctypes.c_int32),
        ("n_instances", ctypes.c_int32),
        ("n_particles", ctypes.c_int32),
        ("quality", ctypes.c_int32),
        ("n_grid_y", ctypes.c_int32),
        ("track_glow", ctypes.c_int32),
        ("path", ctypes.c_int32),
        ("axisym", ctypes.c_int32),
        ("domain_height", ctypes.c_double),
        ("nan_invalid", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class JmTuning(ctypes.Structure):
    _fields_ = [
        ("block_substeps", ctypes.c_int32),
        ("sort_interval", ctypes.c_int32),
        ("threads", ctypes.c_int32),
        ("ppt", ctypes.c_int32),
        ("use_graphs", ctypes.c_int32),
        ("xcd_affinity", ctypes.c_int32),
        ("lazy_vc", ctypes.c_int32),
        ("sort_strip", ctypes.c_int32),
        ("instance_groups", ctypes.c_int32),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class JmDiffInputs(ctypes.Structure):
    _fields_ = [
        ("inst", ctypes.c_void_p),
        ("mat", ctypes.c_void_p),
        ("fiber", ctypes.c_void_p),
        ("amp_scale", ctypes.c_void_p),
        ("params", ctypes.c_void_p),
    ]


class JmError(RuntimeError):
    def __init__(self, func: str, code: int, msg: str):
        super().__init__(f"{func} failed (status {code}): {msg}")
        self.code = code


_lib = None


def load() -> ctypes.CDLL:
    """Load libjellymx.so (raises if it is missing: build it with jellyfih_amd.build)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing; build it with `python -m jellyfih_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    H = ctypes.c_void_p
    P = ctypes.POINTER
    f32p, i32p, f64p = P(ctypes.c_float), P(ctypes.c_int32), P(ctypes.c_double)
    sig = {
        "jm_default_config": [P(JmConfig)],
        "jm_create": [P(JmConfig), P(H)],
        "jm_destroy": [H],
        "jm_set_tuning": [H, P(JmTuning)],
        "jm_get_tuning": [H, P(JmTuning)],
        "jm_load_particles": [H, ctypes.c_int32, f32p, i32p, f32p, ctypes.c_int64],
        "jm_set_instance_param": [H, ctypes.c_int32, f32p],
        "jm_get_instance_param": [H, ctypes.c_int32, f32p],
        "jm_set_time": [H, ctypes.c_float],
        "jm_get_time": [H, f32p],
        "jm_substep": [H, ctypes.c_int64],
        "jm_compute_stats": [H, ctypes.c_int32, f32p],
        "jm_run_batch_headless": [H, ctypes.c_int64, f64p],
        "jm_read_field": [H, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64],
        "jm_write_field": [H, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64],
        "jm_live_count": [H, ctypes.c_int32, P(ctypes.c_int64)],
        "jm_read_device_order": [H, ctypes.c_int32, i32p, f32p, ctypes.c_int64],
        "jm_sort_now": [H],
        "jm_timed_block": [H, ctypes.c_int32, f32p, f32p],
        "jm_sync": [H],
        "jm_water_mask": [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64,
                          ctypes.c_double, ctypes.c_void_p],
        "jm_polygon_self_intersects": [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p],
        "jm_points_in_polygon_grid": [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                      ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64,
                                      ctypes.c_void_p],
        "jm_generate_phenotype": [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                  ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p],
        "jm_fill_tank": [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_int32,
                         ctypes.c_void_p, ctypes.c_double, ctypes.c_int32, ctypes.c_double,
                         ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                         ctypes.c_void_p],
        "jm_diff_create": [P(JmConfig), ctypes.c_int64, ctypes.c_int32, P(H)],
        "jm_diff_destroy": [H],
        "jm_diff_set_water_start": [H, ctypes.c_int64],
        "jm_fwd_block": [H, P(JmDiffInputs), ctypes.c_void_p, ctypes.c_double, ctypes.c_int32,
                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
        "jm_fwd_block_noalloc": [P(JmConfig), ctypes.c_int64, ctypes.c_int32, P(JmDiffInputs),
                                 ctypes.c_void_p, ctypes.c_double, ctypes.c_int32,
                                 ctypes.c_void_p, ctypes.c_void_p],
        "jm_bwd_block": [H, P(JmDiffInputs), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                         ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                         ctypes.c_void_p, ctypes.c_void_p],
        "jm_render_clear": [H, ctypes.c_int32, ctypes.c_float],
        "jm_render_abyss": [H, ctypes.c_int32, ctypes.c_int32, ctypes.c_float, f32p, f32p,
                            ctypes.c_int32],
        "jm_render_flat": [H, ctypes.c_int32, ctypes.c_int32, ctypes.c_float, ctypes.c_int32,
                           ctypes.c_float, ctypes.c_float, ctypes.c_float],
        "jm_render_tone_map": [H],
        "jm_render_vorticity_grid": [H, ctypes.c_float],
        "jm_render_vorticity": [H, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                ctypes.c_float],
        "jm_read_frame": [H, f32p, ctypes.c_int64],
        "jm_read_vorticity": [H, f32p, ctypes.c_int64],
        "jm_read_frame_u8": [H, ctypes.POINTER(ctypes.c_uint8), ctypes.c_int64],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    L.jm_stream.argtypes = [H]
    L.jm_stream.restype = ctypes.c_void_p
    L.jm_last_error.argtypes = []
    L.jm_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


def check(func: str, rc: int) -> None:
    if rc != JM_OK:
        msg = load().jm_last_error().decode(errors="replace")
        raise JmError(func, rc, msg)


def call(func: str, *args) -> None:
    check(func, getattr(load(), func)(*args))
