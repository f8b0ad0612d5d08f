"""Python front of the HIP engine: one :class:`Engine` per batch of instances.

Thin over the C ABI (include/jellymx.h): numpy arrays in the reference's layout go in
and come out; all state stays on the device between calls.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import JmConfig, JmTuning, call, load

FIELD_SHAPES = {
    _lib.JM_FIELD_X: ("p", (2,), np.float32),
    _lib.JM_FIELD_V: ("p", (2,), np.float32),
    _lib.JM_FIELD_C: ("p", (2, 2), np.float32),
    _lib.JM_FIELD_F: ("p", (2, 2), np.float32),
    _lib.JM_FIELD_MATERIAL: ("p", (), np.int32),
    _lib.JM_FIELD_JP: ("p", (), np.float32),
    _lib.JM_FIELD_FIBER: ("p", (2,), np.float32),
    _lib.JM_FIELD_GLOW: ("p", (), np.float32),
    _lib.JM_FIELD_GRID_V: ("g", (2,), np.float32),
    _lib.JM_FIELD_GRID_M: ("g", (), np.float32),
    _lib.JM_FIELD_FITNESS: ("f", (3,), np.float32),
}

FIELD_IDS = {
    "x": _lib.JM_FIELD_X, "v": _lib.JM_FIELD_V, "C": _lib.JM_FIELD_C, "F": _lib.JM_FIELD_F,
    "material": _lib.JM_FIELD_MATERIAL, "Jp": _lib.JM_FIELD_JP,
    "fiber_dir": _lib.JM_FIELD_FIBER, "visual_glow": _lib.JM_FIELD_GLOW,
    "grid_v": _lib.JM_FIELD_GRID_V, "grid_m": _lib.JM_FIELD_GRID_M,
    "fitness_buffer": _lib.JM_FIELD_FITNESS,
}

PARAM_IDS = {
    "actuation": _lib.JM_PARAM_ACTUATION,
    "contraction": _lib.JM_PARAM_CONTRACTION,
    "freq": _lib.JM_PARAM_FREQ,
    "payload_density": _lib.JM_PARAM_PAYLOAD_DENSITY,
    "e_jelly": _lib.JM_PARAM_E_JELLY,
}


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class Engine:
    """A batch of ``n_instances`` simulations of ``n_particles`` slots each."""

    def __init__(self, n_instances: int = 16, n_particles: int = 80000, quality: int = 1,
                 n_grid_y: int | None = None, domain_height: float = 1.0,
                 track_glow: bool = False, path: str = "fused", axisym: bool = False,
                 nan_invalid: bool = False):
        L = load()
        cfg = JmConfig()
        call("jm_default_config", ctypes.byref(cfg))
        cfg.n_instances = int(n_instances)
        cfg.n_particles = int(n_particles)
        cfg.quality = int(quality)
        cfg.n_grid_y = int(n_grid_y or 0)
        cfg.track_glow = int(bool(track_glow))
        cfg.path = {"fused": _lib.JM_PATH_FUSED, "plain": _lib.JM_PATH_PLAIN}[path]
        cfg.domain_height = float(domain_height)
        cfg.axisym = int(bool(axisym))
        cfg.nan_invalid = int(bool(nan_invalid))
        h = ctypes.c_void_p()
        call("jm_create", ctypes.byref(cfg), ctypes.byref(h))
        self._L = L
        self._h = h
        self.M = cfg.n_instances
        self.N = cfg.n_particles
        self.nx = 128 * cfg.quality
        self.ny = cfg.n_grid_y or self.nx
        self.quality = cfg.quality
        self.domain_height = cfg.domain_height
        self.path = path
        self.track_glow = bool(track_glow)
        self.axisym = bool(axisym)

    # -- lifecycle ------------------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            self._L.jm_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- configuration ----------------------------------------------------------------
    def tuning(self) -> dict:
        t = JmTuning()
        call("jm_get_tuning", self._h, ctypes.byref(t))
        return t.as_dict()

    def set_tuning(self, **kw) -> None:
        t = JmTuning()
        call("jm_get_tuning", self._h, ctypes.byref(t))
        for k, v in kw.items():
            if not hasattr(t, k):
                raise KeyError(k)
            setattr(t, k, int(v))
        call("jm_set_tuning", self._h, ctypes.byref(t))

    @property
    def stream(self) -> int:
        return int(self._L.jm_stream(self._h) or 0)

    # -- loading and parameters ---------------------------------------------------------
    def load_particles(self, instance: int, pos, mat, fiber=None) -> None:
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 2)
        mat = np.ascontiguousarray(mat, dtype=np.int32).reshape(-1)
        if pos.shape[0] != mat.shape[0]:
            raise ValueError("pos and mat lengths differ")
        fp = None
        if fiber is not None:
            fiber = np.ascontiguousarray(fiber, dtype=np.float32).reshape(-1, 2)
            if fiber.shape[0] != pos.shape[0]:
                raise ValueError("fiber length differs from pos")
            fp = _fptr(fiber)
        call("jm_load_particles", self._h, int(instance), _fptr(pos),
             mat.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), fp, pos.shape[0])

    def set_param(self, name: str, values) -> None:
        vals = np.ascontiguousarray(np.broadcast_to(np.asarray(values, np.float32), (self.M,)),
                                    dtype=np.float32)
        call("jm_set_instance_param", self._h, PARAM_IDS[name], _fptr(vals))

    def get_param(self, name: str) -> np.ndarray:
        out = np.empty(self.M, np.float32)
        call("jm_get_instance_param", self._h, PARAM_IDS[name], _fptr(out))
        return out

    def set_time(self, t: float) -> None:
        call("jm_set_time", self._h, ctypes.c_float(t))

    def get_time(self) -> float:
        out = ctypes.c_float()
        call("jm_get_time", self._h, ctypes.byref(out))
        return out.value

    # -- stepping -----------------------------------------------------------------------
    def substep(self, n: int = 1) -> None:
        call("jm_substep", self._h, int(n))

    def sync(self) -> None:
        call("jm_sync", self._h)

    def timed_block(self, n: int = 100) -> tuple[float, float]:
        """Run one n-substep block with HIP events around its fused kernels; returns
        (average fused-kernel ms, whole-block ms)."""
        a, b = ctypes.c_float(), ctypes.c_float()
        call("jm_timed_block", self._h, int(n), ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    # -- renderer (mpm_sim.py:343-712) -------------------------------------------------
    def render_clear(self, res: int = 1024, value: float = 0.0) -> None:
        """Fill the res x res x 3 frame buffer with ``value`` (clear_frame_buffer: 0,
        clear_frame_buffer_white: 0.98), allocating it on first use."""
        call("jm_render_clear", self._h, int(res), float(value))
        self._frame_res = int(res)  # only once the library has accepted res

    def render_abyss(self, res_sub: int, grid_side: int, radius: float, hue, muscle_hue,
                     water_mode: int = 0) -> None:
        h = np.ascontiguousarray(hue, np.float32)
        mh = np.ascontiguousarray(muscle_hue, np.float32)
        if h.shape != (self.M,) or mh.shape != (self.M,):
            raise ValueError(f"hue arrays must be ({self.M},)")
        call("jm_render_abyss", self._h, int(res_sub), int(grid_side), float(radius),
             _fptr(h), _fptr(mh), int(water_mode))

    def render_flat(self, res_sub: int, grid_side: int, radius: float, material: int,
                    r: float, g: float, b: float) -> None:
        call("jm_render_flat", self._h, int(res_sub), int(grid_side), float(radius),
             int(material), float(r), float(g), float(b))

    def render_tone_map(self) -> None:
        call("jm_render_tone_map", self._h)

    def render_vorticity_grid(self, scale: float) -> None:
        call("jm_render_vorticity_grid", self._h, float(scale))

    def render_vorticity(self, mode: int, res_sub: int, grid_side: int, param: float) -> None:
        """mode 0 overlay (param vort_scale), 1 RdBu, 2 hue shift, 3 white (param alpha)."""
        call("jm_render_vorticity", self._h, int(mode), int(res_sub), int(grid_side),
             float(param))

    def read_frame(self) -> np.ndarray:
        res = getattr(self, "_frame_res", 0)
        if not res:
            raise RuntimeError("no frame buffer: call render_clear first")
        out = np.empty((res, res, 3), np.float32)
        call("jm_read_frame", self._h, _fptr(out), out.size)
        return out

    def read_frame_u8(self) -> np.ndarray:
        """The frame encoded to uint8 on the device, as the viewers encode it on the host:
        (np.clip(frame, 0, 1) * 255).astype(np.uint8) (evolve.py:396-397)."""
        res = getattr(self, "_frame_res", 0)
        if not res:
            raise RuntimeError("no frame buffer: call render_clear first")
        out = np.empty((res, res, 3), np.uint8)
        call("jm_read_frame_u8", self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
             out.size)
        return out

    def read_vorticity(self) -> np.ndarray:
        out = np.empty((self.M, self.nx, self.ny), np.float32)
        call("jm_read_vorticity", self._h, _fptr(out), out.size)
        return out

    def sort_now(self) -> None:
        call("jm_sort_now", self._h)

    # -- outputs ------------------------------------------------------------------------
    def raw_stats(self, mode: str = "payload") -> np.ndarray:
        m = {"payload": _lib.JM_STATS_PAYLOAD, "body": _lib.JM_STATS_BODY,
             "any": _lib.JM_STATS_ANY}[mode]
        out = np.empty((self.M, 3), np.float32)
        call("jm_compute_stats", self._h, m, _fptr(out))
        return out

    def run_batch_headless(self, steps: int) -> np.ndarray:
        out = np.empty((self.M, 5), np.float64)
        call("jm_run_batch_headless", self._h, int(steps),
             out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return out

    def field_shape(self, field: int) -> tuple:
        kind, tail, _ = FIELD_SHAPES[field]
        if kind == "p":
            return (self.M, self.N) + tail
        if kind == "g":
            return (self.M, self.nx, self.ny) + tail
        return (self.M,) + tail

    def read(self, name: str) -> np.ndarray:
        fid = FIELD_IDS[name]
        out = np.empty(self.field_shape(fid), FIELD_SHAPES[fid][2])
        call("jm_read_field", self._h, fid, out.ctypes.data_as(ctypes.c_void_p), out.nbytes)
        return out

    def write(self, name: str, values) -> None:
        fid = FIELD_IDS[name]
        arr = np.ascontiguousarray(values, dtype=FIELD_SHAPES[fid][2])
        if arr.shape != self.field_shape(fid):
            raise ValueError(f"{name}: expected shape {self.field_shape(fid)}, got {arr.shape}")
        call("jm_write_field", self._h, fid, arr.ctypes.data_as(ctypes.c_void_p), arr.nbytes)

    def live_count(self, instance: int) -> int:
        n = ctypes.c_int64()
        call("jm_live_count", self._h, int(instance), ctypes.byref(n))
        return n.value

    def device_order(self, instance: int):
        """(slot ids, x) of an instance's live particles in the engine's current order."""
        n = self.live_count(instance)
        slot = np.empty(n, np.int32)
        x = np.empty((n, 2), np.float32)
        call("jm_read_device_order", self._h, int(instance),
             slot.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), _fptr(x), n)
        return slot, x
