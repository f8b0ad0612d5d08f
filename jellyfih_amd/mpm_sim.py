"""Drop-in for the simulation core of the reference's ``mpm_sim`` module.

c-m-x-r/jellyfih drives its Taichi MLS-MPM solver through module-level constants, fields
with item access and a few functions (``mpm_sim.py:11-149,153-340,715-834``), called from
``evolve.py:127-155,158-245,474-478`` and ``run_population.py:33-57``. This module keeps
those names and their meaning, backed by the HIP engine (``libjellymx.so``):

* configuration is read from the same environment variables at import time
  (``JELLY_INSTANCES``, ``JELLY_PARTICLES``, ``JELLY_GRID_Y``, ``JELLY_DOMAIN_H``);
* fields (``x``, ``v``, ``C``, ``F``, ``material``, ``Jp``, ``fiber_dir``, ``grid_v``,
  ``grid_m``, ``fitness_buffer``, ``sim_time``, ``instance_*``) support ``to_numpy()``,
  ``from_numpy()`` and item get/set like Taichi fields;
* ``substep()``, ``load_particles()``, ``get_payload_stats()``, ``run_batch_headless()``
  and the fitness-buffer kernels behave as in the reference.

``substep()`` queues its substep, as a Taichi kernel launch does. Queued substeps run as one
block (HIP-graph replay, one fused launch per substep) once a block's worth has
accumulated, and before anything else reaches the engine: a field read or write, a render,
a stats call, ``sync()``/``ti.sync()``. A caller that steps in a Python loop
(``evolve.py:389-390``, ``run_population.py:44-52``) therefore gets the block rate rather
than two launches per call, and sees the same state at every read as with one launch per
call.

The device engine is created on first use, so importing this module never touches a GPU.
The renderer (``mpm_sim.py:343-712``) draws into a device frame buffer read back with
``frame_buffer.to_numpy()``. ``JELLY_AXISYM=1`` selects the axisymmetric variant
(``mpm_sim.py:127-133,163-165,236-304``), which runs on the engine's fused path.
"""
from __future__ import annotations

import math
import os

import numpy as np

# --- configuration (mpm_sim.py:11-58) ----------------------------------------------------
n_instances = int(os.environ.get("JELLY_INSTANCES", "16"))
quality = 1
n_particles = int(os.environ.get("JELLY_PARTICLES", "80000"))
n_grid = 128 * quality
n_grid_y = int(os.environ.get("JELLY_GRID_Y", str(n_grid)))
domain_height = float(os.environ.get("JELLY_DOMAIN_H", "1.0"))
axisym = os.environ.get("JELLY_AXISYM", "0") == "1"
# Opt-in: run_batch_headless marks an instance whose CoM is not finite invalid. Off by
# default, where the reference's `final_y < 0.01` gate lets a NaN CoM count as valid
# (mpm_sim.py:830-833).
nan_invalid = os.environ.get("JELLY_NAN_INVALID", "0") == "1"
dx, inv_dx = 1 / n_grid, float(n_grid)
dt = 2e-5 / quality
p_vol = (dx * 0.5) ** 2
p_rho = 1
p_mass = p_vol * p_rho

E_base = 0.1e4
nu_base = 0.2
mu_0_base = E_base / (2 * (1 + nu_base))
lambda_0_base = E_base * nu_base / ((1 + nu_base) * (1 - 2 * nu_base))
E_jelly = 0.7e3
nu_jelly = 0.3
mu_jelly = E_jelly / (2 * (1 + nu_jelly))
lambda_jelly = E_jelly * nu_jelly / ((1 + nu_jelly) * (1 - 2 * nu_jelly))
E_payload = 2.0e5
nu_payload = 0.2
mu_payload = E_payload / (2 * (1 + nu_payload))
lambda_payload = E_payload * nu_payload / ((1 + nu_payload) * (1 - 2 * nu_payload))
water_lambda = 100000.0
gravity = 10.0
actuation_freq = 1
actuation_strength = 500.0
ACT_CONTRACTION_END = 0.2
ACT_RELAXATION_END = 0.6

WEB_PALETTE = [
    (0, 0.933, 0.949, 0.957),
    (1, 0.678, 0.882, 0.875),
    (3, 0.976, 0.757, 0.765),
    (2, 0.961, 0.824, 0.576),
]
video_res = 1024
grid_side = int(math.ceil(math.sqrt(n_instances)))
res_sub = video_res // grid_side

_engine = None
_pending = 0        # substep() calls queued and not yet sent to the engine
_pending_max = 100  # the engine's block length, read when the engine is created


def _tuning_for(instances: int, slots: int, path: str | None = None):
    """The autotuner's launch parameters (``tuning.json`` beside this module, written by
    ``python -m jellyfih_amd.autotune --write``) when they were measured on this batch
    shape, else None (the engine then picks its launch shape by batch size)."""
    from .autotune import tuning_for
    return tuning_for(instances, slots, path)


def engine():
    """The device engine behind this module (created on first use), with the autotuner's
    launch parameters when they were measured on this batch shape. Queued substeps run
    first, so whatever the caller does next sees them."""
    e = _device()
    _flush(e)
    return e


def _flush(e) -> None:
    global _pending
    if _pending:
        n, _pending = _pending, 0
        e.substep(n)


def _device():
    global _engine, _pending_max
    if _engine is None:
        from .engine import Engine
        _engine = Engine(n_instances=n_instances, n_particles=n_particles, quality=quality,
                         n_grid_y=n_grid_y, domain_height=domain_height,
                         track_glow=os.environ.get("JELLY_GLOW", "0") == "1",
                         path="fused", axisym=axisym, nan_invalid=nan_invalid)
        tuned = None if (axisym or n_grid_y != n_grid) else _tuning_for(n_instances,
                                                                         n_particles)
        if tuned:
            _engine.set_tuning(**{k: v for k, v in tuned.items() if k in _engine.tuning()})
        _pending_max = max(1, int(_engine.tuning().get("block_substeps", 100)))
    return _engine


# --- fields --------------------------------------------------------------------------------
class _Field:
    """Taichi-like field over an engine array: to_numpy / from_numpy / item access."""

    def __init__(self, name: str):
        self._name = name

    @property
    def shape(self):
        from .engine import FIELD_IDS, FIELD_SHAPES
        fid = FIELD_IDS[self._name]
        full = engine().field_shape(fid)
        return full[:len(full) - len(FIELD_SHAPES[fid][1])]

    def to_numpy(self) -> np.ndarray:
        return engine().read(self._name)

    def from_numpy(self, arr) -> None:
        engine().write(self._name, arr)

    def __getitem__(self, idx):
        return self.to_numpy()[idx]

    def __setitem__(self, idx, value) -> None:
        arr = self.to_numpy()
        arr[idx] = value
        self.from_numpy(arr)


class _InstanceField:
    """Per-instance scalar field (``instance_actuation[i] = ...``)."""

    def __init__(self, param: str):
        self._param = param

    @property
    def shape(self):
        return (n_instances,)

    def to_numpy(self) -> np.ndarray:
        return engine().get_param(self._param)

    def from_numpy(self, arr) -> None:
        engine().set_param(self._param, arr)

    def __getitem__(self, i):
        return float(self.to_numpy()[i])

    def __setitem__(self, i, value) -> None:
        vals = self.to_numpy()
        vals[i] = value
        self.from_numpy(vals)


class _TimeField:
    """``sim_time[None]``: the fp32 simulation clock (mpm_sim.py:95,340)."""

    shape = ()

    def __getitem__(self, idx):
        return engine().get_time()

    def __setitem__(self, idx, value) -> None:
        engine().set_time(float(value))

    def to_numpy(self) -> np.ndarray:
        return np.float32(engine().get_time())


class _HostField:
    """Render-only per-instance settings kept on the host (mpm_sim.py:136-138,102)."""

    def __init__(self, values):
        self._v = np.array(values, dtype=np.float64)

    @property
    def shape(self):
        return self._v.shape

    def to_numpy(self):
        return self._v.copy()

    def from_numpy(self, arr):
        self._v[...] = arr

    def __getitem__(self, i):
        return self._v[() if i is None else i]

    def __setitem__(self, i, value):
        self._v[() if i is None else i] = value


class _FitnessBuffer:
    """fitness_buffer (M,3) = [sum_y, count, sum_x] accumulated by the stats kernels."""

    def __init__(self):
        self._buf = np.zeros((n_instances, 3), np.float32)

    shape = (n_instances, 3)

    def to_numpy(self) -> np.ndarray:
        return self._buf.copy()

    def __getitem__(self, idx):
        return self._buf[idx]


x = _Field("x")
v = _Field("v")
C = _Field("C")
F = _Field("F")
material = _Field("material")
Jp = _Field("Jp")
fiber_dir = _Field("fiber_dir")
visual_glow = _Field("visual_glow")
grid_v = _Field("grid_v")
grid_m = _Field("grid_m")
sim_time = _TimeField()
fitness_buffer = _FitnessBuffer()
instance_actuation = _InstanceField("actuation")
instance_act_contraction = _InstanceField("contraction")
instance_freq = _InstanceField("freq")
instance_payload_density = _InstanceField("payload_density")
instance_e_jelly = _InstanceField("e_jelly")
instance_hue = _HostField([0.55] * n_instances)
instance_muscle_hue = _HostField([(0.55 + 0.15) % 1.0] * n_instances)
water_angle_color = _HostField(0)


# --- kernels -------------------------------------------------------------------------------
def substep() -> None:
    """One MLS-MPM substep of every instance (mpm_sim.py:153-340), queued: it runs with the
    next block of queued substeps, at the latest before the next access to the engine."""
    global _pending
    e = _device()
    _pending += 1
    if _pending >= _pending_max:
        _flush(e)


def step(n: int) -> None:
    """``n`` substeps in one call (what ``for _ in range(n): substep()`` does)."""
    engine().substep(int(n))


def load_particles(instance, pos_np, mat_np, fiber_np=None) -> None:
    """Reset one instance with new particles (mpm_sim.py:730-736); fibres default to (0,1)."""
    engine().load_particles(int(instance), pos_np, mat_np, fiber_np)


def clear_fitness_buffer() -> None:
    """mpm_sim.py:740-744"""
    fitness_buffer._buf[...] = 0


def compute_payload_stats() -> None:
    """mpm_sim.py:746-753 (adds into fitness_buffer)."""
    fitness_buffer._buf += engine().raw_stats("payload")


def compute_body_stats() -> None:
    """mpm_sim.py:755-762 (adds into fitness_buffer)."""
    fitness_buffer._buf += engine().raw_stats("body")


def _finish(raw: np.ndarray) -> np.ndarray:
    stats = np.zeros((raw.shape[0], 3))
    for i in range(raw.shape[0]):
        count = raw[i, 1]
        if count > 0:
            stats[i, 0] = raw[i, 0] / count
            stats[i, 1] = raw[i, 2] / count
            stats[i, 2] = count
    return stats


def get_payload_stats() -> np.ndarray:
    """Payload CoM per instance: (M,3) = [com_y, com_x, count] (mpm_sim.py:764-777)."""
    clear_fitness_buffer()
    compute_payload_stats()
    return _finish(fitness_buffer.to_numpy())


def _get_stats_any() -> np.ndarray:
    """Payload CoM, or body CoM where an instance has no payload (mpm_sim.py:779-800)."""
    raw = engine().raw_stats("any")
    fitness_buffer._buf[...] = raw
    return _finish(raw)


def run_batch_headless(steps: int) -> np.ndarray:
    """(M,5) = [init_y, init_x, final_y, final_x, valid] (mpm_sim.py:802-834)."""
    return engine().run_batch_headless(int(steps))


def sync() -> None:
    """ti.sync()"""
    engine().sync()


class _TaichiNamespace:
    """The ``ti`` name callers reach through this module: ``sim.ti.sync()``
    (run_population.py:37,47; evolve.py's viewers) waits for all queued device work."""

    @staticmethod
    def sync() -> None:
        if _engine is not None:  # nothing can be queued before the engine exists
            engine().sync()


ti = _TaichiNamespace()


# --- renderer (mpm_sim.py:343-712) -----------------------------------------------------------
# The frame buffer (video_res x video_res x 3 fp32, indexed [py, px, c]) and the vorticity
# grid live on the device; they are allocated on first use.
_frame_ready = False


def _frame() -> "Engine":
    global _frame_ready
    e = engine()
    if not _frame_ready:
        e.render_clear(video_res, 0.0)
        _frame_ready = True
    return e


def clear_frame_buffer() -> None:
    """mpm_sim.py:361-364: black."""
    global _frame_ready
    engine().render_clear(video_res, 0.0)
    _frame_ready = True


def clear_frame_buffer_white() -> None:
    """mpm_sim.py:483-486: 0.98 (#FAFAFA)."""
    global _frame_ready
    engine().render_clear(video_res, 0.98)
    _frame_ready = True


def render_frame_abyss(p_res_sub: int, p_grid_side: int, radius: float) -> None:
    """Additive glow splat of every live particle (mpm_sim.py:385-477), coloured by material
    with instance_hue, instance_muscle_hue and water_angle_color."""
    _frame().render_abyss(p_res_sub, p_grid_side, radius, instance_hue.to_numpy(),
                          instance_muscle_hue.to_numpy(), int(water_angle_color[None]))


def render_flat_pass(p_res_sub: int, p_grid_side: int, radius: float, target_mat: int,
                     cr: float, cg: float, cb: float) -> None:
    """Flat-colour discs of one material (mpm_sim.py:487-517)."""
    _frame().render_flat(p_res_sub, p_grid_side, radius, target_mat, cr, cg, cb)


def tone_map_and_encode() -> None:
    """Reinhard tone map and gamma of the frame buffer in place (mpm_sim.py:366-383)."""
    _frame().render_tone_map()


def render_vorticity_overlay(p_res_sub: int, p_grid_side: int, vort_scale: float) -> None:
    """Additive warm/cool squares from the grid curl (mpm_sim.py:520-558)."""
    _frame().render_vorticity(0, p_res_sub, p_grid_side, vort_scale)


def compute_vorticity_grid(vort_scale: float) -> None:
    """vort_grid = tanh(curl * vort_scale) (mpm_sim.py:560-578)."""
    engine().render_vorticity_grid(vort_scale)


def render_vorticity_rdbu(p_res_sub: int, p_grid_side: int, alpha_max: float) -> None:
    """mpm_sim.py:581-624"""
    _frame().render_vorticity(1, p_res_sub, p_grid_side, alpha_max)


def render_vorticity_hueshift(p_res_sub: int, p_grid_side: int, strength: float) -> None:
    """mpm_sim.py:627-668"""
    _frame().render_vorticity(2, p_res_sub, p_grid_side, strength)


def render_vorticity_white(p_res_sub: int, p_grid_side: int, alpha_max: float) -> None:
    """mpm_sim.py:671-712 (fills the frame buffer with white first)."""
    _frame().render_vorticity(3, p_res_sub, p_grid_side, alpha_max)


class _FrameBuffer:
    """frame_buffer: (video_res, video_res, 3) fp32, read back with to_numpy() (or
    to_uint8(), already encoded to 8 bits)."""

    shape = (video_res, video_res, 3)

    def to_numpy(self) -> np.ndarray:
        return _frame().read_frame()

    def to_uint8(self) -> np.ndarray:
        """(np.clip(frame_buffer.to_numpy(), 0, 1) * 255).astype(np.uint8), the viewers'
        encode (evolve.py:396-397), done on the device: a quarter of the bytes cross."""
        return _frame().read_frame_u8()


class _VortGrid:
    """vort_grid: (n_instances, n_grid, n_grid_y) fp32, read back with to_numpy()."""

    shape = (n_instances, n_grid, n_grid_y)

    def to_numpy(self) -> np.ndarray:
        return engine().read_vorticity()


frame_buffer = _FrameBuffer()
vort_grid = _VortGrid()
