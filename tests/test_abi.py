"""The C ABI library: loads, exports every symbol include/jellymx.h declares, validates its
arguments before touching a device (no compute calls here: this runs without a GPU)."""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np
import pytest

from jellyfih_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "jellymx.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void\*|const char\*)\s+(jm_\w+)\s*\(", text,
                                 re.M)))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    syms = header_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert sorted(_lib.EXPORTS) == syms


def _header_fields(struct):
    """Field names of ``typedef struct <struct> {...}`` in include/jellymx.h."""
    import re
    text = open(os.path.join(ROOT, "include", "jellymx.h")).read()
    body = text[text.index(f"typedef struct {struct} {{"):]
    body = body[body.index("{") + 1:body.index("}")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    return [m.group(1) for m in re.finditer(r"\b(\w+)\s*;", body)]


def test_struct_layouts_match_header():
    assert [f for f, _ in _lib.JmConfig._fields_] == _header_fields("jm_config")
    assert [f for f, _ in _lib.JmTuning._fields_] == _header_fields("jm_tuning")
    assert [f for f, _ in _lib.JmDiffInputs._fields_] == _header_fields("jm_diff_inputs")
    assert ctypes.sizeof(_lib.JmConfig) == 48
    assert _lib.JmConfig.domain_height.offset == 32
    assert _lib.JmConfig.nan_invalid.offset == 40
    assert ctypes.sizeof(_lib.JmTuning) == 36


def test_default_config_is_the_reference_default():
    cfg = _lib.JmConfig()
    _lib.call("jm_default_config", ctypes.byref(cfg))
    assert cfg.abi_version == _lib.JM_ABI_VERSION
    assert (cfg.n_instances, cfg.n_particles, cfg.quality) == (16, 80000, 1)   # mpm_sim.py:11-13
    assert cfg.n_grid_y == 0 and cfg.domain_height == 1.0 and cfg.path == _lib.JM_PATH_FUSED
    assert cfg.nan_invalid == 0 and cfg.reserved == 0  # the reference's validity gate


@pytest.mark.parametrize("field,value", [("abi_version", 99), ("n_instances", 0),
                                         ("n_particles", -5), ("quality", 0), ("path", 7),
                                         ("n_grid_y", 6), ("domain_height", 0.0),
                                         ("axisym", 2), ("nan_invalid", 3),
                                         ("reserved", 1)])
def test_create_rejects_bad_config_before_device_work(field, value):
    cfg = _lib.JmConfig()
    _lib.call("jm_default_config", ctypes.byref(cfg))
    setattr(cfg, field, value)
    h = ctypes.c_void_p()
    rc = _lib.load().jm_create(ctypes.byref(cfg), ctypes.byref(h))
    assert rc == _lib.JM_ERR_ARG
    assert not h.value
    assert _lib.load().jm_last_error()


def test_null_handle_is_rejected():
    L = _lib.load()
    assert L.jm_substep(None, 1) == _lib.JM_ERR_ARG
    assert L.jm_sync(None) == _lib.JM_ERR_ARG
    assert L.jm_stream(None) is None
    with pytest.raises(_lib.JmError):
        _lib.call("jm_set_time", None, ctypes.c_float(0.0))


def test_renderer_entry_points_reject_a_null_handle():
    L = _lib.load()
    f = (ctypes.c_float * 4)()
    assert L.jm_render_clear(None, 64, 0.0) == _lib.JM_ERR_ARG
    assert L.jm_render_abyss(None, 32, 2, 1.0, f, f, 0) == _lib.JM_ERR_ARG
    assert L.jm_render_flat(None, 32, 2, 1.0, 1, 0.1, 0.2, 0.3) == _lib.JM_ERR_ARG
    assert L.jm_render_tone_map(None) == _lib.JM_ERR_ARG
    assert L.jm_render_vorticity_grid(None, 1.0) == _lib.JM_ERR_ARG
    assert L.jm_render_vorticity(None, 1, 32, 2, 0.5) == _lib.JM_ERR_ARG
    assert L.jm_read_frame(None, f, 4) == _lib.JM_ERR_ARG
    assert L.jm_read_vorticity(None, f, 4) == _lib.JM_ERR_ARG
    b = (ctypes.c_uint8 * 4)()
    assert L.jm_read_frame_u8(None, b, 4) == _lib.JM_ERR_ARG


def test_engine_fails_loudly_without_a_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from jellyfih_amd.engine import Engine
    with pytest.raises(_lib.JmError):
        Engine(2, 100)


def test_axisym_refused_by_the_differentiable_path():
    """The differentiable path has no axisymmetric variant: jm_diff_create refuses it before
    any device work (the engine runs it on both of its paths)."""
    import ctypes
    cfg = _lib.JmConfig()
    _lib.call("jm_default_config", ctypes.byref(cfg))
    cfg.axisym = 1
    h = ctypes.c_void_p()
    with pytest.raises(_lib.JmError, match="axisymmetric"):
        _lib.call("jm_diff_create", ctypes.byref(cfg), ctypes.c_int64(100), 32, ctypes.byref(h))


def test_facade_constants_match_reference():
    from jellyfih_amd import mpm_sim as s
    assert s.n_grid == 128 and s.dt == 2e-5 and s.inv_dx == 128.0
    assert s.p_vol == (1 / 256) ** 2 and s.p_mass == s.p_vol
    assert s.mu_jelly == pytest.approx(269.2307692307692)
    assert s.lambda_jelly == pytest.approx(403.8461538461538)
    assert s.mu_payload == pytest.approx(83333.33333333333)
    assert s.lambda_payload == pytest.approx(55555.555555555555)
    assert s.mu_0_base == pytest.approx(416.6666666666667)
    assert s.lambda_0_base == pytest.approx(277.77777777777777)
    assert s.grid_side == 4 and s.res_sub == 256
    assert s.frame_buffer.shape == (1024, 1024, 3)
    for name in ("clear_frame_buffer", "clear_frame_buffer_white", "render_frame_abyss",
                 "render_flat_pass", "tone_map_and_encode", "compute_vorticity_grid",
                 "render_vorticity_overlay", "render_vorticity_rdbu",
                 "render_vorticity_hueshift", "render_vorticity_white"):
        assert callable(getattr(s, name))


def test_facade_ti_sync_without_an_engine():
    """run_population.py:37,47 call sim.ti.sync(); before the engine exists there is
    nothing queued, and it must not create one (no device work on import paths)."""
    from jellyfih_amd import mpm_sim as s
    had = s._engine
    s.ti.sync()
    assert s._engine is had
    assert s.nan_invalid is False  # the reference's validity gate unless JELLY_NAN_INVALID=1


def test_facade_uses_the_autotuned_parameters_for_their_batch_shape(tmp_path):
    """The facade takes the tuning file's entry for its own batch shape only; the file holds
    one entry per shape (autotune.write_entry replaces the same shape's entry), and the
    single-entry form it had before still reads."""
    import json
    from jellyfih_amd import autotune
    from jellyfih_amd import mpm_sim as s
    t = s._tuning_for(32, 80000)
    assert t is not None and t["threads"] * t["ppt"] > 0 and t["sort_interval"] >= 1
    assert s._tuning_for(32, 160000) is None
    assert s._tuning_for(32, 80000, path=str(tmp_path / "missing.json")) is None
    path = str(tmp_path / "tuning.json")
    entry = lambda m, n, ppt: {"tuning": {"threads": 256, "ppt": ppt}, "rate": 1.0,
                               "workload": {"instances": m, "slots": n, "live": 1}}
    with open(path, "w") as f:
        json.dump(entry(32, 80000, 4), f)                      # the single-entry form
    assert s._tuning_for(32, 80000, path=path)["ppt"] == 4
    autotune.write_entry(entry(16, 80000, 2), path)
    autotune.write_entry(entry(32, 80000, 8), path)             # replaces the 32 x 80K entry
    assert len(autotune.load_entries(path)) == 2
    assert s._tuning_for(16, 80000, path=path)["ppt"] == 2
    assert s._tuning_for(32, 80000, path=path)["ppt"] == 8
    assert s._tuning_for(8, 80000, path=path) is None


def test_facade_queues_substeps_until_the_next_access(monkeypatch):
    """substep() queues (a Taichi launch is asynchronous too): queued substeps go to the
    engine in blocks of the engine's block length, and all of them before any field read,
    parameter write, render or sync, in call order. Host logic, with a stand-in engine."""
    from jellyfih_amd import mpm_sim as s

    class Recorder:
        def __init__(self):
            self.calls = []

        def tuning(self):
            return {"block_substeps": 100}

        def substep(self, n):
            self.calls.append(("substep", n))

        def read(self, name):
            self.calls.append(("read", name))
            return np.zeros(1)

        def get_param(self, name):
            self.calls.append(("get_param", name))
            return np.zeros(s.n_instances, np.float32)

        def set_param(self, name, values):
            self.calls.append(("set_param", name))

        def get_time(self):
            self.calls.append(("get_time",))
            return 0.0

        def sync(self):
            self.calls.append(("sync",))

    rec = Recorder()
    monkeypatch.setattr(s, "_engine", rec)
    monkeypatch.setattr(s, "_pending", 0)
    monkeypatch.setattr(s, "_pending_max", 100)
    for _ in range(250):
        s.substep()
    assert rec.calls == [("substep", 100), ("substep", 100)]
    s.x.to_numpy()
    assert rec.calls[2:] == [("substep", 50), ("read", "x")]
    for _ in range(3):
        s.substep()
    s.instance_actuation[0] = 1.0
    s.step(7)
    s.sim_time[None]
    s.ti.sync()
    s.ti.sync()
    assert rec.calls[4:] == [("substep", 3), ("get_param", "actuation"),
                             ("set_param", "actuation"), ("substep", 7), ("get_time",),
                             ("sync",), ("sync",)], rec.calls
    assert s._pending == 0


def test_water_start_rejects_a_null_handle():
    """jm_diff_set_water_start checks its handle before anything else (its range check
    needs a handle, which needs a device: tests/test_gpu_diff.py)."""
    L = _lib.load()
    assert L.jm_diff_set_water_start(None, 0) == _lib.JM_ERR_ARG
