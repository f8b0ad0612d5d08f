"""Parity of the HIP frame renderer against the renderer oracle (oracle/render.py, a numpy
restatement of mpm_sim.py:343-712) on the engine's own state after a few hundred substeps.

The oracle reads the state the renderer reads (x, v, Jp, material, the grid of the last
substep and the clock) back through the C ABI, so these tests compare the rendering alone.
Tolerances:

* additive splats (abyss, vorticity overlay): fp32 atomics add in an unspecified order, so
  |gpu - ref| <= 1e-4 |ref| + 1e-6 max|ref| per pixel and channel;
* flat passes, tone map inputs' coverage, pixel-to-cell maps: bit-exact (the kernels round
  every product and sum on its own, as the oracle does); the composites: 4 fp32 ulps; the
  tone map: 8 ulps (the device powf against the host libm's);
* vorticity grid: |gpu - ref| <= 1e-6 (tanh of identical curls).
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

from jellyfih_amd.engine import Engine
from oracle import render
from tests.cases import small_tank

pytestmark = pytest.mark.gpu

WEB_PALETTE = [(0, 0.933, 0.949, 0.957), (1, 0.678, 0.882, 0.875), (3, 0.976, 0.757, 0.765),
               (2, 0.961, 0.824, 0.576)]


def say(msg: str) -> None:
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def _engine(M=5, N=12000, substeps=300, path="fused", H=1.0, seed=0, axisym=False):
    ny = int(128 * H)
    eng = Engine(M, N, n_grid_y=ny, domain_height=H, path=path, axisym=axisym)
    for m in range(M):
        region = (0.3, 0.7, 0.25 * H, 0.25 * H + 0.4)
        pos, mat, fib = small_tank(seed + m, N, region=region)
        eng.load_particles(m, pos, mat, fib)
    rng = np.random.default_rng(seed)
    eng.set_param("freq", rng.uniform(0.8, 1.5, M).astype(np.float32))
    eng.set_param("contraction", rng.uniform(0.1, 0.3, M).astype(np.float32))
    eng.set_time(0.03)
    eng.substep(substeps)
    eng.sync()
    return eng


def _state(eng):
    return dict(x=eng.read("x"), v=eng.read("v"), Jp=eng.read("Jp"), mat=eng.read("material"),
                t=np.float32(eng.get_time()), freq=eng.get_param("freq"),
                contraction=eng.get_param("contraction"))


def _consts(eng):
    import oracle
    return oracle.reference_consts(n_grid_y=eng.ny, domain_height=eng.domain_height)


def _close_additive(got, want, what):
    err = np.abs(got.astype(np.float64) - want)
    lim = 1e-4 * np.abs(want) + 1e-6 * np.abs(want).max()
    bad = err > lim
    say(f"{what}: max |d| {err.max():.3e}, max |ref| {np.abs(want).max():.3e}, "
        f"{int(bad.sum())} pixels over")
    assert not bad.any(), what


def _ulps(got, want, n, what, inp=None):
    got = np.asarray(got, np.float32)
    want = np.asarray(want, np.float32)
    d = np.abs(got.astype(np.float64) - want) / np.spacing(np.maximum(np.abs(want), 1e-30))
    k = np.unravel_index(int(np.argmax(d)), d.shape)
    at = f" (input {inp[k]:.6e})" if inp is not None else ""
    say(f"{what}: max {d.max():.1f} ulps at got {got[k]:.8e} ref {want[k]:.8e}{at}, "
        f"{int((d > 4).sum())} values over 4 ulps")
    assert (d <= n).all(), what


@pytest.mark.parametrize("water_mode", [0, 1, 2])
@pytest.mark.parametrize("radius", [1.0, 3.0])
def test_abyss_frame(water_mode, radius):
    say(f"abyss water_mode={water_mode} radius={radius}")
    eng = _engine()
    st = _state(eng)
    rng = np.random.default_rng(water_mode)
    if water_mode:            # give the water speeds on both sides of the 0.02 cut
        v = (rng.normal(size=st["v"].shape) * 0.5).astype(np.float32)
        eng.write("v", v)
        st["v"] = eng.read("v")
    M = eng.M
    res, grid_side = 256, 3
    res_sub = res // grid_side
    hue = rng.uniform(0, 1, M).astype(np.float32)
    mhue = rng.uniform(0, 1, M).astype(np.float32)
    eng.render_clear(res, 0.0)
    eng.render_abyss(res_sub, grid_side, radius, hue, mhue, water_mode)
    got = eng.read_frame()
    want = render.render_frame_abyss(np.zeros((res, res, 3), np.float32), st["x"], st["v"],
                                     st["Jp"], st["mat"], st["t"], st["freq"],
                                     st["contraction"], hue, mhue, water_mode, res_sub,
                                     grid_side, radius, eng.domain_height, _consts(eng))
    assert want.max() > 0
    _close_additive(got, want, "abyss")
    eng.render_tone_map()
    tm = eng.read_frame()
    _ulps(tm, render.tone_map_and_encode(got.copy()), 8, "tone map", inp=got)
    eng.close()


def test_flat_passes_web_palette():
    say("flat passes, web palette")
    eng = _engine()
    st = _state(eng)
    res, grid_side = 256, 3
    res_sub = res // grid_side
    radius = max(1.0, res_sub / 128 * 0.6) * 2.5
    eng.render_clear(res, 0.98)
    want = np.full((res, res, 3), 0.98, np.float32)
    for mat_id, r, g, b in WEB_PALETTE:
        eng.render_flat(res_sub, grid_side, radius, mat_id, r, g, b)
        render.render_flat_pass(want, st["x"], st["mat"], mat_id, (r, g, b), res_sub,
                                grid_side, radius, eng.domain_height)
    got = eng.read_frame()
    bad = (got != want).any(axis=2)
    say(f"flat: {int(bad.sum())} of {res * res} pixels differ")
    assert not bad.any()
    eng.close()


def _grid_refs(eng, scale):
    gv, gm = eng.read("grid_v"), eng.read("grid_m")
    return gv, gm, render.compute_vorticity_grid(gv, gm, 1 / 128, scale)


@pytest.mark.parametrize("path", ["fused", "plain"])
def test_vorticity_grid(path):
    say(f"vorticity grid path={path}")
    eng = _engine(path=path)
    eng.render_vorticity_grid(2e-3)
    got = eng.read_vorticity()
    _, _, want = _grid_refs(eng, 2e-3)
    err = np.abs(got - want).max()
    say(f"vorticity grid max |d| {err:.3e}, max |ref| {np.abs(want).max():.3e}")
    assert np.abs(want).max() > 1e-3
    assert err <= 1e-6
    eng.close()


@pytest.mark.parametrize("path", ["fused", "plain"])
def test_axisym_vorticity_and_abyss(path):
    """The axisymmetric variant's grid after a run holds velocities on both paths (the
    fused path turns the grid its last G2P read into velocities in place, through the
    axisymmetric grid op): the vorticity grid against the oracle's curl of the grid read
    back, and the abyss frame against the oracle's of the state read back (20 substeps,
    tanks at r >= 0.3, before the formulation diverges)."""
    say(f"axisym vorticity grid and abyss frame, path={path}")
    eng = _engine(M=3, substeps=20, path=path, axisym=True)
    eng.render_vorticity_grid(2e-3)
    got = eng.read_vorticity()
    _, _, want = _grid_refs(eng, 2e-3)
    err = np.abs(got - want).max()
    say(f"vorticity grid max |d| {err:.3e}, max |ref| {np.abs(want).max():.3e}")
    assert np.abs(want).max() > 1e-4
    assert err <= 1e-6
    st = _state(eng)
    res, grid_side = 256, 2
    res_sub = res // grid_side
    hue = np.array([0.1, 0.5, 0.9], np.float32)
    eng.render_clear(res, 0.0)
    eng.render_abyss(res_sub, grid_side, 1.0, hue, hue, 0)
    frame = eng.read_frame()
    ref = render.render_frame_abyss(np.zeros((res, res, 3), np.float32), st["x"], st["v"],
                                    st["Jp"], st["mat"], st["t"], st["freq"],
                                    st["contraction"], hue, hue, 0, res_sub, grid_side, 1.0,
                                    eng.domain_height, _consts(eng))
    assert ref.max() > 0
    _close_additive(frame, ref, "abyss (axisymmetric)")
    eng.close()


@pytest.mark.parametrize("H", [1.0, 2.0])
def test_vorticity_overlays(H):
    say(f"vorticity overlays H={H}")
    eng = _engine(M=3, N=12000, H=H)
    st = _state(eng)
    res, grid_side = 256, 2
    res_sub = res // grid_side
    scale = 0.1       # |curl| reaches ~6/s here: vorticity up to ~0.5, overlay squares drawn
    hue = np.full(eng.M, 0.55, np.float32)
    mhue = np.full(eng.M, 0.7, np.float32)
    gv, gm, vort = _grid_refs(eng, scale)

    # additive overlay between the abyss splat and the tone map
    eng.render_clear(res, 0.0)
    eng.render_abyss(res_sub, grid_side, 1.0, hue, mhue, 0)
    eng.render_vorticity(0, res_sub, grid_side, scale)
    got = eng.read_frame()
    want = render.render_frame_abyss(np.zeros((res, res, 3), np.float32), st["x"], st["v"],
                                     st["Jp"], st["mat"], st["t"], st["freq"],
                                     st["contraction"], hue, mhue, 0, res_sub, grid_side, 1.0,
                                     eng.domain_height, _consts(eng))
    splat = want.copy()
    render.render_vorticity_overlay(want, gv, gm, 1 / 128, scale, res_sub, grid_side,
                                    eng.domain_height)
    drawn = (want != splat).any(axis=2)
    say(f"overlay: {int(drawn.sum())} pixels drawn")
    assert drawn.sum() > 0
    _close_additive(got, want, "overlay")

    # per-pixel composites over the tone-mapped frame, and over white, from the engine's own
    # vorticity grid (its parity is test_vorticity_grid's)
    eng.render_vorticity_grid(scale)
    gvort = eng.read_vorticity()
    err = np.abs(gvort - vort).max()
    say(f"vorticity grid max |d| {err:.3e}")
    assert err <= 1e-6
    for mode, name, param in ((1, "rdbu", 0.6), (2, "hueshift", 0.5), (3, "white", 0.65)):
        eng.render_clear(res, 0.0)
        eng.render_abyss(res_sub, grid_side, 1.0, hue, mhue, 0)
        eng.render_vorticity(0, res_sub, grid_side, scale)
        eng.render_tone_map()
        base = eng.read_frame()
        eng.render_vorticity(mode, res_sub, grid_side, param)
        got = eng.read_frame()
        start = np.ones_like(base) if name == "white" else base.copy()
        want = render.render_vorticity_pixels(start.copy(), gvort, name, param, res_sub,
                                              grid_side, eng.domain_height, eng.M)
        changed = (want != start).any(axis=2)
        say(f"{name}: {int(changed.sum())} pixels composited")
        assert changed.sum() > 0
        _ulps(got, want, 4, name)
    eng.close()


def test_render_edge_cases():
    """Edge cases against the oracle: an empty instance, instances whose tiles fall outside
    the frame (grid_side^2 < M and oversized tiles), radius 0 (a disc of no pixels), a disc
    larger than the LDS box of a workgroup (the global-atomic fallback), and re-allocation
    of the frame at another resolution."""
    say("render edge cases")
    M, N = 5, 12000
    eng = Engine(M, N)
    for m in range(M):
        if m == 2:
            continue                                 # instance 2 stays empty
        pos, mat, fib = small_tank(10 + m, N, region=(0.3, 0.7, 0.1, 0.4))
        eng.load_particles(m, pos, mat, fib)
    eng.substep(50)
    eng.sync()
    st = _state(eng)
    hue = np.linspace(0.1, 0.9, M).astype(np.float32)
    mhue = np.linspace(0.9, 0.1, M).astype(np.float32)
    cases = [(128, 2, 96, 1.0), (128, 2, 64, 0.0), (200, 3, 80, 12.0), (96, 1, 96, 2.0)]
    for res, grid_side, res_sub, radius in cases:
        say(f"  res {res} grid_side {grid_side} res_sub {res_sub} radius {radius}")
        eng.render_clear(res, 0.0)
        eng.render_abyss(res_sub, grid_side, radius, hue, mhue, 0)
        got = eng.read_frame()
        want = render.render_frame_abyss(np.zeros((res, res, 3), np.float32), st["x"],
                                         st["v"], st["Jp"], st["mat"], st["t"], st["freq"],
                                         st["contraction"], hue, mhue, 0, res_sub, grid_side,
                                         radius, eng.domain_height, _consts(eng))
        if radius == 0.0:
            assert not want.any() and not got.any()
        else:
            assert want.max() > 0
            _close_additive(got, want, f"abyss res {res} radius {radius}")
        eng.render_clear(res, 0.98)
        eng.render_flat(res_sub, grid_side, radius, 1, 0.1, 0.2, 0.3)
        wflat = render.render_flat_pass(np.full((res, res, 3), 0.98, np.float32), st["x"],
                                        st["mat"], 1, (0.1, 0.2, 0.3), res_sub, grid_side,
                                        radius, eng.domain_height)
        assert (eng.read_frame() == wflat).all()
    eng.close()


def test_frame_encode_u8():
    """jm_read_frame_u8 against the oracle's encode of the fp32 frame read back, bit-exact:
    constant frames on both sides of [0, 1] and NaN, a raw abyss frame (values above 1), its
    tone-mapped frame, and a resolution whose byte count is not a multiple of 4."""
    say("frame encode u8")
    eng = _engine(M=3, N=12000, substeps=100)
    hue = np.full(eng.M, 0.3, np.float32)
    for res in (97, 256):
        for value in (-0.5, 0.0, 1 / 255, 0.3, 0.5, 1.0, 7.0, float("nan")):
            eng.render_clear(res, value)
            got = eng.read_frame_u8()
            assert (got == render.encode_u8(eng.read_frame())).all(), (res, value)
        grid_side = 2
        eng.render_clear(res, 0.0)
        eng.render_abyss(res // grid_side, grid_side, 2.0, hue, hue, 1)
        raw = eng.read_frame()
        assert raw.max() > 1.0 and (raw == 0).any()
        assert (eng.read_frame_u8() == render.encode_u8(raw)).all(), res
        eng.render_tone_map()
        tm = eng.read_frame()
        enc = eng.read_frame_u8()
        say(f"  res {res}: {int((enc > 0).sum())} nonzero bytes, max {int(enc.max())}")
        assert (enc == render.encode_u8(tm)).all(), res
    eng.close()
