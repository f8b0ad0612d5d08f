"""Pins of the renderer oracle (oracle/render.py) against what fixes it independently:
colorsys for hsv2rgb, literal per-pixel loops of the reference kernels on tiny frames
(mpm_sim.py:385-477,487-517,520-558,581-712), the splat's closed-form integral and centroid,
hand-derived tone-map values and the exact curl of linear velocity fields."""
from __future__ import annotations

import colorsys
import math

import numpy as np
import pytest

import oracle
from oracle import render

f32 = np.float32


def _consts():
    return oracle.reference_consts()


def test_hsv2rgb_matches_colorsys():
    rng = np.random.default_rng(0)
    h = np.concatenate([rng.uniform(0, 1, 4000), [0.0, 1 / 6, 0.5, 5 / 6, 1.0]]).astype(f32)
    s = rng.uniform(0, 1, h.size).astype(f32)
    v = rng.uniform(0, 1, h.size).astype(f32)
    got = render.hsv2rgb(h, s, v)
    want = np.array([colorsys.hsv_to_rgb(float(a), float(b), float(c))
                     for a, b, c in zip(h, s, v)])
    np.testing.assert_allclose(got, want, atol=2e-6)


def _state(rows):
    """Slot-layout arrays of instances given as lists of (x, y, vx, vy, Jp, mat)."""
    M = len(rows)
    N = max(1, max(len(r) for r in rows))
    x = np.full((M, N, 2), -1.0, f32)
    v = np.zeros((M, N, 2), f32)
    Jp = np.ones((M, N), f32)
    mat = np.full((M, N), -1, np.int32)
    for m, r in enumerate(rows):
        for p, (px, py, vx, vy, jp, mt) in enumerate(r):
            x[m, p] = px, py
            v[m, p] = vx, vy
            Jp[m, p] = jp
            mat[m, p] = mt
    return x, v, Jp, mat


def _abyss(fb, x, v, Jp, mat, radius, grid_side, res_sub, H=1.0, t=0.0, water_mode=0,
           hue=None, muscle_hue=None):
    M = mat.shape[0]
    hue = np.full(M, 0.55, f32) if hue is None else np.asarray(hue, f32)
    muscle_hue = np.full(M, 0.7, f32) if muscle_hue is None else np.asarray(muscle_hue, f32)
    return render.render_frame_abyss(fb, x, v, Jp, mat, t, np.ones(M, f32),
                                     np.full(M, 0.2, f32), hue, muscle_hue, water_mode,
                                     res_sub, grid_side, radius, H, _consts())


def _literal_splat(res, pts, radius, rgb, res_sub, grid_side, H=1.0):
    """mpm_sim.py:461-480 as written: nested pixel loops per particle, fp32 scalars."""
    fb = np.zeros((res, res, 3), f32)
    for m, (px_, py_) in pts:
        if px_ < 0:
            continue
        row, col = m // grid_side, m % grid_side
        cx = (f32(px_) / f32(H) + f32(col)) * f32(res_sub)
        cy = ((f32(H) - f32(py_)) / f32(H) + f32(row)) * f32(res_sub)
        dr = f32(radius) * f32(1.5)
        for px in range(int(cx - dr), int(cx + dr) + 1):
            for py in range(int(cy - dr), int(cy + dr) + 1):
                if 0 <= px < res and 0 <= py < res:
                    d2 = (f32(px) - cx) ** 2 + (f32(py) - cy) ** 2
                    dn = f32(d2) / (dr * dr)
                    if dn < 1.0:
                        fall = (f32(1.0) - dn) ** 4
                        for c in range(3):
                            fb[py, px, c] += f32(rgb[c]) * f32(0.8) * f32(fall)
    return fb


def test_abyss_splat_matches_literal_loops():
    rng = np.random.default_rng(1)
    res, grid_side = 48, 2
    res_sub = res // grid_side
    rows = []
    pts = []
    for m in range(3):
        r = []
        for _ in range(25):
            px, py = rng.uniform(-0.05, 1.05, 2)   # some discs cross the tile and frame edges
            r.append((px, py, 0.0, 0.0, 1.0, 2))
            pts.append((m, (px, py)))
        rows.append(r)
    x, v, Jp, mat = _state(rows)
    got = _abyss(np.zeros((res, res, 3), f32), x, v, Jp, mat, 2.3, grid_side, res_sub)
    want = _literal_splat(res, [(m, tuple(x[m, p])) for m, p in
                                [(m, p) for m in range(3) for p in range(25)]],
                          2.3, (1.0, 0.2, 0.0), res_sub, grid_side)
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-7)
    assert got[:, :, 0].sum() > 0


def test_abyss_splat_integral_and_centroid():
    """One payload particle: the frame holds 0.8 * (1, 0.2, 0) * sum (1 - d^2/R^2)^4, whose
    sum over the pixel lattice is pi R^2 / 5 to O(1/R), centred on the mapped position."""
    res, grid_side, radius = 512, 2, 20.0
    res_sub = res // grid_side
    R = 1.5 * radius
    x, v, Jp, mat = _state([[], [(0.3125, 0.6875, 0.0, 0.0, 1.0, 2)]])   # instance 1: col 1
    fb = _abyss(np.zeros((res, res, 3), f32), x, v, Jp, mat, radius, grid_side, res_sub)
    tot = fb[:, :, 0].astype(np.float64).sum()
    assert tot == pytest.approx(0.8 * math.pi * R * R / 5, rel=2e-3)
    assert fb[:, :, 1].sum() == pytest.approx(0.2 * tot, rel=1e-5)
    assert fb[:, :, 2].max() == 0.0
    w = fb[:, :, 0].astype(np.float64)
    py, px = np.mgrid[:res, :res]
    cx, cy = (w * px).sum() / tot, (w * py).sum() / tot
    assert cx == pytest.approx((0.3125 + 1) * res_sub, abs=1e-6)   # x to the right, col 1
    assert cy == pytest.approx((1 - 0.6875) * res_sub, abs=1e-6)   # y up, row 0 on top
    # the tall tank (domain height 2) halves both scales
    fb2 = _abyss(np.zeros((res, res, 3), f32), x, v, Jp, mat, radius, grid_side, res_sub, H=2.0)
    w2 = fb2[:, :, 0].astype(np.float64)
    assert (w2 * px).sum() / w2.sum() == pytest.approx((0.3125 / 2 + 1) * res_sub, abs=1e-6)
    assert (w2 * py).sum() / w2.sum() == pytest.approx((2 - 0.6875) / 2 * res_sub, abs=1e-6)


def _centre_pixel(m, grid_side, res_sub, pos=(0.5, 0.5)):
    return int((1 - pos[1] + m // grid_side) * res_sub), int((pos[0] + m % grid_side) * res_sub)


def test_abyss_material_colours():
    """Colour x intensity at a disc's centre pixel (falloff 1) per material."""
    res_sub, grid_side = 64, 3
    res = res_sub * grid_side
    t = 0.05          # phase 0.05 of a 1 s period with contraction 0.2: raised cosine
    act = 0.5 * (1 - math.cos(0.05 / 0.2 * math.pi))
    rows = [
        [(0.5, 0.5, 0.0, 0.0, 0.7, 1)],              # jelly, stress 0.3, row 0
        [(0.5, 0.5, 0.0, 0.0, 1.0, 3)],              # muscle
        [(0.5, 0.5, 0.0, 0.0, 1.0, 2)],              # payload
        [(0.5, 0.5, 0.0, 0.0, 1.3, 1)],              # jelly, stress -0.3, row 1
        [(0.5, 0.5, 0.0, 0.01, 1.0, 0)],             # slow water: not drawn
    ]
    x, v, Jp, mat = _state(rows)
    hue = [0.1, 0.2, 0.3, 0.9, 0.5]
    mhue = [0.4, 0.65, 0.0, 0.0, 0.0]
    fb = _abyss(np.zeros((res, res, 3), f32), x, v, Jp, mat, 4.0, grid_side, res_sub, t=t,
                hue=hue, muscle_hue=mhue)

    def at(m):
        return fb[_centre_pixel(m, grid_side, res_sub)].astype(np.float64)

    for m, norm_y in ((0, 0.0), (3, 0.5)):
        glow = 0.3 * 2 + 0.1
        want = (np.array(colorsys.hsv_to_rgb(hue[m], 0.4 + 0.4 * norm_y, 0.9)) + glow) \
            * (0.05 + 0.2 * glow)
        np.testing.assert_allclose(at(m), want, rtol=1e-5)
    want = (np.array(colorsys.hsv_to_rgb(0.65, 0.2, 1.0)) + act) * (0.4 + 0.4 * act)
    np.testing.assert_allclose(at(1), want, rtol=1e-5)
    np.testing.assert_allclose(at(2), [0.8, 0.16, 0.0], rtol=1e-6)
    r, c = _centre_pixel(4, grid_side, res_sub)
    assert fb[r - 8:r + 8, c - 8:c + 8].max() == 0.0


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_abyss_water_modes(mode):
    res_sub, grid_side = 64, 2
    res = res_sub * grid_side
    vel = [(0.0, 1.0), (-0.5, 0.0), (3.0, -3.0)]
    x, v, Jp, mat = _state([[(0.5, 0.5, vx, vy, 1.0, 0)] for vx, vy in vel])
    fb = _abyss(np.zeros((res, res, 3), f32), x, v, Jp, mat, 4.0, grid_side, res_sub,
                water_mode=mode)
    for m, (vx, vy) in enumerate(vel):
        got = fb[_centre_pixel(m, grid_side, res_sub)].astype(np.float64)
        speed = math.hypot(vx, vy)
        b = min(speed / 2, 1.0)
        if mode == 0:
            want = np.array([0.15, 0.15, 0.18]) * 0.008
        elif mode == 1:
            want = np.array(colorsys.hsv_to_rgb(0.6, 0.55, b)) * (0.025 + 0.075 * b)
        else:
            h = (math.atan2(vy, vx) / (2 * math.pi) + 0.5) % 1.0
            want = np.array(colorsys.hsv_to_rgb(h, 0.8, b)) * (0.025 + 0.075 * b)
        np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-8)


def test_flat_pass_matches_literal_loops_and_disc_area():
    rng = np.random.default_rng(2)
    res, grid_side, radius = 40, 2, 2.7
    res_sub = res // grid_side
    rows = [[(*rng.uniform(-0.05, 1.05, 2), 0.0, 0.0, 1.0, int(rng.integers(0, 4)))
             for _ in range(30)] for _ in range(4)]
    x, v, Jp, mat = _state(rows)
    got = render.render_flat_pass(np.full((res, res, 3), 0.98, f32), x, mat, 3, (0.1, 0.2, 0.3),
                                  res_sub, grid_side, radius, 1.0)
    want = np.full((res, res, 3), 0.98, f32)
    for m in range(4):
        for p in range(30):
            if mat[m, p] != 3 or x[m, p, 0] < 0:
                continue
            cx = (x[m, p, 0] / f32(1.0) + f32(m % grid_side)) * f32(res_sub)
            cy = ((f32(1.0) - x[m, p, 1]) / f32(1.0) + f32(m // grid_side)) * f32(res_sub)
            r = f32(radius)
            for px in range(int(cx - r), int(cx + r) + 1):
                for py in range(int(cy - r), int(cy + r) + 1):
                    if 0 <= px < res and 0 <= py < res:
                        if (f32(px) - cx) ** 2 + (f32(py) - cy) ** 2 <= r * r:
                            want[py, px] = (0.1, 0.2, 0.3)
    np.testing.assert_array_equal(got, want)
    # a lone large disc covers pi r^2 pixels to within its perimeter
    x1, _, _, m1 = _state([[(0.5, 0.5, 0, 0, 1, 1)]])
    fb = render.render_flat_pass(np.zeros((400, 400, 3), f32), x1, m1, 1, (1, 1, 1), 400, 1,
                                 100.0, 1.0)
    assert fb[:, :, 0].sum() == pytest.approx(math.pi * 100 ** 2, rel=2e-3)


def test_tone_map_closed_forms():
    fb = np.array([0.0, 0.4, 1e30, 0.1, 1.0], f32).reshape(1, 5, 1).repeat(3, axis=2)
    out = render.tone_map_and_encode(fb.copy())[0, :, 0].astype(np.float64)
    assert out[0] == 0.0
    # 2.5 * 0.4 = 1 -> 1/2 -> (1/2) / (1/2 + 0.15) = 10/13
    assert out[1] == pytest.approx((10 / 13) ** (1 / 2.2), rel=1e-6)
    assert out[2] == pytest.approx((1 / 1.15) ** (1 / 2.2), rel=1e-6)
    assert np.all(np.diff(out[[0, 3, 1, 4, 2]]) > 0)


def test_encode_u8_against_exact_floor():
    """encode_u8 is floor(255 v) of the clipped value: checked in float64 arithmetic, where
    the fp32 product may only move a value lying within one fp32 ulp of an integer to its
    neighbour; out-of-range values and NaN saturate; the map is monotone."""
    rng = np.random.default_rng(7)
    k = np.arange(256, dtype=np.float64)
    edges = (k / 255).astype(f32)
    v = np.concatenate([rng.random(200000).astype(f32), edges, np.nextafter(edges, f32(2)),
                        np.nextafter(edges, f32(-1))])
    e = render.encode_u8(v.reshape(1, -1, 1)).ravel().astype(np.int64)
    c = np.clip(v.astype(np.float64), 0, 1) * 255
    exact = np.floor(c).astype(np.int64)
    near = np.abs(c - np.round(c)) <= np.spacing(np.float32(255))
    assert np.all((e == exact) | (near & (np.abs(e - exact) <= 1)))
    assert (e[v >= 1] == 255).all() and (e[v <= 0] == 0).all()
    assert e[(v > 0) & (v < 1)].max() <= 254
    order = np.argsort(v, kind="stable")
    assert np.all(np.diff(e[order]) >= 0)
    special = np.array([np.nan, -np.inf, np.inf, -0.0, 1e30, -3.0, 0.5], f32)
    assert render.encode_u8(special).tolist() == [0, 0, 255, 0, 255, 0, 127]


def test_curl_of_linear_fields():
    """v = (a y + c, b x + d) has curl dvx/dy - dvy/dx = a - b at every interior node with mass;
    boundary nodes and massless nodes are skipped (vorticity 0)."""
    M, nx, ny = 2, 16, 24
    dx = 1 / 128
    i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    a, b = np.array([3.0, -1.5]), np.array([1.0, 2.5])
    gv = np.zeros((M, nx, ny, 2), f32)
    for m in range(M):
        gv[m, :, :, 0] = a[m] * j * dx + 0.25
        gv[m, :, :, 1] = b[m] * i * dx - 0.5
    gm = np.ones((M, nx, ny), f32)
    gm[1, 5, 7] = 0.0
    curl = render.grid_curl(gv, gm, dx)
    assert np.isnan(curl[:, 0, :]).all() and np.isnan(curl[:, -1, :]).all()
    assert np.isnan(curl[:, :, 0]).all() and np.isnan(curl[:, :, -1]).all()
    assert np.isnan(curl[1, 5, 7]) and not np.isnan(curl[0, 5, 7])
    for m in range(M):
        inner = curl[m, 1:-1, 1:-1]
        np.testing.assert_allclose(inner[~np.isnan(inner)], a[m] - b[m], rtol=1e-4)
    vort = render.compute_vorticity_grid(gv, gm, dx, 0.1)
    assert vort[1, 5, 7] == 0.0 and (vort[:, 0, :] == 0).all()
    np.testing.assert_allclose(vort[0, 3, 4], math.tanh(0.1 * 2.0), rtol=1e-4)
    np.testing.assert_allclose(vort[1, 3, 4], math.tanh(0.1 * -4.0), rtol=1e-4)


def _literal_pixels(fb, vort, mode, amax, res_sub, grid_side, H, n_inst):
    """mpm_sim.py:597-624 (rdbu), 642-668 (hueshift), 683-712 (white) as written."""
    res = fb.shape[0]
    M, nx, ny = vort.shape
    if mode == "white":
        fb[...] = 1.0
    cols = {"rdbu": (0.84, 0.10, 0.11, 0.13, 0.31, 0.80),
            "white": (0.84, 0.10, 0.11, 0.13, 0.31, 0.80),
            "hueshift": (0.90, 0.18, 0.15, 0.15, 0.35, 0.90)}[mode]
    for px in range(res):
        for py in range(res):
            col, row = px * grid_side // res, py * grid_side // res
            m = row * grid_side + col
            if m >= n_inst:
                continue
            sub_x = f32(px - col * res_sub) / f32(res_sub)
            sub_y = f32(1.0) - f32(py - row * res_sub) / f32(res_sub)
            gi, gj = int(sub_x * f32(nx) * f32(H)), int(sub_y * f32(ny))
            if gi < 0 or gi >= nx or gj < 0 or gj >= ny:
                continue
            val = vort[m, gi, gj]
            if abs(val) < 0.02:
                continue
            alpha = f32(amax) * abs(val)
            t = f32(0.5) + f32(0.5) * f32(np.sign(val))
            for c in range(3):
                tgt = t * f32(cols[c]) + (f32(1.0) - t) * f32(cols[c + 3])
                base = f32(1.0) if mode == "white" else fb[py, px, c]
                fb[py, px, c] = (f32(1.0) - alpha) * base + alpha * tgt
    return fb


@pytest.mark.parametrize("mode", ["rdbu", "hueshift", "white"])
@pytest.mark.parametrize("H", [1.0, 2.0])
def test_vorticity_pixels_match_literal_loops(mode, H):
    rng = np.random.default_rng(3)
    res, grid_side, n_inst = 30, 2, 3
    res_sub = res // grid_side
    nx, ny = 8, int(8 * H)
    vort = np.tanh(rng.normal(0, 1, (n_inst, nx, ny))).astype(f32)
    vort[rng.uniform(size=vort.shape) < 0.2] = 0.01          # below the 0.02 cut
    fb0 = rng.uniform(0, 1, (res, res, 3)).astype(f32)
    if mode == "white":
        fb0[...] = 1.0
    got = render.render_vorticity_pixels(fb0.copy(), vort, mode, 0.6, res_sub, grid_side, H,
                                         n_inst)
    want = _literal_pixels(fb0.copy(), vort, mode, 0.6, res_sub, grid_side, H, n_inst)
    np.testing.assert_allclose(got, want, rtol=1e-6, atol=1e-7)


def test_vorticity_pixel_colour_extremes():
    """Full-strength cells take the end colours; the top pixel row maps to gj = ny, past the
    grid, and is left as it was (the reference's inverse map, mpm_sim.py:603-608)."""
    vort = np.zeros((1, 4, 4), f32)
    vort[0, 0, 3] = 1.0       # top-left cell, CCW at full strength
    vort[0, 3, 0] = -1.0      # bottom-right cell, CW
    fb = render.render_vorticity_pixels(np.zeros((8, 8, 3), f32), vort, "rdbu", 1.0, 8, 1, 1.0, 1)
    np.testing.assert_allclose(fb[1, 0], [0.84, 0.10, 0.11], rtol=1e-6)
    np.testing.assert_allclose(fb[2, 1], [0.84, 0.10, 0.11], rtol=1e-6)
    assert (fb[0] == 0).all()
    np.testing.assert_allclose(fb[7, 7], [0.13, 0.31, 0.80], rtol=1e-6)
    assert (fb[6, 7] == 0).all()
    fbw = render.render_vorticity_pixels(np.ones((8, 8, 3), f32), vort * f32(0.5), "white", 1.0,
                                         8, 1, 1.0, 1)
    np.testing.assert_allclose(fbw[1, 0], 0.5 + 0.5 * np.array([0.84, 0.10, 0.11]), rtol=1e-6)
    assert (fbw[3:5, 3:5] == 1.0).all()


def test_vorticity_overlay_single_node():
    """One interior node with curl a - b > 0: a (2 (cpx // 2) + 1)^2 square of
    (1, 0.4, 0) * min(|curl| scale, 1) * 0.5 centred on the node's mapped pixel."""
    M, nx, ny = 1, 16, 16
    dx = 1 / 128
    gv = np.zeros((M, nx, ny, 2), f32)
    gm = np.zeros((M, nx, ny), f32)
    gm[0, 6, 9] = 1.0
    gv[0, 6, 10, 0], gv[0, 6, 8, 0] = 0.5, -0.5          # dvx/dy = 1 / (2 dx)
    res_sub = 64                                          # cpx = 64 // 16 = 4 -> 5x5 square
    fb = render.render_vorticity_overlay(np.zeros((64, 64, 3), f32), gv, gm, dx, 0.005,
                                         res_sub, 1, 1.0)
    vi = min(1 / (2 * dx) * 0.005, 1.0)
    cx, cy = int(6 / 16 * 64), int((1 - 9 / 16) * 64)
    want = np.zeros((64, 64, 3))
    want[cy - 2:cy + 3, cx - 2:cx + 3] = np.array([1.0, 0.4, 0.0]) * vi * 0.5
    np.testing.assert_allclose(fb, want, rtol=1e-6, atol=1e-8)


def _literal_overlay(fb, gv, gm, dx, scale, res_sub, grid_side, H):
    """mpm_sim.py:537-558 as written: per node, fp32 scalars."""
    res = fb.shape[0]
    M, nx, ny, _ = gv.shape
    for m in range(M):
        for i in range(nx):
            for j in range(ny):
                if i < 1 or i >= nx - 1 or j < 1 or j >= ny - 1 or not gm[m, i, j] > 0:
                    continue
                dvx_dy = (gv[m, i, j + 1, 0] - gv[m, i, j - 1, 0]) / (f32(2.0) * f32(dx))
                dvy_dx = (gv[m, i + 1, j, 1] - gv[m, i - 1, j, 1]) / (f32(2.0) * f32(dx))
                curl = f32(dvx_dy - dvy_dx)
                vi = min(abs(curl) * f32(scale), f32(1.0))
                if vi < 0.01:
                    continue
                col = (np.array([1.0, 0.4, 0.0], f32) if curl > 0
                       else np.array([0.0, 0.4, 1.0], f32)) * f32(vi)
                row, c = m // grid_side, m % grid_side
                cx = int((f32(i) / f32(nx) / f32(H) + f32(c)) * f32(res_sub))
                cy = int(((f32(H) - f32(j) / f32(nx)) / f32(H) + f32(row)) * f32(res_sub))
                cpx = max(1, res_sub // nx)
                for px in range(cx - cpx // 2, cx + cpx // 2 + 1):
                    for py in range(cy - cpx // 2, cy + cpx // 2 + 1):
                        if 0 <= px < res and 0 <= py < res:
                            fb[py, px] += col * f32(0.5)
    return fb


@pytest.mark.parametrize("H", [1.0, 2.0])
def test_vorticity_overlay_matches_literal_loops(H):
    rng = np.random.default_rng(5)
    M, nx = 3, 12
    ny = int(nx * H)
    gv = rng.normal(0, 0.05, (M, nx, ny, 2)).astype(f32)
    gm = (rng.uniform(size=(M, nx, ny)) < 0.7).astype(f32)
    res, grid_side = 72, 2
    res_sub = res // grid_side
    got = render.render_vorticity_overlay(np.zeros((res, res, 3), f32), gv, gm, 1 / 128, 0.02,
                                          res_sub, grid_side, H)
    want = _literal_overlay(np.zeros((res, res, 3), f32), gv, gm, 1 / 128, 0.02, res_sub,
                            grid_side, H)
    assert want.any()
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-7)
