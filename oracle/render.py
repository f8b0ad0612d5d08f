"""TEST INFRASTRUCTURE — the oracle of the frame renderer. Not part of the product path.

numpy restatement of the reference's rendering kernels (c-m-x-r/jellyfih mpm_sim.py:343-712)
in fp32, on the reference's slot-layout arrays (instance m, slot p). Splats loop over the
pixel offsets of the largest disc and add with ``np.add.at``; the reference adds with atomics
in an unspecified order, so a frame is defined up to fp32 summation order.
"""
from __future__ import annotations

import numpy as np

from . import OracleConsts, activation

f32 = np.float32
PI = f32(3.14159265)


def hsv2rgb(h, s, v):
    """mpm_sim.py:345-359, element-wise on fp32 arrays."""
    h, s, v = (np.asarray(a, f32) for a in (h, s, v))
    h, s, v = np.broadcast_arrays(h, s, v)
    i = (h * f32(6.0)).astype(np.int64)               # int(): truncation
    f = (h * f32(6.0)) - i.astype(f32)
    p = v * (f32(1.0) - s)
    q = v * (f32(1.0) - f * s)
    t = v * (f32(1.0) - (f32(1.0) - f) * s)
    r6 = np.mod(i, 6)
    table = [(v, t, p), (q, v, p), (p, v, t), (p, q, v), (t, p, v), (v, p, q)]
    out = np.zeros(h.shape + (3,), f32)
    for k, (a, b, c) in enumerate(table):
        sel = r6 == k
        out[sel, 0], out[sel, 1], out[sel, 2] = a[sel], b[sel], c[sel]
    return out


def _centres(x, inst, res_sub, grid_side, domain_height):
    """Tile-space splat centre of each particle (mpm_sim.py:463-464,500-501)."""
    dh = f32(domain_height)
    row = (inst // grid_side).astype(f32)
    col = (inst % grid_side).astype(f32)
    cx = (x[:, 0] / dh + col) * f32(res_sub)
    cy = ((dh - x[:, 1]) / dh + row) * f32(res_sub)
    return cx, cy


def _splat(fb, cx, cy, draw_r, add_rgb):
    """For every particle and every pixel (px, py) with int(c - r) <= p <= int(c + r) inside
    the frame: call add_rgb(idx, dist_sq) -> (rgb values or None, mask)."""
    res = fb.shape[0]
    lx, hx = (cx - draw_r).astype(np.int64), (cx + draw_r).astype(np.int64)
    ly, hy = (cy - draw_r).astype(np.int64), (cy + draw_r).astype(np.int64)
    span = int(max((hx - lx).max(initial=0), (hy - ly).max(initial=0))) + 1
    for ox in range(span):
        px = lx + ox
        okx = px <= hx
        for oy in range(span):
            py = ly + oy
            ok = okx & (py <= hy) & (px >= 0) & (px < res) & (py >= 0) & (py < res)
            if not ok.any():
                continue
            dxp = px.astype(f32) - cx
            dyp = py.astype(f32) - cy
            dist_sq = dxp * dxp + dyp * dyp
            vals, keep = add_rgb(dist_sq)
            sel = ok & keep
            if not sel.any():
                continue
            yield py[sel], px[sel], vals[sel] if vals is not None else None, sel


def render_frame_abyss(fb, x, v, Jp, mat, t, freq, contraction, hue, muscle_hue, water_mode,
                       res_sub, grid_side, radius, domain_height, consts: OracleConsts):
    """mpm_sim.py:385-477 on slot-layout arrays x, v (M, N, 2), Jp, mat (M, N); freq,
    contraction, hue, muscle_hue (M,); t the fp32 sim_time. Adds into fb (res, res, 3)."""
    M, N = mat.shape
    inst = np.repeat(np.arange(M), N)
    x = np.asarray(x, f32).reshape(-1, 2)
    v = np.asarray(v, f32).reshape(-1, 2)
    Jp = np.asarray(Jp, f32).reshape(-1)
    mat = np.asarray(mat).reshape(-1)
    live = (mat >= 0) & ~(x[:, 0] < 0)
    inst, x, v, Jp, mat = inst[live], x[live], v[live], Jp[live], mat[live]
    act = activation(consts, np.full(M, t, f32), np.asarray(freq, f32),
                     np.asarray(contraction, f32))[inst]
    vel = np.sqrt(v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1])
    stress = f32(1.0) - Jp
    color = np.zeros((len(mat), 3), f32)
    intensity = np.zeros(len(mat), f32)
    gnf = f32(grid_side - 1) if grid_side > 1 else f32(1.0)
    norm_y = (inst // grid_side).astype(f32) / gnf
    w = (mat == 0) & (vel > f32(0.02))
    if water_mode == 0:
        color[w] = np.array([0.15, 0.15, 0.18], f32)
        intensity[w] = f32(0.008)
    else:
        bright = np.minimum(vel / f32(2.0), f32(1.0))
        if water_mode == 1:
            color[w] = hsv2rgb(f32(0.60), f32(0.55), bright[w])
        else:
            ang = np.arctan2(v[:, 1], v[:, 0]).astype(f32)
            h = ang / (f32(2.0) * PI) + f32(0.5)
            h = h - np.floor(h / f32(1.0)) * f32(1.0)
            color[w] = hsv2rgb(h[w], f32(0.8), bright[w])
        intensity[w] = f32(0.025) + f32(0.075) * bright[w]
    j = mat == 1
    base = hsv2rgb(np.asarray(hue, f32)[inst[j]], f32(0.4) + (norm_y[j] * f32(0.4)), f32(0.9))
    glow = np.abs(stress[j]) * f32(2.0) + f32(0.1)
    color[j] = base + glow[:, None]
    intensity[j] = f32(0.05) + (glow * f32(0.2))
    mu = mat == 3
    base = hsv2rgb(np.asarray(muscle_hue, f32)[inst[mu]], f32(0.2), f32(1.0))
    color[mu] = base + act[mu][:, None]
    intensity[mu] = f32(0.4) + (act[mu] * f32(0.4))
    pl = mat == 2
    color[pl] = np.array([1.0, 0.2, 0.0], f32)
    intensity[pl] = f32(0.8)
    draw = intensity > f32(0.001)
    inst, x, color, intensity = inst[draw], x[draw], color[draw], intensity[draw]
    cx, cy = _centres(x, inst, res_sub, grid_side, domain_height)
    draw_r = f32(radius) * f32(1.5)
    r2 = draw_r * draw_r
    ci = color * intensity[:, None]

    def add(dist_sq):
        # radius 0: d^2 / 0 is inf or NaN, so no pixel passes dn < 1, as in the reference
        with np.errstate(divide="ignore", invalid="ignore"):
            dn = dist_sq / r2
            keep = dn < f32(1.0)
            u = f32(1.0) - dn
            u2 = u * u
            return ci * (u2 * u2)[:, None], keep

    for py, px, vals, _ in _splat(fb, cx, cy, draw_r, add):
        for c in range(3):
            np.add.at(fb[:, :, c], (py, px), vals[:, c])
    return fb


def render_flat_pass(fb, x, mat, target_mat, rgb, res_sub, grid_side, radius, domain_height):
    """mpm_sim.py:487-517: pixels of the discs of `target_mat` particles set to rgb."""
    M, N = mat.shape
    inst = np.repeat(np.arange(M), N)
    x = np.asarray(x, f32).reshape(-1, 2)
    mat = np.asarray(mat).reshape(-1)
    sel = (mat == target_mat) & ~(x[:, 0] < 0)
    cx, cy = _centres(x[sel], inst[sel], res_sub, grid_side, domain_height)
    draw_r = f32(radius)

    def inside(dist_sq):
        return None, dist_sq <= draw_r * draw_r

    for py, px, _, _ in _splat(fb, cx, cy, draw_r, inside):
        fb[py, px] = np.asarray(rgb, f32)
    return fb


def tone_map_and_encode(fb):
    """mpm_sim.py:366-383, in place."""
    e = fb * f32(2.5)
    mapped = e / (e + f32(1.0))
    mapped = mapped / (mapped + f32(0.15))
    fb[...] = np.power(mapped, f32(1.0 / 2.2)).astype(f32)
    return fb


def encode_u8(fb):
    """The viewers' 8-bit encode of a frame, (np.clip(img, 0, 1) * 255).astype(np.uint8)
    (evolve.py:396-397, helpers/view_single.py:257): fp32 clip, fp32 product, truncation.
    NaN is given the value 0 here (numpy's cast of NaN is unspecified; 0 on x86-64)."""
    img = np.asarray(fb, f32)
    c = np.clip(np.nan_to_num(img, nan=0.0, posinf=1.0, neginf=0.0), f32(0), f32(1))
    return (c * f32(255)).astype(np.uint8)


def grid_curl(grid_v, grid_m, dx):
    """curl = dvx/dy - dvy/dx at interior nodes with mass (mpm_sim.py:528-533,575-579);
    grid_v (M, nx, ny, 2) post-grid-op velocity, grid_m (M, nx, ny). NaN where skipped."""
    gv = np.asarray(grid_v, f32)
    curl = np.full(np.asarray(grid_m).shape, np.nan, f32)
    two_dx = f32(2.0) * f32(dx)
    dvx_dy = (gv[:, 1:-1, 2:, 0] - gv[:, 1:-1, :-2, 0]) / two_dx
    dvy_dx = (gv[:, 2:, 1:-1, 1] - gv[:, :-2, 1:-1, 1]) / two_dx
    inner = dvx_dy - dvy_dx
    has = np.asarray(grid_m, f32)[:, 1:-1, 1:-1] > 0
    curl[:, 1:-1, 1:-1] = np.where(has, inner, np.nan)
    return curl


def compute_vorticity_grid(grid_v, grid_m, dx, scale):
    """mpm_sim.py:560-578: tanh(curl * scale), 0 where skipped; (M, nx, ny)."""
    curl = grid_curl(grid_v, grid_m, dx)
    return np.where(np.isnan(curl), f32(0.0), np.tanh(curl * f32(scale))).astype(f32)


def render_vorticity_overlay(fb, grid_v, grid_m, dx, scale, res_sub, grid_side,
                             domain_height):
    """mpm_sim.py:520-558: additive squares of (1, .4, 0) or (0, .4, 1) x intensity x 0.5."""
    curl = grid_curl(grid_v, grid_m, dx)
    M, nx, ny = curl.shape
    res = fb.shape[0]
    vi = np.minimum(np.abs(curl) * f32(scale), f32(1.0))
    m, i, j = np.nonzero(~np.isnan(curl) & (vi >= f32(0.01)))
    c = curl[m, i, j]
    s = vi[m, i, j]
    col = np.where((c > 0)[:, None], np.array([1.0, 0.4, 0.0], f32) * s[:, None],
                   np.array([0.0, 0.4, 1.0], f32) * s[:, None]).astype(f32)
    row, cc = m // grid_side, m % grid_side
    dh, ng = f32(domain_height), f32(nx)
    cx = ((i.astype(f32) / ng / dh + cc.astype(f32)) * f32(res_sub)).astype(np.int64)
    cy = (((dh - j.astype(f32) / ng) / dh + row.astype(f32)) * f32(res_sub)).astype(np.int64)
    cpx = max(1, res_sub // nx)
    for ox in range(-(cpx // 2), cpx // 2 + 1):
        for oy in range(-(cpx // 2), cpx // 2 + 1):
            px, py = cx + ox, cy + oy
            ok = (px >= 0) & (px < res) & (py >= 0) & (py < res)
            for ch in range(3):
                np.add.at(fb[:, :, ch], (py[ok], px[ok]), col[ok, ch] * f32(0.5))
    return fb


def render_vorticity_pixels(fb, vort, mode, amax, res_sub, grid_side, domain_height, n_inst):
    """mpm_sim.py:581-712: mode 'rdbu', 'hueshift' or 'white' (the caller fills fb with 1.0
    for 'white'); vort (M, nx, ny)."""
    res = fb.shape[0]
    M, nx, ny = vort.shape
    py, px = np.meshgrid(np.arange(res), np.arange(res), indexing="ij")
    col = px * grid_side // res
    row = py * grid_side // res
    m = row * grid_side + col
    sub_x = (px - col * res_sub).astype(f32) / f32(res_sub)
    sub_y = f32(1.0) - (py - row * res_sub).astype(f32) / f32(res_sub)
    gi = (sub_x * f32(nx) * f32(domain_height)).astype(np.int64)
    gj = (sub_y * f32(ny)).astype(np.int64)
    ok = (m < n_inst) & (gi >= 0) & (gi < nx) & (gj >= 0) & (gj < ny)
    val = np.zeros(fb.shape[:2], f32)
    val[ok] = vort[m[ok], gi[ok], gj[ok]]
    mag = np.abs(val)
    ok &= mag >= f32(0.02)
    alpha = f32(amax) * mag
    t = f32(0.5) + f32(0.5) * np.sign(val).astype(f32)
    if mode == "hueshift":
        cr, cg, cb = (0.90, 0.15), (0.18, 0.35), (0.15, 0.90)
    else:
        cr, cg, cb = (0.84, 0.13), (0.10, 0.31), (0.11, 0.80)
    for ch, (a, b) in enumerate((cr, cg, cb)):
        target = t * f32(a) + (f32(1.0) - t) * f32(b)
        if mode == "white":
            new = (f32(1.0) - alpha) + alpha * target
        else:
            new = (f32(1.0) - alpha) * fb[:, :, ch] + alpha * target
        fb[:, :, ch] = np.where(ok, new, fb[:, :, ch])
    return fb
