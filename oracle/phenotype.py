"""TEST INFRASTRUCTURE — the oracle of the input generator. Not part of the product path.

A numpy restatement of the reference's ``make_jelly.generate_phenotype`` and ``fill_tank``
(c-m-x-r/jellyfih ``make_jelly.py:31-402``) in the reference's own operation order and with
the same library calls where the reference makes them (matplotlib ``Path.contains_points``
for the rasterised masks, scipy ``cKDTree`` for the fibre lookup and the water exclusion,
a Python loop for the self-intersection test), so that it checks the engine's native
generator (``jm_generate_phenotype`` / ``jm_fill_tank``) independently of the native
helpers that generator calls. It is itself pinned by the fixtures written from the
reference (``tests/golden/phenotype_golden.npz``). Used by ``tests/`` only.
"""
from __future__ import annotations

import numpy as np

PAYLOAD_WIDTH = 0.08                    # make_jelly.py:8
PAYLOAD_HEIGHT = 0.05                   # make_jelly.py:9
DEFAULT_SPAWN = np.array([0.5, 0.40])   # make_jelly.py:10
AXISYM_SPAWN = np.array([0.0, 0.40])    # make_jelly.py:11


def _segments_cross(a, b, c, d) -> bool:
    # make_jelly.py:31-43
    r, s = b - a, d - c
    rxs = r[0] * s[1] - r[1] * s[0]
    if abs(rxs) < 1e-12:
        return False
    qp = c - a
    t = (qp[0] * s[1] - qp[1] * s[0]) / rxs
    u = (qp[0] * r[1] - qp[1] * r[0]) / rxs
    return 0 < t < 1 and 0 < u < 1


def polygon_self_intersects(poly) -> bool:
    # make_jelly.py:46-58
    n = len(poly)
    for i in range(n):
        a, b = poly[i], poly[(i + 1) % n]
        for j in range(i + 2, n):
            if j == n - 1 and i == 0:
                continue
            if _segments_cross(a, b, poly[j], poly[(j + 1) % n]):
                return True
    return False


def _bezier(p0, p1, p2, p3, t):
    # make_jelly.py:61-63
    return (1 - t) ** 3 * p0 + 3 * (1 - t) ** 2 * t * p1 + 3 * (1 - t) * t ** 2 * p2 + t ** 3 * p3


def _unit_normals(points):
    # make_jelly.py:65-73
    tan = np.gradient(points, axis=0)
    nrm = np.zeros_like(tan)
    nrm[:, 0] = -tan[:, 1]
    nrm[:, 1] = tan[:, 0]
    length = np.linalg.norm(nrm, axis=1, keepdims=True)
    length[length == 0] = 1
    return nrm / length


def generate_phenotype(genome, spawn_offset=None, grid_res=128, with_payload=True,
                       axisym=False):
    """make_jelly.py:74-334: (pos f64 (n,2), mat int (n,), fiber f32 (n,2), self_int)."""
    from matplotlib.path import Path
    from scipy.spatial import cKDTree
    if spawn_offset is None:
        spawn_offset = DEFAULT_SPAWN
    g = np.asarray(genome, dtype=np.float64)
    start = np.array([PAYLOAD_WIDTH / 2.0, 0.0])
    cp1 = start + np.array([abs(g[0]), g[1]])
    cp2 = start + np.array([abs(g[2]), g[3]])
    end = start + np.array([abs(g[4]), g[5]])
    t_base, t_mid, t_tip = abs(g[6]), abs(g[7]), abs(g[8])

    ts = np.linspace(0, 1, 50)
    spine = np.array([_bezier(start, cp1, cp2, end, t) for t in ts])
    tan = np.gradient(spine, axis=0)
    tan_len = np.linalg.norm(tan, axis=1, keepdims=True)
    tan_len[tan_len < 1e-10] = 1.0
    spine_tan = tan / tan_len

    nrm = _unit_normals(spine)
    half = np.interp(ts, [0, 0.5, 1], [t_base, t_mid, t_tip])[:, None] / 2.0
    outer = spine + nrm * half
    inner = spine - nrm * half

    spacing = 1.0 / (grid_res * 2.0)
    muscle_ratio = 0.2
    muscle_thick = np.minimum(spacing * 1.2 + half * muscle_ratio, half * 0.9)
    muscle_edge = spine - nrm * (half - muscle_thick)

    body_poly = np.vstack([outer, inner[::-1]])
    muscle_poly = np.vstack([muscle_edge, inner[::-1]])
    self_int = polygon_self_intersects(body_poly) or polygon_self_intersects(muscle_poly)

    raster = grid_res * 2
    lo_x = min(np.min(body_poly[:, 0]), -PAYLOAD_WIDTH)
    hi_x = max(np.max(body_poly[:, 0]), PAYLOAD_WIDTH)
    lo_y = min(np.min(body_poly[:, 1]), -0.05)
    hi_y = max(np.max(body_poly[:, 1]), PAYLOAD_HEIGHT)
    pad = 0.05
    xs = np.linspace(lo_x - pad, hi_x + pad, int((hi_x - lo_x + 2 * pad) * raster))
    ys = np.linspace(lo_y - pad, hi_y + pad, int((hi_y - lo_y + 2 * pad) * raster))
    gx, gy = np.meshgrid(xs, ys)
    cand = np.vstack([gx.ravel(), gy.ravel()]).T

    in_bell = Path(body_poly).contains_points(cand)
    in_muscle = Path(muscle_poly).contains_points(cand)
    payload_box = (cand[:, 0] <= PAYLOAD_WIDTH / 2.0) & (cand[:, 1] <= PAYLOAD_HEIGHT)

    # collar around the payload socket (make_jelly.py:164-223)
    collar_top = PAYLOAD_HEIGHT * 0.50
    collar_thick = t_base * 0.35
    cts = np.linspace(0, 1, 20)
    q0 = np.array([PAYLOAD_WIDTH / 2.0, collar_top])
    q3 = inner[0]
    q1 = np.array([PAYLOAD_WIDTH / 2.0, collar_top - collar_top * 0.4])
    q2 = q3 + np.array([0.0, min(collar_top * 0.3, abs(q3[1] - q0[1]) * 0.4)])
    collar_in = np.array([_bezier(q0, q1, q2, q3, t) for t in cts])
    outer0 = outer[0]
    if outer0[0] > PAYLOAD_WIDTH / 2.0:
        p0 = np.array([PAYLOAD_WIDTH / 2.0 + collar_thick, collar_top])
        p3 = outer0
        p1 = np.array([p0[0], p0[1] - collar_top * 0.5])
        bt = outer[min(2, len(outer) - 1)] - outer[0]
        bt_len = np.linalg.norm(bt)
        p2 = p3 - (bt / bt_len) * (collar_top * 0.6) if bt_len > 1e-10 else p3 + np.array([0, 0.01])
        collar_out = np.array([_bezier(p0, p1, p2, p3, t) for t in cts])
        collar_mus = collar_in + muscle_ratio * (collar_out - collar_in)
        in_collar = (Path(np.vstack([collar_out, collar_in[::-1]])).contains_points(cand)
                     & ~payload_box)
        in_collar_mus = (Path(np.vstack([collar_mus, collar_in[::-1]])).contains_points(cand)
                         & ~payload_box)
        bridge_hw = max(outer0[0], PAYLOAD_WIDTH / 2.0 + collar_thick)
    else:
        in_collar = np.zeros(len(cand), dtype=bool)
        in_collar_mus = np.zeros(len(cand), dtype=bool)
        bridge_hw = PAYLOAD_WIDTH / 2.0 + 0.015

    # transverse bridge (make_jelly.py:225-238)
    if not with_payload:
        bridge_hw = max(bridge_hw, PAYLOAD_WIDTH / 2.0 + 0.02)
    bridge_top = PAYLOAD_HEIGHT if not with_payload else 0.0
    in_bridge = ((cand[:, 0] >= -bridge_hw) & (cand[:, 0] <= bridge_hw)
                 & (cand[:, 1] >= -0.03) & (cand[:, 1] <= bridge_top))

    # materials, later masks win (make_jelly.py:240-248)
    mats = np.zeros(len(cand), dtype=int)
    mats[in_bridge] = 1
    mats[in_collar] = 1
    mats[in_collar_mus] = 3
    mats[in_bell] = 1
    mats[in_bell & in_muscle] = 3

    right = (cand[:, 0] >= 0) & (mats > 0)
    rpos = cand[right]
    rmat = mats[right]
    rfib = np.zeros((len(rpos), 2), dtype=np.float32)
    rfib[:, 1] = 1.0
    if len(rpos) > 0 and np.any(rmat == 3):
        is_mus = rmat == 3
        _, idx = cKDTree(spine).query(rpos[is_mus])
        rfib[is_mus] = spine_tan[idx]

    if axisym:                      # make_jelly.py:276-280
        soft_pos, soft_mat, soft_fib = rpos, rmat, rfib
    elif len(rpos) > 0:             # make_jelly.py:281-291
        lpos = rpos.copy()
        lpos[:, 0] *= -1
        lfib = rfib.copy()
        lfib[:, 0] *= -1
        soft_pos = np.vstack([rpos, lpos])
        soft_mat = np.concatenate([rmat, rmat.copy()])
        soft_fib = np.vstack([rfib, lfib])
    else:
        soft_pos = np.zeros((0, 2))
        soft_mat = np.zeros(0, dtype=int)
        soft_fib = np.zeros((0, 2), dtype=np.float32)

    parts_pos, parts_mat, parts_fib = [], [], []
    offset = np.array(spawn_offset)
    if len(soft_pos) > 0:
        parts_pos.append(soft_pos + offset)
        parts_mat.append(soft_mat)
        parts_fib.append(soft_fib)
    if with_payload:                # make_jelly.py:298-325
        if axisym:
            px = np.linspace(0, PAYLOAD_WIDTH / 2, int(PAYLOAD_WIDTH / 2 * raster))
        else:
            px = np.linspace(-PAYLOAD_WIDTH / 2, PAYLOAD_WIDTH / 2, int(PAYLOAD_WIDTH * raster))
        py = np.linspace(0, PAYLOAD_HEIGHT, int(PAYLOAD_HEIGHT * raster))
        pgx, pgy = np.meshgrid(px, py)
        pay = np.vstack([pgx.ravel(), pgy.ravel()]).T
        if len(pay) > 0:
            parts_pos.append(pay + offset)
            parts_mat.append(np.full(len(pay), 2, dtype=int))
            pf = np.zeros((len(pay), 2), dtype=np.float32)
            pf[:, 1] = 1.0
            parts_fib.append(pf)
    if not parts_pos:
        return np.zeros((0, 2)), np.zeros(0, dtype=int), np.zeros((0, 2), np.float32), True
    return (np.vstack(parts_pos), np.concatenate(parts_mat).astype(int),
            np.vstack(parts_fib).astype(np.float32), self_int)


def fill_tank(genome, max_particles, grid_res=128, spawn_offset=None, water_margin=0.005,
              with_payload=True, domain_height=1.0, axisym=False):
    """make_jelly.py:336-402: (pos (N,2) f32, mat (N,) i32, fiber (N,2) f32, info)."""
    from scipy.spatial import cKDTree
    if spawn_offset is None:
        spawn_offset = AXISYM_SPAWN if axisym else DEFAULT_SPAWN
    robot_pos, robot_mat, robot_fib, self_int = generate_phenotype(
        genome, spawn_offset, grid_res=grid_res, with_payload=with_payload, axisym=axisym)
    n_robot = len(robot_pos)
    spacing = 1.0 / (grid_res * 2.0)
    margin = spacing * 3
    wx = np.arange(margin, 1.0 - margin, spacing)
    wy = np.arange(margin, domain_height - margin, spacing)
    wgx, wgy = np.meshgrid(wx, wy)
    cands = np.vstack([wgx.ravel(), wgy.ravel()]).T
    if n_robot > 0:
        dist, _ = cKDTree(robot_pos).query(cands, k=1)
        water = cands[dist > water_margin]
    else:
        water = cands
    n_water = len(water)
    if n_robot + n_water > max_particles:
        n_water = max_particles - n_robot
    pos = np.full((max_particles, 2), -1.0, dtype=np.float32)
    mat = np.full(max_particles, -1, dtype=np.int32)
    fib = np.zeros((max_particles, 2), dtype=np.float32)
    fib[:, 1] = 1.0
    if n_robot > 0:
        pos[:n_robot] = robot_pos
        mat[:n_robot] = robot_mat
        fib[:n_robot] = robot_fib
    if n_water > 0:
        pos[n_robot:n_robot + n_water] = water[:n_water]
        mat[n_robot:n_robot + n_water] = 0
    return pos, mat, fib, {
        "n_robot": n_robot, "n_water": n_water, "n_total": n_robot + n_water,
        "n_dead": max_particles - (n_robot + n_water),
        "muscle_count": int(np.sum(mat[:n_robot] == 3)),
        "n_jelly": int(np.sum(mat[:n_robot] == 1)), "self_intersecting": self_int,
    }
