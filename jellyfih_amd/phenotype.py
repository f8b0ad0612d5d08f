"""Jellyfish phenotype and water-tank fill: the engine's input generator.

The engine consumes fixed-size particle arrays in the format of the reference's
``make_jelly.fill_tank`` (c-m-x-r/jellyfih ``make_jelly.py:336-402``): ``pos (N,2) f32``,
``mat (N,) i32`` (0 water, 1 jelly, 2 payload, 3 muscle, -1 dead slot), ``fiber (N,2) f32``,
robot particles first, then water, then dead slots at ``(-1,-1)``. The reference's own
generator stays on the host of whoever drives the engine; this module restates it so
that benchmarks and tests can build identical inputs where the reference is absent.
It is checked bit-for-bit against fixtures written from the reference
(``tests/golden``).

Steps follow ``make_jelly.py:74-334`` (``generate_phenotype``):
Bezier spine -> normals -> thickness envelope -> muscle band -> rasterised bell, collar,
bridge masks -> right half mirrored -> payload block -> spawn offset; and
``make_jelly.py:336-402`` (``fill_tank``): 4-per-cell water lattice minus a KD-tree
exclusion zone around the robot, padded with dead slots.

Axisymmetric mode (``make_jelly.py:11-14``, ``JELLY_AXISYM=1`` at import, or ``axisym=True``
per call): the right half of the profile is the whole cross-section (no mirror), the
payload block spans ``[0, PAYLOAD_WIDTH/2]``, and the default spawn is the axis
``(0, 0.40)`` (``make_jelly.py:276-280,301-304,343``).
"""
from __future__ import annotations

import os

import numpy as np
from scipy.spatial import cKDTree

PAYLOAD_WIDTH = 0.08                    # make_jelly.py:8
PAYLOAD_HEIGHT = 0.05                   # make_jelly.py:9
DEFAULT_SPAWN = np.array([0.5, 0.40])   # make_jelly.py:10
AXISYM_SPAWN = np.array([0.0, 0.40])    # make_jelly.py:11
AXISYM = os.environ.get("JELLY_AXISYM", "0") == "1"   # make_jelly.py:14
TALL_SPAWN = np.array([0.5, 0.60])      # evolve.py:136

# Genome bounds (evolve.py:58-59): 9 morphology genes, contraction_frac, freq_mult.
GENOME_LOWER = np.array([0.0, -0.15, 0.0, -0.2, 0.05, -0.30, 0.025, 0.025, 0.01, 0.05, 0.5])
GENOME_UPPER = np.array([0.25, 0.15, 0.3, 0.15, 0.35, +0.10, 0.08, 0.1, 0.04, 0.60, 2.0])

# make_jelly.py:19-29
AURELIA_GENOME = np.array([0.05, 0.04, 0.18, -0.03, 0.22, -0.12, 0.04, 0.05, 0.015])


def seeded_genome(seed: int) -> np.ndarray:
    """11-gene genome drawn uniformly within the evolve.py bounds from ``default_rng(seed)``."""
    return np.random.default_rng(seed).uniform(GENOME_LOWER, GENOME_UPPER)


def _bezier(p0, p1, p2, p3, t):
    # make_jelly.py:61-63 (same operation order, so results are bit-identical)
    return (1 - t) ** 3 * p0 + 3 * (1 - t) ** 2 * t * p1 + 3 * (1 - t) * t ** 2 * p2 + t ** 3 * p3


def polygon_self_intersects(poly: np.ndarray) -> bool:
    """Any proper crossing between non-adjacent edges of a closed polygon
    (make_jelly.py:31-58), computed by jm_polygon_self_intersects in libjellymx.so."""
    from . import _lib
    _lib.load()
    poly = np.ascontiguousarray(poly, dtype=np.float64).reshape(-1, 2)
    out = np.zeros(1, dtype=np.int32)
    _lib.call("jm_polygon_self_intersects", poly.ctypes.data, len(poly), out.ctypes.data)
    return bool(out[0])


def points_in_polygon_grid(poly, xs, ys) -> np.ndarray:
    """Inside flags of the meshgrid of ascending ``xs`` and ``ys`` (ravel order: rows of
    ys), as ``Path(poly).contains_points`` decides them (make_jelly.py:150-157,209-215),
    computed row by row by jm_points_in_polygon_grid in libjellymx.so."""
    from . import _lib
    _lib.load()
    poly = np.ascontiguousarray(poly, dtype=np.float64).reshape(-1, 2)
    xs = np.ascontiguousarray(xs, dtype=np.float64)
    ys = np.ascontiguousarray(ys, dtype=np.float64)
    out = np.empty(len(xs) * len(ys), dtype=np.uint8)
    _lib.call("jm_points_in_polygon_grid", poly.ctypes.data, len(poly), xs.ctypes.data,
              len(xs), ys.ctypes.data, len(ys), out.ctypes.data)
    return out.astype(bool)


def _unit_normals(points: np.ndarray) -> np.ndarray:
    # make_jelly.py:65-73
    tan = np.gradient(points, axis=0)
    nrm = np.zeros_like(tan)
    nrm[:, 0] = -tan[:, 1]
    nrm[:, 1] = tan[:, 0]
    length = np.linalg.norm(nrm, axis=1, keepdims=True)
    length[length == 0] = 1
    return nrm / length


def generate_phenotype(genome, spawn_offset=None, grid_res: int = 128, with_payload: bool = True,
                       axisym: bool | None = None):
    """Robot particles of one genome: ``(pos f64 (n,2), mat int (n,), fiber f32 (n,2),
    self_intersecting)`` (make_jelly.py:74-334). ``axisym`` None follows ``AXISYM``."""
    if axisym is None:
        axisym = AXISYM
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

    in_bell = points_in_polygon_grid(body_poly, xs, ys)
    in_muscle = points_in_polygon_grid(muscle_poly, xs, ys)
    payload_box = (cand[:, 0] <= PAYLOAD_WIDTH / 2.0) & (cand[:, 1] <= PAYLOAD_HEIGHT)

    # Collar around the payload socket (make_jelly.py:164-223)
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
        collar_poly = np.vstack([collar_out, collar_in[::-1]])
        collar_mus = collar_in + muscle_ratio * (collar_out - collar_in)
        collar_mus_poly = np.vstack([collar_mus, collar_in[::-1]])
        in_collar = points_in_polygon_grid(collar_poly, xs, ys) & ~payload_box
        in_collar_mus = points_in_polygon_grid(collar_mus_poly, xs, ys) & ~payload_box
        bridge_hw = max(outer0[0], PAYLOAD_WIDTH / 2.0 + collar_thick)
    else:
        in_collar = np.zeros(len(cand), dtype=bool)
        in_collar_mus = np.zeros(len(cand), dtype=bool)
        bridge_hw = PAYLOAD_WIDTH / 2.0 + 0.015

    # Transverse bridge (make_jelly.py:225-238)
    if not with_payload:
        bridge_hw = max(bridge_hw, PAYLOAD_WIDTH / 2.0 + 0.02)
    bridge_top = PAYLOAD_HEIGHT if not with_payload else 0.0
    in_bridge = ((cand[:, 0] >= -bridge_hw) & (cand[:, 0] <= bridge_hw)
                 & (cand[:, 1] >= -0.03) & (cand[:, 1] <= bridge_top))

    # Material assignment, later masks win (make_jelly.py:240-248)
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

    if axisym:                      # the right half is the cross-section (make_jelly.py:276-280)
        soft_pos, soft_mat, soft_fib = rpos, rmat, rfib
    elif len(rpos) > 0:
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
    if with_payload:
        if axisym:                  # make_jelly.py:301-302
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


def water_lattice(grid_res: int = 128, domain_height: float = 1.0) -> np.ndarray:
    """Candidate water positions: 4 per cell, 3-spacing margin (make_jelly.py:349-357)."""
    spacing = 1.0 / (grid_res * 2.0)
    margin = spacing * 3
    wx = np.arange(margin, 1.0 - margin, spacing)
    wy = np.arange(margin, domain_height - margin, spacing)
    wgx, wgy = np.meshgrid(wx, wy)
    return np.vstack([wgx.ravel(), wgy.ravel()]).T


def water_mask(body, cands, margin: float) -> np.ndarray:
    """Candidates farther than `margin` from every body point: the nearest-neighbour
    distance test of make_jelly.py:359-366 (a k-d tree query, then `distance > margin`),
    computed by jm_water_mask in libjellymx.so (cell buckets, same distance arithmetic)."""
    from . import _lib
    _lib.load()
    body = np.ascontiguousarray(body, dtype=np.float64).reshape(-1, 2)
    cands = np.ascontiguousarray(cands, dtype=np.float64).reshape(-1, 2)
    keep = np.empty(len(cands), dtype=np.uint8)
    _lib.call("jm_water_mask", body.ctypes.data, len(body), cands.ctypes.data, len(cands),
              float(margin), keep.ctypes.data)
    return keep.astype(bool)


def fill_tank(genome, max_particles: int, grid_res: int = 128, spawn_offset=None,
              water_margin: float = 0.005, with_payload: bool = True,
              domain_height: float = 1.0, axisym: bool | None = None):
    """Fixed-size particle arrays for one instance (make_jelly.py:336-402).

    Returns ``pos (N,2) f32, mat (N,) i32, fiber (N,2) f32, info dict``. ``axisym`` None
    follows ``AXISYM``; the default spawn is then the axis (make_jelly.py:342-343)."""
    if axisym is None:
        axisym = AXISYM
    if spawn_offset is None:
        spawn_offset = AXISYM_SPAWN if axisym else DEFAULT_SPAWN
    robot_pos, robot_mat, robot_fib, self_int = generate_phenotype(
        genome, spawn_offset, grid_res=grid_res, with_payload=with_payload, axisym=axisym)
    n_robot = len(robot_pos)
    cands = water_lattice(grid_res, domain_height)
    if n_robot > 0:
        water = cands[water_mask(robot_pos, cands, water_margin)]
    else:
        water = cands
    n_water = len(water)
    if n_robot + n_water > max_particles:
        n_water = max_particles - n_robot     # truncate water to fit (make_jelly.py:371-374)
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
    muscle = int(np.sum(mat[:n_robot] == 3))
    return pos, mat, fib, {
        "n_robot": n_robot, "n_water": n_water, "n_total": n_robot + n_water,
        "n_dead": max_particles - (n_robot + n_water), "muscle_count": muscle,
        "n_jelly": int(np.sum(mat[:n_robot] == 1)), "self_intersecting": self_int,
    }


def population(seeds, max_particles: int = 80000, grid_res: int = 128,
               domain_height: float = 1.0, spawn_offset=None, axisym: bool | None = None):
    """Fill one tank per seeded genome; returns (pos (M,N,2), mat (M,N), fiber (M,N,2),
    genomes (M,11), infos)."""
    seeds = list(seeds)
    M = len(seeds)
    pos = np.empty((M, max_particles, 2), np.float32)
    mat = np.empty((M, max_particles), np.int32)
    fib = np.empty((M, max_particles, 2), np.float32)
    genomes = np.empty((M, 11))
    infos = []
    for i, s in enumerate(seeds):
        g = seeded_genome(s)
        genomes[i] = g
        pos[i], mat[i], fib[i], info = fill_tank(g, max_particles, grid_res, spawn_offset,
                                                 domain_height=domain_height, axisym=axisym)
        infos.append(info)
    return pos, mat, fib, genomes, infos
