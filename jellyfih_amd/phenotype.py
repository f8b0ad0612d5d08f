"""Jellyfish phenotype and water-tank fill: the engine's input generator.

The engine consumes fixed-size particle arrays in the format of the reference's
``make_jelly.fill_tank`` (c-m-x-r/jellyfih ``make_jelly.py:336-402``): ``pos (N,2) f32``,
``mat (N,) i32`` (0 water, 1 jelly, 2 payload, 3 muscle, -1 dead slot), ``fiber (N,2) f32``,
robot particles first, then water, then dead slots at ``(-1,-1)``. ``generate_phenotype``
and ``fill_tank`` keep the reference's names and arguments and run natively
(``jm_generate_phenotype`` / ``jm_fill_tank`` in libjellymx.so, host code): Bezier spine ->
normals -> thickness envelope -> muscle band -> rasterised bell, collar and bridge masks ->
right half mirrored -> payload block -> spawn offset (``make_jelly.py:61-334``), then the
4-per-cell water lattice minus the exclusion zone around the robot, padded with dead slots
(``make_jelly.py:336-402``), bit-identical to the reference's numpy evaluation (checked
against fixtures written from the reference, ``tests/golden``, and against a numpy
restatement over many genomes). ``population`` builds a batch of tanks on host threads.

Axisymmetric mode (``make_jelly.py:11-14``, ``JELLY_AXISYM=1`` at import, or ``axisym=True``
per call): the right half of the profile is the whole cross-section (no mirror), the
payload block spans ``[0, PAYLOAD_WIDTH/2]``, and the default spawn is the axis
``(0, 0.40)`` (``make_jelly.py:276-280,301-304,343``).
"""
from __future__ import annotations

import os

import numpy as np

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


def generate_phenotype(genome, spawn_offset=None, grid_res: int = 128, with_payload: bool = True,
                       axisym: bool | None = None):
    """Robot particles of one genome: ``(pos f64 (n,2), mat int (n,), fiber f32 (n,2),
    self_intersecting)`` (make_jelly.py:74-334), by jm_generate_phenotype. ``axisym`` None
    follows ``AXISYM``."""
    from . import _lib
    _lib.load()
    if axisym is None:
        axisym = AXISYM
    if spawn_offset is None:
        spawn_offset = DEFAULT_SPAWN
    g = np.ascontiguousarray(genome, dtype=np.float64).ravel()
    sp = np.ascontiguousarray(spawn_offset, dtype=np.float64).ravel()
    n = np.zeros(1, np.int64)
    si = np.zeros(1, np.int32)

    def run(cap, checked):
        pos = np.empty((cap, 2), np.float64)
        mat = np.empty(cap, np.int32)
        fib = np.empty((cap, 2), np.float32)
        args = (g.ctypes.data, len(g), sp.ctypes.data, int(grid_res), int(bool(with_payload)),
                int(bool(axisym)), pos.ctypes.data, mat.ctypes.data, fib.ctypes.data, cap,
                n.ctypes.data, si.ctypes.data)
        if checked:
            _lib.call("jm_generate_phenotype", *args)
            return pos, mat, fib
        return (pos, mat, fib) if _lib.load().jm_generate_phenotype(*args) == 0 else None

    # a genome within the evolve.py bounds makes well under 2^16 robot particles; a larger
    # one is told its count by the first call and runs again with room for it
    out = run(1 << 16, False)
    if out is None:
        out = run(max(int(n[0]), 1), True)
    pos, mat, fib = out
    k = int(n[0])
    return pos[:k].copy(), mat[:k].astype(int), fib[:k].copy(), bool(si[0])


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
    """Fixed-size particle arrays for one instance (make_jelly.py:336-402), by jm_fill_tank.

    Returns ``pos (N,2) f32, mat (N,) i32, fiber (N,2) f32, info dict``. ``axisym`` None
    follows ``AXISYM``; the default spawn is then the axis (make_jelly.py:342-343)."""
    from . import _lib
    _lib.load()
    if axisym is None:
        axisym = AXISYM
    g = np.ascontiguousarray(genome, dtype=np.float64).ravel()
    sp = None if spawn_offset is None else np.ascontiguousarray(spawn_offset,
                                                                dtype=np.float64).ravel()
    pos = np.empty((max_particles, 2), np.float32)
    mat = np.empty(max_particles, np.int32)
    fib = np.empty((max_particles, 2), np.float32)
    info = np.zeros(5, np.int64)
    _lib.call("jm_fill_tank", g.ctypes.data, len(g), int(max_particles), int(grid_res),
              None if sp is None else sp.ctypes.data, float(water_margin),
              int(bool(with_payload)), float(domain_height), int(bool(axisym)),
              pos.ctypes.data, mat.ctypes.data, fib.ctypes.data, info.ctypes.data)
    n_robot, n_water, muscle, jelly, self_int = (int(v) for v in info)
    return pos, mat, fib, {
        "n_robot": n_robot, "n_water": n_water, "n_total": n_robot + n_water,
        "n_dead": max_particles - (n_robot + n_water), "muscle_count": muscle,
        "n_jelly": jelly, "self_intersecting": bool(self_int),
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
    genomes = np.stack([seeded_genome(s) for s in seeds]) if M else np.empty((0, 11))

    def one(i):
        # the native call releases the GIL, so tanks build in parallel
        pos[i], mat[i], fib[i], info = fill_tank(genomes[i], max_particles, grid_res,
                                                 spawn_offset, domain_height=domain_height,
                                                 axisym=axisym)
        return info

    from concurrent.futures import ThreadPoolExecutor
    workers = max(1, min(M, 8, len(os.sched_getaffinity(0))))
    with ThreadPoolExecutor(workers) as ex:
        infos = list(ex.map(one, range(M)))
    return pos, mat, fib, genomes, infos
