"""The input generator against fixtures written from the reference's make_jelly.fill_tank
(tests/golden/make_golden.py). Bit-exact: positions, materials, fibres, water mask, counts,
in Cartesian and axisymmetric mode (JELLY_AXISYM=1: right half only, make_jelly.py:276-304)."""
from __future__ import annotations

import os

import numpy as np
import pytest

from jellyfih_amd import phenotype

CASES = ["aurelia", "seed0", "seed1", "seed2", "tall_seed0", "axisym_seed0",
         "axisym_view_aurelia"]


def _fixture_args(golden, name):
    dom = golden[f"{name}/domain"]
    counts = golden[f"{name}/counts"]
    return (golden[f"{name}/genome"], int(counts[0]), None if dom[1] < 0 else dom[1:],
            float(dom[0]), bool(golden[f"{name}/axisym"]))


@pytest.mark.parametrize("name", CASES)
def test_oracle_fill_tank_matches_reference_fixture(golden, name):
    """The numpy restatement (oracle/phenotype.py, the check for the native generator) is
    itself pinned by the reference's outputs."""
    from oracle import phenotype as ref
    g, nmax, spawn, height, axisym = _fixture_args(golden, name)
    pos, mat, fib, info = ref.fill_tank(g, nmax, 128, spawn_offset=spawn,
                                        domain_height=height, axisym=axisym)
    nr = info["n_robot"]
    counts = golden[f"{name}/counts"]
    assert (nr, info["n_water"], info["muscle_count"], info["n_jelly"]) == tuple(
        int(c) for c in counts[1:5])
    np.testing.assert_array_equal(pos[:nr], golden[f"{name}/robot_pos"])
    np.testing.assert_array_equal(mat[:nr], golden[f"{name}/robot_mat"].astype(np.int32))
    np.testing.assert_array_equal(fib[:nr], golden[f"{name}/robot_fiber"])


@pytest.mark.parametrize("mode", ["square", "tall", "axisym", "no_payload"])
def test_native_fill_tank_matches_the_numpy_restatement(mode):
    """jm_fill_tank against the numpy restatement, bit for bit (positions, materials, fibre
    bits, counts), over seeded genomes spanning the evolve.py bounds, in the four fill
    modes the reference's callers use."""
    from oracle import phenotype as ref
    kw = {"square": {}, "tall": {"domain_height": 2.0, "spawn_offset": phenotype.TALL_SPAWN},
          "axisym": {"axisym": True}, "no_payload": {"with_payload": False}}[mode]
    nmax = 160000 if mode == "tall" else 80000
    seeds = range(40) if mode == "square" else range(100, 112)
    for seed in seeds:
        g = phenotype.seeded_genome(seed)
        a = ref.fill_tank(g, nmax, 128, **kw)
        b = phenotype.fill_tank(g, nmax, 128, axisym=kw.get("axisym", False),
                                **{k: v for k, v in kw.items() if k != "axisym"})
        np.testing.assert_array_equal(a[0], b[0])
        np.testing.assert_array_equal(a[1], b[1])
        np.testing.assert_array_equal(a[2].view(np.uint32), b[2].view(np.uint32))
        assert a[3] == b[3], (seed, a[3], b[3])


def test_native_generate_phenotype_matches_the_numpy_restatement():
    from oracle import phenotype as ref
    for seed in range(200, 216):
        g = phenotype.seeded_genome(seed)
        for axisym in (False, True):
            a = ref.generate_phenotype(g, axisym=axisym)
            b = phenotype.generate_phenotype(g, axisym=axisym)
            assert a[0].dtype == b[0].dtype == np.float64
            np.testing.assert_array_equal(a[0], b[0])
            np.testing.assert_array_equal(a[1], b[1])
            np.testing.assert_array_equal(a[2].view(np.uint32), b[2].view(np.uint32))
            assert a[3] == b[3]


def test_population_matches_one_tank_at_a_time():
    pos, mat, fib, genomes, infos = phenotype.population(range(5, 13), 80000, 128)
    for i, s in enumerate(range(5, 13)):
        p1, m1, f1, info = phenotype.fill_tank(phenotype.seeded_genome(s), 80000, 128)
        np.testing.assert_array_equal(pos[i], p1)
        np.testing.assert_array_equal(mat[i], m1)
        np.testing.assert_array_equal(fib[i], f1)
        assert infos[i] == info


def test_fill_tank_rejects_a_robot_larger_than_the_tank():
    with pytest.raises(RuntimeError, match="max_particles"):
        phenotype.fill_tank(phenotype.AURELIA_GENOME, 100, 128)


@pytest.mark.parametrize("name", CASES)
def test_fill_tank_matches_reference_fixture(golden, name):
    g = golden[f"{name}/genome"]
    counts = golden[f"{name}/counts"]
    dom = golden[f"{name}/domain"]
    spawn = None if dom[1] < 0 else dom[1:]
    nmax = int(counts[0])
    axisym = bool(golden[f"{name}/axisym"])
    pos, mat, fib, info = phenotype.fill_tank(g, nmax, 128, spawn_offset=spawn,
                                              domain_height=float(dom[0]), axisym=axisym)
    nr, nw = info["n_robot"], info["n_water"]
    assert (nr, nw, info["muscle_count"], info["n_jelly"]) == tuple(int(c) for c in counts[1:5])
    assert int(info["self_intersecting"]) == int(counts[5])
    np.testing.assert_array_equal(pos[:nr], golden[f"{name}/robot_pos"])
    np.testing.assert_array_equal(mat[:nr], golden[f"{name}/robot_mat"].astype(np.int32))
    np.testing.assert_array_equal(fib[:nr], golden[f"{name}/robot_fiber"])
    cands = phenotype.water_lattice(128, float(dom[0])).astype(np.float32)
    keep = np.unpackbits(golden[f"{name}/water_keep"])[:len(cands)].astype(bool)
    np.testing.assert_array_equal(pos[nr:nr + nw], cands[keep])
    assert np.all(mat[nr:nr + nw] == 0)
    assert np.all(mat[nr + nw:] == -1) and np.all(pos[nr + nw:] == -1)


def test_axisym_mode_from_the_environment():
    """JELLY_AXISYM=1 at import selects the axisymmetric branch by default, as the reference's
    module flag does (make_jelly.py:14,342-343): same robot as axisym=True, spawn at the axis,
    every robot particle at r >= 0, and the payload block only on the right half."""
    import subprocess
    import sys
    code = ("import numpy as np; from jellyfih_amd import phenotype as p; "
            "a = p.fill_tank(p.AURELIA_GENOME, 80000); "
            "b = p.fill_tank(p.AURELIA_GENOME, 80000, axisym=True); "
            "assert p.AXISYM; n = a[3]['n_robot']; "
            "assert n == b[3]['n_robot'] and np.array_equal(a[0], b[0]); "
            "pay = a[0][:n][a[1][:n] == 2]; "
            "assert a[0][:n, 0].min() >= 0 and pay[:, 0].min() == 0 and "
            "abs(pay[:, 0].max() - 0.04) < 1e-7; print('ok')")
    env = dict(os.environ, JELLY_AXISYM="1")
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                       cwd=os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stderr
    full = phenotype.fill_tank(phenotype.AURELIA_GENOME, 80000, axisym=False)
    half = phenotype.fill_tank(phenotype.AURELIA_GENOME, 80000, axisym=True)
    assert half[3]["n_robot"] < full[3]["n_robot"]


def test_seeded_genomes_within_bounds():
    for s in range(32):
        g = phenotype.seeded_genome(s)
        assert g.shape == (11,)
        assert np.all(g >= phenotype.GENOME_LOWER) and np.all(g <= phenotype.GENOME_UPPER)


def test_truncation_when_too_many_particles():
    pos, mat, fib, info = phenotype.fill_tank(phenotype.AURELIA_GENOME, 30000, 128)
    assert info["n_total"] == 30000 and info["n_dead"] == 0
    assert np.all(mat >= 0)


def test_self_intersection_test():
    square = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], float)
    bowtie = np.array([[0, 0], [1, 1], [1, 0], [0, 1]], float)
    assert not phenotype.polygon_self_intersects(square)
    assert phenotype.polygon_self_intersects(bowtie)


def test_water_mask_matches_kdtree():
    """The native water exclusion equals the k-d tree test it replaces
    (make_jelly.py:359-366), including candidates at exactly the margin."""
    from scipy.spatial import cKDTree
    from jellyfih_amd.phenotype import water_mask
    rng = np.random.default_rng(3)
    for n_body, spread in ((1, 0.2), (50, 0.05), (3000, 0.3)):
        body = rng.uniform(0.35, 0.35 + spread, size=(n_body, 2))
        cands = np.concatenate([rng.uniform(0.2, 0.8, size=(20000, 2)),
                                body + [0.005, 0.0], body - [0.0, 0.005],
                                body + rng.normal(size=body.shape) * 0.004])
        ref = cKDTree(body).query(cands, k=1)[0] > 0.005
        np.testing.assert_array_equal(water_mask(body, cands, 0.005), ref)
    assert water_mask(np.zeros((0, 2)), cands[:10], 0.005).all()


def _segments_cross_py(a, b, c, d):
    """make_jelly.py:31-43 restated in Python (the check for the native version)."""
    cross = lambda u, v: u[0] * v[1] - u[1] * v[0]
    r, s = b - a, d - c
    rxs = cross(r, s)
    if abs(rxs) < 1e-12:
        return False
    qp = c - a
    t = cross(qp, s) / rxs
    u = cross(qp, r) / rxs
    return 0 < t < 1 and 0 < u < 1


def _self_intersects_py(poly):
    """make_jelly.py:46-58 restated in Python."""
    n = len(poly)
    for i in range(n):
        for j in range(i + 2, n):
            if i == 0 and j == n - 1:
                continue
            if _segments_cross_py(poly[i], poly[(i + 1) % n], poly[j], poly[(j + 1) % n]):
                return True
    return False


def test_polygon_self_intersects_matches_python():
    from jellyfih_amd.phenotype import polygon_self_intersects
    rng = np.random.default_rng(4)
    hits = 0
    for trial in range(300):
        n = int(rng.integers(3, 40))
        ang = np.sort(rng.uniform(0, 2 * np.pi, n))
        rad = rng.uniform(0.5, 1.0, n) if trial % 2 else rng.uniform(0.0, 1.0, n)
        poly = np.stack([rad * np.cos(ang), rad * np.sin(ang)], 1)
        if trial % 3 == 0:
            poly = rng.permutation(poly)          # mostly self-intersecting
        ref = _self_intersects_py(poly)
        hits += ref
        assert polygon_self_intersects(poly) == ref, trial
    assert 0 < hits < 300
    square = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]])
    bow = np.array([[0.0, 0.0], [1.0, 1.0], [1.0, 0.0], [0.0, 1.0]])
    assert not polygon_self_intersects(square) and polygon_self_intersects(bow)


def test_points_in_polygon_grid_matches_matplotlib():
    """The native lattice fill against matplotlib's Path.contains_points (the reference's
    call, make_jelly.py:150-157,209-215) bit for bit: star-shaped and self-intersecting
    polygons, vertices and edges on lattice lines, lattices of 1 to 90 points per side."""
    from matplotlib.path import Path
    rng = np.random.default_rng(0)
    total = 0
    for trial in range(200):
        n = int(rng.integers(3, 60))
        ang = np.sort(rng.uniform(0, 2 * np.pi, n))
        r = rng.uniform(0.2, 1, n)
        poly = np.stack([r * np.cos(ang), r * np.sin(ang)], 1)
        if trial % 3 == 0:
            poly = rng.uniform(-1, 1, (n, 2))
        if trial % 5 == 0:
            poly = np.round(poly * 8) / 8
        xs = np.linspace(-1.1, 1.1, int(rng.integers(1, 90)))
        ys = np.linspace(-1.1, 1.1, int(rng.integers(1, 90)))
        if trial % 7 == 0:
            xs, ys = np.unique(np.round(xs * 8) / 8), np.unique(np.round(ys * 8) / 8)
        gx, gy = np.meshgrid(xs, ys)
        ref = Path(poly).contains_points(np.vstack([gx.ravel(), gy.ravel()]).T)
        np.testing.assert_array_equal(phenotype.points_in_polygon_grid(poly, xs, ys), ref)
        total += ref.size
    assert total > 100000
