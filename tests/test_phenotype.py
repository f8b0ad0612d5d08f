"""The input generator against fixtures written from the reference's make_jelly.fill_tank
(tests/golden/make_golden.py). Bit-exact: positions, materials, fibres, water mask, counts."""
from __future__ import annotations

import numpy as np
import pytest

from jellyfih_amd import phenotype

CASES = ["aurelia", "seed0", "seed1", "seed2", "tall_seed0"]


@pytest.mark.parametrize("name", CASES)
def test_fill_tank_matches_reference_fixture(golden, name):
    g = golden[f"{name}/genome"]
    counts = golden[f"{name}/counts"]
    dom = golden[f"{name}/domain"]
    spawn = None if dom[1] < 0 else dom[1:]
    nmax = int(counts[0])
    pos, mat, fib, info = phenotype.fill_tank(g, nmax, 128, spawn_offset=spawn,
                                              domain_height=float(dom[0]))
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


def test_seeded_genomes_within_bounds():
    for s in range(32):
        g = phenotype.seeded_genome(s)
        assert g.shape == (11,)
        assert np.all(g >= phenotype.GENOME_LOWER) and np.all(g <= phenotype.GENOME_UPPER)


def test_truncation_when_too_many_particles():
    pos, mat, fib, info = phenotype.fill_tank(phenotype.AURELIA_GENOME, 30000, 128)
    assert info["n_total"] == 30000 and info["n_dead"] == 0
    assert np.all(mat >= 0)


def test_self_intersection_checker():
    square = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], float)
    bowtie = np.array([[0, 0], [1, 1], [1, 0], [0, 1]], float)
    assert not phenotype.polygon_self_intersects(square)
    assert phenotype.polygon_self_intersects(bowtie)
