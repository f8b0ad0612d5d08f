"""Pins of the oracle's axisymmetric variant (JELLY_AXISYM=1, mpm_sim.py:127-133,163-165,
236-304): the whole substep against the independent fp64 model (oracle/brute.py) on small
states that reach the axis, and what the mathematics fixes: water's hoop term cancels,
ring weighting multiplies mass but not the stress momentum, the hoop force of a stretched
ring at rest, and the axis condition replacing the left wall clamp."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from oracle import brute


@pytest.fixture
def c():
    return oracle.reference_consts()


def _instance(rng, n, c, mats=(0, 1, 2, 3), lo=0.004, hi=0.6, dtype=np.float64,
              axisym=True, stretch=0.1):
    s = oracle.OracleSim(1, n, c, dtype=dtype, axisym=axisym)
    pos = np.stack([rng.uniform(lo, hi, n), rng.uniform(0.2, 0.8, n)], 1)
    mat = rng.choice(mats, size=n).astype(np.int32)
    ang = rng.uniform(0, 2 * np.pi, n)
    s.load(0, pos, mat, np.stack([np.cos(ang), np.sin(ang)], 1))
    s.x[0] = pos
    s.r_ref[0] = pos[:, 0] * (1 + rng.uniform(-stretch, stretch, n))   # hoop stretch != 1
    s.v[0] = rng.normal(size=(n, 2)) * 0.2
    s.C[0] = rng.normal(size=(n, 2, 2)) * 0.5
    s.F[0] = np.eye(2) + rng.normal(size=(n, 2, 2)) * 0.03
    s.t[0] = 0.07
    return s


def _state(s):
    return {"x": s.x[0].copy(), "v": s.v[0].copy(), "C": s.C[0].copy(), "F": s.F[0].copy(),
            "Jp": s.Jp[0].copy(), "mat": s.material[0].copy(), "fiber": s.fiber[0].copy(),
            "t": float(s.t[0]), "r_ref": s.r_ref[0].copy()}


def _act(s):
    return {"amp": 500.0, "contraction": 0.2, "freq": 1.0, "pdens": 2.5,
            "mu_j": float(s.mu_j[0]), "la_j": float(s.la_j[0])}


@pytest.mark.parametrize("seed,n", [(0, 1), (1, 7), (2, 30), (3, 60)])
def test_axisym_substep_matches_brute_force(c, seed, n):
    rng = np.random.default_rng(100 + seed)
    s = _instance(rng, n, c)
    if n >= 7:
        s.x[0, :2, 0] = [0.0045, 0.009]          # base cell 0: nodes on the axis (i = 0, 1)
    ref = brute.substep(_state(s), c, _act(s), axisym=True)
    s.run(1)
    for k in ("x", "v", "C", "F", "Jp"):
        np.testing.assert_allclose(getattr(s, k)[0], ref[k], rtol=1e-9, atol=1e-12, err_msg=k)
    np.testing.assert_allclose(s.grid_m[0], ref["grid_m"], rtol=1e-12, atol=1e-22)
    np.testing.assert_allclose(s.grid_v[0], ref["grid_v"], rtol=1e-9, atol=1e-14)


def test_axisym_brute_force_sees_the_hoop_term(c):
    """The comparison is sensitive to the hoop term: brute force given r_ref 5% larger (a
    different circumferential stretch) disagrees with the oracle."""
    rng = np.random.default_rng(7)
    s = _instance(rng, 20, c, mats=(1,), lo=0.2, hi=0.4, stretch=0.0)
    st = _state(s)
    st["r_ref"] = st["r_ref"] * 1.05
    ref = brute.substep(st, c, _act(s), axisym=True)
    s.run(1)
    assert not np.allclose(s.grid_v[0], ref["grid_v"], rtol=1e-6, atol=1e-14)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_axisym_water_hoop_term_cancels(c, dtype):
    """Water has mu = 0, so sigma_rr = sigma_tt = la J (J - 1) for every particle: the two
    stress grids agree bit for bit and the hoop correction is exactly zero."""
    rng = np.random.default_rng(8)
    s = _instance(rng, 400, c, mats=(0,), lo=0.05, hi=0.9, dtype=dtype)
    s.run(1, stages=oracle.ST_RESET | oracle.ST_P2G)
    np.testing.assert_array_equal(s.grid_srr, s.grid_shoop)
    assert np.abs(s.grid_srr).max() > 0


def test_axisym_ring_weighting(c):
    """Particles on one line x = r0 at rest: the ring weighting multiplies every node mass by
    r0 and leaves the stress part of the momentum as in the planar substep."""
    rng = np.random.default_rng(9)
    n, r0 = 40, 0.3125
    out = {}
    pos = np.stack([np.full(n, r0), rng.uniform(0.3, 0.7, n)], 1)
    for ax in (False, True):
        s = oracle.OracleSim(1, n, c, dtype=np.float64, axisym=ax)
        s.load(0, pos, np.ones(n, np.int32))
        s.x[0] = pos
        s.r_ref[0] = r0
        s.F[0] = np.eye(2) + np.random.default_rng(10).normal(size=(n, 2, 2)) * 0.05
        s.run(1, stages=oracle.ST_RESET | oracle.ST_P2G)
        out[ax] = (s.grid_m[0].copy(), s.grid_v[0].copy())
    np.testing.assert_allclose(out[True][0], r0 * out[False][0], rtol=1e-15, atol=0)
    np.testing.assert_array_equal(out[True][1], out[False][1])
    assert np.abs(out[True][1]).max() > 0


def test_axisym_hoop_force_of_a_stretched_ring(c):
    """One jelly particle at rest with F = I on a ring stretched by lam = r / r_ref: sigma_rr
    = 0, sigma_tt = 2 mu (lam - 1), so every node it reaches gets the radial velocity
    dt * (-2 mu (lam - 1) / r) / r_node (the damping band and walls are far away)."""
    lam, r = 1.1, 0.3
    s = oracle.OracleSim(1, 1, c, dtype=np.float64, axisym=True)
    s.load(0, np.array([[r, 0.5]]), np.ones(1, np.int32))
    s.x[0] = [[r, 0.5]]
    s.r_ref[0] = r / lam
    s.run(1, stages=oracle.ST_RESET | oracle.ST_P2G | oracle.ST_GRIDOP)
    mu = float(s.mu_j[0])
    i, j = np.nonzero(s.grid_m[0] > 0)
    assert len(i) == 9
    r_node = (i + 0.5) * c.dx
    want = c.dt * (-2 * mu * (lam - 1) / r) / r_node
    np.testing.assert_allclose(s.grid_v[0][i, j, 0], want, rtol=1e-12)
    np.testing.assert_array_equal(s.grid_v[0][i, j, 1], 0.0)
    assert (want < 0).all()                     # a stretched ring pulls toward the axis


def test_axisym_axis_condition_replaces_the_left_wall(c):
    """A particle moving toward x = 0 next to it: the planar grid op clamps v_x < 0 on nodes
    i < 3; the axisymmetric one zeroes v_x on the axis node i = 0 only."""
    out = {}
    for ax in (False, True):
        s = oracle.OracleSim(1, 1, c, dtype=np.float64, axisym=ax)
        s.load(0, np.array([[1.8 * c.dx, 0.5]]), np.zeros(1, np.int32))
        s.x[0] = [[1.8 * c.dx, 0.5]]
        s.v[0] = [[-1.0, 0.0]]
        s.run(1, stages=oracle.ST_RESET | oracle.ST_P2G | oracle.ST_GRIDOP)
        out[ax] = s.grid_v[0][:, 64, 0].copy()
    planar, axi = out[False], out[True]
    assert (planar[:3] == 0).all()
    assert axi[0] == 0 and (axi[1:3] < 0).all()
