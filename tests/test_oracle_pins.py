"""Pins of the oracle parts that the substep tests do not reach: the render-only glow pass
(mpm_sim.py:335-338), the material constants (mpm_sim.py:24-48) and the jelly Young's
modulus map, and the host logic around the stats (mpm_sim.py:793-800, 821-834).

The glow pass is pinned by its closed form: with Jp held fixed, glow after n passes is
max(0.9^n g0, |1 - Jp|), with the switch from decay to the floor where 0.9^n g0 falls
below |1 - Jp|; with Jp changing, glow_n = max_k 0.9^(n-k) |1 - Jp_k|. The constants are
pinned against the reference's own numbers (mu, lambda to the digits mpm_sim.py's
comments and the reference's docs state them), so a wrong E or nu fails here.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle


def _glow_sim(Jp, glow0):
    n = len(Jp)
    s = oracle.OracleSim(1, n)
    s.material[0] = 0
    s.x[0] = 0.5
    s.Jp[0] = np.asarray(Jp, np.float32)
    s.glow[0] = np.asarray(glow0, np.float32)
    return s


def test_glow_decay_closed_form():
    # |1 - Jp| = 0 everywhere: pure decay g0 * 0.9^n, rounded once per pass in fp32
    g0 = np.array([1.0, 0.5, 3.0e-3, 7.25, 1e-30], np.float32)
    s = _glow_sim(np.ones(5), g0)
    for n in range(1, 61):
        s.run(1, oracle.ST_GLOW)
        exp = g0.astype(np.float64) * 0.9 ** n
        rel = np.abs(s.glow[0] - exp) / exp
        assert np.all(rel < (n + 1) * 2.0 ** -23), (n, rel)


def test_glow_floor_and_switch():
    # glow = max(0.9^n g0, |1 - Jp|): it decays while above the floor, then sits on it
    # exactly (the fp32 value of |1 - Jp|)
    Jp = np.array([0.9, 1.05, 1.0 - 0.25, 2.0, 1.0], np.float32)
    g0 = np.array([0.5, 0.5, 0.1, 0.0, 0.2], np.float32)
    floor = np.abs(np.float32(1.0) - Jp).astype(np.float32)
    s = _glow_sim(Jp, g0)
    switched = np.zeros(5, bool)
    for n in range(1, 80):
        s.run(1, oracle.ST_GLOW)
        decay = g0.astype(np.float64) * 0.9 ** n
        above = decay > floor.astype(np.float64) * (1 + 1e-5)
        below = decay < floor.astype(np.float64) * (1 - 1e-5)
        g = s.glow[0]
        assert np.all(g >= floor)
        np.testing.assert_array_equal(g[below], floor[below])
        assert np.all(np.abs(g[above] - decay[above]) <= decay[above] * (n + 1) * 2.0 ** -23)
        switched |= below
    # the first instance switches from decay to its floor 0.1 after 16 passes
    assert switched[:3].all() and switched[3]
    # a particle at Jp = 1 with glow 0.2 decays towards 0 (floor 0)
    assert 0.0 < s.glow[0, 4] < 0.2 * 0.9 ** 70


def test_glow_tracks_largest_past_floor():
    # Jp changing between passes: glow_n = max_k 0.9^(n-k) |1 - Jp_k| (k = 1..n), g0 = 0
    rng = np.random.default_rng(7)
    n_p, steps = 16, 40
    jps = 1.0 + rng.uniform(-0.3, 0.3, (steps, n_p)).astype(np.float32)
    jps[rng.uniform(size=(steps, n_p)) < 0.5] = 1.0
    s = _glow_sim(np.ones(n_p), np.zeros(n_p))
    for k in range(steps):
        s.Jp[0] = jps[k]
        s.run(1, oracle.ST_GLOW)
    dev = np.abs(1.0 - jps.astype(np.float64))
    exp = np.max(dev * 0.9 ** (steps - 1 - np.arange(steps))[:, None], axis=0)
    np.testing.assert_allclose(s.glow[0], exp, rtol=steps * 2.0 ** -22, atol=0)


def test_glow_reads_jp_after_the_substep():
    # In a full substep the pass comes after P2G has updated Jp (mpm_sim.py:216 then
    # :335-338): a stiff payload under strain gets glow = max(0.9 g, |1 - Jp_new|) with the
    # new Jp of that same substep.
    from tests.cases import small_tank
    p, mt, f = small_tank(3, 6000, region=(0.3, 0.6, 0.2, 0.5))
    s = oracle.OracleSim(1, 6000)
    s.load(0, p, mt, f)
    assert (mt == 2).sum() > 20
    s.C[0][mt == 2] = np.array([[300.0, 100.0], [-50.0, -250.0]], np.float32)
    s.t[0] = np.float32(0.05)
    for _ in range(3):
        g_prev = s.glow[0].copy()
        s.run(1, oracle.ST_SUBSTEP | oracle.ST_GLOW)
        exp = np.maximum(g_prev * np.float32(0.9), np.abs(np.float32(1) - s.Jp[0]))
        np.testing.assert_array_equal(s.glow[0], exp)
    assert np.max(s.glow[0][mt == 2]) > 1e-3  # the payload has deformed plastically


# --- constants -------------------------------------------------------------------------------
# The reference's numbers (mpm_sim.py:24-48 evaluated; the values its docs quote):
# base (416.67, 277.78), jelly/muscle E = 700, nu = 0.3 -> (269.23, 403.85),
# payload E = 2e5, nu = 0.2 -> (83333.3, 55555.6), water lambda 1e5.
REF = {"mu_base": 416.67, "la_base": 277.78, "mu_payload": 83333.3, "la_payload": 55555.6}


def test_reference_constants():
    c = oracle.reference_consts()
    for name, val in REF.items():
        assert abs(getattr(c, name) - val) <= 2e-5 * val, (name, getattr(c, name))
    assert c.water_la == 1e5
    assert c.gravity == 10.0 and c.act_freq == 1.0
    assert c.nx == 128 and c.ny == 128
    assert c.dx == 1 / 128 and c.inv_dx == 128.0
    assert c.dt == 2e-5 and c.p_vol == (1 / 256) ** 2 and c.p_mass == c.p_vol
    t = oracle.reference_consts(n_grid_y=256, domain_height=2.0, quality=2)
    assert t.nx == 256 and t.ny == 256 and t.dt == 1e-5 and t.domain_height == 2.0


def test_jelly_lame_map():
    mu, la = oracle.jelly_lame(700.0)
    assert abs(mu - 269.23) <= 2e-5 * 269.23 and abs(la - 403.85) <= 2e-5 * 403.85
    # E != 700 (the engine's runtime E_jelly): mu = E / (2 * 1.3), la = E * 0.3 / (1.3 * 0.4)
    E = np.array([100.0, 350.0, 1400.0, 2500.0, 1e4])
    mu, la = oracle.jelly_lame(E)
    np.testing.assert_allclose(mu, E / 2.6, rtol=1e-14)
    np.testing.assert_allclose(la, E * 0.3 / 0.52, rtol=1e-14)
    # an oracle instance takes the map of its E_jelly, rounded once to its precision
    s = oracle.OracleSim(3, 1)
    s.set_e_jelly([700.0, 1400.0, 350.0])
    np.testing.assert_array_equal(s.mu_j, np.float32([700 / 2.6, 1400 / 2.6, 350 / 2.6]))
    np.testing.assert_array_equal(s.la_j, np.float32([700 * 0.3 / 0.52, 1400 * 0.3 / 0.52,
                                                      350 * 0.3 / 0.52]))


# --- host logic around the stats -------------------------------------------------------------
def test_finish_stats_divides_in_fp32():
    # raw is the fp32 fitness buffer; raw[i, 0] / count is an fp32 quotient
    raw = np.array([[1.0, 3.0, 2.0], [250.17, 241.0, 120.3], [0.0, 0.0, 0.0]], np.float32)
    st = oracle.finish_stats(raw)
    assert st[0, 0] == float(np.float32(1.0) / np.float32(3.0))
    assert st[0, 0] != 1.0 / 3.0  # the double quotient differs in the last bits
    assert st[1, 1] == float(np.float32(120.3) / np.float32(241.0))
    assert st[0, 2] == 3.0 and np.all(st[2] == 0.0)


def test_nan_com_validity():
    init = np.array([[0.4, 0.5, 240.0], [0.4, 0.5, 240.0], [0.4, 0.5, 240.0]])
    final = np.array([[np.nan, 0.5, 240.0], [0.005, 0.5, 240.0], [0.45, np.inf, 240.0]])
    ref = oracle.assemble_results(init, final)
    # the reference's gate: NaN < 0.01 is False, so a NaN CoM counts as valid (H12)
    assert ref[0, 4] == 1.0 and ref[1, 4] == 0.0 and ref[2, 4] == 1.0
    opt = oracle.assemble_results(init, final, nan_invalid=True)
    assert opt[0, 4] == 0.0 and opt[1, 4] == 0.0 and opt[2, 4] == 0.0
    np.testing.assert_array_equal(opt[:, :4], ref[:, :4])
    # a diverged instance (non-finite particle state, finite CoM) is invalid with the flag
    fin = np.array([[0.45, 0.5, 240.0]] * 3)
    div = oracle.assemble_results(init, fin, nan_invalid=True, diverged=[False, True, False])
    np.testing.assert_array_equal(div[:, 4], [1.0, 0.0, 1.0])
    np.testing.assert_array_equal(oracle.assemble_results(init, fin, diverged=[0, 1, 0])[:, 4],
                                  [1.0, 1.0, 1.0])


def test_nan_velocity_clamps_into_the_corner():
    # A NaN velocity makes x NaN, and the clamp ti.max(ti.min(x, hi), lo) returns hi
    # (mpm_sim.py:331-333): the particle lands in the top-right corner with a finite
    # position, so the instance's CoM stays finite while its state has diverged.
    from tests.cases import small_tank
    p, mt, f = small_tank(5, 6000, region=(0.3, 0.6, 0.2, 0.5))
    s = oracle.OracleSim(1, 6000)
    s.load(0, p, mt, f)
    live = np.flatnonzero(mt >= 0)
    bad = live[[10, 500, 1500]]
    s.v[0, bad] = np.nan
    s.run(3)
    np.testing.assert_array_equal(s.x[0, bad], np.float32([[0.999, 0.999]] * 3))
    assert np.isfinite(s.x).all() and s.diverged()[0]
    # the NaN reached the grid nodes of those particles and the particles gathering them
    nanv = np.isnan(s.v[0]).any(-1)
    assert nanv[bad].all() and nanv.sum() > 3


def test_fitness_penalises_invalid():
    # evolve.py:90-124: invalid instances get max(valid) + 1
    res = np.array([[0.4, 0.5, 0.5, 0.5, 1.0], [0.4, 0.5, np.nan, 0.5, 0.0],
                    [0.4, 0.5, 0.45, 0.5, 1.0]])
    fit = oracle.compute_fitness(res, [500, 500, 500], mode="displacement")
    assert fit[0] == pytest.approx(-0.1) and fit[2] == pytest.approx(-0.05)
    assert fit[1] == pytest.approx(-0.05 + 1.0)
