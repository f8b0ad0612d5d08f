"""The adjoint math (jellyfih_amd/csrc/adjoint.h), compiled for the host in fp64
(tests/adjoint_harness.cpp), against:

* the oracle's forward (oracle/liboracle.so, the independent restatement of
  mpm_sim.py:153-340): the differentiable forward must be the reference substep;
* central finite differences (fp64) of the oracle forward and of the harness forward,
  along seeded random directions in the initial velocities, the per-particle actuation
  multiplier, the actuation amplitude and timing (contraction, frequency) and the jelly
  Young's modulus.
"""
from __future__ import annotations

import numpy as np
import pytest

from tests.harness import Case, load_harness


@pytest.fixture(scope="module")
def harness():
    return load_harness()


def test_forward_is_the_reference_substep(harness):
    cs = Case(0)
    cs.amp_scale[:] = 1.0
    traj = cs.forward(harness)
    ref = cs.oracle_forward()
    assert np.max(np.abs(traj[-1] - ref)) < 1e-9   # fp64 rounding order only


DIRECTIONS = ["x", "v", "C", "F", "amp_scale", "amp", "contraction", "freq", "E"]


def _direction(cs, what, rng):
    ds = np.zeros((cs.P, 12))
    da = np.zeros(cs.P)
    dp = np.zeros_like(cs.params)
    if what == "x":
        ds[:, 0:2] = rng.normal(size=(cs.P, 2)) * 1e-3
    elif what == "v":
        ds[:, 2:4] = rng.normal(size=(cs.P, 2))
    elif what == "C":
        ds[:, 4:8] = rng.normal(size=(cs.P, 4))
    elif what == "F":
        ds[:, 8:12] = rng.normal(size=(cs.P, 4)) * 1e-2
    elif what == "amp_scale":
        da = rng.normal(size=cs.P)
    elif what == "amp":
        dp[:, 0] = rng.normal(size=cs.M) * 50
    elif what == "contraction":
        dp[:, 1] = rng.normal(size=cs.M) * 0.01
    elif what == "freq":
        dp[:, 2] = rng.normal(size=cs.M) * 0.01
    elif what == "E":
        dp[:, 4] = rng.normal(size=cs.M) * 50
    return ds, da, dp


@pytest.mark.parametrize("steps", [1, 12])
@pytest.mark.parametrize("what", DIRECTIONS)
def test_gradient_vs_central_differences(harness, what, steps):
    """Every adjoint input and parameter against fp64 central differences of the same
    forward (all four materials, perturbed state so that every branch is exercised)."""
    cs = Case(1, steps=steps)
    rng = np.random.default_rng(100 + DIRECTIONS.index(what))
    wout = rng.normal(size=(cs.P, 12))
    traj = cs.forward(harness)
    gin, gamp, gpar = cs.backward(harness, traj, wout)
    objective = lambda tr: float(np.sum(wout * tr[-1]))
    ds, da, dp = _direction(cs, what, rng)
    ad = float(np.sum(gin * ds) + np.sum(gamp * da) + np.sum(gpar * dp))
    # central differences at three step sizes: the best one bounds truncation (large eps)
    # and round-off (small eps) error
    errs = []
    for eps in (1e-3, 1e-4, 1e-5):
        vals = [objective(cs.forward(harness, cs.state + sg * eps * ds,
                                     cs.amp_scale + sg * eps * da, cs.params + sg * eps * dp))
                for sg in (1, -1)]
        fd = (vals[0] - vals[1]) / (2 * eps)
        errs.append(abs(ad - fd) / (abs(fd) + 1e-8))
    assert min(errs) <= 1e-5, (what, ad, errs)


def test_gradient_detects_planted_mistake(harness):
    """A wrong adjoint fails the same check: flipping the sign of one adjoint output."""
    cs = Case(1, steps=3)
    rng = np.random.default_rng(0)
    wout = rng.normal(size=(cs.P, 12))
    gin, _, _ = cs.backward(harness, cs.forward(harness), wout)
    ds, _, _ = _direction(cs, "C", rng)
    ad = -float(np.sum(gin * ds))
    eps = 1e-5
    vals = [float(np.sum(wout * cs.forward(harness, cs.state + sg * eps * ds)[-1]))
            for sg in (1, -1)]
    fd = (vals[0] - vals[1]) / (2 * eps)
    assert abs(ad - fd) > 1e-3 * abs(fd)


def test_gradient_vs_oracle_central_differences(harness):
    """Directional derivatives in v0, amplitude, timing and E_jelly against fp64 central
    differences of the independent oracle forward (eps = 1e-4 relative scale)."""
    cs = Case(7, steps=10)
    cs.amp_scale[:] = 1.0
    rng = np.random.default_rng(7)
    wout = rng.normal(size=(cs.P, 12))
    traj = cs.forward(harness)
    gin, gamp, gpar = cs.backward(harness, traj, wout)
    for trial in range(8):
        dv = rng.normal(size=(cs.P, 2))
        dp = np.zeros_like(cs.params)
        dp[:, [0, 1, 2, 4]] = rng.normal(size=(cs.M, 4)) * [500, 0.25, 1.1, 700]
        ad = float(np.sum(gin[:, 2:4] * dv) + np.sum(gpar * dp))
        eps = 1e-4
        vals = []
        for sgn in (1, -1):
            st = cs.state.copy()
            st[:, 2:4] += sgn * eps * dv
            vals.append(float(np.sum(wout * cs.oracle_forward(st, cs.params + sgn * eps * dp))))
        fd = (vals[0] - vals[1]) / (2 * eps)
        assert abs(ad - fd) <= 1e-3 * max(1e-3, abs(fd)), (trial, ad, fd)


@pytest.mark.parametrize("comp", ["v", "C", "F"])
def test_wall_region_gradient_vs_central_differences(harness, comp):
    """Water of a full reference tank, including the damped wall band and the wall clamps
    (mpm_sim.py:282-309) that the small cases never reach. The tank's water starts exactly
    at rest, where the clamps' kink sits at zero velocity; small random initial velocities
    move it off the kink so that the derivative is unique."""
    from tests.harness import reference_tank_case
    cs, _, _, _ = reference_tank_case(3)
    rng = np.random.default_rng(1)
    cs.state[:, 2:4] += rng.normal(size=(cs.P, 2)) * 1e-2
    cs.round_f32()
    wout = rng.normal(size=(cs.P, 12))
    gin, _, _ = cs.backward(harness, cs.forward(harness), wout)
    sl, sc = {"v": (slice(2, 4), 1.0), "C": (slice(4, 8), 1.0), "F": (slice(8, 12), 1e-2)}[comp]
    ds = np.zeros((cs.P, 12))
    ds[:, sl] = rng.normal(size=(cs.P, sl.stop - sl.start)) * sc * (cs.mat == 0)[:, None]
    ad = float(np.sum(gin * ds))
    errs = []
    for eps in (1e-5, 1e-6, 1e-7):
        vals = [float(np.sum(wout * cs.forward(harness, cs.state + sg * eps * ds)[-1]))
                for sg in (1, -1)]
        fd = (vals[0] - vals[1]) / (2 * eps)
        errs.append(abs(ad - fd) / abs(fd))
    assert min(errs) <= 1e-5, errs
