"""The batched forward + backward configuration at its real shape, through the C ABI
(DiffTank over jm_fwd_block / jm_bwd_block): 32 seeded genomes (0-31) in tall tanks (128x256
grid, domain height 2, 160K slots, spawn (0.5, 0.60); evolve.py:32-36,136; ~126K live
particles each, ~4.03M in all), 4000 substeps from rest with a checkpoint every 200,
per-instance gradients of the payload CoM y with respect to actuation amplitude,
contraction, frequency and E_jelly in one batched adjoint pass.

* Gradients of instances 7 and 20 against fp64 central differences of the oracle
  (tests/golden/grad_fd_golden_tall.json, eps 1e-5 and 1e-6; how those instances were
  chosen is in tests/golden/make_grad_golden_tall.py): fp64 within 1e-4; fp32 buffers with
  fp64 arithmetic within 1e-2 (best over eps): over 4000 substeps the flow amplifies the
  fp32 rounding of the stored states (the payload CoM itself moves by ~3e-6), and the
  gradients move with it (2.4e-3 measured for instance 7's amplitude, 1e-3 at 2000
  substeps on the single-tank configuration).
* Every instance's objective and gradients are finite, in every precision; the
  reference's fp32 arithmetic is reported against fp64 (not bounded: see test_gpu_diff).
* Mass: every grid of a stored 20-substep trajectory from the final state holds, per
  instance, the mass of that instance's in-range particles (B-spline weights sum to one).
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "grad_fd_golden_tall.json")
NAMES = ["amp", "contraction", "freq", "E"]
P_MASS = (1 / 256) ** 2          # p_vol * p_rho at quality 1 (mpm_sim.py:20-22)


def say(msg: str) -> None:
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def tall():
    from jellyfih_amd import phenotype
    with open(GOLDEN) as f:
        gg = json.load(f)
    pos, mat, fib, genomes, _ = phenotype.population(
        gg["seeds"], gg["slots"], 128, domain_height=2.0, spawn_offset=phenotype.TALL_SPAWN)
    assert [int((mat[m] >= 0).sum()) for m in range(len(gg["seeds"]))] == gg["n_live"]
    return gg, pos, mat, fib, genomes


def run_batch(gg, pos, mat, fib, genomes, dtype, math):
    """All tanks' payload CoM y after the episode and d/d (amp, contraction, freq, E)
    (M, 4), plus the final state and the tank."""
    import torch
    from jellyfih_amd.diff import DiffTank
    td = getattr(torch, dtype)
    tank, st = DiffTank.from_tanks(pos, mat, fib, n_grid_y=256, domain_height=2.0, dtype=td,
                                   math=math)
    M = pos.shape[0]
    base = np.array([[500.0, np.clip(g[9], 0.05, 0.6), np.clip(g[10], 0.5, 2.0), 700.0]
                     for g in genomes]).astype(np.float32).astype(np.float64)
    for m in gg["instances"]:
        p = gg["params"][str(m)]
        np.testing.assert_array_equal(base[m], [p[0], p[1], p[2], p[4]])
    leaves = [torch.tensor(base[:, i], dtype=td, device=tank.device).requires_grad_(True)
              for i in range(4)]
    out = tank.run(st, None, leaves[0], leaves[1], leaves[2], leaves[3], 2.5, 0.0,
                   gg["steps"], 200)
    com = tank.payload_com_y(out)
    com.sum().backward()
    grads = np.stack([t.grad.detach().cpu().numpy().astype(np.float64) for t in leaves], 1)
    assert grads.shape == (M, 4)
    return com.detach().cpu().numpy().astype(np.float64), grads, out.detach(), tank, base


@pytest.mark.parametrize("dtype,math,bound", [("float64", "fp64", 1e-4),
                                              ("float32", "fp64", 1e-2),
                                              ("float32", "fp32", None)])
def test_tall_batch_gradients(require_gpu, tall, dtype, math, bound):
    gg, pos, mat, fib, genomes = tall
    say(f"32 tall tanks x 160K, 4000 substeps, checkpoint 200, {dtype}/{math}: batched adjoint")
    com, grads, _, tank, _ = run_batch(gg, pos, mat, fib, genomes, dtype, math)
    tank.close()
    assert np.isfinite(com).all() and np.isfinite(grads).all()
    assert np.all(com > 0.1) and np.all(com < 2.0)
    worst = 0.0
    for m in gg["instances"]:
        c_ref = gg["payload_com_y"][str(m)]
        say(f"instance {m}: payload CoM y {com[m]:.9f} (oracle at +-eps {c_ref})")
        assert abs(com[m] - 0.5 * (c_ref[0] + c_ref[1])) <= 5e-4
        for k, name in enumerate(NAMES):
            fds = gg["oracle_fd"][str(m)][name]
            rel = min(abs(grads[m, k] - f) / abs(f) for f in fds)
            worst = max(worst, rel)
            say(f"  d/d {name:11s} AD={grads[m, k]: .9e} FD={fds} |AD-FD|={rel:.1e}")
            if bound is None:
                assert np.sign(grads[m, k]) == np.sign(fds[0]), (m, name)
            else:
                assert rel <= bound, (m, name, rel)
    say(f"worst |AD-FD| {worst:.2e} (bound {bound})")


def test_tall_batch_mass_per_stored_grid(require_gpu, tall):
    """After the 4000-substep episode (fp32 buffers), a stored trajectory of 20 more
    substeps: every grid's mass per instance equals the mass of that instance's particles
    inside the transfer range at that substep (mpm_sim.py:184), within fp32 summation."""
    import torch
    gg, pos, mat, fib, genomes = tall
    say("32 tall tanks: mass of every stored grid after 4000 substeps")
    from jellyfih_amd.diff import DiffTank, clock
    with torch.no_grad():
        tank, st = DiffTank.from_tanks(pos, mat, fib, n_grid_y=256, domain_height=2.0)
        M = pos.shape[0]
        dev = tank.device
        par = torch.tensor(np.stack([np.full(M, 500.0), np.clip(genomes[:, 9], 0.05, 0.6),
                                     np.clip(genomes[:, 10], 0.5, 2.0), np.full(M, 2.5),
                                     np.full(M, 700.0)], 1), dtype=torch.float32, device=dev)
        amp = torch.ones(tank.P, dtype=torch.float32, device=dev)
        state = st
        t = 0.0
        for _ in range(gg["steps"] // 200):
            state = tank.forward_raw(state.contiguous(), amp, par, t, 200)
            t = clock(t, 200)
        assert torch.isfinite(state).all()
        n = 20
        traj, grids = tank.trajectory(state.contiguous(), amp, par, t, n)
        mass_p = torch.full((tank.P,), P_MASS, dtype=torch.float64, device=dev)
        pay = tank.mat == 2
        mass_p[pay] *= 2.5
        for s in range(n):
            x = traj[s, :, 0:2].double()
            b = torch.trunc(x * 128.0 - 0.5)
            inr = (b[:, 0] >= 0) & (b[:, 1] >= 0) & (b[:, 0] < 126) & (b[:, 1] < 254)
            expect = tank.instance_sums(mass_p * inr)
            got = grids[s, :, 2].double().sum(1)
            rel = ((got - expect).abs() / expect).max().item()
            assert rel <= 1e-5, (s, rel)
        say(f"grid mass of {n} stored substeps x {M} instances within {rel:.1e} (last)")
        tank.close()
