"""GPU tests of the differentiable path (``jm_fwd_block`` / ``jm_bwd_block`` through
``jellyfih_amd.diff``).

* the fp32 HIP forward against the oracle in fp32 and fp64 (as accurate as an fp32 build
  of the reference substep);
* the fp32 HIP adjoint against the fp64 host build of the same adjoint math
  (tests/adjoint_harness.cpp, itself pinned to fp64 central differences on the CPU);
* a seeded reference tank (one instance, ~62K live particles), 2000 substeps with a
  checkpoint every 100: the gradient of the payload CoM y-displacement against fp64
  central differences (eps = 1e-4, 1e-5, 1e-6; tests/golden/grad_fd_golden.json) of the
  oracle, along 8 seeded directions in the body's initial velocities, actuation amplitude,
  contraction fraction, frequency and E_jelly, and of the fp64 harness along the
  per-particle actuation multiplier;
* gradients do not depend on the checkpoint interval.
"""
from __future__ import annotations

import numpy as np
import pytest

from tests.harness import (Case, amp_scale_direction, gradient_directions, load_harness,
                           reference_tank_case)

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _gpu(cs):
    import torch
    from jellyfih_amd.diff import DiffTank
    tank = DiffTank(cs.inst, cs.mat, cs.fiber, cs.M)
    f = lambda a: torch.tensor(np.asarray(a), dtype=torch.float32, device=tank.device)
    return tank, f(cs.state), f(cs.amp_scale), f(cs.params)


GROUPS = {"x": slice(0, 2), "v": slice(2, 4), "C": slice(4, 8), "F": slice(8, 12)}


def test_diff_forward_vs_oracle(require_gpu):
    """One seeded reference tank (~62K live particles), 200 substeps from rest: the final
    state, and the last state of a stored trajectory (a second run, other atomics order),
    are as accurate as an fp32 build of the reference substep."""
    say("diff forward: reference tank seed0, 200 substeps, vs oracle f32/f64")
    cs, _, _, _ = reference_tank_case(200)
    tank, st, amp, par = _gpu(cs)
    out = tank.forward_raw(st, amp, par, cs.t0, cs.steps).cpu().numpy().astype(np.float64)
    traj = tank.forward_raw(st, amp, par, cs.t0, cs.steps, store_all=True).cpu().numpy()
    assert traj.shape == (cs.steps, cs.P, 12)
    o64 = cs.oracle_forward()
    o32 = cs.oracle_forward(dtype=np.float32).astype(np.float64)
    for g, sl in GROUPS.items():
        e_32 = np.abs(o32[:, sl] - o64[:, sl]).max()
        for name, res in (("final", out), ("trajectory", traj[-1])):
            e_gpu = np.abs(res[:, sl] - o64[:, sl]).max()
            say(f"{g} {name}: |gpu-f64|={e_gpu:.3e} |oracle32-f64|={e_32:.3e}")
            assert e_gpu <= 2.0 * e_32 + 1e-6, (g, name)
    assert np.abs(out[:, 0:2] - o64[:, 0:2]).max() <= 1e-4


def test_diff_adjoint_vs_fp64_adjoint(require_gpu):
    say("diff adjoint: 2x160 perturbed particles, 12 substeps, vs fp64 harness adjoint")
    import torch
    cs = Case(1, steps=12).round_f32()
    L = load_harness()
    rng = np.random.default_rng(5)
    wout = rng.normal(size=(cs.P, 12))
    gin64, gamp64, gpar64 = cs.backward(L, cs.forward(L), wout)
    tank, st, amp, par = _gpu(cs)
    gin, gamp, gpar = tank.backward_raw(st, amp, par, cs.t0, cs.steps,
                                        torch.tensor(wout, dtype=torch.float32,
                                                     device=tank.device))
    gin, gamp, gpar = (t.cpu().numpy().astype(np.float64) for t in (gin, gamp, gpar))
    checks = [(g, gin[:, sl], gin64[:, sl]) for g, sl in GROUPS.items()]
    checks += [("amp_scale", gamp, gamp64)]
    checks += [(n, gpar[:, i], gpar64[:, i])
               for i, n in ((0, "amp"), (1, "contraction"), (2, "freq"), (4, "E"))]
    for name, a, b in checks:
        err = np.linalg.norm(a - b) / np.linalg.norm(b)
        say(f"{name}: |g32-g64|/|g64| = {err:.3e}")
        assert err <= 1e-3, name
    assert np.all(gpar[:, 3] == 0)


def _gpu_gradients(cs, checkpoint):
    """Autograd through DiffTank.run: d(payload CoM y displacement) w.r.t. the state,
    the per-particle multiplier and the per-instance parameters."""
    import torch
    tank, st, amp, par = _gpu(cs)
    st.requires_grad_(True)
    amp.requires_grad_(True)
    leaves = [par[:, i].clone().requires_grad_(True) for i in (0, 1, 2, 4)]
    out = tank.run(st, amp, *leaves[:3], E_jelly=leaves[3], payload_density=par[:, 3],
                   t0=cs.t0, nsteps=cs.steps, checkpoint=checkpoint)
    y = tank.payload_com_y(out) - tank.payload_com_y(st.detach())
    y.sum().backward()
    g = lambda t: t.grad.detach().cpu().numpy().astype(np.float64)
    return float(y.item()), g(st), g(amp), np.array([g(t)[0] for t in leaves])


def _best(ad, fds):
    """Relative error of ``ad`` against the best of the central differences at several eps
    (large eps: truncation and kink crossings; small eps: round-off)."""
    return min(abs(ad - f) / abs(f) for f in fds)


def test_reference_tank_gradients_vs_fd(require_gpu, golden_grad):
    """One seeded reference tank, 2000 substeps, checkpoint every 100: the HIP adjoint's
    directional derivatives of the payload CoM y against fp64 central differences of the
    oracle (body initial velocities, amplitude, contraction, frequency, E_jelly) and of
    the fp64 harness (per-particle actuation multiplier), from the golden fixture."""
    say("reference tank seed0, 2000 substeps, checkpoint 100: GPU adjoint vs fp64 FD")
    gg = golden_grad
    cs, pos, mat, fib = reference_tank_case(gg["steps"])
    assert cs.P == gg["n_live"]
    np.testing.assert_array_equal(cs.params[0], np.array(gg["params"]))
    y, gst, gamp, gpar = _gpu_gradients(cs, checkpoint=100)
    say(f"payload displacement {y:.6e}; dJ/d(amp, contraction, freq, E) = {gpar}")
    kinds, dvs, dps = gradient_directions(cs.mat, gg["direction_seed"])
    assert kinds == gg["kinds"]
    for k, kind in enumerate(kinds):
        ad = float(np.sum(gst[:, 2:4] * dvs[k]) + np.sum(gpar * dps[k][[0, 1, 2, 4]]))
        fds = gg["oracle_fd"][k]
        rel = _best(ad, fds)
        say(f"direction {k} ({kind}): AD={ad:.6e} FD(eps={gg['eps']})={fds} best rel={rel:.2e}")
        assert rel <= 1e-3, (k, kind, ad, fds)
    d = amp_scale_direction(cs.mat, gg["amp_scale_seed"])
    ad = float(np.sum(gamp * d))
    fds = gg["harness_amp_scale_fd"]
    rel = _best(ad, fds)
    say(f"amp_scale direction: AD={ad:.6e} FD={fds} best rel={rel:.2e}")
    assert rel <= 1e-3


def test_gradients_independent_of_checkpoint_interval(require_gpu):
    """Recomputing from checkpoints every 100 or 64 substeps gives the gradients of one
    300-substep block, up to the run-to-run noise of fp32 atomics order (measured by
    repeating the 300-substep run). The full state gradient is reported, not asserted:
    water resting against the walls sits on the wall clamps' kink (zero grid velocity), so
    the branch its adjoint takes follows that noise."""
    say("reference tank seed0, 300 substeps: checkpoint 300 (twice) vs 100 vs 64")
    cs, _, _, _ = reference_tank_case(300)
    runs = {k: _gpu_gradients(cs, checkpoint=k) for k in (300, 100, 64)}
    again = _gpu_gradients(cs, checkpoint=300)
    body = cs.mat > 0

    def groups(r):
        return {"state": r[1], "body v": r[1][body, 2:4], "amp_scale": r[2], "params": r[3]}

    base = groups(runs[300])
    floor = {n: np.linalg.norm(v - base[n]) / np.linalg.norm(base[n])
             for n, v in groups(again).items()}
    for k in (100, 64):
        for n, v in groups(runs[k]).items():
            err = np.linalg.norm(v - base[n]) / np.linalg.norm(base[n])
            say(f"checkpoint {k} vs 300, {n}: {err:.2e} (repeat of 300: {floor[n]:.2e})")
            if n != "state":
                assert err <= max(10 * floor[n], 1e-5) and err <= 1e-3, (k, n, err)
