"""GPU tests of the differentiable path (``jm_fwd_block`` / ``jm_bwd_block`` through
``jellyfih_amd.diff``), in fp32 and fp64.

* the HIP forward against the oracle: in fp32 as accurate as an fp32 build of the
  reference substep, in fp64 equal to the fp64 oracle up to summation order;
* the HIP adjoint against the fp64 host build of the same adjoint math
  (tests/adjoint_harness.cpp, itself pinned to fp64 central differences on the CPU);
* a seeded reference tank (one instance, ~62K live particles), 2000 substeps with a
  checkpoint every 100: the gradient of the payload CoM y against fp64 central differences
  (eps = 1e-4, 1e-5, 1e-6; tests/golden/grad_fd_golden.json) of the oracle, along 8 seeded
  directions in the body's initial velocities, actuation amplitude, contraction fraction,
  frequency and E_jelly, and of the fp64 harness along the per-particle actuation
  multiplier; the same at 200 substeps (tests/golden/grad_fd_golden_200.json). fp64: within
  1e-4. fp32 buffers with fp64 arithmetic inside each substep (DiffTank's default for fp32):
  within 1e-3 at 200 substeps, 5e-3 at 2000 (GRAD_BOUNDS). fp32 arithmetic (the
  reference's): small strains near F = I are differences of numbers near 1 that fp32
  resolves to ~1e-7, so its gradients are off by percents (a sequential fp32 build of the
  same algorithm, in the fixtures: up to 4% at 200 substeps, 5% at 2000); they must agree in
  sign and magnitude (within a factor of 2), and the spread is reported;
* gradients do not depend on the checkpoint interval or on recomputation.
"""
from __future__ import annotations

import numpy as np
import pytest

from tests.harness import (Case, amp_scale_direction, gradient_directions, load_harness,
                           reference_tank_case)

pytestmark = pytest.mark.gpu

DTYPES = ["float32", "float64"]
# (buffers, arithmetic): fp32 buffers with fp64 arithmetic (DiffTank's default for fp32),
# the reference's fp32 arithmetic, fp64 throughout
PRECISIONS = [("float32", "fp64"), ("float32", "fp32"), ("float64", "fp64")]


def say(*a):
    print(*a, flush=True)


def _gpu(cs, dtype="float32", math="fp64"):
    import torch
    from jellyfih_amd.diff import DiffTank
    td = getattr(torch, dtype)
    tank = DiffTank(cs.inst, cs.mat, cs.fiber, cs.M, dtype=td, math=math)
    f = lambda a: torch.tensor(np.asarray(a), dtype=td, device=tank.device)
    return tank, tank.encode_state(f(cs.state)), f(cs.amp_scale), f(cs.params)


GROUPS = {"x": slice(0, 2), "v": slice(2, 4), "C": slice(4, 8), "F": slice(8, 12)}


@pytest.mark.parametrize("dtype,math", PRECISIONS)
def test_diff_forward_vs_oracle(require_gpu, dtype, math):
    """One seeded reference tank (~62K live particles), 200 substeps from rest: the final
    state, and the last state of a stored trajectory (a second run, other atomics order).
    fp32 buffers (either arithmetic): as accurate as the fp32 oracle; fp64: equal to the
    fp64 oracle up to summation order."""
    import torch
    say(f"diff forward {dtype}/{math}: reference tank seed0, 200 substeps, vs oracle f32/f64")
    cs, _, _, _ = reference_tank_case(200)
    tank, st, amp, par = _gpu(cs, dtype, math)
    out = tank.forward_raw(st, amp, par, cs.t0, cs.steps)
    out = tank.decode_state(out.double()).cpu().numpy()
    traj = torch.empty((cs.steps, cs.P, 12), dtype=tank.dtype, device=tank.device)
    grids = torch.empty((cs.steps, 1, 3, tank.G), dtype=tank.dtype, device=tank.device)
    tank.forward_raw(st, amp, par, cs.t0, cs.steps, out=traj, grids=grids)
    traj = tank.decode_state(traj.double()).cpu().numpy()
    # B-spline weights sum to one: every stored grid holds the total particle mass,
    # p_mass * (non-payload count + payload density * payload count)
    n_pay = int(np.sum(cs.mat == 2))
    expect = cs.c.p_mass * (cs.P - n_pay + cs.params[0, 3] * n_pay)
    masses = grids[:, 0, 2].double().sum(1).cpu().numpy()
    say(f"grid mass per substep {masses.min():.12e} .. {masses.max():.12e}, "
        f"expected {expect:.12e}")
    assert np.all(np.abs(masses - expect) <= (1e-5 if dtype == "float32" else 1e-12) * expect)
    o64 = cs.oracle_forward()
    o32 = cs.oracle_forward(dtype=np.float32).astype(np.float64)
    for g, sl in GROUPS.items():
        e_32 = np.abs(o32[:, sl] - o64[:, sl]).max()
        scale = np.abs(o64[:, sl]).max()
        for name, res in (("final", out), ("trajectory", traj[-1])):
            e_gpu = np.abs(res[:, sl] - o64[:, sl]).max()
            say(f"{g} {name}: |gpu-f64|={e_gpu:.3e} |oracle32-f64|={e_32:.3e}")
            if dtype == "float32":
                assert e_gpu <= 2.0 * e_32 + 1e-6, (g, name)
            else:
                assert e_gpu <= 1e-9 * max(scale, 1.0), (g, name)
    assert np.abs(out[:, 0:2] - o64[:, 0:2]).max() <= 1e-4


@pytest.mark.parametrize("dtype", DTYPES)
def test_diff_forward_tall_tank(require_gpu, dtype):
    """The tall configuration (evolve.py:33-36,136: 128x256 grid, domain height 2, spawn
    (0.5, 0.60), 160K slots; one seeded tank, ~126K live), 200 substeps of the
    differentiable forward from rest, packed in cell order, against the oracle in fp32 and
    fp64: fp32 as accurate as the fp32 oracle (fp32_like), fp64 equal to the fp64 oracle up
    to summation order."""
    import torch
    import oracle
    from jellyfih_amd import phenotype
    from jellyfih_amd.diff import DiffTank
    from tests.cases import fp32_like, reference_tank
    say(f"diff forward {dtype}: tall tank seed3, 128x256 H=2, 200 substeps, vs oracle")
    steps = 200
    pos, mat, fib, g, info = reference_tank(3, n_slots=160000, domain_height=2.0,
                                            spawn=phenotype.TALL_SPAWN)
    con, frq = float(np.clip(g[9], 0.05, 0.6)), float(np.clip(g[10], 0.5, 2.0))
    td = getattr(torch, dtype)
    tank, st = DiffTank.from_tanks(pos[None], mat[None], fib[None], n_grid_y=256,
                                   domain_height=2.0, dtype=td)
    params = torch.tensor([[500.0, con, frq, 2.5, 700.0]], dtype=torch.float32).to(
        tank.device, td)
    amp = torch.ones(tank.P, dtype=td, device=tank.device)
    out = tank.decode_state(tank.forward_raw(st, amp, params, 0.0, steps).double())
    out = out.cpu().numpy()
    slot = tank.slots[:, 1]
    consts = oracle.reference_consts(n_grid_y=256, domain_height=2.0)
    o32 = oracle.OracleSim(1, 160000, consts, dtype=np.float32)
    o64 = oracle.OracleSim(1, 160000, consts, dtype=np.float64)
    for o in (o32, o64):
        o.load(0, pos, mat, fib)
        o.contraction[:] = np.float32(con)
        o.freq[:] = np.float32(frq)
        o.run(steps, nthreads=8)
    fields = {"x": (0, 2, "x"), "v": (2, 4, "v"), "C": (4, 8, "C"), "F": (8, 12, "F")}
    for name, (a, b, attr) in fields.items():
        gpu = out[:, a:b]
        r32 = getattr(o32, attr)[0][slot].reshape(len(slot), -1).astype(np.float64)
        r64 = getattr(o64, attr)[0][slot].reshape(len(slot), -1)
        if dtype == "float32":
            fp32_like(name, gpu, r32, r64)
        else:
            err = np.abs(gpu - r64).max()
            say(f"{name}: |gpu-f64| = {err:.3e}")
            assert err <= 1e-9 * max(np.abs(r64).max(), 1.0), name


@pytest.mark.parametrize("dtype,math", PRECISIONS)
def test_diff_adjoint_vs_fp64_adjoint(require_gpu, dtype, math):
    """Perturbed states of every material (2 x 160 particles), 12 substeps: the whole
    adjoint (state, multiplier, parameters) against the fp64 host build."""
    import torch
    say(f"diff adjoint {dtype}/{math}: 2x160 perturbed particles, 12 substeps, vs fp64 harness")
    cs = Case(1, steps=12).round_f32()
    L = load_harness()
    rng = np.random.default_rng(5)
    wout = rng.normal(size=(cs.P, 12))
    gin64, gamp64, gpar64 = cs.backward(L, cs.forward(L), wout)
    tank, st, amp, par = _gpu(cs, dtype, math)
    gin, gamp, gpar = tank.backward_raw(st, amp, par, cs.t0, cs.steps,
                                        torch.tensor(wout, dtype=tank.dtype,
                                                     device=tank.device))
    gin, gamp, gpar = (t.cpu().numpy().astype(np.float64) for t in (gin, gamp, gpar))
    checks = [(g, gin[:, sl], gin64[:, sl]) for g, sl in GROUPS.items()]
    checks += [("amp_scale", gamp, gamp64)]
    checks += [(n, gpar[:, i], gpar64[:, i])
               for i, n in ((0, "amp"), (1, "contraction"), (2, "freq"), (4, "E"))]
    tol = 1e-3 if dtype == "float32" else 1e-9
    for name, a, b in checks:
        err = np.linalg.norm(a - b) / np.linalg.norm(b)
        say(f"{name}: |g-g64|/|g64| = {err:.3e}")
        assert err <= tol, name
    assert np.all(gpar[:, 3] == 0)


def _gpu_gradients(cs, checkpoint, dtype="float32", recompute=True, resort_every=0,
                   math="fp64", prefetch=True, keep_last=None):
    """Autograd through DiffTank.run: d(payload CoM y displacement) w.r.t. the state,
    the per-particle multiplier and the per-instance parameters."""
    tank, st, amp, par = _gpu(cs, dtype, math)
    st.requires_grad_(True)
    amp.requires_grad_(True)
    leaves = [par[:, i].clone().requires_grad_(True) for i in (0, 1, 2, 4)]
    out = tank.run(st, amp, *leaves[:3], E_jelly=leaves[3], payload_density=par[:, 3],
                   t0=cs.t0, nsteps=cs.steps, checkpoint=checkpoint, recompute=recompute,
                   resort_every=resort_every, prefetch=prefetch, keep_last=keep_last)
    if keep_last is not None and recompute:
        assert tank.last_kept == min(keep_last, -(-cs.steps // checkpoint))
    y = tank.payload_com_y(out) - tank.payload_com_y(st.detach())
    y.sum().backward()
    g = lambda t: t.grad.detach().cpu().numpy().astype(np.float64)
    return float(y.item()), g(st), g(amp), np.array([g(t)[0] for t in leaves])


def _best(ad, fds):
    """Relative error of ``ad`` against the best of the central differences at several eps
    (large eps: truncation and kink crossings; small eps: round-off)."""
    return min(abs(ad - f) / abs(f) for f in fds)


def _directional(cs, gg, gst, gamp, gpar):
    """(kind, AD, fp64 harness AD, fp32 harness AD, FD list) per fixture direction."""
    kinds, dvs, dps = gradient_directions(cs.mat, gg["direction_seed"])
    assert kinds == gg["kinds"]
    rows = []
    for k, kind in enumerate(kinds):
        ad = float(np.sum(gst[:, 2:4] * dvs[k]) + np.sum(gpar * dps[k][[0, 1, 2, 4]]))
        rows.append((kind, ad, gg["harness_ad"][k], gg["harness_ad_fp32"][k],
                     gg["oracle_fd"][k]))
    d = amp_scale_direction(cs.mat, gg["amp_scale_seed"])
    rows.append(("amp_scale", float(np.sum(gamp * d)), gg["harness_amp_scale_ad"],
                 gg["harness_amp_scale_ad_fp32"], gg["harness_amp_scale_fd"]))
    return rows


def _golden(steps: int):
    import json
    import os
    name = "grad_fd_golden.json" if steps == 2000 else f"grad_fd_golden_{steps}.json"
    with open(os.path.join(os.path.dirname(__file__), "golden", name)) as f:
        return json.load(f)


# Bounds on |AD - FD| / |FD| (best over eps) per (horizon, buffers, arithmetic). fp64: the
# fp64 adjoint is the derivative of the fp64 model (1e-4). fp32 buffers (F stored as F - I)
# with fp64 arithmetic: 1e-3 at 200 substeps (measured 7.5e-5); over 2000 substeps the
# fp32 rounding of the stored states, compounded, moves the linearisation point: 5e-3
# (measured 1.0e-3, as the host build of the same scheme; the multiplier direction, whose
# differences come from the fp64 host build, 1.4e-2 against AMP_SCALE_BOUND_2000). fp32
# arithmetic (the reference's): None = each direction within a factor 2 of FD, and the
# worst direction within twice the worst of the sequential fp32 host build of the same
# algorithm (4.3e-2 at 200 substeps, 5.2e-2 at 2000, in the fixtures) plus 1e-3: fp32
# rounding of the small strains, not the adjoint, sets that error (measured 4.3e-2, 4.7e-2).
GRAD_BOUNDS = {(200, "float64", "fp64"): 1e-4, (200, "float32", "fp64"): 1e-3,
               (200, "float32", "fp32"): None, (2000, "float64", "fp64"): 1e-4,
               (2000, "float32", "fp64"): 5e-3, (2000, "float32", "fp32"): None}
AMP_SCALE_BOUND_2000 = 3e-2


@pytest.mark.parametrize("steps", [200, 2000])
@pytest.mark.parametrize("dtype,math", PRECISIONS)
def test_reference_tank_gradients_vs_fd(require_gpu, steps, dtype, math):
    """One seeded reference tank, 200 or 2000 substeps, checkpoint every 100 (the
    differentiable single-instance workload):
    directional derivatives of the payload CoM y along 8 seeded directions and the
    per-particle multiplier direction against fp64 central differences."""
    say(f"reference tank seed0, {steps} substeps, checkpoint 100, {dtype}/{math}: "
        "adjoint vs fp64 FD")
    gg = _golden(steps)
    cs, pos, mat, fib = reference_tank_case(gg["steps"])
    assert cs.P == gg["n_live"]
    np.testing.assert_array_equal(cs.params[0], np.array(gg["params"]))
    y, gst, gamp, gpar = _gpu_gradients(cs, checkpoint=100, dtype=dtype, math=math, keep_last=0)
    say(f"payload displacement {y:.9e}; dJ/d(amp, contraction, freq, E) = {gpar}")
    bound = GRAD_BOUNDS[(steps, dtype, math)]
    worst, worst32 = 0.0, 0.0
    for kind, ad, ad64, ad32, fds in _directional(cs, gg, gst, gamp, gpar):
        r64 = _best(ad64, fds)
        rel = _best(ad, fds)
        say(f"{kind:12s} AD={ad:.9e} AD64(host)={ad64:.9e} |AD64-FD|={r64:.1e} "
            f"|AD-FD|={rel:.1e} |AD32(host)-FD|={_best(ad32, fds):.1e}")
        # the multiplier's differences come from the fp64 host build (2e-4 at 200 substeps)
        fd_tol = 2e-4 if kind == "amp_scale" else 1e-4
        assert r64 <= fd_tol, (kind, ad64, fds)
        worst = max(worst, rel)
        worst32 = max(worst32, _best(ad32, fds))
        if bound is None:
            ratio = ad / fds[0]
            assert 0.5 <= ratio <= 2.0, (kind, ad, fds)
        elif kind == "amp_scale" and steps == 2000 and dtype == "float32":
            assert rel <= AMP_SCALE_BOUND_2000, (kind, ad, fds)
        else:
            assert rel <= max(bound, 2 * r64), (kind, rel, bound)
    say(f"worst |AD-FD| {worst:.2e} (bound {bound}); sequential fp32 build: {worst32:.2e}")
    if bound is None:
        assert worst <= 2 * worst32 + 1e-3, (worst, worst32)


def test_gradients_independent_of_checkpoint_interval(require_gpu):
    """Recomputing from checkpoints every 100 or 64 substeps (each block's recomputation
    beside the next block's adjoint on a second stream, or both on one stream), keeping the
    last 2 of 5 blocks' trajectories from the forward pass and recomputing the other 3, or
    keeping every block's trajectory, gives the gradients of one 300-substep block, up to
    the run-to-run noise of atomics order (measured by repeating the 300-substep run). fp64,
    so that the noise is small. The full state gradient is reported, not asserted: water
    resting against the walls sits on the wall clamps' kink (zero grid velocity), so the
    branch its adjoint takes follows that noise."""
    say("reference tank seed0, 300 substeps, fp64: checkpoint 300 (twice) vs 100 vs 64 vs "
        "64 on one stream vs 64 keeping the last 2 blocks vs stored")
    cs, _, _, _ = reference_tank_case(300)
    runs = {k: _gpu_gradients(cs, checkpoint=k, dtype="float64", keep_last=0)
            for k in (300, 100, 64)}
    runs["stored"] = _gpu_gradients(cs, checkpoint=100, dtype="float64", recompute=False)
    runs["64, one stream"] = _gpu_gradients(cs, checkpoint=64, dtype="float64", prefetch=False,
                                            keep_last=0)
    runs["64, last 2 kept"] = _gpu_gradients(cs, checkpoint=64, dtype="float64", keep_last=2)
    again = _gpu_gradients(cs, checkpoint=300, dtype="float64", keep_last=0)
    body = cs.mat > 0

    def groups(r):
        return {"state": r[1], "body v": r[1][body, 2:4], "amp_scale": r[2], "params": r[3]}

    base = groups(runs[300])
    floor = {n: np.linalg.norm(v - base[n]) / np.linalg.norm(base[n])
             for n, v in groups(again).items()}
    for k in (100, 64, "64, one stream", "64, last 2 kept", "stored"):
        for n, v in groups(runs[k]).items():
            err = np.linalg.norm(v - base[n]) / np.linalg.norm(base[n])
            say(f"checkpoint {k} vs 300, {n}: {err:.2e} (repeat of 300: {floor[n]:.2e})")
            if n != "state":
                assert err <= max(10 * floor[n], 1e-9) and err <= 1e-6, (k, n, err)


def test_gradients_with_resorting_between_blocks(require_gpu):
    """Re-sorting the particles by cell before every checkpoint block (a gather of the state
    and multiplier; autograd scatters the adjoints back) gives the objective and gradients
    of the unsorted run, in the caller's particle order, up to the run-to-run noise of
    atomics order (measured by repeating the unsorted run), also with the last 2 blocks
    keeping their trajectories from the forward pass (kept on the re-ordered views). fp64,
    300 substeps, blocks of 100."""
    say("reference tank seed0, 300 substeps, fp64: re-sorted every block (recomputed, or the "
        "last 2 blocks kept) vs not")
    cs, _, _, _ = reference_tank_case(300)
    base = _gpu_gradients(cs, checkpoint=100, dtype="float64", keep_last=0)
    again = _gpu_gradients(cs, checkpoint=100, dtype="float64", keep_last=0)
    rs = _gpu_gradients(cs, checkpoint=100, dtype="float64", resort_every=1,
                        keep_last=0)
    rs_kept = _gpu_gradients(cs, checkpoint=100, dtype="float64", resort_every=1,
                             keep_last=2)
    body = cs.mat > 0

    def groups(r):
        return {"body v": r[1][body, 2:4], "amp_scale": r[2], "params": r[3]}

    b = groups(base)
    floor = {n: np.linalg.norm(v - b[n]) / np.linalg.norm(b[n]) for n, v in groups(again).items()}
    for label, r in (("re-sorted", rs), ("re-sorted, last 2 kept", rs_kept)):
        assert abs(r[0] - base[0]) <= 1e-9 * abs(base[0]) + 1e-15, (label, r[0], base[0])
        for n, v in groups(r).items():
            err = np.linalg.norm(v - b[n]) / np.linalg.norm(b[n])
            say(f"{label} vs not, {n}: {err:.2e} (repeat: {floor[n]:.2e})")
            if n != "body v":
                assert err <= max(10 * floor[n], 1e-9) and err <= 1e-6, (label, n, err)


def test_batched_gradients_match_per_instance(require_gpu):
    """The batched adjoint keeps instances apart: three tall tanks (128x256, H = 2, seeds
    0-2) in one DiffTank give, per instance, the payload CoM and the gradients w.r.t.
    amplitude, contraction and frequency of the same tank run alone (fp64, 200 substeps,
    checkpoint 100), up to fp64 summation order."""
    import torch
    from jellyfih_amd import phenotype
    from jellyfih_amd.diff import DiffTank
    say("3 tall tanks batched vs one at a time, fp64, 200 substeps")
    pos, mat, fib, genomes, infos = phenotype.population(
        range(3), 160000, 128, domain_height=2.0, spawn_offset=phenotype.TALL_SPAWN)

    def run(P, Mt, Fb, G):
        tank, st = DiffTank.from_tanks(P, Mt, Fb, n_grid_y=256, domain_height=2.0,
                                       dtype=torch.float64)
        dev = tank.device
        vec = lambda a: torch.tensor(np.asarray(a, np.float64), device=dev).requires_grad_(True)
        amp = vec(np.full(len(G), 500.0))
        con, frq = vec(np.clip(G[:, 9], 0.05, 0.6)), vec(np.clip(G[:, 10], 0.5, 2.0))
        out = tank.run(st, None, amp, con, frq, 700.0, 2.5, 0.0, 200, 100)
        com = tank.payload_com_y(out)
        com.sum().backward()
        res = [t.detach().cpu().numpy() for t in (com, amp.grad, con.grad, frq.grad)]
        tank.close()
        return res

    both = run(pos, mat, fib, genomes)
    for m in range(3):
        one = run(pos[m:m + 1], mat[m:m + 1], fib[m:m + 1], genomes[m:m + 1])
        for name, a, b in zip(("com", "d/d amp", "d/d contraction", "d/d freq"), both, one):
            err = abs(a[m] - b[0]) / max(abs(b[0]), 1e-30)
            say(f"instance {m} {name}: batched {a[m]:.10e} alone {b[0]:.10e} rel {err:.1e}")
            assert err <= 1e-8, (m, name)


@pytest.mark.parametrize("dtype,math", PRECISIONS)
def test_split_packing_gives_the_same_gradients(require_gpu, dtype, math):
    """Packing every instance's body first and the water after it (``order="split"``), so
    that the adjoint runs the water with its water-only kernel, gives the payload CoM and
    the gradients w.r.t. amplitude, contraction and frequency of the cell-ordered packing
    (three tall tanks, seeds 0-2, 200 substeps, checkpoint 100), also re-sorted before every
    block, up to the run-to-run noise of atomics order (measured by repeating the
    cell-ordered run): fp64 within 1e-9, fp32 buffers within ten times that noise or 1e-4."""
    import torch
    from jellyfih_amd import phenotype
    from jellyfih_amd.diff import DiffTank
    say(f"3 tall tanks, 200 substeps, {dtype}/{math}: split packing vs cell packing")
    pos, mat, fib, genomes, infos = phenotype.population(
        range(3), 160000, 128, domain_height=2.0, spawn_offset=phenotype.TALL_SPAWN)
    td = getattr(torch, dtype)

    def run(order, resort=0):
        tank, st = DiffTank.from_tanks(pos, mat, fib, order=order, n_grid_y=256,
                                       domain_height=2.0, dtype=td, math=math)
        if order == "split":
            ws = tank.water_start
            assert 0 < ws < tank.P and bool((tank.mat[ws:] == 0).all())
            assert bool((tank.mat[:ws] > 0).all())
        dev = tank.device
        vec = lambda a: torch.tensor(np.asarray(a), dtype=td, device=dev).requires_grad_(True)
        amp = vec(np.full(3, 500.0))
        con, frq = vec(np.clip(genomes[:, 9], 0.05, 0.6)), vec(np.clip(genomes[:, 10], 0.5, 2.0))
        out = tank.run(st, None, amp, con, frq, 700.0, 2.5, 0.0, 200, 100, resort_every=resort)
        com = tank.payload_com_y(out)
        com.sum().backward()
        res = np.concatenate([t.detach().double().cpu().numpy()
                              for t in (com, amp.grad, con.grad, frq.grad)])
        tank.close()
        return res

    def rel(a, b):
        return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-30)))

    base, again = run("cell"), run("cell")
    floor = rel(again, base)
    bound = 1e-9 if dtype == "float64" else max(10 * floor, 1e-4)
    for name, r in (("split", run("split")), ("split, re-sorted every block", run("split", 1))):
        err = rel(r, base)
        say(f"{name} vs cell: max rel {err:.2e} over CoM and 9 gradients "
            f"(cell repeated: {floor:.2e}); bound {bound:.1e}")
        assert err <= bound, (name, err)


def test_kept_block_count(require_gpu):
    """DiffTank.blocks_to_keep: a block's bytes are its stored states and grids; no block for
    a zero fraction or an empty run; every block of a short run of one tank for the default
    fraction; as many as fit in the given fraction of free memory otherwise (a block of a
    third to two thirds of the free memory: none at 0.3, at least one at 0.9); and run()
    keeps that many (tank.last_kept), keep_last overriding it; without recomputation
    every block keeps its trajectory."""
    import torch
    say("reference tank seed0: kept block counts against block size and free memory")
    cs, _, _, _ = reference_tank_case(300)
    tank, st, amp, par = _gpu(cs, "float32", "fp32")
    assert tank.block_bytes(100) == (101 * tank.P * 12 + 100 * tank.M * 3 * tank.G) * 4
    assert tank.blocks_to_keep(300, 100, 0.0) == 0
    assert tank.blocks_to_keep(0, 100, 0.8) == 0
    assert tank.blocks_to_keep(300, 100, 0.8) == 3
    assert tank.blocks_to_keep(250, 100, 0.8) == 3
    free, _ = torch.cuda.mem_get_info(tank.device)
    free += torch.cuda.memory_reserved(tank.device) - torch.cuda.memory_allocated(tank.device)
    n = 1
    while tank.block_bytes(n) < free / 3:   # the block doubles, less its shared first state
        n *= 2
    say(f"free {free / 2**30:.1f} GiB; block of {n} substeps {tank.block_bytes(n) / 2**30:.1f} GiB")
    assert tank.blocks_to_keep(20 * n, n, 0.3) == 0
    assert tank.blocks_to_keep(20 * n, n, 0.9) >= 1
    for kw, want in (({}, 3), ({"keep_last": 1}, 1), ({"keep_memory": 0.0}, 0),
                     ({"keep_last": 7}, 3), ({"recompute": False, "keep_last": 0}, 3)):
        tank.run(st, amp, par[:, 0], par[:, 1], par[:, 2], par[:, 4], par[:, 3], cs.t0, 300,
                 100, **kw)
        assert tank.last_kept == want, (kw, tank.last_kept)
    tank.close()


@pytest.mark.parametrize("keep_last", [0, 1])
def test_second_backward_through_a_retained_graph(require_gpu, keep_last):
    """A second backward pass through the same graph (``retain_graph=True``) recomputes every
    block again (the first pass consumed the pipelined recomputations and released the kept
    trajectories) and adds the same gradients once more: fp64, one reference tank, 300
    substeps, checkpoint 100, with every block recomputed in the first pass or the last one
    kept from the forward pass."""
    import torch
    say(f"reference tank seed0, 300 substeps, fp64, last {keep_last} kept: backward twice "
        "through a retained graph")
    cs, _, _, _ = reference_tank_case(300)
    tank, st, amp, par = _gpu(cs, "float64", "fp64")
    st.requires_grad_(True)
    leaves = [par[:, i].clone().requires_grad_(True) for i in (0, 1, 2)]
    out = tank.run(st, amp, *leaves, E_jelly=par[:, 4], payload_density=par[:, 3], t0=cs.t0,
                   nsteps=cs.steps, checkpoint=100, keep_last=keep_last)
    y = tank.payload_com_y(out).sum()
    y.backward(retain_graph=True)
    first = [t.grad.clone() for t in [st] + leaves]
    y.backward()
    for name, g1, t in zip(("state", "amp", "contraction", "freq"), first, [st] + leaves):
        err = float((t.grad - 2 * g1).norm() / (2 * g1).norm())
        say(f"{name}: |g(2 passes) - 2 g(1 pass)| / |2 g| = {err:.2e}")
        assert err <= 1e-9, name
    tank.close()


@pytest.mark.parametrize("dtype,math", PRECISIONS)
def test_diff_quality_two(require_gpu, dtype, math):
    """quality = 2 (mpm_sim.py:12,14,19: a 256x256 grid and dt = 1e-5, which the reference
    fixes at 1 and the engine takes as a parameter): a small tank with a jelly, muscle and
    payload patch at 4 particles per cell of the finer grid, 200 substeps from t = 0.05 (a
    contraction phase). The differentiable forward against the oracle built with the same
    constants (fp32 buffers as accurate as the fp32 oracle; fp64 equal to the fp64 oracle
    up to summation order), and the gradients of the payload CoM y w.r.t. amplitude,
    contraction, frequency and E_jelly (every block recomputed, checkpoint 100) against
    fp64 central differences of that oracle: fp64 within 1e-4, fp32 buffers with fp64
    arithmetic within 1e-3, fp32 arithmetic within 2e-3 (measured 3e-7, 9e-5 and 2e-4: on
    this small tank over 0.002 s the fp32 rounding of the water's small strains, which
    sets fp32's error on the reference tank, stays small)."""
    import oracle
    import torch
    from jellyfih_amd.diff import DiffTank
    from tests.cases import small_tank
    from tests.harness import oracle_payload_fd
    say(f"quality 2 (256x256, dt 1e-5), small tank, 200 substeps, {dtype}/{math}: forward "
        "and gradients vs the oracle")
    pos, mat, fib = small_tank(7, 9000, region=(0.30, 0.50, 0.30, 0.45), ppc_spacing=1 / 512)
    live = mat >= 0
    P = int(live.sum())
    state = np.zeros((P, 12))
    state[:, 0:2] = pos[live]
    state[:, 8] = state[:, 11] = 1.0
    params = np.array([[500.0, 0.25, 1.1, 2.5, 700.0]])
    cs = Case.from_arrays(np.zeros(P, np.int32), mat[live], fib[live], state, np.ones(P),
                          params, steps=200, t0=0.05).round_f32()
    cs.c = oracle.reference_consts(quality=2)
    td = getattr(torch, dtype)
    tank = DiffTank(cs.inst, cs.mat, cs.fiber, 1, quality=2, dtype=td, math=math)
    f = lambda a: torch.tensor(np.asarray(a), dtype=td, device=tank.device)
    st, amp, par = tank.encode_state(f(cs.state)), f(cs.amp_scale), f(cs.params)
    out = tank.decode_state(tank.forward_raw(st, amp, par, cs.t0, cs.steps).double())
    out = out.cpu().numpy()
    o64 = cs.oracle_forward()
    o32 = cs.oracle_forward(dtype=np.float32).astype(np.float64)
    for g, sl in GROUPS.items():
        e_gpu = np.abs(out[:, sl] - o64[:, sl]).max()
        e_32 = np.abs(o32[:, sl] - o64[:, sl]).max()
        say(f"{g}: |gpu-f64|={e_gpu:.3e} |oracle32-f64|={e_32:.3e}")
        if dtype == "float32":
            assert e_gpu <= 2.0 * e_32 + 1e-6, g
        else:
            assert e_gpu <= 1e-9 * max(np.abs(o64[:, sl]).max(), 1.0), g
    leaves = [par[:, i].clone().requires_grad_(True) for i in (0, 1, 2, 4)]
    res = tank.run(st, amp, *leaves[:3], E_jelly=leaves[3], payload_density=par[:, 3],
                   t0=cs.t0, nsteps=cs.steps, checkpoint=100, keep_last=0)
    tank.payload_com_y(res).sum().backward()
    scales = [500.0, 0.25, 1.1, 700.0]
    ad = [float(t.grad[0]) * s for t, s in zip(leaves, scales)]
    dps = [np.eye(5)[i] * s for i, s in zip((0, 1, 2, 4), scales)]
    fd = oracle_payload_fd(cs, pos, mat, fib, [np.zeros((P, 2))] * 4, dps, [1e-4, 1e-5],
                           nthreads=8)
    bound = {("float64", "fp64"): 1e-4, ("float32", "fp64"): 1e-3,
             ("float32", "fp32"): 2e-3}[(dtype, math)]
    for name, a, fds in zip(("amp", "contraction", "freq", "E_jelly"), ad, fd):
        rel = min(abs(a - x) / abs(x) for x in fds)
        say(f"d/d {name:11s} (x scale) AD={a: .6e} FD={fds} |AD-FD|/|FD|={rel:.1e}")
        assert rel <= bound, (name, rel)
    tank.close()


def _clustered_case(M, N, steps, seed=3):
    """M instances of N particles of random materials, each instance a patch of the
    4-per-cell lattice (2 particles per cell side) at (0.43, 0.43) with small jitter, mildly
    perturbed states (v, C ~ N(0, 0.05), F = I + N(0, 0.003)): particles share grid nodes,
    so the few-particle dynamics stay well conditioned (scattered particles of water's
    stiffness blow up within a dozen substeps, and a lone particle's position adjoint is a
    sum of cancelling terms, rounding-dominated at any precision)."""
    rng = np.random.default_rng(seed)
    P = M * N
    side = int(np.ceil(np.sqrt(N)))
    k = np.tile(np.arange(N), M)
    pos = 0.43 + np.stack([k % side, k // side], 1) / 256.0
    pos += rng.uniform(-0.1, 0.1, (P, 2)) / 256.0
    mat = rng.choice([0, 0, 1, 2, 3, 3], size=P).astype(np.int32)
    ang = rng.uniform(0, 2 * np.pi, P)
    fib = np.stack([np.cos(ang), np.sin(ang)], 1)
    state = np.zeros((P, 12))
    state[:, 0:2] = pos
    state[:, 2:4] = rng.normal(size=(P, 2)) * 0.05
    state[:, 4:8] = rng.normal(size=(P, 4)) * 0.05
    state[:, 8:12] = (np.eye(2) + rng.normal(size=(P, 2, 2)) * 0.003).reshape(P, 4)
    params = np.tile([500.0, 0.25, 1.1, 2.5, 700.0], (M, 1))
    params[:, 0] *= rng.uniform(0.8, 1.2, M)
    params[:, 4] *= rng.uniform(0.8, 1.2, M)
    return Case.from_arrays(np.repeat(np.arange(M), N), mat, fib, state,
                            rng.uniform(0.5, 1.5, P), params, steps=steps,
                            t0=0.05).round_f32()


@pytest.mark.parametrize("shape", [(40, 3), (5, 7), (2, 300)])
@pytest.mark.parametrize("dtype,math", PRECISIONS)
def test_diff_small_and_ragged_batches(require_gpu, dtype, math, shape):
    """Batches whose workgroups hold few particles or several instances: 40 instances of 3
    particles and 5 of 7 (one workgroup each), 2 of 300 (the second workgroup spans both),
    clustered particles of every material (_clustered_case), 12 substeps. A workgroup
    spanning instances takes the kernels' global-memory paths (no LDS tile; per-particle
    parameter-adjoint atomics), so these check those paths: the forward's last state and the
    whole adjoint (state, multiplier, parameters) against the fp64 host build, fp64 within
    1e-9 and fp32 buffers within three times the fp32 host build's own error, + 1e-6
    (relative norms; the GPU's global fp32 atomics add in a run-dependent order)."""
    import torch
    M, N = shape
    say(f"diff small/ragged {dtype}/{math}: {M} x {N} particles, 12 substeps, vs fp64 harness")
    cs = _clustered_case(M, N, 12)
    L64, L32 = load_harness(), load_harness(fp32=True)
    rng = np.random.default_rng(9)
    wout = rng.normal(size=(cs.P, 12))
    ref, f32 = {}, {}
    for L, res in ((L64, ref), (L32, f32)):
        traj = cs.forward(L)
        gin, gamp, gpar = cs.backward(L, traj, wout)
        for g, sl in GROUPS.items():
            res["state " + g] = traj[-1][:, sl]
            res["adjoint " + g] = gin[:, sl]
        res["amp_scale"] = gamp
        for i, n in ((0, "amp"), (1, "contraction"), (2, "freq"), (4, "E")):
            res[n] = gpar[:, i]
    tank, st, amp, par = _gpu(cs, dtype, math)
    out = tank.decode_state(tank.forward_raw(st, amp, par, cs.t0, cs.steps).double())
    out = out.cpu().numpy()
    gin, gamp, gpar = tank.backward_raw(st, amp, par, cs.t0, cs.steps,
                                        torch.tensor(wout, dtype=tank.dtype,
                                                     device=tank.device))
    gin, gamp, gpar = (t.cpu().numpy().astype(np.float64) for t in (gin, gamp, gpar))
    got = {"amp_scale": gamp}
    for g, sl in GROUPS.items():
        got["state " + g] = out[:, sl]
        got["adjoint " + g] = gin[:, sl]
    for i, n in ((0, "amp"), (1, "contraction"), (2, "freq"), (4, "E")):
        got[n] = gpar[:, i]
    for name, b in ref.items():
        nb = np.linalg.norm(b)
        if nb == 0.0:     # no muscle (or no jelly) in this draw: the adjoint must be 0 too
            assert np.linalg.norm(got[name]) == 0.0, name
            continue
        err = np.linalg.norm(got[name] - b) / nb
        e32 = np.linalg.norm(f32[name] - b) / nb
        tol = 3.0 * e32 + 1e-6 if dtype == "float32" else 1e-9
        say(f"{name}: |gpu-f64|/|f64| = {err:.3e} (fp32 host build {e32:.3e})")
        assert err <= tol, name
    assert np.all(gpar[:, 3] == 0)
    tank.close()
