"""Parity of the engine's axisymmetric variant (JELLY_AXISYM=1) on both paths, the plain
one (k_axisym_p2g) and the fused one (k_fused's axisymmetric instantiation), against the
oracle's (mpm_sim.py:127-133,163-165,236-304), through the C ABI.

The reference weights the grid mass and the particle-velocity momentum by the radius r but
not the affine (stress + APIC) momentum, so the affine part of every grid velocity carries
a factor 1/r; runs diverge within ~10 substeps next to the axis and within ~100 at
r >= 0.6 (the oracle shows the same). The runs compared here stop before that.

Tolerances as for the planar parity tests: one substep's grid mass and momentum within
1e-5 of their maximum, grid velocity within 1e-5 where the node mass is at least 1e-3 of
the maximum; after several substeps x, v, C and F within twice the fp32 oracle's own error
against the fp64 oracle (C: mean within 2x, tail quantiles 3x, maximum 5x; see fp32_like)."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle
from jellyfih_amd.engine import Engine
from tests.cases import fp32_like, reference_tank, small_tank

pytestmark = pytest.mark.gpu

PATHS = ["plain", "fused"]


def say(msg: str) -> None:
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def relmax(a, b):
    den = max(float(np.max(np.abs(b))), 1e-30)
    return float(np.max(np.abs(a.astype(np.float64) - b.astype(np.float64)))) / den


def axis_cases(n_slots=20000):
    """Tanks against the axis (x = 0) and one away from it."""
    return [
        small_tank(1, n_slots, region=(0.004, 0.40, 0.25, 0.55)),
        small_tank(2, n_slots, region=(0.004, 0.30, 0.05, 0.30), dead_every=7),
        small_tank(3, n_slots, region=(0.45, 0.85, 0.40, 0.70)),
    ]


def make_pair(cases, consts=None, dtypes=(np.float32,), path="plain", **eng_kw):
    M, N = len(cases), cases[0][0].shape[0]
    eng = Engine(M, N, path=path, axisym=True, **eng_kw)
    orcs = [oracle.OracleSim(M, N, consts, dtype=d, axisym=True) for d in dtypes]
    for m, (pos, mat, fib) in enumerate(cases):
        eng.load_particles(m, pos, mat, fib)
        for o in orcs:
            o.load(m, pos, mat, fib)
    return eng, orcs


def perturb(eng, orcs, seed, sv=0.05, sc=0.5, sf=0.02, t=0.1):
    o0 = orcs[0]
    rng = np.random.default_rng(seed)
    live = o0.material >= 0
    v = (rng.normal(size=o0.v.shape) * sv).astype(np.float32) * live[..., None]
    C = (rng.normal(size=o0.C.shape) * sc).astype(np.float32) * live[..., None, None]
    F = (np.eye(2) + rng.normal(size=o0.F.shape) * sf).astype(np.float32)
    F[~live] = np.eye(2, dtype=np.float32)
    eng.write("v", v)
    eng.write("C", C)
    eng.write("F", F)
    eng.set_time(t)
    for o in orcs:
        o.v[...] = v
        o.C[...] = C
        o.F[...] = F
        o.t[0] = np.float32(t)


@pytest.mark.parametrize("path", PATHS)
def test_axisym_one_substep(path):
    say(f"axisym one substep, tanks at the axis, {path} path")
    eng, (orc,) = make_pair(axis_cases(), path=path)
    perturb(eng, [orc], seed=21)
    eng.substep(1)
    orc.run(1)
    gm, gv = eng.read("grid_m"), eng.read("grid_v")
    assert relmax(gm, orc.grid_m) <= 1e-5, relmax(gm, orc.grid_m)
    mom = gv * gm[..., None]
    assert relmax(mom, orc.grid_v * orc.grid_m[..., None]) <= 1e-5
    heavy = orc.grid_m >= 1e-3 * orc.grid_m.max()
    dv = np.abs(gv - orc.grid_v)[heavy]
    say(f"grid velocity max |d| {dv.max():.3e} of {np.abs(orc.grid_v[heavy]).max():.3e}")
    assert dv.max() <= 1e-5 * np.abs(orc.grid_v[heavy]).max(), dv.max()
    # the axis condition: no radial velocity on the axis column
    assert (gv[:, 0, :, 0] == 0).all()
    x, v, F = eng.read("x"), eng.read("v"), eng.read("F")
    assert np.max(np.abs(x - orc.x)) <= 2e-7
    assert relmax(v, orc.v) <= 1e-5
    assert np.max(np.abs(F - orc.F)) <= 1e-6
    assert eng.get_time() == orc.t[0]
    eng.close()


@pytest.mark.parametrize("path", PATHS)
def test_axisym_40_substeps_with_a_sort(path):
    """40 substeps away from the axis (r >= 0.55) with a re-sort by cell after 20: the spawn
    radii must follow their particles through the permutation."""
    say(f"axisym 40 substeps at r >= 0.55, sort after 20, {path} path")
    cases = [small_tank(1, 20000, region=(0.55, 0.95, 0.30, 0.60)),
             small_tank(2, 20000, region=(0.60, 0.99, 0.05, 0.30), dead_every=7)]
    eng, (o32, o64) = make_pair(cases, dtypes=(np.float32, np.float64), path=path)
    eng.substep(20)
    eng.sort_now()
    eng.substep(20)
    o32.run(40, nthreads=8)
    o64.run(40, nthreads=8)
    live = o32.material >= 0
    for n, a32, a64 in (("x", o32.x, o64.x), ("v", o32.v, o64.v), ("F", o32.F, o64.F),
                        ("C", o32.C, o64.C)):
        fp32_like(n, eng.read(n)[live], a32[live], a64[live])
    assert np.max(np.abs(eng.read("x")[live] - o64.x[live])) <= 1e-5
    eng.close()


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("case", ["axis", "evolve", "view"])
def test_axisym_5_substeps(case, path):
    """Five substeps of tanks that reach the axis: small tanks; two axisymmetric phenotypes
    as evolve.py fills them under JELLY_AXISYM=1 (right-half profile at the axis, spawn
    (0, 0.40), make_jelly.py:276-304,343); and helpers/view_axisym.py's tank (Aurelia's
    right-half profile at spawn (0.5, 0.40) in the tall domain, 128x256, H = 2, 160K slots,
    view_axisym.py:19-25,129-135).

    The axis-spawned phenotypes put payload and bell particles at r = 0, where the
    reference's formulation gives them no grid mass but full affine momentum: the fp64
    oracle's largest |v| is 1.8e3 after one substep, 6.4e5 after two, 9e17 after three, and
    fp32 overflows at the fourth. That case compares two substeps."""
    steps = 2 if case == "evolve" else 5
    say(f"axisym {steps} substeps, {case}, {path} path")
    params = None
    if case == "axis":
        cases, consts, kw = axis_cases(), None, {}
    elif case == "evolve":
        tanks = [reference_tank(s, n_slots=80000, axisym=True) for s in (0, 1)]
        cases = [t[:3] for t in tanks]
        consts, kw = None, {}
        params = [(float(np.clip(t[3][9], 0.05, 0.6)), float(np.clip(t[3][10], 0.5, 2.0)))
                  for t in tanks]
    else:
        pos, mat, fib, g, info = reference_tank("aurelia", n_slots=160000, domain_height=2.0,
                                                spawn=np.array([0.5, 0.40]), axisym=True)
        cases = [(pos, mat, fib)]
        consts = oracle.reference_consts(n_grid_y=256, domain_height=2.0)
        kw = {"n_grid_y": 256, "domain_height": 2.0}
    eng, (o32, o64) = make_pair(cases, consts=consts, dtypes=(np.float32, np.float64),
                                path=path, **kw)
    if params:
        con = np.array([p[0] for p in params], np.float32)
        frq = np.array([p[1] for p in params], np.float32)
        eng.set_param("contraction", con)
        eng.set_param("freq", frq)
        for o in (o32, o64):
            o.contraction[:] = con
            o.freq[:] = frq
    eng.substep(steps)
    o32.run(steps, nthreads=8)
    o64.run(steps, nthreads=8)
    live = o32.material >= 0
    assert np.isfinite(o64.v[live]).all() and np.isfinite(o32.v[live]).all()
    # C next to the axis (spawn radius below 0.05, 6.4 cells), where the formulation's 1/r
    # amplifies every rounding, is compared by its mean error only; elsewhere as x, v, F.
    # (Measured on the view tank after 5 substeps: beyond r = 0.05 both paths' C errors
    # equal the fp32 oracle's in every radius band; the 5060 particles below it set the
    # whole-tank 99th percentile, 2.5x the oracle's on the plain path, 4.7x on the fused.)
    r0 = np.stack([c[0] for c in cases])[..., 0][live]
    far = r0 >= 0.05
    for n, a32, a64 in (("x", o32.x, o64.x), ("v", o32.v, o64.v), ("F", o32.F, o64.F),
                        ("C", o32.C, o64.C)):
        gpu, r32, r64 = eng.read(n)[live], a32[live], a64[live]
        if n != "C":
            fp32_like(n, gpu, r32, r64)
            continue
        fp32_like("C (r >= 0.05)", gpu[far], r32[far], r64[far])
        if (~far).any():
            e_gpu = np.abs(gpu[~far] - r64[~far]).mean()
            e_32 = np.abs(r32[~far] - r64[~far]).mean()
            say(f"C (r < 0.05, {int((~far).sum())} particles): mean |gpu-f64| {e_gpu:.3e}, "
                f"|oracle32-f64| {e_32:.3e}")
            assert e_gpu <= 2.0 * e_32 + 1e-6
    eng.close()


def test_axisym_fused_launch_shapes():
    """The fused axisymmetric kernel is built for 256 x 2 and 256 x 4 chunks:
    jm_set_tuning refuses other shapes on such an engine, and 256 x 2 with one instance group
    gives the 256 x 4 two-group run's state after 5 substeps (fp32-like against the oracle)."""
    from jellyfih_amd import _lib
    say("axisym fused: launch shapes 256x2 / 256x4, refused 256x8")
    cases = axis_cases()
    runs = {}
    for ppt, groups in ((4, 2), (2, 1)):
        eng, (o32, o64) = make_pair(cases, dtypes=(np.float32, np.float64), path="fused")
        if ppt == 4:
            with pytest.raises(_lib.JmError, match="256 x 2 or 256 x 4"):
                eng.set_tuning(ppt=8)
        eng.set_tuning(threads=256, ppt=ppt, instance_groups=groups)
        eng.substep(5)
        runs[ppt] = {n: eng.read(n) for n in ("x", "v", "F", "C")}
        eng.close()
    o32.run(5, nthreads=8)
    o64.run(5, nthreads=8)
    live = o32.material >= 0
    for ppt, r in runs.items():
        for n in ("x", "v", "F"):
            fp32_like(f"{n} (256 x {ppt})", r[n][live], getattr(o32, n)[live],
                      getattr(o64, n)[live])


def test_axisym_fused_headless_through_the_divergence():
    """evolve.py's headless run under JELLY_AXISYM=1 on the fused path (32 axis-spawned
    tanks, 80K slots, 20,000 substeps through run_batch_headless): the reference's
    formulation throws nearly every particle against the walls within ~10 substeps, with
    velocities that overflow fp32; the run still completes, every position stays inside
    the domain (the advection clamp, mpm_sim.py:331-333), and the clock is the fp32
    accumulation of dt."""
    from jellyfih_amd import phenotype
    from jellyfih_amd.diff import clock
    say("axisym fused, 32 axis-spawned tanks x 80K, run_batch_headless(20000)")
    pos, mat, fib, genomes, infos = phenotype.population(range(32), 80000, 128, axisym=True)
    eng = Engine(32, 80000, path="fused", axisym=True)
    for m in range(32):
        eng.load_particles(m, pos[m], mat[m], fib[m])
    res = eng.run_batch_headless(20000)
    x = eng.read("x")
    live = mat >= 0
    say(f"results finite rows {int(np.isfinite(res).all(1).sum())}/32, valid "
        f"{int(res[:, 4].sum())}, x in [{x[live].min():.4f}, {x[live].max():.4f}]")
    assert res.shape == (32, 5)
    assert np.isfinite(x[live]).all()
    assert x[live].min() >= 0.001 - 1e-7 and x[live].max() <= 0.999 + 1e-7
    assert eng.get_time() == np.float32(clock(0.0, 20000))
    eng.close()
