"""Failure handling of the HIP engine, through its C ABI, against the oracle.

* NaN propagation: a NaN velocity injected into a few particles reaches exactly the grid
  nodes of those particles and the particles that gather from them, on both step paths, as
  in the oracle (bit-exact NaN masks). The fused path's fixed-point LDS scatter must send
  such particles to the global grid (a NaN encoded as fixed point would be a large finite
  number instead).
* Instance isolation: a NaN in one instance leaves every other instance's positions and
  headless results bit-identical to a clean run. Each instance fits one fused chunk here, so
  every grid node receives one fp32 global add per substep and clean runs are bit-for-bit
  reproducible (larger instances flush several chunks per node in a run-dependent order).
* The opt-in NaN->invalid flag of jm_run_batch_headless, against the oracle's.
* Finite contributions beyond the fused kernel's fixed-point bound (particles at 2e5)
  scatter through global atomics beside their chunk's LDS window, against the oracle.
* The clock of a batch with no live particles still advances one fp32 dt per substep.
* Limits: the largest batch jm_create accepts (2^28 slots) against the oracle; one whole
  150,000-substep fitness episode of the 32 x 80K generation batch, checked by what holds
  at any length (clock, validity, finiteness, domain, the results' own statistics).
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle
from jellyfih_amd.engine import Engine
from tests.cases import small_tank

pytestmark = pytest.mark.gpu


def say(msg: str) -> None:
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


REGIONS = [(0.10, 0.22, 0.10, 0.22), (0.40, 0.52, 0.30, 0.42), (0.70, 0.82, 0.15, 0.27),
           (0.45, 0.57, 0.60, 0.72), (0.15, 0.27, 0.55, 0.67), (0.75, 0.87, 0.70, 0.82),
           (0.30, 0.42, 0.80, 0.92), (0.60, 0.72, 0.45, 0.57)]
N_SLOTS = 1100


def one_chunk_cases():
    return [small_tank(10 + m, N_SLOTS, region=r) for m, r in enumerate(REGIONS)]


def make_engine(cases, path="fused", nan_invalid=False):
    eng = Engine(len(cases), N_SLOTS, path=path, nan_invalid=nan_invalid)
    if path == "fused":
        eng.set_tuning(threads=256, ppt=4, instance_groups=2, sort_interval=100)
    for m, (p, mt, f) in enumerate(cases):
        eng.load_particles(m, p, mt, f)
    return eng


def relmax(a, b):
    """max |a - b| relative to max |b|."""
    den = max(float(np.max(np.abs(b))), 1e-30)
    return float(np.max(np.abs(a.astype(np.float64) - b.astype(np.float64)))) / den


def nan_slots(cases, m, picks=(5, 300, 700)):
    live = np.flatnonzero(cases[m][1] >= 0)
    return live[list(picks)]


@pytest.mark.parametrize("path", ["fused", "plain"])
def test_nan_propagation_matches_oracle(require_gpu, path):
    say(f"NaN propagation path={path}")
    cases = one_chunk_cases()
    eng = make_engine(cases, path)
    orc = oracle.OracleSim(len(cases), N_SLOTS)
    for m, (p, mt, f) in enumerate(cases):
        orc.load(m, p, mt, f)
    bad = nan_slots(cases, 3)
    v = np.zeros((len(cases), N_SLOTS, 2), np.float32)
    v[3, bad] = np.nan
    eng.write("v", v)
    orc.v[...] = v
    eng.set_time(0.05)
    orc.t[0] = np.float32(0.05)
    for steps in (1, 4):
        eng.substep(steps)
        orc.run(steps)
        vg = eng.read("v")
        nan_g = np.isnan(vg).any(-1)
        nan_o = np.isnan(orc.v).any(-1)
        np.testing.assert_array_equal(nan_g, nan_o)
        np.testing.assert_array_equal(np.isnan(eng.read("C")).any((-2, -1)),
                                      np.isnan(orc.C).any((-2, -1)))
        assert nan_g[3].sum() > len(bad) and not nan_g[np.arange(8) != 3].any()
        x = eng.read("x")
        assert np.isfinite(x).all()
        np.testing.assert_array_equal(x[3, nan_g[3]], np.float32([[0.999, 0.999]]) *
                                      np.ones((nan_g[3].sum(), 1), np.float32))
        clean = orc.material >= 0
        clean[3] = False
        assert float(np.max(np.abs(x[clean] - orc.x[clean]))) < 1e-5
    if path == "fused":
        # the grid the last G2P read: NaN at the same nodes as the oracle's grid
        gv = eng.read("grid_v")
        np.testing.assert_array_equal(np.isnan(gv).any(-1), np.isnan(orc.grid_v).any(-1))
    eng.close()


@pytest.mark.parametrize("path", ["fused", "plain"])
def test_contributions_beyond_the_fixed_point_bound(require_gpu, path):
    """Three particles of one instance moving at 2e5 (momentum m v ~ 3 against the fused
    kernel's fixed-point bound of 0.75, so they scatter through global fp32 atomics while the
    rest of their chunk uses the LDS window), one substep from t = 0.05: the grid's mass
    within 1e-5 of max of the fp32 oracle's; grid momentum and the particles' v within twice
    the fp32 oracle's error against fp64 (relative to max, + 1e-6); the three particles'
    x (on the walls or near them) within twice the fp32 oracle's error against fp64, + 1e-6;
    x of every other particle within 1e-5."""
    say(f"contributions beyond the fixed-point bound, path={path}")
    cases = one_chunk_cases()
    eng = make_engine(cases, path)
    o32 = oracle.OracleSim(len(cases), N_SLOTS)
    o64 = oracle.OracleSim(len(cases), N_SLOTS, dtype=np.float64)
    for m, (p, mt, f) in enumerate(cases):
        o32.load(m, p, mt, f)
        o64.load(m, p, mt, f)
    fast = nan_slots(cases, 2)
    v = np.zeros((len(cases), N_SLOTS, 2), np.float32)
    v[2, fast] = np.float32([[2e5, -1e5], [-2e5, 5e4], [1e5, 2e5]])
    eng.write("v", v)
    eng.set_time(0.05)
    for o in (o32, o64):
        o.v[...] = v
        o.t[0] = np.float32(0.05)
    eng.substep(1)
    o32.run(1)
    o64.run(1)
    gm, gv = eng.read("grid_m"), eng.read("grid_v")
    assert relmax(gm, o32.grid_m) <= 1e-5, relmax(gm, o32.grid_m)
    mom = gv.astype(np.float64) * gm[..., None]
    mom64 = o64.grid_v * o64.grid_m[..., None]
    mom32 = o32.grid_v.astype(np.float64) * o32.grid_m[..., None]
    e_gpu, e_32 = relmax(mom, mom64), relmax(mom32, mom64)
    say(f"grid momentum: |gpu-f64| {e_gpu:.3e}, |oracle32-f64| {e_32:.3e} (relative to max)")
    assert e_gpu <= 2.0 * e_32 + 1e-6
    live = o64.material >= 0
    vg = eng.read("v")
    e_gpu, e_32 = relmax(vg[live], o64.v[live]), relmax(o32.v[live], o64.v[live])
    say(f"v: |gpu-f64| {e_gpu:.3e}, |oracle32-f64| {e_32:.3e} (relative to max)")
    assert e_gpu <= 2.0 * e_32 + 1e-6
    x = eng.read("x")
    e_gpu = float(np.max(np.abs(x[2, fast] - o64.x[2, fast])))
    e_32 = float(np.max(np.abs(o32.x[2, fast] - o64.x[2, fast])))
    say(f"x of the fast particles {x[2, fast].tolist()}: |gpu-f64| {e_gpu:.3e}, "
        f"|oracle32-f64| {e_32:.3e}")
    assert e_gpu <= 2.0 * e_32 + 1e-6
    others = live.copy()
    others[2, fast] = False
    assert float(np.max(np.abs(x[others] - o64.x[others]))) < 1e-5
    eng.close()


def test_nan_instance_is_isolated(require_gpu):
    say("NaN instance isolation, 8 one-chunk instances, 300 substeps")
    cases = one_chunk_cases()
    bad = nan_slots(cases, 3)
    runs = {}
    for name, inject, flag in (("clean", False, False), ("clean2", False, False),
                               ("nan", True, False), ("nan_flag", True, True)):
        eng = make_engine(cases, "fused", nan_invalid=flag)
        if inject:
            v = eng.read("v")
            v[3, bad] = np.nan
            eng.write("v", v)
        res = eng.run_batch_headless(300)
        runs[name] = (res, eng.read("x"), eng.read("v"))
        eng.close()
    others = np.arange(8) != 3
    # clean runs of one-chunk instances are reproducible bit for bit
    np.testing.assert_array_equal(runs["clean"][0], runs["clean2"][0])
    np.testing.assert_array_equal(runs["clean"][1], runs["clean2"][1])
    for name in ("nan", "nan_flag"):
        res, x, v = runs[name]
        np.testing.assert_array_equal(res[others], runs["clean"][0][others])
        np.testing.assert_array_equal(x[others], runs["clean"][1][others])
        np.testing.assert_array_equal(v[others], runs["clean"][2][others])
        assert np.isnan(v[3]).any() and np.isfinite(x[3]).all()
    # the reference's gate keeps the diverged instance valid (its CoM is finite); the opt-in
    # flag marks it invalid and leaves the others as they were
    assert runs["nan"][0][3, 4] == runs["clean"][0][3, 4] == 1.0
    assert runs["nan_flag"][0][3, 4] == 0.0
    np.testing.assert_array_equal(runs["nan_flag"][0][others, 4], runs["clean"][0][others, 4])
    # the oracle agrees on the validity flags, with and without the flag
    orc = oracle.OracleSim(len(cases), N_SLOTS)
    for m, (p, mt, f) in enumerate(cases):
        orc.load(m, p, mt, f)
    orc.v[3, bad] = np.nan
    ref = orc.run_batch_headless(300, nan_invalid=True)
    np.testing.assert_array_equal(ref[:, 4], runs["nan_flag"][0][:, 4])
    assert float(np.max(np.abs(ref[others, :4] - runs["nan"][0][others, :4]))) < 1e-5


@pytest.mark.parametrize("path", ["fused", "plain"])
def test_empty_batch_clock(require_gpu, path):
    say(f"empty batch clock path={path}")
    eng = Engine(2, 100, path=path)
    eng.set_time(0.25)
    eng.substep(1000)
    t = np.float32(0.25)
    for _ in range(1000):
        t = np.float32(t + np.float32(2e-5))
    assert eng.get_time() == t
    # a run_batch_headless of an empty batch restarts the clock and reports nothing valid
    res = eng.run_batch_headless(150)
    assert np.all(res == 0.0)
    t = np.float32(0.0)
    for _ in range(150):
        t = np.float32(t + np.float32(2e-5))
    assert eng.get_time() == t
    eng.close()


def test_maximum_slot_count():
    """The largest batch jm_create accepts, 2^28 slots (1024 instances x 262,144 slots, each
    with a ~1,000-particle tank, the rest dead): the particles of the last instances sit at
    the top of the particle arrays (F and C byte offsets just below 2^32). Five substeps on
    the fused path; instances 0, 511 and 1023 against the oracle (x within twice the fp32
    oracle's own error against fp64, + 1e-6), every live position finite and in the
    domain."""
    M, N = 1024, 262144
    say(f"maximum slot count: {M} x {N} = {M * N} slots, 5 substeps")
    eng = Engine(M, N)
    keep = {0: None, 511: None, 1023: None}
    for m in range(M):
        p, mt, f = small_tank(100 + m, N, region=REGIONS[m % len(REGIONS)])
        eng.load_particles(m, p, mt, f)
        if m in keep:
            keep[m] = (p, mt, f)
    eng.substep(5)
    x = eng.read("x")
    mat = eng.read("material")
    live = mat >= 0
    assert np.isfinite(x[live]).all()
    assert x[live].min() >= 0.001 - 1e-7 and x[live].max() <= 0.999 + 1e-7
    orcs = [oracle.OracleSim(len(keep), N, dtype=d) for d in (np.float32, np.float64)]
    for o in orcs:
        for k, (p, mt, f) in enumerate(keep.values()):
            o.load(k, p, mt, f)
        o.run(5, nthreads=8)
    o32, o64 = orcs
    for k, m in enumerate(keep):
        lv = o32.material[k] >= 0
        np.testing.assert_array_equal(mat[m], o32.material[k])
        e_gpu = np.abs(x[m][lv] - o64.x[k][lv]).max()
        e_32 = np.abs(o32.x[k][lv] - o64.x[k][lv]).max()
        say(f"instance {m}: {int(lv.sum())} live, |x-x64| {e_gpu:.2e} (fp32 oracle {e_32:.2e})")
        assert e_gpu <= 2 * e_32 + 1e-6, m
    eng.close()


def test_full_fitness_episode():
    """One whole fitness episode of the 32 x 80K generation batch (seeds 0-31 through the
    fill_tank restatement, the actuation genes and the autotuned launch parameters of the
    throughput benchmark): run_batch_headless(150000), as evolve.py:48 runs it
    (mpm_sim.py:802-834). The oracle cannot run this far, so the test checks what holds at
    any length: the clock equals the fp32 sum of 150,000 steps; every result is finite and
    every instance valid; materials and live counts are unchanged and every live position
    is finite and inside the advection clamp [0.001, 0.999]; the initial and final payload
    CoM equal, bit for bit, the fp32 quotients of jm_compute_stats' sums of the state they
    were taken from; the fitness of evolve.py:90-124 is finite for every genome."""
    from jellyfih_amd import phenotype
    from jellyfih_amd.autotune import tuning_for
    from jellyfih_amd.diff import clock
    say("full fitness episode: 32 x 80K, run_batch_headless(150000)")
    pos, mat, fib, genomes, infos = phenotype.population(range(32), 80000, 128)
    eng = Engine(32, 80000)
    tuning = tuning_for(32, 80000)
    assert tuning is not None
    eng.set_tuning(**{k: v for k, v in tuning.items() if k in eng.tuning()})
    for m in range(32):
        eng.load_particles(m, pos[m], mat[m], fib[m])
    eng.set_param("contraction", np.clip(genomes[:, 9], 0.05, 0.60))
    eng.set_param("freq", np.clip(genomes[:, 10], 0.5, 2.0))
    live = [eng.live_count(m) for m in range(32)]
    init = oracle.finish_stats(eng.raw_stats("payload"))
    res = eng.run_batch_headless(150000)
    final = oracle.finish_stats(eng.raw_stats("payload"))
    say(f"clock {eng.get_time()!r}; valid {int(res[:, 4].sum())}/32; "
        f"final y {res[:, 2].min():.4f}..{res[:, 2].max():.4f}")
    assert eng.get_time() == np.float32(clock(0.0, 150000))
    assert np.isfinite(res).all()
    assert np.all(res[:, 4] == 1.0)
    np.testing.assert_array_equal(res[:, 0], init[:, 0])
    np.testing.assert_array_equal(res[:, 1], init[:, 1])
    np.testing.assert_array_equal(res[:, 2], final[:, 0])
    np.testing.assert_array_equal(res[:, 3], final[:, 1])
    assert [eng.live_count(m) for m in range(32)] == live
    np.testing.assert_array_equal(eng.read("material"), mat)
    x = eng.read("x")[mat >= 0]
    assert np.isfinite(x).all()
    assert x.min() >= np.float32(0.001) and x.max() <= np.float32(0.999)
    muscle = [info["muscle_count"] for info in infos]
    fit = oracle.compute_fitness(res, muscle, genomes)
    assert all(f is not None and np.isfinite(f) for f in fit)
    eng.close()
