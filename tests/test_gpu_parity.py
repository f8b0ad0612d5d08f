"""Parity of the HIP engine against the oracle (CPU restatement of mpm_sim.py:153-340).

Every test calls the engine through its C ABI (libjellymx.so via ctypes) and the oracle
(oracle/liboracle.so) on the same seeded inputs. Tolerances:

* integer outputs (materials, live counts, sort permutation, clock bits): bit-exact;
* grid mass and momentum of one substep: max |gpu - ref| <= 1e-5 * max |ref|;
* particle x after 1000 substeps of a full reference tank: within 1e-4 of the domain size
  (1.0) of both oracles; for x, v and F the GPU's error against the fp64 oracle is at
  most twice the fp32 oracle's own error against the fp64 oracle, and for C the same on
  the 99.9th percentile with the maximum within 5x (see fp32_like; both are reported).
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle
from jellyfih_amd.engine import Engine
from tests.cases import fp32_like, reference_tank, small_tank

pytestmark = pytest.mark.gpu


def say(msg: str) -> None:
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def make_pair(cases, path="fused", dtype=np.float32, **tuning):
    M = len(cases)
    N = cases[0][0].shape[0]
    eng = Engine(M, N, path=path)
    if tuning:
        eng.set_tuning(**tuning)
    orc = oracle.OracleSim(M, N, dtype=dtype)
    for m, (pos, mat, fib) in enumerate(cases):
        eng.load_particles(m, pos, mat, fib)
        orc.load(m, pos, mat, fib)
    return eng, orc


def perturb(eng, orc, seed=0, sv=0.05, sc=0.5, sf=0.02, t=0.1, orcs=()):
    """Give both engines the same non-trivial state: random v, C, F near I, mid-contraction."""
    rng = np.random.default_rng(seed)
    live = (orc.material >= 0)
    v = (rng.normal(size=orc.v.shape) * sv).astype(np.float32) * live[..., None]
    C = (rng.normal(size=orc.C.shape) * sc).astype(np.float32) * live[..., None, None]
    F = (np.eye(2) + rng.normal(size=orc.F.shape) * sf).astype(np.float32)
    F[~live] = np.eye(2, dtype=np.float32)
    eng.write("v", v)
    eng.write("C", C)
    eng.write("F", F)
    eng.set_time(t)
    for o in (orc,) + tuple(orcs):
        o.v[...] = v
        o.C[...] = C
        o.F[...] = F
        o.t[0] = np.float32(t)


def relmax(a, b):
    den = max(float(np.max(np.abs(b))), 1e-30)
    return float(np.max(np.abs(a.astype(np.float64) - b.astype(np.float64)))) / den


def ragged_cases(n_slots=20000):
    return [
        small_tank(1, n_slots, region=(0.30, 0.70, 0.25, 0.55)),
        small_tank(2, n_slots, region=(0.05, 0.45, 0.02, 0.20), dead_every=7),
        small_tank(3, n_slots, region=(0.55, 0.995, 0.70, 0.995)),   # touches the skip strip
    ]


@pytest.mark.parametrize("path", ["fused", "plain"])
def test_one_substep_grid_and_particles(require_gpu, path):
    say(f"one substep grid+particles path={path}")
    eng, orc = make_pair(ragged_cases(), path=path)
    perturb(eng, orc, seed=11)
    eng.substep(1)
    orc.run(1)
    gm, gv = eng.read("grid_m"), eng.read("grid_v")
    # grid mass and momentum per node (momentum after the grid op: velocity x mass)
    assert relmax(gm, orc.grid_m) <= 1e-5, relmax(gm, orc.grid_m)
    mom = gv * gm[..., None]
    assert relmax(mom, orc.grid_v * orc.grid_m[..., None]) <= 1e-5
    # velocity where the node carries mass (a near-massless node's p/m is ill-conditioned;
    # G2P weighs it by the same vanishing weight)
    heavy = orc.grid_m >= 1e-3 * orc.grid_m.max()
    dv = np.abs(gv - orc.grid_v)[heavy]
    assert dv.max() <= 1e-5 * np.abs(orc.grid_v[heavy]).max(), dv.max()
    x, v, C, F, Jp = (eng.read(n) for n in ("x", "v", "C", "F", "Jp"))
    assert np.max(np.abs(x - orc.x)) <= 2e-7
    assert relmax(v, orc.v) <= 1e-5
    assert relmax(C, orc.C) <= 1e-4
    assert np.max(np.abs(F - orc.F)) <= 1e-6
    assert np.max(np.abs(Jp - orc.Jp)) <= 1e-6
    np.testing.assert_array_equal(eng.read("material"), orc.material)
    assert eng.get_time() == orc.t[0]


@pytest.mark.parametrize("path", ["fused", "plain"])
def test_ragged_batch_200_substeps(require_gpu, path):
    say(f"ragged batch 200 substeps path={path}")
    eng, orc = make_pair(ragged_cases(), path=path)
    perturb(eng, orc, seed=5, sv=0.02, sc=0.2, sf=0.01, t=0.05)
    eng.substep(200)
    orc.run(200)
    x = eng.read("x")
    assert np.max(np.abs(x - orc.x)) <= 1e-5
    assert relmax(eng.read("v"), orc.v) <= 1e-3
    for m in range(3):
        assert eng.live_count(m) == int((orc.material[m] >= 0).sum())


@pytest.mark.parametrize("strip", [8, 0, 5])
def test_sort_permutation_bitexact(require_gpu, strip):
    say(f"sort permutation vs stable counting sort, strip {strip}")
    eng, orc = make_pair(ragged_cases())
    perturb(eng, orc, seed=3, sv=0.5)
    # no automatic sort inside the steps below (after the one a new order asks for)
    eng.set_tuning(sort_interval=1000000, sort_strip=strip)
    eng.substep(50)
    c = oracle.reference_consts()
    for m in range(3):
        slot0, x0 = eng.device_order(m)
        eng.sort_now()
        slot1, x1 = eng.device_order(m)
        perm = oracle.stable_cell_sort(c, x0, strip)
        np.testing.assert_array_equal(slot1, slot0[perm])
        np.testing.assert_array_equal(x1, x0[perm])
        keys = oracle.cell_keys(c, x1, strip)
        assert np.all(np.diff(keys) >= 0)


def test_unsorted_chunks_grid(require_gpu):
    """Grid of one substep from a state whose particles have left their sorted cells
    (no sort for 300 substeps of fast random motion), so that cells recur with gaps along
    the chunk order: the fused path's cross-lane pre-sums must count each particle once.
    The oracles (fp32 and fp64) start from the engine's state. Grid mass has no
    cancellation: within 1e-5 of the fp32 oracle (a particle counted twice or dropped
    moves a node by its whole weight). Momentum of this state is a sum of large stress
    terms that cancel, so its error against the fp64 oracle is bounded by twice the fp32
    oracle's own error, as for the long runs."""
    say("unsorted chunks: 300 substeps without sort, then one substep vs oracle f32/f64")
    cases = ragged_cases()
    eng, orc = make_pair(cases)
    perturb(eng, orc, seed=8, sv=0.5, sc=0.2, sf=0.01)
    eng.set_tuning(sort_interval=1000000)
    eng.substep(300)
    orc64 = oracle.OracleSim(len(cases), cases[0][0].shape[0], dtype=np.float64)
    for m, (pos, mat, fib) in enumerate(cases):
        orc64.load(m, pos, mat, fib)
    for name in ("x", "v", "C", "F", "Jp"):
        val = eng.read(name)
        for o in (orc, orc64):
            getattr(o, name)[...] = val
    for o in (orc, orc64):
        o.t[0] = eng.get_time()
    # the chunk order really is unsorted now
    c = oracle.reference_consts()
    _, x0 = eng.device_order(0)
    keys = oracle.cell_keys(c, x0)
    say(f"descending key steps in instance 0: {int(np.sum(np.diff(keys) < 0))}")
    assert np.sum(np.diff(keys) < 0) > 100
    eng.substep(1)
    orc.run(1)
    orc64.run(1)
    gm, gv = eng.read("grid_m"), eng.read("grid_v")
    assert relmax(gm, orc.grid_m) <= 1e-5, relmax(gm, orc.grid_m)
    mom = gv.astype(np.float64) * gm[..., None]
    mom64 = orc64.grid_v * orc64.grid_m[..., None]
    mom32 = orc.grid_v.astype(np.float64) * orc.grid_m[..., None]
    e_gpu, e_32 = relmax(mom, mom64), relmax(mom32, mom64)
    say(f"momentum: |gpu-f64| {e_gpu:.3e}, |oracle32-f64| {e_32:.3e} (relative to max)")
    assert e_gpu <= 2.0 * e_32 + 1e-6
    assert np.max(np.abs(eng.read("x") - orc64.x)) <= 2e-7


def test_reference_tank_1000_substeps(require_gpu):
    """The single-instance workload: one seeded reference tank (~62K live of 80K slots),
    1000 substeps."""
    say("reference tank seed0, 1000 substeps, fused vs oracle f32/f64")
    pos, mat, fib, g, info = reference_tank(0)
    eng = Engine(1, 80000)
    eng.load_particles(0, pos, mat, fib)
    eng.set_param("contraction", float(np.clip(g[9], 0.05, 0.6)))
    eng.set_param("freq", float(np.clip(g[10], 0.5, 2.0)))
    o32 = oracle.OracleSim(1, 80000, dtype=np.float32)
    o64 = oracle.OracleSim(1, 80000, dtype=np.float64)
    for o in (o32, o64):
        o.load(0, pos, mat, fib)
        o.contraction[:] = np.float32(np.clip(g[9], 0.05, 0.6))
        o.freq[:] = np.float32(np.clip(g[10], 0.5, 2.0))
    eng.substep(1000)
    o32.run(1000, nthreads=8)
    o64.run(1000, nthreads=8)
    out = {n: eng.read(n) for n in ("x", "v", "C", "F")}
    ref32 = {"x": o32.x, "v": o32.v, "C": o32.C, "F": o32.F}
    ref64 = {"x": o64.x, "v": o64.v, "C": o64.C, "F": o64.F}
    live = mat >= 0
    for n in out:
        fp32_like(n, out[n][0][live], ref32[n][0][live], ref64[n][0][live])
    # positions within 1e-4 of the domain size (1.0) after 1000 substeps
    assert np.max(np.abs(out["x"][0][live] - ref32["x"][0][live])) <= 1e-4
    assert np.max(np.abs(out["x"][0][live] - ref64["x"][0][live])) <= 1e-4
    assert eng.get_time() == o32.t[0]


def test_tall_tank_500_substeps(require_gpu):
    """The tall configuration (evolve.py:33-36,136): 128x256 grid, domain height 2, 160K
    slots, spawn (0.5, 0.60); one seeded tank (~126K live), 500 substeps, fused path
    against the oracle in fp32 and fp64."""
    from jellyfih_amd import phenotype
    say("tall tank seed3, 128x256 H=2, 500 substeps, fused vs oracle f32/f64")
    pos, mat, fib, g, info = reference_tank(3, n_slots=160000, domain_height=2.0,
                                            spawn=phenotype.TALL_SPAWN)
    eng = Engine(1, 160000, n_grid_y=256, domain_height=2.0)
    eng.load_particles(0, pos, mat, fib)
    con, frq = float(np.clip(g[9], 0.05, 0.6)), float(np.clip(g[10], 0.5, 2.0))
    eng.set_param("contraction", con)
    eng.set_param("freq", frq)
    consts = oracle.reference_consts(n_grid_y=256, domain_height=2.0)
    o32 = oracle.OracleSim(1, 160000, consts, dtype=np.float32)
    o64 = oracle.OracleSim(1, 160000, consts, dtype=np.float64)
    for o in (o32, o64):
        o.load(0, pos, mat, fib)
        o.contraction[:] = np.float32(con)
        o.freq[:] = np.float32(frq)
    eng.substep(500)
    o32.run(500, nthreads=8)
    o64.run(500, nthreads=8)
    live = mat >= 0
    assert int(live.sum()) == eng.live_count(0)
    for n, a32, a64 in (("x", o32.x, o64.x), ("v", o32.v, o64.v), ("C", o32.C, o64.C),
                        ("F", o32.F, o64.F)):
        fp32_like(n, eng.read(n)[0][live], a32[0][live], a64[0][live])
    assert np.max(np.abs(eng.read("x")[0][live] - o64.x[0][live])) <= 1e-4
    assert eng.get_time() == o32.t[0]


def test_headless_matches_oracle(require_gpu):
    say("run_batch_headless vs oracle")
    cases = [reference_tank(s)[:3] for s in (1, 2)]
    eng, orc = make_pair(cases)
    r_gpu = eng.run_batch_headless(300)
    r_ref = orc.run_batch_headless(300, nthreads=8)
    np.testing.assert_array_equal(r_gpu[:, 4], r_ref[:, 4])
    assert np.max(np.abs(r_gpu[:, :4] - r_ref[:, :4])) <= 1e-5


def test_sixteen_uneven_instances(require_gpu):
    """The batched-forward configuration: 16 seeded genomes (seeds 1-16, uneven live counts)
    in one launch per stage, 1000 substeps through run_batch_headless: validity flags equal
    and per-instance payload CoM (start and end) within what the fp32 oracle itself shows
    against the fp64 oracle (2x, + 1e-6); the x of every particle likewise (fp32_like)."""
    from jellyfih_amd import phenotype
    say("16 uneven instances (seeds 1-16), 1000 substeps, fused vs oracle f32/f64")
    pos, mat, fib, genomes, infos = phenotype.population(range(1, 17), 80000, 128)
    live = (mat >= 0).sum(1)
    say(f"live particles per instance {live.min()}-{live.max()}")
    cases = [(pos[m], mat[m], fib[m]) for m in range(16)]
    eng, o32 = make_pair(cases)
    o64 = oracle.OracleSim(16, 80000, dtype=np.float64)
    for m, (p_, mt_, f_) in enumerate(cases):
        o64.load(m, p_, mt_, f_)
    con = np.clip(genomes[:, 9], 0.05, 0.60).astype(np.float32)
    frq = np.clip(genomes[:, 10], 0.5, 2.0).astype(np.float32)
    eng.set_param("contraction", con)
    eng.set_param("freq", frq)
    for o in (o32, o64):
        o.contraction[:] = con
        o.freq[:] = frq
    r_gpu = eng.run_batch_headless(1000)
    r_32 = o32.run_batch_headless(1000, nthreads=16)
    r_64 = o64.run_batch_headless(1000, nthreads=16)
    np.testing.assert_array_equal(r_gpu[:, 4], r_64[:, 4])
    e_gpu = np.abs(r_gpu[:, :4] - r_64[:, :4]).max()
    e_32 = np.abs(r_32[:, :4] - r_64[:, :4]).max()
    say(f"CoM: |gpu-f64| {e_gpu:.3e}, |oracle32-f64| {e_32:.3e}")
    assert e_gpu <= 2.0 * e_32 + 1e-6
    live_mask = mat >= 0
    fp32_like("x", eng.read("x")[live_mask], o32.x[live_mask], o64.x[live_mask])


def test_generation_batch_with_bench_tuning(require_gpu):
    """The measured configuration itself: the 32 seeded genomes x 80K slots of the
    throughput benchmark, with the autotuned launch parameters it runs (tuning.json), 600
    substeps through run_batch_headless and the fitness of evolve.py:90-124: validity flags
    equal, payload CoM and x within what the fp32 oracle shows against the fp64 oracle
    (2x, + 1e-6), fitness within 1e-4 of the fp64 oracle's."""
    from jellyfih_amd import phenotype
    from jellyfih_amd.autotune import tuning_for
    say("32-genome generation batch with the bench tuning, 600 substeps")
    pos, mat, fib, genomes, infos = phenotype.population(range(32), 80000, 128)
    cases = [(pos[m], mat[m], fib[m]) for m in range(32)]
    eng, o32 = make_pair(cases)
    tuning = tuning_for(32, 80000)
    assert tuning is not None
    eng.set_tuning(**{k: v for k, v in tuning.items() if k in eng.tuning()})
    say(f"tuning {eng.tuning()}")
    o64 = oracle.OracleSim(32, 80000, dtype=np.float64)
    for m, (p_, mt_, f_) in enumerate(cases):
        o64.load(m, p_, mt_, f_)
    con = np.clip(genomes[:, 9], 0.05, 0.60).astype(np.float32)
    frq = np.clip(genomes[:, 10], 0.5, 2.0).astype(np.float32)
    eng.set_param("contraction", con)
    eng.set_param("freq", frq)
    for o in (o32, o64):
        o.contraction[:] = con
        o.freq[:] = frq
    r_gpu = eng.run_batch_headless(600)
    r_32 = o32.run_batch_headless(600, nthreads=16)
    r_64 = o64.run_batch_headless(600, nthreads=16)
    np.testing.assert_array_equal(r_gpu[:, 4], r_64[:, 4])
    e_gpu = np.abs(r_gpu[:, :4] - r_64[:, :4]).max()
    e_32 = np.abs(r_32[:, :4] - r_64[:, :4]).max()
    say(f"CoM: |gpu-f64| {e_gpu:.3e}, |oracle32-f64| {e_32:.3e}")
    assert e_gpu <= 2.0 * e_32 + 1e-6
    muscle = [info["muscle_count"] for info in infos]
    f_gpu = np.array(oracle.compute_fitness(r_gpu, muscle, genomes), dtype=np.float64)
    f_64 = np.array(oracle.compute_fitness(r_64, muscle, genomes), dtype=np.float64)
    say(f"fitness max |gpu-f64| {np.abs(f_gpu - f_64).max():.3e}")
    np.testing.assert_allclose(f_gpu, f_64, atol=1e-4)
    live_mask = mat >= 0
    fp32_like("x", eng.read("x")[live_mask], o32.x[live_mask], o64.x[live_mask])


def test_quality_two_grid(require_gpu):
    """quality = 2 (mpm_sim.py:12,14,19: a 256x256 grid, dt = 1e-5; the reference fixes
    quality = 1, the engine takes it as a parameter): two small tanks at 4 particles per
    cell of the finer grid, 300 substeps, fused path against the oracle in fp32 and fp64."""
    say("quality 2 (256x256, dt 1e-5), 300 substeps, fused vs oracle f32/f64")
    cases = [small_tank(5, 12000, region=(0.30, 0.50, 0.30, 0.45), ppc_spacing=1 / 512),
             small_tank(6, 12000, region=(0.55, 0.75, 0.10, 0.25), ppc_spacing=1 / 512)]
    consts = oracle.reference_consts(quality=2)
    eng = Engine(2, 12000, quality=2)
    o32 = oracle.OracleSim(2, 12000, consts, dtype=np.float32)
    o64 = oracle.OracleSim(2, 12000, consts, dtype=np.float64)
    for m, (p_, mt_, f_) in enumerate(cases):
        eng.load_particles(m, p_, mt_, f_)
        o32.load(m, p_, mt_, f_)
        o64.load(m, p_, mt_, f_)
    eng.set_time(0.05)
    for o in (o32, o64):
        o.t[0] = np.float32(0.05)
    eng.substep(300)
    o32.run(300, nthreads=2)
    o64.run(300, nthreads=2)
    live = o32.material >= 0
    for n, a32, a64 in (("x", o32.x, o64.x), ("v", o32.v, o64.v), ("F", o32.F, o64.F),
                        ("C", o32.C, o64.C)):
        fp32_like(n, eng.read(n)[live], a32[live], a64[live])
    assert eng.get_time() == o32.t[0]


@pytest.mark.parametrize("path", ["fused", "plain"])
def test_glow_pass(require_gpu, path):
    """The render-only glow pass (mpm_sim.py:335-338, engine option track_glow): glow =
    max(0.9 glow, |1 - Jp|) after every substep, on tanks with a stiff payload whose Jp
    moves; 200 substeps against the oracle with its glow stage, in fp32 and fp64: glow and
    Jp within what the fp32 oracle shows against the fp64 oracle (fp32_like)."""
    say(f"glow pass path={path}, 200 substeps")
    cases = ragged_cases(12000)[:2]
    M, N = len(cases), cases[0][0].shape[0]
    eng = Engine(M, N, path=path, track_glow=True)
    o32 = oracle.OracleSim(M, N)
    o64 = oracle.OracleSim(M, N, dtype=np.float64)
    for m, (p_, mt_, f_) in enumerate(cases):
        eng.load_particles(m, p_, mt_, f_)
        o32.load(m, p_, mt_, f_)
        o64.load(m, p_, mt_, f_)
    perturb(eng, o32, seed=4, sv=0.05, sc=0.5, sf=0.02, orcs=(o64,))
    eng.substep(200)
    for o in (o32, o64):
        o.run(200, stages=oracle.ST_SUBSTEP | oracle.ST_GLOW, nthreads=2)
    live = o32.material >= 0
    assert o64.glow[live].max() > 1e-3
    fp32_like("glow", eng.read("visual_glow")[live], o32.glow[live], o64.glow[live])
    fp32_like("Jp", eng.read("Jp")[live], o32.Jp[live], o64.Jp[live])


def test_variants_agree(require_gpu):
    """Graphs on/off, lazy v/C on/off, XCD affinity, 1-4 concurrent instance groups,
    several launch shapes, and block lengths that do not divide the 60 substeps (37: a block
    of 37 then one of 23, with sorts every 13 substeps, off the block boundaries) or graph
    one substep at a time."""
    say("fused variants")
    cases = ragged_cases(12000)[:2] * 5          # 10 instances (affinity needs >= 8)
    base = None
    variants = [dict(), dict(use_graphs=0), dict(lazy_vc=0), dict(xcd_affinity=0),
                dict(threads=256, ppt=4), dict(threads=512, ppt=8), dict(threads=128, ppt=16),
                dict(threads=256, ppt=2), dict(threads=512, ppt=2),
                dict(sort_interval=7, block_substeps=5), dict(instance_groups=1),
                dict(instance_groups=3, use_graphs=0), dict(instance_groups=4),
                dict(block_substeps=37, sort_interval=13), dict(block_substeps=1)]
    for var in variants:
        say(f"variant {var}")
        eng, orc = make_pair(cases, **var)
        perturb(eng, orc, seed=2, sv=0.02, sc=0.2, sf=0.01)
        eng.substep(60)
        x = eng.read("x")
        if base is None:
            orc.run(60)
            base = orc.x.copy()
        assert np.max(np.abs(x - base)) <= 1e-5, var


def test_clock_and_params(require_gpu):
    say("clock accumulation and per-instance params")
    eng, orc = make_pair(ragged_cases(12000)[:1])
    eng.set_time(0.25)
    eng.substep(333)
    t = np.float32(0.25)
    for _ in range(333):
        t = np.float32(t + np.float32(2e-5))
    assert eng.get_time() == t
    eng.set_param("actuation", [123.0])
    assert eng.get_param("actuation")[0] == np.float32(123.0)


def test_empty_and_tiny_instances(require_gpu):
    say("empty and tiny instances")
    N = 5000
    pos, mat, fib = small_tank(4, N, region=(0.4, 0.6, 0.4, 0.5))
    dead = (np.full((N, 2), -1, np.float32), np.full(N, -1, np.int32), None)
    tiny_pos = pos.copy()
    tiny_mat = np.full(N, -1, np.int32)
    tiny_mat[:3] = mat[:3]
    eng = Engine(3, N)
    orc = oracle.OracleSim(3, N)
    for m, (p, mt, f) in enumerate([(pos, mat, fib), dead, (tiny_pos, tiny_mat, fib)]):
        eng.load_particles(m, p, mt, f)
        orc.load(m, p, mt, f)
    eng.substep(20)
    orc.run(20)
    assert eng.live_count(1) == 0
    assert np.max(np.abs(eng.read("x") - orc.x)) <= 1e-6
    st = eng.raw_stats("any")
    ref = orc.raw_stats(False)
    assert st[1, 1] == 0
    assert abs(st[0, 0] - ref[0, 0]) <= 1e-4 * max(1.0, abs(ref[0, 0]))


def test_field_write_roundtrip(require_gpu):
    say("field write/read round trip")
    eng, orc = make_pair(ragged_cases(12000)[:2])
    rng = np.random.default_rng(9)
    for name in ("x", "v", "C", "F", "Jp", "fiber_dir"):
        shape = eng.field_shape({"x": 0, "v": 1, "C": 2, "F": 3, "Jp": 5, "fiber_dir": 6}[name])
        val = rng.uniform(0.1, 0.9, size=shape).astype(np.float32)
        eng.write(name, val)
        back = eng.read(name)
        live = orc.material >= 0
        np.testing.assert_array_equal(back[live], val[live])


def scattered_case(seed=21, n_slots=4000, n=3000):
    """n jelly and muscle particles scattered uniformly over 80% of the tank (~0.2 per cell;
    water or payload this sparse is unstable at the reference's dt: its stiffness acts on
    the mass of one or two particles per node)."""
    rng = np.random.default_rng(seed)
    pos = np.full((n_slots, 2), -1.0, np.float32)
    mat = np.full(n_slots, -1, np.int32)
    fib = np.zeros((n_slots, 2), np.float32)
    fib[:, 1] = 1.0
    pos[:n] = rng.uniform(0.1, 0.9, (n, 2))
    mat[:n] = rng.choice([1, 1, 1, 3], n)
    mus = np.flatnonzero(mat == 3)
    ang = rng.uniform(0, 2 * np.pi, len(mus))
    fib[mus] = np.stack([np.cos(ang), np.sin(ang)], 1)
    return pos, mat, fib


@pytest.mark.parametrize("path", ["fused", "plain"])
def test_scattered_instance_through_the_global_grid(require_gpu, path):
    """An instance whose 3,000 particles are scattered over the whole tank, beside a dense
    small tank: after the cell sort a 1024-particle chunk of the scattered instance still
    spans about a third of the tank, a window larger than the LDS holds, so the fused kernel
    gathers from and scatters to the global grid for it (the plain path always does). One
    substep's grid (mass within 1e-5 of max against the fp32 oracle), then 100 substeps:
    x, v, F, C against the fp32/fp64 oracles (fp32_like)."""
    say(f"scattered instance beside a dense tank, {path}: 1 + 99 substeps vs oracle f32/f64")
    cases = [scattered_case(), small_tank(22, 4000, region=(0.35, 0.55, 0.35, 0.5))]
    eng, o32 = make_pair(cases, path=path)
    o64 = oracle.OracleSim(2, 4000, dtype=np.float64)
    for m, c in enumerate(cases):
        o64.load(m, *c)
    eng.substep(1)
    o32.run(1)
    o64.run(1)
    gm = eng.read("grid_m")
    assert relmax(gm, o32.grid_m) <= 1e-5, relmax(gm, o32.grid_m)
    if path == "fused":
        # the premise: some chunk's window (bounding box of its base cells + 2-node halo)
        # exceeds the kernel's LDS window (1280 nodes for chunks of up to 1024 particles)
        tun = eng.tuning()
        chunk = tun["threads"] * tun["ppt"]
        assert chunk <= 1024, chunk
        _, x0 = eng.device_order(0)
        b = oracle.base_cells(oracle.reference_consts(), x0)
        boxes = [(np.ptp(b[i:i + chunk, 0]) + 3) * (np.ptp(b[i:i + chunk, 1]) + 3)
                 for i in range(0, len(b), chunk)]
        say(f"chunk of {chunk}: window nodes {boxes}")
        assert max(boxes) > 1280, boxes
    eng.substep(99)
    o32.run(99, nthreads=8)
    o64.run(99, nthreads=8)
    live = o32.material >= 0
    for n, a32, a64 in (("x", o32.x, o64.x), ("v", o32.v, o64.v), ("F", o32.F, o64.F),
                        ("C", o32.C, o64.C)):
        fp32_like(n, eng.read(n)[live], a32[live], a64[live])
    eng.close()


@pytest.mark.parametrize("path", ["fused", "plain"])
def test_per_instance_parameters(require_gpu, path):
    """Per-instance inputs of evolve.py and the facade (mpm_sim.py:116-125,142-149,197-201;
    E_jelly, a module constant in the reference, taken per instance at run time): four small
    tanks with actuation 500 / 0 / 1000 / 250, payload density 2.5 / 1 / 5 / 2.5 and E_jelly
    700 / 300 / 1400 / 700 (the payload of density 5 is heavier than the fixed-point mass
    bound, so it scatters through global atomics), 300 substeps from t = 0.05 (a contraction
    phase): x, v, F, C against the fp32/fp64 oracles with the same parameters (fp32_like)."""
    say(f"per-instance actuation, payload density, E_jelly, {path}: 300 substeps vs oracle")
    regions = [(0.30, 0.50, 0.30, 0.45), (0.55, 0.75, 0.10, 0.25), (0.10, 0.30, 0.55, 0.70),
               (0.60, 0.80, 0.60, 0.75)]
    cases = [small_tank(30 + m, 4000, region=r) for m, r in enumerate(regions)]
    amp = np.float32([500.0, 0.0, 1000.0, 250.0])
    pdens = np.float32([2.5, 1.0, 5.0, 2.5])
    e_jelly = np.float32([700.0, 300.0, 1400.0, 700.0])
    eng, o32 = make_pair(cases, path=path)
    o64 = oracle.OracleSim(4, 4000, dtype=np.float64)
    for m, c in enumerate(cases):
        o64.load(m, *c)
    eng.set_param("actuation", amp)
    eng.set_param("payload_density", pdens)
    eng.set_param("e_jelly", e_jelly)
    eng.set_time(0.05)
    for o in (o32, o64):
        o.amp[:] = amp
        o.pdens[:] = pdens
        o.set_e_jelly(e_jelly)
        o.t[0] = np.float32(0.05)
    np.testing.assert_array_equal(eng.get_param("payload_density"), pdens)
    np.testing.assert_array_equal(eng.get_param("e_jelly"), e_jelly)
    eng.substep(300)
    o32.run(300, nthreads=4)
    o64.run(300, nthreads=4)
    live = o32.material >= 0
    for n, a32, a64 in (("x", o32.x, o64.x), ("v", o32.v, o64.v), ("F", o32.F, o64.F),
                        ("C", o32.C, o64.C)):
        fp32_like(n, eng.read(n)[live], a32[live], a64[live])
    # the parameters move the particles: instance 0 (the defaults) equals a default run,
    # each other instance differs from it
    d = oracle.OracleSim(4, 4000, dtype=np.float64)
    for m, c in enumerate(cases):
        d.load(m, *c)
    d.t[0] = np.float32(0.05)
    d.run(300, nthreads=4)
    moved = [float(np.abs(o64.x[m][live[m]] - d.x[m][live[m]]).max()) for m in range(4)]
    say(f"max |x - x(default parameters)| per instance: {moved}")
    assert moved[0] == 0.0 and all(v > 1e-6 for v in moved[1:]), moved
    eng.close()


@pytest.mark.parametrize("path", ["fused", "plain"])
def test_headless_without_payload(require_gpu, path):
    """run_batch_headless's statistics when an instance has no payload (mpm_sim.py:764-800,
    826-834): one small tank with a payload, one whose body has none (its CoM falls back to
    the body's: jelly and muscle), one of water only (no body: count 0, invalid); 200
    substeps (the run resets the clock to 0, mpm_sim.py:808). Validity flags equal to the
    oracle's, CoM within 1e-5, and the
    payload and body statistics of the final state equal the oracle's sums within fp32
    rounding."""
    say(f"headless without payload, {path}: 3 tanks, 200 substeps vs oracle")
    with_payload = small_tank(40, 4000, region=(0.30, 0.50, 0.30, 0.45))
    no_payload = small_tank(41, 4000, region=(0.55, 0.75, 0.10, 0.25))
    no_payload[1][no_payload[1] == 2] = 1           # payload block becomes jelly
    water = small_tank(42, 4000, region=(0.10, 0.30, 0.55, 0.70), robot=False)
    cases = [with_payload, no_payload, water]
    eng, orc = make_pair(cases, path=path)
    r_gpu = eng.run_batch_headless(200)
    r_ref = orc.run_batch_headless(200, nthreads=3)
    say(f"gpu {r_gpu.tolist()}\noracle {r_ref.tolist()}")
    np.testing.assert_array_equal(r_gpu[:, 4], r_ref[:, 4])
    np.testing.assert_array_equal(r_gpu[:, 4], [1.0, 1.0, 0.0])
    assert np.max(np.abs(r_gpu[:, :4] - r_ref[:, :4])) <= 1e-5
    for mode in ("payload", "body"):
        g = eng.raw_stats(mode).astype(np.float64)
        o = np.asarray(orc.raw_stats(mode == "body"), np.float64)
        np.testing.assert_array_equal(g[:, 1], o[:, 1])          # counts
        assert np.max(np.abs(g[:, [0, 2]] - o[:, [0, 2]])) <= 1e-5 * max(o[:, 1].max(), 1.0)
    eng.close()
