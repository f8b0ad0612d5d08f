"""The ``mpm_sim`` facade driven as the reference's callers drive it (tests/facade_driver.py,
an evolve.py-style generation plus view-mode frames, in a subprocess with its own JELLY_*
environment), checked against the oracle on the same population:

* ``run_batch_headless`` results: validity flags equal, CoM within 1e-5 (as the engine's own
  headless test), fitness (evolve.py:90-124) from both within 1e-4;
* the frames the facade read back: the abyss HDR frame against the renderer oracle on the
  state the facade read back (splat tolerance of tests/test_gpu_render.py), the tone-mapped
  frame against the oracle's tone map of that HDR frame (8 ulps) and its 8-bit encode
  (frame_buffer.to_uint8) bit-exact, the web palette bit-exact;
  ``vort_grid`` against the oracle's vorticity of the ``grid_v``/``grid_m`` read back (1e-6)
  and the RdBu composite against the oracle's over the same tone-mapped frame (4 ulps);
* 137 queued ``substep()`` calls against the same substeps asked of the engine directly:
  the clock bit-exact, positions within 1e-5.
"""
from __future__ import annotations

import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
from jellyfih_amd import phenotype
from oracle import render

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_evolve_style_driver(tmp_path):
    M, N = 4, 80000
    out = tmp_path / "facade.npz"
    env = dict(os.environ, JELLY_INSTANCES=str(M), JELLY_PARTICLES=str(N))
    print(f"[gpu-case] facade driver, {M} x {N}", flush=True)
    subprocess.run([sys.executable, "-u", "-m", "tests.facade_driver", str(out)], cwd=ROOT,
                   env=env, check=True, timeout=300)
    d = np.load(out)

    orc = oracle.OracleSim(M, N)
    for i, g in enumerate(d["genomes"]):
        pos, mat, fib, _ = phenotype.fill_tank(g, N, grid_res=128)
        orc.load(i, pos, mat, fib)
    orc.contraction[:] = d["contraction"]
    orc.freq[:] = d["freq"]
    ref = orc.run_batch_headless(400, nthreads=8)
    np.testing.assert_array_equal(d["results"][:, 4], ref[:, 4])
    assert np.max(np.abs(d["results"][:, :4] - ref[:, :4])) <= 1e-5
    f_gpu = oracle.compute_fitness(d["results"], d["muscle_counts"], d["genomes"])
    f_ref = oracle.compute_fitness(ref, d["muscle_counts"], d["genomes"])
    np.testing.assert_allclose(f_gpu, f_ref, atol=1e-4)

    res = 1024
    want = render.render_frame_abyss(np.zeros((res, res, 3), np.float32), d["x"], d["v"],
                                     d["Jp"], d["material"], np.float32(d["t"]), d["freq"],
                                     d["contraction"], d["hue"], d["muscle_hue"], 0,
                                     int(d["res_sub"]), int(d["grid_side"]), float(d["radius"]),
                                     1.0, oracle.reference_consts())
    err = np.abs(d["hdr"].astype(np.float64) - want)
    lim = 1e-4 * np.abs(want) + 1e-6 * np.abs(want).max()
    print(f"[gpu-case] facade abyss frame max |d| {err.max():.3e} of {want.max():.3e}",
          flush=True)
    assert want.max() > 0 and (err <= lim).all()
    tm = render.tone_map_and_encode(d["hdr"].copy())
    ulps = np.abs(d["ldr"].astype(np.float64) - tm) / np.spacing(np.maximum(np.abs(tm), 1e-30))
    assert ulps.max() <= 8, ulps.max()
    np.testing.assert_array_equal(d["ldr_u8"], render.encode_u8(d["ldr"]))
    web = np.full((res, res, 3), 0.98, np.float32)
    for mat_id, r, g, b in [(0, 0.933, 0.949, 0.957), (1, 0.678, 0.882, 0.875),
                            (3, 0.976, 0.757, 0.765), (2, 0.961, 0.824, 0.576)]:
        render.render_flat_pass(web, d["x"], d["material"], mat_id, (r, g, b),
                                int(d["res_sub"]), int(d["grid_side"]), float(d["radius"]), 1.0)
    np.testing.assert_array_equal(d["web"], web)
    # queued substep() calls: the same clock, bit for bit, and the same positions up to the
    # order of the fp32 global atomics across workgroups
    assert d["t_queued"] == d["t_direct"]
    live = d["material"] >= 0
    dq = np.abs(d["x_queued"][live] - d["x_direct"][live]).max()
    print(f"[gpu-case] 137 queued substeps vs direct: max |dx| {dq:.2e}", flush=True)
    assert dq <= 1e-5
    vort = render.compute_vorticity_grid(d["grid_v"], d["grid_m"], 1 / 128, 0.1)
    assert np.abs(vort).max() > 0.02
    assert np.abs(d["vort"] - vort).max() <= 1e-6
    rdbu = render.render_vorticity_pixels(d["base"].copy(), d["vort"], "rdbu", 0.6,
                                          int(d["res_sub"]), int(d["grid_side"]), 1.0, M)
    assert (rdbu != d["base"]).any()
    ulps = np.abs(d["rdbu"].astype(np.float64) - rdbu) / np.spacing(
        np.maximum(np.abs(rdbu), 1e-30))
    assert ulps.max() <= 4, ulps.max()


AXISYM_DRIVER = r"""
import sys
import numpy as np
import jellyfih_amd.mpm_sim as sim
from jellyfih_amd import phenotype
out = sys.argv[1]
for i in range(sim.n_instances):
    pos, mat, fib, info = phenotype.fill_tank(phenotype.seeded_genome(i), sim.n_particles,
                                              grid_res=128, axisym=True)
    sim.load_particles(i, pos, mat, fib)
for _ in range(2):
    sim.substep()
sim.ti.sync()
np.savez(out, x=sim.x.to_numpy(), v=sim.v.to_numpy(), F=sim.F.to_numpy(),
         grid_m=sim.grid_m.to_numpy(), t=sim.sim_time[None],
         path=sim.engine().path)
"""


def test_axisym_facade(tmp_path):
    """JELLY_AXISYM=1 through the facade (two axisymmetric phenotypes as evolve.py fills
    them, make_jelly.py:276-304,343; two substep() calls, the horizon before the reference's
    formulation overflows fp32 next to the axis): the facade runs the fused path, and x, v, F
    are as accurate as the fp32 axisymmetric oracle against the fp64 one."""
    from tests.cases import fp32_like
    M, N = 2, 80000
    out = tmp_path / "axi.npz"
    env = dict(os.environ, JELLY_INSTANCES=str(M), JELLY_PARTICLES=str(N), JELLY_AXISYM="1")
    print(f"[gpu-case] axisymmetric facade, {M} x {N}, 2 substeps", flush=True)
    subprocess.run([sys.executable, "-u", "-c", AXISYM_DRIVER, str(out)], cwd=ROOT, env=env,
                   check=True, timeout=300)
    d = np.load(out)
    assert str(d["path"]) == "fused"
    orcs = [oracle.OracleSim(M, N, dtype=dt, axisym=True) for dt in (np.float32, np.float64)]
    for i in range(M):
        pos, mat, fib, _ = phenotype.fill_tank(phenotype.seeded_genome(i), N, grid_res=128,
                                               axisym=True)
        for o in orcs:
            o.load(i, pos, mat, fib)
    for o in orcs:
        o.run(2, nthreads=8)
    o32, o64 = orcs
    assert float(d["t"]) == float(o32.t[0])
    live = o32.material >= 0
    for n in ("x", "v", "F"):
        fp32_like(n, d[n][live], getattr(o32, n)[live], getattr(o64, n)[live])
    e_gpu = np.abs(d["grid_m"] - o64.grid_m).max()
    e_32 = np.abs(o32.grid_m - o64.grid_m).max()
    print(f"[gpu-case] grid_m: |gpu-f64| {e_gpu:.3e}, |oracle32-f64| {e_32:.3e}", flush=True)
    assert e_gpu <= 2 * e_32 + 1e-5 * o64.grid_m.max()
