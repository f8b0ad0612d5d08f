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
