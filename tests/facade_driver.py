"""An evolve.py-style caller of the ``jellyfih_amd.mpm_sim`` facade, run as a subprocess so
that its JELLY_* environment is read at import, as the reference's callers do.

It does what ``evolve.py`` does with the simulator, with the same calls:
``load_batch`` (evolve.py:127-155), ``run_batch_headless`` (evolve.py:474-478), then a few
view-mode frames (evolve.py:78-87,375-398: ``sim_time[None] = 0``, ``substep()`` loops,
``clear_frame_buffer`` / ``render_frame_abyss`` / ``tone_map_and_encode`` /
``frame_buffer.to_numpy()``, and the web palette's flat passes), then 137 queued
``substep()`` calls against the same substeps asked of the engine directly, and saves what
it read back.

Usage: python -m tests.facade_driver OUT.npz   (JELLY_INSTANCES / JELLY_PARTICLES from env)
"""
from __future__ import annotations

import sys

import numpy as np


def main(out_path: str) -> int:
    import jellyfih_amd.mpm_sim as sim
    from jellyfih_amd import phenotype

    genomes = [phenotype.seeded_genome(s) for s in range(sim.n_instances)]
    print(f"[driver] {sim.n_instances} instances x {sim.n_particles} slots", flush=True)
    muscle_counts = []
    for i, genome in enumerate(genomes):                      # evolve.py:135-153
        pos, mat, fiber, stats = phenotype.fill_tank(
            genome, sim.n_particles, grid_res=int(sim.n_grid), domain_height=sim.domain_height)
        sim.load_particles(i, pos, mat, fiber)
        muscle_counts.append(stats["muscle_count"])
        sim.instance_act_contraction[i] = float(np.clip(genome[9], 0.05, 0.60))
        sim.instance_freq[i] = float(np.clip(genome[10], 0.5, 2.0))
    sim.sync()
    results = sim.run_batch_headless(400)                    # evolve.py:477
    print("[driver] headless done", flush=True)

    # view mode: reload, reset the clock, step and draw (evolve.py:375-398)
    for i, genome in enumerate(genomes):
        pos, mat, fiber, _ = phenotype.fill_tank(
            genome, sim.n_particles, grid_res=int(sim.n_grid), domain_height=sim.domain_height)
        sim.load_particles(i, pos, mat, fiber)
        sim.instance_hue[i] = 0.8 * i / max(sim.n_instances - 1, 1)
        sim.instance_muscle_hue[i] = (sim.instance_hue[i] + 0.15) % 1.0
    sim.sim_time[None] = 0.0
    radius = max(1.0, sim.res_sub / sim.n_grid * 0.6)        # evolve.py:380
    for _ in range(150):
        sim.substep()
    state = {k: getattr(sim, k).to_numpy() for k in ("x", "v", "Jp", "material")}
    t = float(sim.sim_time[None])
    sim.clear_frame_buffer()
    sim.render_frame_abyss(sim.res_sub, sim.grid_side, radius)
    hdr = sim.frame_buffer.to_numpy()
    sim.tone_map_and_encode()
    ldr = sim.frame_buffer.to_numpy()
    ldr_u8 = sim.frame_buffer.to_uint8()       # evolve.py:396-397's encode, on the device
    sim.clear_frame_buffer_white()
    for mat_id, r, g, b in sim.WEB_PALETTE:
        sim.render_flat_pass(sim.res_sub, sim.grid_side, radius, mat_id, r, g, b)
    web = sim.frame_buffer.to_numpy()
    # vorticity (helpers/view_random.py:116 overlay; RdBu composite after the tone map)
    grid_v, grid_m = sim.grid_v.to_numpy(), sim.grid_m.to_numpy()
    sim.compute_vorticity_grid(0.1)
    vort = sim.vort_grid.to_numpy()
    sim.clear_frame_buffer()
    sim.render_frame_abyss(sim.res_sub, sim.grid_side, radius)
    sim.tone_map_and_encode()
    base = sim.frame_buffer.to_numpy()
    sim.render_vorticity_rdbu(sim.res_sub, sim.grid_side, 0.6)
    rdbu = sim.frame_buffer.to_numpy()
    print("[driver] frames done", flush=True)

    # 137 queued substep() calls (a block of 100, then 37 at the read) against the same
    # substeps asked of the engine directly
    def reload():
        for i, genome in enumerate(genomes):
            pos, mat, fiber, _ = phenotype.fill_tank(
                genome, sim.n_particles, grid_res=int(sim.n_grid),
                domain_height=sim.domain_height)
            sim.load_particles(i, pos, mat, fiber)
        sim.sim_time[None] = 0.0

    reload()
    for _ in range(137):
        sim.substep()
    x_queued, t_queued = sim.x.to_numpy(), float(sim.sim_time[None])
    reload()
    sim.engine().substep(100)
    sim.engine().substep(37)
    x_direct, t_direct = sim.x.to_numpy(), float(sim.sim_time[None])
    print("[driver] queued substeps done", flush=True)
    np.savez(out_path, results=results, x_queued=x_queued, t_queued=t_queued,
             x_direct=x_direct, t_direct=t_direct, muscle_counts=np.array(muscle_counts),
             genomes=np.array(genomes), hdr=hdr, ldr=ldr, ldr_u8=ldr_u8, web=web, t=t,
             radius=radius,
             grid_v=grid_v, grid_m=grid_m, vort=vort, base=base, rdbu=rdbu,
             hue=sim.instance_hue.to_numpy(), muscle_hue=sim.instance_muscle_hue.to_numpy(),
             freq=sim.instance_freq.to_numpy(), contraction=sim.instance_act_contraction.to_numpy(),
             res_sub=sim.res_sub, grid_side=sim.grid_side, **state)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1]))
