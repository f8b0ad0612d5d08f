"""Benchmark: batched 2D MLS-MPM forward on one MI355X.

Workload (the 32 x 80K generation evaluation): 32 seeded 11-gene genomes (seeds 0..31,
uniform within the evolve.py bounds) filled into 80,000-slot water tanks on 128x128 grids
by the fill_tank restatement (about 62-65K live particles each, ~2.0M total), actuation
timing genes applied as evolve.load_batch does. One "step" = one 100-substep block of the
whole batch (HIP-graph replay), a slice of the 150,000-substep fitness episode. After at
least 1000 warm-up substeps (so the particles have moved off their fill order), K steps
are timed 5 times, each repetition bracketed by device synchronisation; the line reports
the median repetition. The rate counts live particles:
value = sum(live) * 100 * K / seconds of the median repetition.

Also reported on the same line:
* roofline: algorithmic 108 B per live particle-substep (x, v, C, F read and written, 48 B
  each way; material and fibre read, 12 B) over the fused kernel's average duration with
  one launch per substep covering every live particle (instance_groups = 1), timed with HIP
  events on the engine's stream; peak 8.0 TB/s HBM3E. Beside it, event-timed: the wall time
  of one substep with the tuned concurrent instance-group launches, and the HBM traffic the
  counters measured for one substep (with the commit it was measured at);
* episode: one full fitness episode, run_batch_headless(150000) on the freshly loaded batch
  (mpm_sim.py:802-834, evolve.py:48), its validity flags and finiteness, and the substep
  time at the end of the episode against the start;
* batch16: the 16-instance forward configuration (seeds 1-16, 20,000 substeps);
* cpu_baseline: the oracle (CPU restatement, fp32, one instance per OpenMP thread) on a
  bounded sample of the same workload on this host, and the single-instance configuration
  (seed 0, 1000 substeps) in full on one thread;
* plain_value: the same workload stepped by the plain path (one launch per stage per
  instance, global atomics, unsorted);
* diff: one differentiable episode of one seeded reference tank (seed 0, ~62K live
  particles, 2000 substeps, checkpoint every 100), in fp32 and fp64: forward through the
  autograd blocks, forward + backward of the payload CoM y as DiffTank.run does it by
  default (the last blocks keep their trajectories from the forward pass as device memory
  allows, the others are recomputed from their checkpoints), the same with every block
  recomputed (beside the next block's adjoint, and on one stream), the same keeping every
  block's stored trajectory (recompute=False), and the fused (non-differentiable, fp32)
  forward of the same episode for comparison; and the batched configuration: 32 tall tanks
  (128x256, H = 2, 160K slots), 4000 substeps, checkpoint every 200, per-instance gradients
  in one adjoint pass, and the adjoint kernels alone over one stored block;
* render: one 1024x1024 video frame of the batch per renderer pipeline (abyss splat + tone
  map, web-palette flat passes, vorticity grid + RdBu composite) and its read-back;
* view_mode: evolve.py's view loop through the mpm_sim facade (16 x 80K, 500 substep()
  calls from Python, then one abyss frame read back and encoded), against the same loop
  with one engine call per substep;
* axisym: the axisymmetric variant's rate on the fused and on the plain path, on 32
  axisymmetric phenotypes (80K slots) and on the tall view configuration (one substep of
  the intact state after each fresh load: the reference's formulation throws nearly every
  particle out of the transfer range within ~10 substeps);
* tank_build_s: host seconds to build the 32 tanks (phenotype.population, the fill_tank
  restatement with its native steps).

Usage: python bench.py [--gpus 1] [--steps 20] [--warmup 10]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-substeps/s (32x80K batch, fp32)"
ALGO_BYTES = 108.0          # per live particle-substep
HBM_PEAK = 8.0e12           # B/s, MI355X HBM3E
BLOCK = 100                 # substeps per step
REPS = 5                    # timed repetitions of K steps (the median is reported)
MIN_WARM = 1000             # warm-up substeps at least
EPISODE = 150000            # substeps of one fitness episode (evolve.py:48)


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_workload(M: int, N: int, seeds):
    from jellyfih_amd import phenotype
    pos, mat, fib, genomes, infos = phenotype.population(seeds, N, 128)
    return pos, mat, fib, genomes, infos


def load_engine(eng, pos, mat, fib, genomes):
    for m in range(pos.shape[0]):
        eng.load_particles(m, pos[m], mat[m], fib[m])
    eng.set_param("contraction", np.clip(genomes[:, 9], 0.05, 0.60))   # evolve.py:150
    eng.set_param("freq", np.clip(genomes[:, 10], 0.5, 2.0))           # evolve.py:152


def load_tuning(eng, instances: int, slots: int) -> dict:
    """The autotuner's launch parameters for this batch shape (jellyfih_amd/tuning.json),
    when it holds an entry for it; else the engine's own choice."""
    from jellyfih_amd.autotune import tuning_for
    t = tuning_for(instances, slots)
    if t:
        eng.set_tuning(**{k: v for k, v in t.items() if k in eng.tuning()})
    return eng.tuning()


def traffic_from_profiles():
    """HBM bytes of one fused substep from the FETCH_SIZE / WRITE_SIZE counter runs
    (rocprofv3 cannot run inside this process), with the commit and command they were
    measured with."""
    path = os.path.join(ROOT, "profiles", "fused_kernel_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def roofline_leg(eng, live: int, groups: int) -> dict:
    """One launch per substep over every live particle (instance_groups = 1): the fused
    kernel's average duration from HIP events on the engine's stream around the n - 1
    fused G2P->P2G kernels of 3 directly launched blocks; then, when `groups` > 1, the tuned
    concurrent instance groups' wall time per substep, event-timed the same way."""
    eng.set_tuning(instance_groups=1)
    single = [eng.timed_block(BLOCK)[0] for _ in range(3)]
    single_ms = float(np.mean(single))
    algo = ALGO_BYTES * live
    achieved = algo / (single_ms * 1e-3)
    tr = traffic_from_profiles()
    traffic = tr.get("bytes_per_launch") if tr and tr.get("live_particles") == live else None
    roofline = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": traffic,
                "kernel": "k_fused<256,4,true,true>, one launch per substep over all "
                          f"{live} live particles (instance_groups=1)",
                "kernel_ms": single_ms, "kernel_ms_blocks": single,
                "algorithmic_bytes_per_launch": algo,
                "traffic_frac": (traffic / (single_ms * 1e-3) / HBM_PEAK) if traffic else None,
                "traffic_measured": ({k: tr.get(k) for k in ("commit", "command", "source")}
                                     if traffic else None)}
    msg = f"fused kernel, one launch {single_ms * 1e3:.1f} us -> {achieved / 1e9:.0f} GB/s"
    if groups > 1:
        eng.set_tuning(instance_groups=groups)
        conc_ms = float(np.mean([eng.timed_block(BLOCK)[0] for _ in range(3)]))
        roofline["event_timed_concurrent"] = {
            "instance_groups": groups,
            "substep_ms": conc_ms, "achieved": algo / (conc_ms * 1e-3) / 1e9,
            "frac": algo / (conc_ms * 1e-3) / HBM_PEAK,
            "note": "wall time of one substep whose instance groups' launches run "
                    "concurrently on separate streams (HIP events on the main stream); not a "
                    "kernel duration"}
        msg += f" algorithmic; concurrent groups {conc_ms * 1e3:.1f} us per substep"
    log(msg)
    return roofline


def cpu_baseline(pos, mat, fib, genomes, budget_s: float = 12.0) -> dict:
    """The oracle fp32 on host cores, one instance per thread, sized to ~budget_s."""
    import oracle
    cores = len(os.sched_getaffinity(0))
    M = min(pos.shape[0], max(cores, 1))
    N = pos.shape[1]
    orc = oracle.OracleSim(M, N)
    for m in range(M):
        orc.load(m, pos[m], mat[m], fib[m])
    orc.contraction[:] = np.clip(genomes[:M, 9], 0.05, 0.6).astype(np.float32)
    orc.freq[:] = np.clip(genomes[:M, 10], 0.5, 2.0).astype(np.float32)
    live = int((mat[:M] >= 0).sum())
    t0 = time.perf_counter()
    orc.run(2, nthreads=cores)
    probe = time.perf_counter() - t0
    steps = int(max(4, min(2000, budget_s / max(probe / 2, 1e-6))))
    t0 = time.perf_counter()
    orc.run(steps, nthreads=cores)
    dt = time.perf_counter() - t0
    # the single-instance configuration in full: seed 0, 1000 substeps, one thread
    one = oracle.OracleSim(1, N)
    one.load(0, pos[0], mat[0], fib[0])
    one.contraction[:] = orc.contraction[0]
    one.freq[:] = orc.freq[0]
    live1 = int((mat[0] >= 0).sum())
    t1 = time.perf_counter()
    one.run(1000, nthreads=1)
    dt1 = time.perf_counter() - t1
    return {"value": live * steps / dt, "unit": "particle-substeps/s", "cores": cores,
            "kind": "port",
            "sample": f"{M} instances of the same workload ({live} live particles), "
                      f"{steps} substeps, oracle fp32, one instance per OpenMP thread",
            "single_instance_full": {"live_particles": live1, "substeps": 1000,
                                     "seconds": dt1, "value": live1 * 1000 / dt1,
                                     "cores": 1}}


def diff_episode(steps: int = 2000, checkpoint: int = 100) -> dict:
    """Timing of a differentiable episode (wall clock around GPU-synchronised work), in
    fp32 and fp64, beside the fused fp32 forward of the same episode."""
    import torch
    from jellyfih_amd import phenotype
    from jellyfih_amd.diff import DiffTank
    from jellyfih_amd.engine import Engine

    g = phenotype.seeded_genome(0)
    pos, mat, fib, info = phenotype.fill_tank(g, 80000, 128)
    con = float(np.clip(g[9], 0.05, 0.60))
    freq = float(np.clip(g[10], 0.5, 2.0))

    eng = Engine(1, 80000)
    eng.load_particles(0, pos, mat, fib)
    eng.set_param("contraction", con)
    eng.set_param("freq", freq)
    eng.substep(steps)
    eng.sync()
    eng.set_time(0.0)
    eng.load_particles(0, pos, mat, fib)
    t0 = time.perf_counter()
    eng.substep(steps)
    eng.sync()
    fused = time.perf_counter() - t0
    eng.close()
    res = {"fused_forward_ms": fused * 1e3}

    for key, dtype, math in (("fp32", torch.float32, "fp32"),
                             ("fp32_buffers_fp64_math", torch.float32, "fp64"),
                             ("fp64", torch.float64, "fp64")):
        tank, st0 = DiffTank.from_tanks(pos[None], mat[None], fib[None], dtype=dtype,
                                        math=math)
        dev = tank.device

        def episode(grad: bool, recompute: bool = True, prefetch: bool = True,
                    keep_last=None) -> float:
            st = st0.clone().requires_grad_(grad)
            leaf = lambda v: torch.full((1,), v, dtype=dtype, device=dev, requires_grad=grad)
            amp_scale = torch.ones(tank.P, dtype=dtype, device=dev, requires_grad=grad)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            with torch.set_grad_enabled(grad):
                out = tank.run(st, amp_scale, leaf(500.0), leaf(con), leaf(freq), leaf(700.0),
                               2.5, 0.0, steps, checkpoint, recompute, prefetch=prefetch,
                               keep_last=keep_last)
                y = tank.payload_com_y(out).sum()
                if grad:
                    y.backward()
            torch.cuda.synchronize()
            return time.perf_counter() - t0

        episode(True)                              # warm-up (allocations, code objects)
        fwd = min(episode(False) for _ in range(2))
        fb = min(episode(True) for _ in range(2))
        kept = tank.last_kept
        fb_rec = min(episode(True, keep_last=0) for _ in range(2))
        fb_serial = min(episode(True, prefetch=False, keep_last=0) for _ in range(2))
        episode(True, recompute=False)
        fb_stored = min(episode(True, recompute=False) for _ in range(2))
        res["workload"] = (f"1 reference tank (seed 0, {tank.P} live), {steps} substeps, "
                           f"checkpoint every {checkpoint}, payload CoM y objective")
        # fwd_bwd: DiffTank.run's defaults (the last blocks keep their trajectories as
        # memory allows: here all of them); recompute: every block recomputed beside the
        # next block's adjoint; serial: the same on one stream; stored: recompute=False
        res[key] = {
            "forward_ms": fwd * 1e3, "fwd_bwd_ms": fb * 1e3, "kept_blocks": kept,
            "fwd_bwd_recompute_ms": fb_rec * 1e3, "fwd_bwd_serial_ms": fb_serial * 1e3,
            "fwd_bwd_stored_ms": fb_stored * 1e3,
            "fwd_bwd_over_forward": fb / fwd, "fwd_bwd_recompute_over_forward": fb_rec / fwd,
            "fwd_bwd_over_fused_forward": fb / fused,
            "fwd_bwd_recompute_over_fused_forward": fb_rec / fused}
        tank.close()
    return res


def diff_batched(M: int = 32, slots: int = 160000, steps: int = 4000,
                 checkpoint: int = 200, resort: bool = False) -> dict:
    """Batched forward + backward at the tall-tank shape: M seeded genomes in 128x256 tanks
    of height 2 (spawn (0.5, 0.60), evolve.py:33-36,136), per-instance gradients of the
    payload CoM y w.r.t. actuation amplitude, contraction and frequency in one adjoint
    pass, with fp32 and with fp64 arithmetic over fp32 buffers; beside the fused forward of
    the same batch."""
    from jellyfih_amd import phenotype
    from jellyfih_amd.engine import Engine

    pos, mat, fib, genomes, infos = phenotype.population(
        range(M), slots, 128, domain_height=2.0, spawn_offset=phenotype.TALL_SPAWN)
    con = np.clip(genomes[:, 9], 0.05, 0.60)
    freq = np.clip(genomes[:, 10], 0.5, 2.0)
    live = int((mat >= 0).sum())
    log(f"batched diff: {M} tall tanks, {live} live particles, {steps} substeps")
    eng = Engine(M, slots, n_grid_y=256, domain_height=2.0)
    load_engine(eng, pos, mat, fib, genomes)
    eng.substep(200)
    eng.sync()
    t0 = time.perf_counter()
    eng.substep(steps)
    eng.sync()
    fused = time.perf_counter() - t0
    eng.close()
    out = {"workload": f"{M} tall tanks 128x256 (H=2, {slots} slots, {live} live), {steps} "
                       f"substeps, checkpoint every {checkpoint}",
           "live_particles": live, "fused_forward_s": fused, "fused_rate": live * steps / fused}
    for key, math in (("fp32", "fp32"), ("fp32_buffers_fp64_math", "fp64")):
        out[key] = _batched_precision(pos, mat, fib, con, freq, M, steps, checkpoint, fused,
                                      math, resort and key == "fp32")
    import torch
    torch.cuda.empty_cache()   # the kept blocks (~200 GB) back to the device for later legs
    return out


def _batched_precision(pos, mat, fib, con, freq, M, steps, checkpoint, fused, math,
                       resort) -> dict:
    import torch
    from jellyfih_amd.diff import DiffTank
    tank, st0 = DiffTank.from_tanks(pos, mat, fib, n_grid_y=256, domain_height=2.0,
                                    math=math)
    dev = tank.device

    def episode(grad: bool, n: int = steps, resort: int = 0, prefetch: bool = True,
                keep_last=None) -> float:
        st = st0.clone()
        vec = lambda a: torch.tensor(np.asarray(a, np.float32), device=dev).requires_grad_(grad)
        amp, c, f = vec(np.full(M, 500.0)), vec(con), vec(freq)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        with torch.set_grad_enabled(grad):
            out = tank.run(st, None, amp, c, f, 700.0, 2.5, 0.0, n, checkpoint,
                           resort_every=resort, prefetch=prefetch, keep_last=keep_last)
            y = tank.payload_com_y(out).sum()
            if grad:
                y.backward()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    episode(True)   # warm-up: first launches, and the allocator caches the kept blocks
    fwd = episode(False)
    log(f"batched diff forward ({math} arithmetic) {fwd:.2f} s")
    fb = episode(True)
    kept = tank.last_kept
    log(f"batched diff forward + backward ({math} arithmetic) {fb:.2f} s, last {kept} of "
        f"{-(-steps // checkpoint)} blocks kept from the forward pass")
    fb_rec = episode(True, keep_last=0)
    log(f"batched diff forward + backward ({math} arithmetic), every block recomputed "
        f"{fb_rec:.2f} s")
    # fwd_bwd: DiffTank.run's defaults (the last blocks keep their trajectories as memory
    # allows); recompute: every block recomputed beside the next block's adjoint
    out = {"forward_s": fwd, "fwd_bwd_s": fb, "kept_blocks": kept,
           "fwd_bwd_recompute_s": fb_rec, "fwd_bwd_over_forward": fb / fwd,
           "fwd_bwd_recompute_over_forward": fb_rec / fwd,
           "fwd_bwd_over_fused_forward": fb / fused}
    if math == "fp32":
        fb_serial = episode(True, prefetch=False, keep_last=0)
        log(f"batched diff forward + backward ({math} arithmetic), one stream {fb_serial:.2f} s")
        out["fwd_bwd_serial_s"] = fb_serial
    out["adjoint"] = _adjoint_alone(tank, st0, con, freq, M, checkpoint)
    log(f"batched adjoint alone ({math} arithmetic) {out['adjoint']['us_per_substep']:.1f} us "
        f"per substep, {out['adjoint']['roofline']['frac']:.2f} of HBM peak")
    if resort:
        episode(True, 2 * checkpoint, 1)
        fwd_rs = episode(False, steps, 1)
        fb_rs = episode(True, steps, 1)
        log(f"re-sorted every block: forward {fwd_rs:.2f} s, forward + backward {fb_rs:.2f} s")
        out["resorted_every_block"] = {"forward_s": fwd_rs, "fwd_bwd_s": fb_rs,
                                       "fwd_bwd_over_forward": fb_rs / fwd_rs,
                                       "fwd_bwd_over_fused_forward": fb_rs / fused}
    tank.close()
    return out


ADJOINT_BYTES = 156   # per live particle-substep: primal re-read 60 B + carried adjoint 2 x 48 B


def adjoint_only(M: int = 32, slots: int = 160000, checkpoint: int = 200) -> dict:
    """The adjoint kernels alone on the batched configuration's 32 tall tanks: one stored
    block (``_adjoint_alone``) with fp32 arithmetic (cell packing) and with fp64 arithmetic
    over fp32 buffers (split packing: body kernel, then water-only kernel), and nothing
    else of the adjoint's in the process, so that a kernel trace's average per adjoint
    kernel is its time per substep."""
    from jellyfih_amd import phenotype
    from jellyfih_amd.diff import DiffTank
    pos, mat, fib, genomes, infos = phenotype.population(
        range(M), slots, 128, domain_height=2.0, spawn_offset=phenotype.TALL_SPAWN)
    con = np.clip(genomes[:, 9], 0.05, 0.60)
    freq = np.clip(genomes[:, 10], 0.5, 2.0)
    out = {}
    for key, math in (("fp32", "fp32"), ("fp32_buffers_fp64_math", "fp64")):
        tank, st0 = DiffTank.from_tanks(pos, mat, fib, n_grid_y=256, domain_height=2.0,
                                        math=math)
        out[key] = _adjoint_alone(tank, st0, con, freq, M, checkpoint)
        log(f"adjoint alone ({math} arithmetic) {out[key]['us_per_substep']:.1f} us per "
            f"substep, {out[key]['roofline']['frac']:.2f} of HBM peak")
        tank.close()
    return out


def _adjoint_alone(tank, st0, con, freq, M, n, reps: int = 3) -> dict:
    """The adjoint kernels alone over one block of ``n`` substeps whose trajectory is
    stored (no recomputation), from the state after 400 substeps (an episode's middle,
    not the rest state): HIP events on the launching stream around ``backward_raw``, the
    median of ``reps`` after one untimed; per live particle-substep 156 B algorithmic (the
    primal state re-read, 60 B, and the carried adjoint read and written, 96 B) against
    the 8 TB/s HBM peak."""
    import torch
    dev = tank.device
    par = torch.tensor(np.stack([np.full(M, 500.0), con, freq, np.full(M, 2.5),
                                 np.full(M, 700.0)], 1), dtype=tank.dtype, device=dev)
    amp = torch.ones(tank.P, dtype=tank.dtype, device=dev)
    s1 = tank.forward_raw(st0, amp, par, 0.0, 400)
    t1 = 400 * tank.dt
    stored = tank.trajectory(s1, amp, par, t1, n)
    g = torch.zeros_like(s1)
    g[:, 1] = 1.0
    ts = []
    for r in range(reps + 1):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        tank.backward_raw(s1, amp, par, t1, n, g, stored=stored)
        e1.record()
        torch.cuda.synchronize()
        if r:
            ts.append(e0.elapsed_time(e1) * 1e-3 / n)
    del stored
    sub = float(np.median(ts))
    live = int((tank.mat >= 0).sum())
    ach = ADJOINT_BYTES * live / sub / 1e9
    return {"substeps": n, "us_per_substep": sub * 1e6,
            "roofline": {"bound": "hbm", "achieved": ach, "peak": 8000.0, "unit": "GB/s",
                         "frac": ach / 8000.0, "bytes_per_particle_substep": ADJOINT_BYTES}}


WEB_PALETTE = [(0, 0.933, 0.949, 0.957), (1, 0.678, 0.882, 0.875), (3, 0.976, 0.757, 0.765),
               (2, 0.961, 0.824, 0.576)]     # mpm_sim.py:60-65


def view_mode(frames: int = 4, every: int = 500) -> dict:
    """evolve.py's view mode through the mpm_sim facade (evolve.py:375-398): 16 instances
    (the reference's JELLY_INSTANCES default) x 80K slots, seeds 0-15; per frame `every`
    substep() calls from a Python loop, then the abyss frame (clear, splat, tone map),
    ti.sync(), frame_buffer.to_numpy() and the 8-bit encode on the host (evolve.py:396).
    One untimed frame, then `frames` timed ones (wall clock). Reported beside it: the same
    substep() loop with one engine call per substep (no queueing), as it ran before the
    facade queued its substeps."""
    import math
    import jellyfih_amd.mpm_sim as sim
    from jellyfih_amd import phenotype
    M = sim.n_instances
    genomes = []
    for i in range(M):
        g = phenotype.seeded_genome(i)
        genomes.append(g)
        pos, mat, fib, _ = phenotype.fill_tank(g, sim.n_particles, 128)
        sim.load_particles(i, pos, mat, fib)
        sim.instance_act_contraction[i] = float(np.clip(g[9], 0.05, 0.60))
        sim.instance_freq[i] = float(np.clip(g[10], 0.5, 2.0))
    live = int((sim.material.to_numpy() >= 0).sum())
    sim.sim_time[None] = 0.0
    radius = max(1.0, sim.res_sub / sim.n_grid * 0.6)       # evolve.py:380

    def frame():
        t0 = time.perf_counter()
        for _ in range(every):
            sim.substep()
        sim.ti.sync()
        t1 = time.perf_counter()
        sim.clear_frame_buffer()
        sim.render_frame_abyss(sim.res_sub, sim.grid_side, radius)
        sim.tone_map_and_encode()
        sim.ti.sync()
        img = sim.frame_buffer.to_numpy()
        (np.clip(img, 0, 1) * 255).astype(np.uint8)
        return t1 - t0, time.perf_counter() - t1

    frame()
    parts = [frame() for _ in range(frames)]
    step_s = sum(p[0] for p in parts) / frames
    draw_s = sum(p[1] for p in parts) / frames
    e = sim.engine()
    e.substep(1)
    e.sync()
    t0 = time.perf_counter()
    for _ in range(every):
        e.substep(1)
    e.sync()
    percall_s = time.perf_counter() - t0
    out = {"instances": M, "live_particles": live, "substeps_per_frame": every,
           "frames": frames, "frame_s": step_s + draw_s, "substeps_s": step_s,
           "render_readback_encode_s": draw_s,
           "particle_substeps_per_s": live * every / step_s,
           "with_frames_particle_substeps_per_s": live * every / (step_s + draw_s),
           "unqueued_substeps_s": percall_s,
           "unqueued_particle_substeps_per_s": live * every / percall_s}
    e.close()
    sim._engine = None
    return out


def render_frames(eng, live: int, reps: int = 20) -> dict:
    """One 1024x1024 video frame of the whole batch per pipeline, as evolve.py's view mode
    and the helpers draw it (mpm_sim.py:343-712): the abyss glow splat + tone map, the web
    palette's four flat passes over white, and the vorticity grid + RdBu composite; each
    timed with HIP events on the engine's stream over `reps` frames, plus the read-back."""
    import math
    import torch
    M = eng.M
    grid_side = int(math.ceil(math.sqrt(M)))
    res_sub = 1024 // grid_side
    radius = max(1.0, res_sub / 128 * 0.6)                  # evolve.py:380
    hue = np.linspace(0.0, 0.8, M).astype(np.float32)
    mhue = ((hue + 0.15) % 1.0).astype(np.float32)
    stream = torch.cuda.ExternalStream(eng.stream)

    def abyss():
        eng.render_clear(1024, 0.0)
        eng.render_abyss(res_sub, grid_side, radius, hue, mhue, 0)
        eng.render_tone_map()

    def web():
        eng.render_clear(1024, 0.98)
        for mat_id, r, g, b in WEB_PALETTE:
            eng.render_flat(res_sub, grid_side, radius, mat_id, r, g, b)

    def vorticity():
        eng.render_vorticity_grid(0.002)
        eng.render_vorticity(1, res_sub, grid_side, 0.6)

    out = {"frame": [1024, 1024], "grid_side": grid_side, "res_sub": res_sub,
           "radius": radius, "live_particles": live}
    for name, fn in (("abyss_ms", abyss), ("web_ms", web), ("vorticity_rdbu_ms", vorticity)):
        fn()
        eng.sync()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        eng.sync()
        out[name] = e0.elapsed_time(e1) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        eng.read_frame()
    out["read_frame_ms"] = (time.perf_counter() - t0) / reps * 1e3
    t0 = time.perf_counter()
    for _ in range(reps):
        eng.read_frame_u8()
    out["read_frame_u8_ms"] = (time.perf_counter() - t0) / reps * 1e3
    t0 = time.perf_counter()
    for _ in range(reps):
        (np.clip(eng.read_frame(), 0, 1) * 255).astype(np.uint8)
    out["read_frame_host_encode_ms"] = (time.perf_counter() - t0) / reps * 1e3
    return out


def axisym_rates(reps: int = 10) -> dict:
    """The axisymmetric variant (JELLY_AXISYM=1) on the fused path (k_fused's axisymmetric
    instantiation) and on the plain path (one launch per stage for the batch): live
    particle-substeps/s on 32 seeded 80K-slot tanks filled as evolve.py fills them under
    JELLY_AXISYM=1 (right-half profiles at the axis, make_jelly.py:276-304,343), and on
    helpers/view_axisym.py's tank (Aurelia's right-half profile at spawn (0.5, 0.40),
    128x256, H=2, 160K slots, view_axisym.py:19-25,129-135; default actuation parameters).

    The reference's variant does not weight the affine momentum by r, so particles next to
    the axis get grid velocities of order 1/r, and within ~10 substeps nearly every particle
    is thrown against the walls, outside the transfer range (measured: 6,671 of the batch's
    1,995,267 particles in range after 10 substeps, 12 of the view tank's 126,276 after
    100), where it does no work. So the rate is taken on the intact state right after a
    load: per window, a fresh load, then on the fused path one 2-substep block with HIP
    events around its middle kernel (G2P of substep 0 and P2G of substep 1, every instance
    group) and on the plain path one untimed substep and then one timed substep (P2G, grid
    op, G2P); `reps` windows after one warm-up window. The fused path runs without graphs
    here (a fresh load drops its captured graphs)."""
    import torch
    from jellyfih_amd import phenotype
    from jellyfih_amd.engine import Engine
    out = {"windows": reps}
    batch = phenotype.population(range(32), 80000, 128, axisym=True)
    view = phenotype.fill_tank(phenotype.AURELIA_GENOME, 160000, 128,
                               spawn_offset=phenotype.DEFAULT_SPAWN, domain_height=2.0,
                               axisym=True)
    cases = {"batch_32x80k": (batch[0], batch[1], batch[2], batch[3], {}),
             "view_tall_1x160k": (view[0][None], view[1][None], view[2][None], None,
                                  {"n_grid_y": 256, "domain_height": 2.0})}
    for name, (P, Mt, Fb, G, kw) in cases.items():
        live = int((Mt >= 0).sum())
        out[name] = {"live_particles": live}
        for path in ("fused", "plain"):
            eng = Engine(P.shape[0], P.shape[1], path=path, axisym=True, **kw)
            if path == "fused":
                eng.set_tuning(use_graphs=0)
            stream = torch.cuda.ExternalStream(eng.stream)
            times = []
            for r in range(reps + 1):
                if G is None:
                    eng.load_particles(0, P[0], Mt[0], Fb[0])
                else:
                    load_engine(eng, P, Mt, Fb, G)
                eng.set_time(0.0)
                if path == "fused":
                    ms = eng.timed_block(2)[0]
                else:
                    eng.substep(1)               # commits the load; not timed
                    eng.sync()
                    e0 = torch.cuda.Event(enable_timing=True)
                    e1 = torch.cuda.Event(enable_timing=True)
                    e0.record(stream)
                    eng.substep(1)
                    e1.record(stream)
                    eng.sync()
                    ms = e0.elapsed_time(e1)
                if r:                            # the first window warms up
                    times.append(ms / 1e3)
            eng.close()
            sec = float(np.median(times))
            out[name][path] = {"particle_substeps_per_s": live / sec,
                               "ms_per_substep": sec * 1e3}
    return out


def episode_run(eng, pos, mat, fib, genomes, live: int) -> dict:
    """One full fitness episode as evolve.py runs it (load, then run_batch_headless(150000),
    mpm_sim.py:802-834), timed by wall clock around the synchronising call; then the
    substep time at the episode's end state against the start (HIP events)."""
    load_engine(eng, pos, mat, fib, genomes)
    eng.live_count(0)   # commits the load (uploads, chunk tables)
    eng.set_time(0.0)
    start_ms = float(np.mean([eng.timed_block(BLOCK)[0] for _ in range(2)]))
    load_engine(eng, pos, mat, fib, genomes)
    eng.live_count(0)   # the load's upload stays outside the timed call, as in the reference
    eng.sync()
    t0 = time.perf_counter()
    res = eng.run_batch_headless(EPISODE)
    secs = time.perf_counter() - t0
    end_ms = float(np.mean([eng.timed_block(BLOCK)[0] for _ in range(2)]))
    return {"substeps": EPISODE, "seconds": secs, "rate": live * EPISODE / secs,
            "valid": int(np.sum(res[:, 4] == 1.0)), "instances": int(res.shape[0]),
            "finite": bool(np.isfinite(res).all()),
            "final_y": [round(float(v), 6) for v in res[:, 2]],
            "substep_us_start": start_ms * 1e3, "substep_us_end": end_ms * 1e3,
            "clock_end": eng.get_time()}


def batch16(reps: int = 3) -> dict:
    """The 16-instance batched forward configuration: seeds 1-16, 80K slots, 20,000
    substeps in 100-substep graph blocks after 1000 warm-up substeps; median of reps."""
    from jellyfih_amd.engine import Engine
    pos, mat, fib, genomes, _ = build_workload(16, 80000, range(1, 17))
    live = int((mat >= 0).sum())
    eng = Engine(16, 80000)
    tuning = load_tuning(eng, 16, 80000)
    load_engine(eng, pos, mat, fib, genomes)
    eng.substep(MIN_WARM)
    eng.sync()
    walls = []
    for _ in range(reps):
        t0 = time.perf_counter()
        eng.substep(20000)
        eng.sync()
        walls.append(time.perf_counter() - t0)
    eng.close()
    wall = float(np.median(walls))
    return {"instances": 16, "seeds": [1, 16], "live_particles": live, "substeps": 20000,
            "seconds": wall, "value": live * 20000 / wall, "tuning": tuning}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20,
                    help="timed 100-substep blocks per repetition (5 repetitions)")
    ap.add_argument("--warmup", type=int, default=10, help="untimed 100-substep blocks")
    ap.add_argument("--instances", type=int, default=32)
    ap.add_argument("--slots", type=int, default=80000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-diff", action="store_true")
    ap.add_argument("--no-diff-batched", action="store_true",
                    help="skip the batched tall-tank forward + backward")
    ap.add_argument("--diff-resort", action="store_true",
                    help="batched leg: also time re-sorting before every block")
    ap.add_argument("--no-episode", action="store_true")
    ap.add_argument("--no-batch16", action="store_true")
    ap.add_argument("--only-diff", action="store_true",
                    help="run only the differentiable legs (no throughput line fields)")
    ap.add_argument("--only-batched", action="store_true",
                    help="run only the batched tall-tank forward + backward")
    ap.add_argument("--only-adjoint", action="store_true",
                    help="run only the adjoint kernels over one stored block of the batch")
    ap.add_argument("--only-render", action="store_true",
                    help="run only the renderer leg (after the warm-up blocks)")
    ap.add_argument("--no-view", action="store_true")
    ap.add_argument("--only-view", action="store_true",
                    help="run only the facade's view-mode leg")
    ap.add_argument("--only-roofline", action="store_true",
                    help="warm up and time the fused kernel with one instance group only")
    ap.add_argument("--only-axisym", action="store_true",
                    help="run only the axisymmetric variant's leg")
    ap.add_argument("--batched-steps", type=int, default=4000)
    ap.add_argument("--batched-checkpoint", type=int, default=200)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="also run one device copy of MIB MiB (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); instances are "
              "independent, so a multi-GPU run is one process per GPU on its own batch.",
              file=sys.stderr)
        return 2

    import torch
    from jellyfih_amd.engine import Engine

    assert torch.cuda.is_available(), "no GPU visible"
    if args.only_batched:
        print(json.dumps({"diff": {"batched": diff_batched(
            steps=args.batched_steps, checkpoint=args.batched_checkpoint,
            resort=args.diff_resort)}}), flush=True)
        return 0
    if args.only_adjoint:
        print(json.dumps({"diff": {"adjoint_only": adjoint_only(
            checkpoint=args.batched_checkpoint)}}), flush=True)
        return 0
    if args.only_diff:
        out = {"diff": diff_episode()}
        if not args.no_diff_batched:
            out["diff"]["batched"] = diff_batched(steps=args.batched_steps,
                                                  checkpoint=args.batched_checkpoint,
                                                  resort=args.diff_resort)
        print(json.dumps(out), flush=True)
        return 0
    if args.only_view:
        print(json.dumps({"view_mode": view_mode()}), flush=True)
        return 0
    M, N = args.instances, args.slots
    log(f"building workload: {M} genomes x {N} slots")
    t_build = time.perf_counter()
    pos, mat, fib, genomes, infos = build_workload(M, N, range(M))
    t_build = time.perf_counter() - t_build
    live = int((mat >= 0).sum())
    log(f"live particles {live} (slots {M * N})")

    if args.only_axisym:
        print(json.dumps({"axisym": axisym_rates()}), flush=True)
        return 0

    if args.copy_probe:
        src = torch.ones(args.copy_probe << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: {args.copy_probe} MiB read + {args.copy_probe} MiB written")
        del src, dst

    eng = Engine(M, N)
    load_engine(eng, pos, mat, fib, genomes)
    tuning = load_tuning(eng, M, N)
    if args.only_roofline:
        # The warm-up runs with another launch shape (k_fused<256, 8, ...>) and the tuned
        # shape returns for the timed blocks, so the only k_fused<256,4,true,true> dispatches
        # of this process are the measured ones, each one launch over all live particles:
        # a kernel trace's average for that kernel is the figure reported.
        eng.set_tuning(instance_groups=1, threads=256, ppt=8)
    log(f"tuning {eng.tuning()}")
    warm = max(args.warmup * BLOCK, MIN_WARM)
    log(f"warmup {warm} substeps")
    eng.substep(warm)
    eng.sync()
    if args.only_roofline:
        eng.set_tuning(threads=tuning["threads"], ppt=tuning["ppt"])
        print(json.dumps({"roofline": roofline_leg(eng, live, 1), "warmup_substeps": warm,
                          "live_particles": live, "tuning": eng.tuning()}), flush=True)
        eng.close()
        return 0

    if args.only_render:
        print(json.dumps({"render": render_frames(eng, live)}), flush=True)
        return 0

    stream = torch.cuda.ExternalStream(eng.stream)
    substeps = args.steps * BLOCK
    walls, events = [], []
    for r in range(REPS):
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        eng.sync()
        t0 = time.perf_counter()
        ev0.record(stream)
        for _ in range(args.steps):
            eng.substep(BLOCK)
        ev1.record(stream)
        eng.sync()
        torch.cuda.synchronize()
        walls.append(time.perf_counter() - t0)
        events.append(ev0.elapsed_time(ev1) / 1e3)
        log(f"rep {r}: {args.steps} blocks wall {walls[-1]:.4f}s events {events[-1]:.4f}s "
            f"-> {live * substeps / walls[-1]:.3e} particle-substeps/s")
    med = int(np.argsort(walls)[len(walls) // 2])
    wall, gpu_s = walls[med], events[med]
    value = live * substeps / wall
    log(f"median of {REPS}: wall {wall:.4f}s -> {value:.3e} particle-substeps/s")

    roofline = roofline_leg(eng, live, tuning["instance_groups"])

    episode = None
    if not args.no_episode:
        log(f"one fitness episode: run_batch_headless({EPISODE})")
        episode = episode_run(eng, pos, mat, fib, genomes, live)
        log(f"episode {episode['seconds']:.2f} s, {episode['rate']:.3e}, valid "
            f"{episode['valid']}/{episode['instances']}, substep "
            f"{episode['substep_us_start']:.1f} -> {episode['substep_us_end']:.1f} us")
        eng.substep(warm)
        eng.sync()

    b16 = None
    if not args.no_batch16:
        log("16-instance batched forward (seeds 1-16, 20000 substeps)")
        b16 = batch16()
        log(f"batch16 {b16['value']:.3e}")

    log("renderer (one video frame per pipeline)")
    rend = render_frames(eng, live)
    view = None
    if not args.no_view:
        log("view mode through the facade (500 substep() calls + one frame)")
        view = view_mode()
        log(f"view {view}")
    log(f"render {rend}")

    log("axisymmetric variant (fused and plain paths)")
    axisym = axisym_rates()
    log(f"axisym {axisym}")

    plain_value = None
    if not args.no_plain:
        log("plain path (one launch per stage per instance)")
        pe = Engine(M, N, path="plain")
        load_engine(pe, pos, mat, fib, genomes)
        pe.substep(20)
        pe.sync()
        n_plain = 200
        t0 = time.perf_counter()
        pe.substep(n_plain)
        pe.sync()
        plain_value = live * n_plain / (time.perf_counter() - t0)
        pe.close()
        log(f"plain {plain_value:.3e}")

    diff = None
    if not args.no_diff:
        log("differentiable episode (forward, forward + backward)")
        diff = diff_episode()
        log(f"diff {diff}")
    if not args.no_diff_batched:
        diff = dict(diff or {}, batched=diff_batched(steps=args.batched_steps,
                                                     checkpoint=args.batched_checkpoint,
                                                     resort=args.diff_resort))
        log(f"diff batched {diff['batched']}")

    cpu = None
    if not args.no_cpu:
        log("cpu baseline (oracle)")
        cpu = cpu_baseline(pos, mat, fib, genomes)
        log(f"cpu {cpu['value']:.3e} on {cpu['cores']} cores")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "warmup_substeps": warm,
        "repetitions": REPS,
        "rep_values": [live * substeps / w for w in walls],
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded genomes -> fill_tank restatement (reference 4-per-cell "
                "lattice, no jitter)",
        "config": {"workload": f"{M} instances x {N} slots, 128x128 grids; a step is one "
                               f"{BLOCK}-substep graph block of the {EPISODE}-substep "
                               "fitness episode",
                   "instances": M, "slots_per_instance": N, "live_particles": live,
                   "grid": [128, 128], "substeps_per_step": BLOCK, "tuning": tuning},
        "slot_rate": M * N * substeps / wall,
        "substeps_per_s": substeps / wall,
        "gpu_event_s": gpu_s,
        "plain_value": plain_value,
        "speedup_vs_plain": (value / plain_value) if plain_value else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "episode": episode,
        "batch16": b16,
        "diff": diff,
        "render": rend,
        "view_mode": view,
        "axisym": axisym,
        "tank_build_s": t_build,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
