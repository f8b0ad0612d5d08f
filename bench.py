"""Benchmark: batched 2D MLS-MPM forward on one MI355X.

Workload (the 32 x 80K generation evaluation): 32 seeded 11-gene genomes (seeds 0..31,
uniform within the evolve.py bounds) filled into 80,000-slot water tanks on 128x128 grids
by the fill_tank restatement (about 62-65K live particles each, ~2.0M total), actuation
timing genes applied as evolve.load_batch does. One "step" = one 100-substep block of the
whole batch (HIP-graph replay). The rate counts live particles:
value = sum(live) * substeps / seconds of the timed region (GPU-synchronised).

Also reported on the same line:
* roofline: algorithmic 108 B per live particle-substep (x, v, C, F read and written, 48 B
  each way; material and fibre read, 12 B) over the fused kernel's average duration, timed
  with HIP events on the engine's stream; peak 8.0 TB/s HBM3E;
* cpu_baseline: the oracle (CPU restatement, fp32, one instance per OpenMP thread) on a
  bounded sample of the same workload on this host;
* plain_value: the same workload stepped by the plain path (one launch per stage per
  instance, global atomics, unsorted);
* diff: one differentiable episode of one seeded reference tank (seed 0, ~62K live
  particles, 2000 substeps, checkpoint every 100), in fp32 and fp64: forward through the
  autograd blocks, forward + backward of the payload CoM y with recomputation from the
  checkpoints, the same keeping every block's stored trajectory (no recomputation), and the
  fused (non-differentiable, fp32) forward of the same episode for comparison;
* render: one 1024x1024 video frame of the batch per renderer pipeline (abyss splat + tone
  map, web-palette flat passes, vorticity grid + RdBu composite) and its read-back;
* axisym: the axisymmetric variant's rate on the same batch and on the tall view
  configuration (windows of 8 substeps from fresh loads);
* tank_build_s: host seconds to build the 32 tanks (phenotype.population, the fill_tank
  restatement with its native steps).

Usage: python bench.py [--gpus 1] [--steps 100] [--warmup 10]
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


def load_tuning(eng) -> dict:
    path = os.path.join(ROOT, "jellyfih_amd", "tuning.json")
    if os.path.exists(path):
        with open(path) as f:
            t = json.load(f)
        eng.set_tuning(**{k: v for k, v in t.get("tuning", t).items()
                          if k in eng.tuning()})
    return eng.tuning()


def traffic_from_profiles():
    path = os.path.join(ROOT, "profiles", "fused_kernel_traffic.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        d = json.load(f)
    return d.get("bytes_per_launch"), d.get("live_particles")


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
    return {"value": live * steps / dt, "unit": "particle-substeps/s", "cores": cores,
            "kind": "port",
            "sample": f"{M} instances of the same workload ({live} live particles), "
                      f"{steps} substeps, oracle fp32, one instance per OpenMP thread"}


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

    for dtype in (torch.float32, torch.float64):
        tank, st0 = DiffTank.from_tanks(pos[None], mat[None], fib[None], dtype=dtype)
        dev = tank.device

        def episode(grad: bool, recompute: bool = True) -> float:
            st = st0.clone().requires_grad_(grad)
            leaf = lambda v: torch.full((1,), v, dtype=dtype, device=dev, requires_grad=grad)
            amp_scale = torch.ones(tank.P, dtype=dtype, device=dev, requires_grad=grad)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            with torch.set_grad_enabled(grad):
                out = tank.run(st, amp_scale, leaf(500.0), leaf(con), leaf(freq), leaf(700.0),
                               2.5, 0.0, steps, checkpoint, recompute)
                y = tank.payload_com_y(out).sum()
                if grad:
                    y.backward()
            torch.cuda.synchronize()
            return time.perf_counter() - t0

        episode(True)                              # warm-up (allocations, code objects)
        fwd = min(episode(False) for _ in range(2))
        fb = min(episode(True) for _ in range(2))
        episode(True, recompute=False)
        fb_stored = min(episode(True, recompute=False) for _ in range(2))
        res["workload"] = (f"1 reference tank (seed 0, {tank.P} live), {steps} substeps, "
                           f"checkpoint every {checkpoint}, payload CoM y objective")
        res["fp32" if dtype == torch.float32 else "fp64"] = {
            "forward_ms": fwd * 1e3, "fwd_bwd_ms": fb * 1e3, "fwd_bwd_stored_ms": fb_stored * 1e3,
            "fwd_bwd_over_forward": fb / fwd, "fwd_bwd_stored_over_forward": fb_stored / fwd,
            "fwd_bwd_over_fused_forward": fb / fused,
            "fwd_bwd_stored_over_fused_forward": fb_stored / fused}
        tank.close()
    return res


def diff_batched(M: int = 32, slots: int = 160000, steps: int = 4000,
                 checkpoint: int = 200) -> dict:
    """Batched forward + backward at the tall-tank shape: M seeded genomes in 128x256 tanks
    of height 2 (spawn (0.5, 0.60), evolve.py:33-36,136), per-instance gradients of the
    payload CoM y w.r.t. actuation amplitude, contraction and frequency in one adjoint
    pass; beside the fused forward of the same batch."""
    import torch
    from jellyfih_amd import phenotype
    from jellyfih_amd.diff import DiffTank
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
    tank, st0 = DiffTank.from_tanks(pos, mat, fib, n_grid_y=256, domain_height=2.0)
    dev = tank.device

    def episode(grad: bool, n: int = steps, resort: int = 0) -> float:
        st = st0.clone()
        vec = lambda a: torch.tensor(np.asarray(a, np.float32), device=dev).requires_grad_(grad)
        amp, c, f = vec(np.full(M, 500.0)), vec(con), vec(freq)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        with torch.set_grad_enabled(grad):
            out = tank.run(st, None, amp, c, f, 700.0, 2.5, 0.0, n, checkpoint,
                           resort_every=resort)
            y = tank.payload_com_y(out).sum()
            if grad:
                y.backward()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    episode(True, checkpoint)   # warm-up: first launches and the allocator's buffers
    fwd = episode(False)
    log(f"batched diff forward {fwd:.2f} s")
    fb = episode(True)
    log(f"batched diff forward + backward {fb:.2f} s")
    episode(True, 2 * checkpoint, 1)
    fwd_rs = episode(False, steps, 1)
    fb_rs = episode(True, steps, 1)
    log(f"re-sorted every block: forward {fwd_rs:.2f} s, forward + backward {fb_rs:.2f} s")
    tank.close()
    return {"workload": f"{M} tall tanks 128x256 (H=2, {slots} slots, {live} live), {steps} "
                        f"substeps, checkpoint every {checkpoint}, fp32",
            "live_particles": live, "fused_forward_s": fused, "forward_s": fwd,
            "fwd_bwd_s": fb, "fwd_bwd_over_forward": fb / fwd,
            "fwd_bwd_over_fused_forward": fb / fused,
            "resorted_every_block": {"forward_s": fwd_rs, "fwd_bwd_s": fb_rs,
                                     "fwd_bwd_over_forward": fb_rs / fwd_rs,
                                     "fwd_bwd_over_fused_forward": fb_rs / fused},
            "fused_rate": live * steps / fused}


WEB_PALETTE = [(0, 0.933, 0.949, 0.957), (1, 0.678, 0.882, 0.875), (3, 0.976, 0.757, 0.765),
               (2, 0.961, 0.824, 0.576)]     # mpm_sim.py:60-65


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


def axisym_rates(pos, mat, fib, genomes, reps: int = 10, n: int = 8) -> dict:
    """The axisymmetric variant (JELLY_AXISYM=1, plain path, one launch per stage for the
    batch): live particle-substeps/s on the same 32x80K population, and on
    helpers/view_axisym.py's configuration (one seeded tank, 128x256, H=2, 160K slots).
    The reference's variant diverges within ~10 substeps next to the axis (its affine
    momentum is not weighted by r), so each timed window is `n` substeps from a fresh load,
    `reps` windows, HIP events on the engine's stream."""
    import torch
    from jellyfih_amd import phenotype
    from jellyfih_amd.engine import Engine
    out = {"window_substeps": n, "windows": reps}
    cases = {"batch_32x80k": (pos, mat, fib, genomes, {}),
             "view_tall_1x160k": None}
    tall = phenotype.population([0], 160000, 128, domain_height=2.0)
    cases["view_tall_1x160k"] = (tall[0], tall[1], tall[2], tall[3],
                                 {"n_grid_y": 256, "domain_height": 2.0})
    for name, (P, Mt, Fb, G, kw) in cases.items():
        eng = Engine(P.shape[0], P.shape[1], path="plain", axisym=True, **kw)
        stream = torch.cuda.ExternalStream(eng.stream)
        live = int((Mt >= 0).sum())
        total = 0.0
        for r in range(reps + 1):
            load_engine(eng, P, Mt, Fb, G)
            eng.set_time(0.0)
            eng.substep(1)               # commits the load; not timed
            eng.sync()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            eng.substep(n)
            e1.record(stream)
            eng.sync()
            if r:                        # the first window warms up
                total += e0.elapsed_time(e1) / 1e3
        eng.close()
        out[name] = {"live_particles": live, "particle_substeps_per_s": live * n * reps / total,
                     "ms_per_substep": total / (n * reps) * 1e3}
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100, help="timed 100-substep blocks")
    ap.add_argument("--warmup", type=int, default=10, help="untimed 100-substep blocks")
    ap.add_argument("--instances", type=int, default=32)
    ap.add_argument("--slots", type=int, default=80000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-diff", action="store_true")
    ap.add_argument("--diff-batched", action="store_true",
                    help="also time the batched tall-tank forward + backward (~1 minute)")
    ap.add_argument("--only-diff", action="store_true",
                    help="run only the differentiable legs (no throughput line fields)")
    ap.add_argument("--only-batched", action="store_true",
                    help="run only the batched tall-tank forward + backward")
    ap.add_argument("--only-render", action="store_true",
                    help="run only the renderer leg (after the warm-up blocks)")
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
            steps=args.batched_steps, checkpoint=args.batched_checkpoint)}}), flush=True)
        return 0
    if args.only_diff:
        out = {"diff": diff_episode()}
        if args.diff_batched:
            out["diff"]["batched"] = diff_batched(steps=args.batched_steps,
                                                  checkpoint=args.batched_checkpoint)
        print(json.dumps(out), flush=True)
        return 0
    M, N = args.instances, args.slots
    log(f"building workload: {M} genomes x {N} slots")
    t_build = time.perf_counter()
    pos, mat, fib, genomes, infos = build_workload(M, N, range(M))
    t_build = time.perf_counter() - t_build
    live = int((mat >= 0).sum())
    log(f"live particles {live} (slots {M * N})")

    if args.only_axisym:
        print(json.dumps({"axisym": axisym_rates(pos, mat, fib, genomes)}), flush=True)
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
    tuning = load_tuning(eng)
    log(f"tuning {tuning}")
    log(f"warmup {args.warmup} blocks")
    eng.substep(args.warmup * BLOCK)
    eng.sync()

    if args.only_render:
        print(json.dumps({"render": render_frames(eng, live)}), flush=True)
        return 0

    stream = torch.cuda.ExternalStream(eng.stream)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    eng.sync()
    log(f"timing {args.steps} blocks")
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        eng.substep(BLOCK)
    ev1.record(stream)
    eng.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_s = ev0.elapsed_time(ev1) / 1e3
    substeps = args.steps * BLOCK
    value = live * substeps / wall
    log(f"wall {wall:.3f}s events {gpu_s:.3f}s -> {value:.3e} particle-substeps/s")

    # fused-substep duration with HIP events (3 blocks, not graph-captured): the wall time
    # of one substep of all instances, whose instance groups' launches run concurrently
    fused = [eng.timed_block(BLOCK) for _ in range(3)]
    groups = eng.tuning()["instance_groups"]
    fused_ms = float(np.mean([f[0] for f in fused]))
    achieved = ALGO_BYTES * live / (fused_ms * 1e-3)
    traffic, traffic_live = traffic_from_profiles()
    if traffic is not None and traffic_live and traffic_live != live:
        traffic = None
    roofline = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": traffic,
                "kernel": f"k_fused<G2P,P2G>, one substep = {groups} concurrent "
                          f"instance-group launches", "kernel_ms": fused_ms,
                "algorithmic_bytes_per_launch": ALGO_BYTES * live}
    log(f"fused kernel {fused_ms * 1e3:.1f} us -> {achieved / 1e9:.0f} GB/s algorithmic")

    log("renderer (one video frame per pipeline)")
    rend = render_frames(eng, live)
    log(f"render {rend}")

    log("axisymmetric variant (plain path)")
    axisym = axisym_rates(pos, mat, fib, genomes)
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
    if args.diff_batched:
        diff = dict(diff or {}, batched=diff_batched(steps=args.batched_steps,
                                                     checkpoint=args.batched_checkpoint))
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
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded genomes -> fill_tank restatement (reference 4-per-cell "
                "lattice, no jitter)",
        "config": {"workload": f"{M} instances x {N} slots, 128x128 grids, one fitness "
                               f"episode's substeps in {BLOCK}-substep graph blocks",
                   "instances": M, "slots_per_instance": N, "live_particles": live,
                   "grid": [128, 128], "substeps_per_step": BLOCK, "tuning": tuning},
        "slot_rate": M * N * substeps / wall,
        "substeps_per_s": substeps / wall,
        "gpu_event_s": gpu_s,
        "plain_value": plain_value,
        "speedup_vs_plain": (value / plain_value) if plain_value else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "diff": diff,
        "render": rend,
        "axisym": axisym,
        "tank_build_s": t_build,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
