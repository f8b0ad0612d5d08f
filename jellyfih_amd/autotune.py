"""Autotuner for the fused path: picks workgroup size, particles per thread (the chunk of
cell-sorted particles one workgroup owns, which sets the LDS grid-window shape), the sort
interval K and cell order (strip height), the graph block length, the number of
concurrently launched instance groups, and the XCD-affinity / lazy-v,C switches.

It measures each candidate on the caller's own workload (an Engine already loaded with it)
with HIP-event-timed blocks, keeps the fastest, and can write it to ``tuning.json`` beside
this file, which ``bench.py`` and the facade load: one entry per batch shape (instances,
slots per instance), looked up by ``tuning_for``.

Usage: python -m jellyfih_amd.autotune [--instances 32] [--slots 80000] [--write]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TUNING_PATH = os.path.join(HERE, "tuning.json")

SHAPES = [(256, 4), (256, 8), (512, 4), (512, 8), (128, 8), (128, 16), (256, 16), (256, 2),
          (512, 2)]
SORT_INTERVALS = [50, 100, 200, 500, 1000]
SORT_STRIPS = [0, 4, 8, 16]
BLOCKS = [50, 100, 200]


def load_entries(path: str | None = None) -> list:
    """The entries of a tuning file: ``{"entries": [...]}``, or a single entry at the top
    level (the file's form when it held one batch shape); [] when it is missing."""
    try:
        with open(path or TUNING_PATH) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return []
    if isinstance(d, dict) and "entries" in d:
        return list(d["entries"])
    return [d] if isinstance(d, dict) and "tuning" in d else []


def _shape(entry: dict):
    w = entry.get("workload", {})
    return w.get("instances"), w.get("slots")


def tuning_for(instances: int, slots: int, path: str | None = None):
    """The launch parameters measured on a batch of ``instances`` x ``slots``, or None."""
    for e in load_entries(path):
        if _shape(e) == (instances, slots):
            return dict(e.get("tuning", {})) or None
    return None


def write_entry(res: dict, path: str | None = None) -> None:
    """Adds (or replaces) the entry of ``res["workload"]``'s batch shape."""
    entries = [e for e in load_entries(path) if _shape(e) != _shape(res)]
    entries.append(res)
    entries.sort(key=lambda e: tuple(v or 0 for v in _shape(e)))
    with open(path or TUNING_PATH, "w") as f:
        json.dump({"entries": entries}, f, indent=1)


def _rate(eng, live: int, substeps: int, reps: int = 3) -> float:
    """Particle-substeps/s of `substeps` graph-replayed substeps (median of reps)."""
    eng.substep(substeps)          # capture/warm the graphs for this configuration
    eng.sync()
    rates = []
    for _ in range(reps):
        t0 = time.perf_counter()
        eng.substep(substeps)
        eng.sync()
        rates.append(live * substeps / (time.perf_counter() - t0))
    return float(np.median(rates))


def tune(eng, live: int, substeps: int = 1000, log=print) -> dict:
    best = dict(eng.tuning())
    best_rate = _rate(eng, live, substeps)
    log(f"start {best} -> {best_rate:.3e}")

    def trial(**kw):
        nonlocal best, best_rate
        cand = dict(best, **kw)
        try:
            eng.set_tuning(**cand)
        except Exception as e:  # unsupported combination
            log(f"skip {kw}: {e}")
            eng.set_tuning(**best)
            return
        r = _rate(eng, live, substeps)
        log(f"{kw} -> {r:.3e}")
        if r > best_rate * 1.01:
            best, best_rate = dict(eng.tuning()), r
        eng.set_tuning(**best)

    for t, p in SHAPES:
        trial(threads=t, ppt=p)
    for k in SORT_INTERVALS:
        trial(sort_interval=k)
    for st in SORT_STRIPS:
        trial(sort_strip=st)
    for b in BLOCKS:
        trial(block_substeps=b)
    for ng in (1, 2, 3):
        trial(instance_groups=ng)
    for flag in ("lazy_vc", "xcd_affinity"):
        trial(**{flag: 1 - best[flag]})
    log(f"best {best} -> {best_rate:.3e}")
    return {"tuning": best, "rate": best_rate}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--instances", type=int, default=32)
    ap.add_argument("--slots", type=int, default=80000)
    ap.add_argument("--substeps", type=int, default=1000)
    ap.add_argument("--write", action="store_true")
    args = ap.parse_args()
    from .engine import Engine
    from . import phenotype
    pos, mat, fib, genomes, _ = phenotype.population(range(args.instances), args.slots, 128)
    eng = Engine(args.instances, args.slots)
    for m in range(args.instances):
        eng.load_particles(m, pos[m], mat[m], fib[m])
    eng.set_param("contraction", np.clip(genomes[:, 9], 0.05, 0.60))
    eng.set_param("freq", np.clip(genomes[:, 10], 0.5, 2.0))
    live = int((mat >= 0).sum())
    eng.substep(1000)   # disorder the particles a little before measuring
    res = tune(eng, live, args.substeps,
               log=lambda s: print(f"[autotune] {s}", file=sys.stderr, flush=True))
    res["workload"] = {"instances": args.instances, "slots": args.slots, "live": live}
    print(json.dumps(res))
    if args.write:
        write_entry(res)
    return 0


if __name__ == "__main__":
    sys.exit(main())
