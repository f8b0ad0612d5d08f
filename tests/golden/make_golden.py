"""Write golden fixtures of the reference phenotype generator (this container only).

The reference's ``make_jelly.fill_tank`` (c-m-x-r/jellyfih ``make_jelly.py:336-402``) is
pinned by nothing else (its raster, polygon edge semantics and KD-tree exclusion), so its
outputs are recorded here as data. The script imports ``make_jelly`` from the reference
checkout given by ``--reference`` (default ``/root/reference``), runs it on a few seeded
genomes, and writes ``phenotype_golden.npz``: per case the genome, the robot particles
(pos, mat, fiber exactly as fill_tank emits them), the water-lattice keep mask (packed
bits) and the counts. Nothing on the GPU side reads the reference; tests only read the
``.npz``.

The axisymmetric cases (``make_jelly._AXISYM``, read from ``JELLY_AXISYM`` when the module
is imported) reload ``make_jelly`` with ``JELLY_AXISYM=1``: one at the default spawn (the
axis), one as helpers/view_axisym.py builds its tank (Aurelia, 160K slots, H = 2, spawn
(0.5, 0.40)).

Usage: python tests/golden/make_golden.py [--reference /root/reference]
"""
from __future__ import annotations

import argparse
import importlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

LOWER = np.array([0.0, -0.15, 0.0, -0.2, 0.05, -0.30, 0.025, 0.025, 0.01, 0.05, 0.5])
UPPER = np.array([0.25, 0.15, 0.3, 0.15, 0.35, +0.10, 0.08, 0.1, 0.04, 0.60, 2.0])

# (name, genome source, max_particles, domain_height, spawn, axisymmetric)
CASES = [
    ("aurelia", "aurelia", 80000, 1.0, None, False),
    ("seed0", 0, 80000, 1.0, None, False),
    ("seed1", 1, 80000, 1.0, None, False),
    ("seed2", 2, 80000, 1.0, None, False),
    ("tall_seed0", 0, 160000, 2.0, (0.5, 0.60), False),
    ("axisym_seed0", 0, 80000, 1.0, None, True),
    ("axisym_view_aurelia", "aurelia", 160000, 2.0, (0.5, 0.40), True),
]


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    ap.add_argument("--out", default=os.path.join(HERE, "phenotype_golden.npz"))
    args = ap.parse_args()
    sys.path.insert(0, args.reference)
    os.environ["JELLY_AXISYM"] = "0"
    import make_jelly  # noqa: E402  (reference generator, run here only)

    out = {}
    for name, src, nmax, height, spawn, axisym in CASES:
        if make_jelly._AXISYM != axisym:
            os.environ["JELLY_AXISYM"] = "1" if axisym else "0"
            make_jelly = importlib.reload(make_jelly)
            assert make_jelly._AXISYM == axisym
        if src == "aurelia":
            genome = np.asarray(make_jelly.AURELIA_GENOME, dtype=np.float64)
        else:
            genome = np.random.default_rng(src).uniform(LOWER, UPPER)
        spawn_arr = None if spawn is None else np.array(spawn)
        pos, mat, fib, info = make_jelly.fill_tank(genome, nmax, grid_res=128,
                                                   spawn_offset=spawn_arr,
                                                   domain_height=height)
        n_robot = info["n_robot"]
        n_water = info["n_water"]
        # Recover the water keep mask over the lattice candidates.
        spacing = 1.0 / 256.0
        margin = spacing * 3
        wx = np.arange(margin, 1.0 - margin, spacing)
        wy = np.arange(margin, height - margin, spacing)
        wgx, wgy = np.meshgrid(wx, wy)
        cands = np.vstack([wgx.ravel(), wgy.ravel()]).T.astype(np.float32)
        water = pos[n_robot:n_robot + n_water]
        # water rows are an ordered subset of the candidates: match by exact value
        keyc = np.ascontiguousarray(cands).view(np.uint64).ravel()
        keyw = np.ascontiguousarray(water).view(np.uint64).ravel()
        keep = np.isin(keyc, keyw)
        assert keep.sum() == n_water, (name, keep.sum(), n_water)
        out[f"{name}/genome"] = genome
        out[f"{name}/robot_pos"] = pos[:n_robot].copy()
        out[f"{name}/robot_mat"] = mat[:n_robot].astype(np.int8)
        out[f"{name}/robot_fiber"] = fib[:n_robot].copy()
        out[f"{name}/water_keep"] = np.packbits(keep)
        out[f"{name}/counts"] = np.array([nmax, n_robot, n_water, info["muscle_count"],
                                          info["n_jelly"], int(info["self_intersecting"])],
                                         dtype=np.int64)
        out[f"{name}/domain"] = np.array([height] + ([-1.0, -1.0] if spawn is None
                                                     else list(spawn)))
        out[f"{name}/axisym"] = np.array(int(axisym))
        print(f"{name}: robot={n_robot} water={n_water} muscle={info['muscle_count']} "
              f"self_int={info['self_intersecting']}")
    np.savez_compressed(args.out, **out)
    print("wrote", args.out, os.path.getsize(args.out), "bytes")


if __name__ == "__main__":
    main()
