"""Writes tests/golden/grad_fd_golden_tall.json: fp64 central differences of the payload CoM
y after 4000 substeps for two instances of the batched tall-tank configuration (32 seeded
genomes 0-31 filled into 128x256 tanks of domain height 2 with 160K slots and the tall
spawn point, evolve.py:32-36,136), with respect to each instance's actuation amplitude,
contraction fraction, actuation frequency and jelly Young's modulus, from the oracle.

The GPU test runs all 32 tanks in one batched adjoint pass (checkpoint every 200 substeps)
and compares the two instances' gradients with these values. Each difference perturbs one
parameter of one tank by eps times its scale (500, 0.2, 1, 700), eps = 1e-5 and 1e-6; the
runs are instances of one oracle call (one per thread).

Why these instances and eps: over 4000 substeps water resting against the walls sits on the
wall clamps' kink at zero velocity (mpm_sim.py:305-309), and the clamps, the payload's
singular-value clamp and the mass > 0 test make the objective piecewise smooth. Differences
at eps = 1e-4 cross such kinks for most tanks (their derivatives then change by up to 100x
between eps = 1e-4 and 1e-5), while at 1e-5 and 1e-6 they agree with each other and with the
adjoint. Instances 7 and 20 are two whose differences for amplitude, contraction and
frequency agree across eps 1e-4..1e-6 to 1e-4, i.e. whose objective is smooth on the scale of
the differences. Takes about five minutes on 8 cores:

    python tests/golden/make_grad_golden_tall.py
"""
from __future__ import annotations

import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import oracle  # noqa: E402
from jellyfih_amd import phenotype  # noqa: E402

STEPS = 4000
SLOTS = 160000
SEEDS = list(range(32))
PICK = [7, 20]                       # the instances whose gradients are pinned
PARAMS = ["amp", "contraction", "freq", "E"]
SCALES = {"amp": 500.0, "contraction": 0.2, "freq": 1.0, "E": 700.0}
EPS = [1e-5, 1e-6]
OUT = os.path.join(ROOT, "tests", "golden", "grad_fd_golden_tall.json")


def instance_params(genome):
    """(amp, contraction, freq, payload density, E) of one tank as the batched test sets
    them: the genome's contraction and frequency clipped as evolve.py:150-152 does, and the
    reference defaults otherwise, rounded to fp32 like the reference's fields."""
    p = [500.0, float(np.clip(genome[9], 0.05, 0.6)), float(np.clip(genome[10], 0.5, 2.0)),
         2.5, 700.0]
    return [float(np.float32(v)) for v in p]


def main() -> None:
    t0 = time.time()
    pos, mat, fib, genomes, _ = phenotype.population(
        SEEDS, SLOTS, 128, domain_height=2.0, spawn_offset=phenotype.TALL_SPAWN)
    consts = oracle.reference_consts(n_grid_y=256, domain_height=2.0)
    runs = [(m, name, eps, sg) for m in PICK for name in PARAMS for eps in EPS
            for sg in (1.0, -1.0)]
    o = oracle.OracleSim(len(runs), SLOTS, consts, dtype=np.float64)
    Es = np.empty(len(runs))
    for r, (m, name, eps, sg) in enumerate(runs):
        o.load(r, pos[m], mat[m], fib[m])
        amp, con, frq, pd, E = instance_params(genomes[m])
        d = sg * eps * SCALES[name]
        amp += d if name == "amp" else 0.0
        con += d if name == "contraction" else 0.0
        frq += d if name == "freq" else 0.0
        E += d if name == "E" else 0.0
        o.amp[r], o.contraction[r], o.freq[r], o.pdens[r] = amp, con, frq, pd
        Es[r] = E
    o.set_e_jelly(Es)
    o.t[:] = 0.0
    nthreads = len(os.sched_getaffinity(0))
    o.run(STEPS, nthreads=nthreads)
    print(f"oracle runs done ({time.time() - t0:.0f} s)", flush=True)
    ys = {}
    for r, key in enumerate(runs):
        pay = mat[key[0]] == 2
        ys[key] = float(o.x[r][pay][:, 1].mean())
    fd = {}
    for m in PICK:
        fd[str(m)] = {name: [(ys[(m, name, eps, 1.0)] - ys[(m, name, eps, -1.0)]) /
                             (2 * eps * SCALES[name]) for eps in EPS] for name in PARAMS}
    rec = {
        "case": "32 seeded genomes 0-31, tall tanks (128x256 grid, domain height 2, 160000 "
                "slots, spawn (0.5, 0.60)), 4000 substeps from t=0 at rest, objective = "
                "payload CoM y of each instance after the run; d/d parameter per unit",
        "steps": STEPS,
        "slots": SLOTS,
        "seeds": SEEDS,
        "instances": PICK,
        "n_live": [int((mat[m] >= 0).sum()) for m in range(len(SEEDS))],
        "params": {str(m): instance_params(genomes[m]) for m in PICK},
        "eps": EPS,
        "scales": SCALES,
        "payload_com_y": {str(m): [ys[(m, "amp", EPS[0], 1.0)], ys[(m, "amp", EPS[0], -1.0)]]
                          for m in PICK},
        "oracle_fd": fd,
    }
    with open(OUT, "w") as f:
        json.dump(rec, f, indent=1)
    print(f"wrote {OUT} ({time.time() - t0:.0f} s)")


if __name__ == "__main__":
    main()
