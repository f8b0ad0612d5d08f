"""Writes tests/golden/grad_fd_golden.json (2000 substeps) or grad_fd_golden_<steps>.json:
fp64 central differences of the payload CoM y
of one seeded reference tank (seed 0, ~62K live particles) after the run, along the
seeded directions of tests/harness.py (body initial velocities, actuation amplitude,
contraction fraction, frequency, E_jelly, mixtures) from the oracle, and along a
per-particle actuation multiplier direction from the fp64 harness (the multiplier has no
counterpart in the reference), at eps = 1e-4, 1e-5, 1e-6; and the directional
derivatives of the harness adjoint (checkpointed every 100 substeps) along the same
directions in fp64, which pins the adjoint against the differences at this horizon, and in
fp32 (the same algorithm in sequential fp32 arithmetic), which measures how far fp32
rounding of a 2000-substep trajectory moves the gradient.

The GPU test recomputes the same directions from their seeds and compares the HIP
adjoint's directional derivatives with these values. Takes a few minutes on 8 cores:

    python tests/golden/make_grad_golden.py          # 2000 substeps
    python tests/golden/make_grad_golden.py 200      # 200 substeps
"""
from __future__ import annotations

import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from tests.harness import (amp_scale_direction, gradient_directions,  # noqa: E402
                           harness_amp_scale_fd, harness_payload_gradient, load_harness,
                           oracle_payload_fd, reference_tank_case)

STEPS = int(sys.argv[1]) if len(sys.argv) > 1 else 2000
EPS = [1e-4, 1e-5, 1e-6]
OUT = os.path.join(ROOT, "tests", "golden", "grad_fd_golden.json" if STEPS == 2000
                   else f"grad_fd_golden_{STEPS}.json")


def main() -> None:
    t0 = time.time()
    cs, pos, mat, fib = reference_tank_case(STEPS)
    kinds, dvs, dps = gradient_directions(cs.mat)
    nthreads = len(os.sched_getaffinity(0))
    fd = oracle_payload_fd(cs, pos, mat, fib, dvs, dps, EPS, nthreads=nthreads)
    print(f"oracle FD done ({time.time() - t0:.0f} s)", flush=True)
    L = load_harness()
    fd_amp = harness_amp_scale_fd(L, cs, amp_scale_direction(cs.mat), EPS)
    print(f"harness FD done ({time.time() - t0:.0f} s)", flush=True)
    gst, gamp, gpar = harness_payload_gradient(L, cs, block=100)
    gp = gpar[0][[0, 1, 2, 4]]
    ad = [float(np.sum(gst[:, 2:4] * dvs[k]) + np.sum(gp * dps[k][[0, 1, 2, 4]]))
          for k in range(len(dvs))]
    ad_amp = float(np.sum(gamp * amp_scale_direction(cs.mat)))
    print(f"harness adjoint done ({time.time() - t0:.0f} s)", flush=True)
    gst, gamp, gpar = harness_payload_gradient(load_harness(fp32=True), cs, block=100)
    gp = gpar[0][[0, 1, 2, 4]]
    ad32 = [float(np.sum(gst[:, 2:4] * dvs[k]) + np.sum(gp * dps[k][[0, 1, 2, 4]]))
            for k in range(len(dvs))]
    ad32_amp = float(np.sum(gamp * amp_scale_direction(cs.mat)))
    print(f"fp32 harness adjoint done ({time.time() - t0:.0f} s)", flush=True)
    rec = {
        "case": f"reference tank seed 0, 80000 slots, {STEPS} substeps from t=0, objective = "
                "payload CoM y after the run",
        "steps": STEPS,
        "n_live": int(cs.P),
        "params": [float(v) for v in cs.params[0]],
        "eps": EPS,
        "direction_seed": 2024,
        "amp_scale_seed": 11,
        "kinds": kinds,
        "oracle_fd": fd.tolist(),
        "harness_amp_scale_fd": fd_amp.tolist(),
        "harness_ad": ad,
        "harness_amp_scale_ad": ad_amp,
        "harness_ad_fp32": ad32,
        "harness_amp_scale_ad_fp32": ad32_amp,
    }
    with open(OUT, "w") as f:
        json.dump(rec, f, indent=1)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
