"""Writes tests/golden/driver_trajectory_32.json (10 outer steps) or driver_trajectory_32_<N>.json (N outer steps):
the oracle driver's trajectory on the 32^3 CPU-sized pair.

The pair is the 32x32x32 recipe (12 Gaussian blobs, sigma 2-6 voxels, 3 sinusoid modes per component, true JD >= 0.6,
noise 0.01, seed 0) built with the oracle's spline sampler; the solve is oracle/driver.py's restatement of the
reference's outer proximal loop and inner PG solve (sequential.py:441-689, poisson_gradient.py:1090-1646) with the
reference's defaults, certified outer steps from u = 0. The oracle's NumPy sweeps take about a minute per outer
step at this size (50 steps: 51 minutes on one core), so the GPU tests (tests/test_gpu_solver.py) read the stored
trajectories instead of recomputing them. Floats are stored as float.hex strings so the file carries every bit.
Usage:

    python tests/golden/make_driver_trajectory.py [--outer 10]     # driver_trajectory_32.json
    python tests/golden/make_driver_trajectory.py --outer 50       # driver_trajectory_32_50.json
"""
import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np
import torch

OUT = Path(__file__).resolve().parent
ROOT = OUT.parent.parent
sys.path.insert(0, str(ROOT))

from oracle import driver as od, q1 as Q, spline as osp  # noqa: E402
from seq_vp_3d_amd.problem import native_weights  # noqa: E402
from seq_vp_3d_amd.synthetic import CONFIG_SMALL, make_pair  # noqa: E402

OUTER_KEYS = ("outer_i", "reason", "inner_status", "inner_accepted_steps", "E_start", "F_start", "GF_before",
              "forcing_target", "E_candidate", "F_candidate", "GF_candidate", "GD_candidate", "GE_candidate")
INNER_KEYS = ("outer", "iteration", "attempt", "reason", "accepted", "step_rule", "tau", "step_proposal",
              "energy_base", "energy_trial", "gradient_norm", "gradient_dot_direction")
THETA_SAMPLES = 97


def enc(v):
    if isinstance(v, (bool, np.bool_)):
        return bool(v)
    if isinstance(v, (int, np.integer)):
        return int(v)
    if isinstance(v, (float, np.floating)):
        return {"hex": float(v).hex()}
    return v


def sampler(img, grid, pts):
    s = osp.build(img.numpy(), Q.Grid(grid.shape, grid.spacing))
    return torch.tensor(osp.sample(s, pts.numpy(), with_gradient=False)[0])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--outer", type=int, default=10)
    args = ap.parse_args()
    t0 = time.perf_counter()
    pair = make_pair(CONFIG_SMALL, "cpu", sampler)
    shape = CONFIG_SMALL.shape
    w = native_weights(shape)
    opair = od.Pair(shape, (1., 1., 1.), w, pair.fixed.numpy(), pair.moving.numpy())
    theta0 = np.zeros(3 * opair.grid.n_free_scalar)
    out = od.solve_sequential(opair, theta0, od.SequentialConfig(max_iterations=args.outer), od.PGConfig(),
                              observer=lambda r: print("outer", r["outer_i"], r["reason"], flush=True))
    theta = np.asarray(out.state.theta, dtype=np.float64)
    index = np.linspace(0, theta.size - 1, THETA_SAMPLES).astype(np.int64)
    doc = dict(
        config=dict(outer=args.outer, shape=list(shape), pair_seed=CONFIG_SMALL.seed, solver="oracle/driver.py",
                    pg="defaults (jd_step_global off, T73 metric, resolution exit)"),
        pair=dict(fixed_sum=enc(float(pair.fixed.sum())), moving_sum=enc(float(pair.moving.sum()))),
        status=out.status,
        history=[{k: enc(r.get(k)) for k in OUTER_KEYS} for r in out.history],
        inner_history=[{k: enc(r.get(k)) for k in INNER_KEYS} for r in out.inner_history],
        theta=dict(index=index.tolist(), values=[enc(float(theta[i])) for i in index],
                   norm=enc(float(np.linalg.norm(theta)))),
    )
    path = OUT / ("driver_trajectory_32.json" if args.outer == 10 else f"driver_trajectory_32_{args.outer}.json")
    path.write_text(json.dumps(doc, separators=(",", ":")) + "\n")
    print(f"wrote {path} ({path.stat().st_size} B): {out.status}, {len(out.history)} outer rows, "
          f"{len(out.inner_history)} attempts, {time.perf_counter() - t0:.1f} s")


if __name__ == "__main__":
    main()
