"""One certified outer step of the oracle's restatement of the reference's CPU path, timed (the CPU baseline of
bench.py; test infrastructure, not the product).

A synthetic pair on a small grid (Gaussian blobs, the moving image warped by a smooth zero-boundary displacement
through the oracle's cubic spline), the native weights (lam = 0.1/L*^2, mu = nu = 0.03, kappa = 0.002), u = 0, and
one outer step of oracle/driver.py with the native route: the T73 metric (GGN diagonal, image shift, T62 + T69 node
measure, DST-I Poisson inverse), up to 5 accepted PG steps each with the gradient, the direction, BB2, the witness
and T66 root predictions and a fully certified trial (edge hull, prescreen, Bernstein fallback, spline sampling,
energies, stable comparisons), then the outer E/F comparison and norms. Single-threaded NumPy.
"""
import time
from types import SimpleNamespace

import numpy as np

from . import driver as od, q1 as Q, spline as osp


def synthetic_pair(shape, seed):
    rng = np.random.default_rng(seed)
    g = Q.Grid(shape, (1., 1., 1.))
    z, y, x = np.meshgrid(*(np.arange(n, dtype=float) for n in shape), indexing="ij")
    fixed = np.zeros(shape)
    for _ in range(12):
        c = rng.uniform([0, 0, 0], shape)
        s = rng.uniform(2, 5)
        fixed += rng.uniform(.2, 1.) * np.exp(-((z - c[0]) ** 2 + (y - c[1]) ** 2 + (x - c[2]) ** 2) / (2 * s * s))
    fixed = (fixed - fixed.min()) / (fixed.max() - fixed.min())
    v = np.zeros(shape + (3,))
    for c in range(3):
        v[..., c] = (1.2 * np.sin(np.pi * z / (shape[0] - 1)) * np.sin(np.pi * y / (shape[1] - 1))
                     * np.sin(np.pi * (c + 1) * x / (shape[2] - 1)))
    moving = osp.sample(osp.build(fixed, g), g.identity() + v, with_gradient=False)[0]
    L = float(np.prod([n - 1. for n in shape])) ** (1 / 3)
    w = SimpleNamespace(lam=.1 / L ** 2, mu=.03, nu=.03, kappa=.002)
    return od.Pair(shape, (1., 1., 1.), w, fixed, moving)


def outer_step_seconds(shape, seed=0, inner=5):
    """(seconds of one outer step, cells, outer reason): the start state's evaluation, reference and norms are set
    up first and not counted."""
    from threadpoolctl import threadpool_limits
    with threadpool_limits(1):
        pair = synthetic_pair(tuple(shape), seed)
        theta = np.zeros(3 * pair.grid.n_free_scalar)
        t0 = time.perf_counter()
        start = pair.evaluate(theta, Q.inject_free(theta, pair.grid))
        pair.outer_norms(start, start.u)
        setup = time.perf_counter() - t0
        t0 = time.perf_counter()
        out = od.solve_sequential(pair, theta, od.SequentialConfig(max_iterations=1),
                                  od.PGConfig(max_iterations=inner, jd_step_global=True))
        total = time.perf_counter() - t0
    return total - setup, pair.grid.n_cells, out.history[0]["reason"] if out.history else out.status
