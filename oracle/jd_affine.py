"""Oracle: the T62/T69 node JD measure of the JD-affine metric -- TEST INFRASTRUCTURE ONLY.

Restates jd_curl/three_d/solvers/jd_affine.py in NumPy:
- the per-cell measure m_c = min(closed27 J, exact edge-quadratic minima) (jd_affine.py:17-46);
- its T69 face/interior extension on candidate cells (jd_affine.py:49-188): the Bernstein hull of the cell's
  triquadratic J decides the candidates (bernstein_barrier.py:26-38), and mu_c = min(9^3 grid values, bounded
  Newton from the 4 best grid points, 30 iterations, 40 halvings);
- the node minimum over the cells around a node (jd_affine.py:205-225).

J on a cell is the triquadratic through its 27 closed-point values (point p = 9 iz + 3 iy + ix at the reference
nodes 0, 1/2, 1). mu_c is a value J attains in the cell, so it is an estimate for the metric and never a bound.

Pinned (tests/test_oracle_jd_affine.py) by what the mathematics fixes, not by retyped formulas:
- quadratic bowls J = c + (x-s)^T A (x-s) with s inside the cell (mu = c exactly after one Newton step);
- bowls centred outside the cell on one axis (the minimum on a face, found with that coordinate fixed);
- random triquadratic cells against a bounded quasi-Newton minimiser (SciPy L-BFGS-B) from many starts, evaluated
  through the tensor Lagrange interpolant (independent of the monomial conversion): mu is never below the
  attained minimum, and equals it where the grid starts reach the global basin;
- the Bernstein coefficients reproduce J at random points and bound it from below.
"""
import numpy as np

from .q1 import cell_indices, geometry

# Closed27 points on the 12 cell edges, three per edge at t = 0, 1/2, 1 (jd_affine.py:17-21).
EDGES = np.array([[9 * z + 3 * y + x for x in range(3)] for z in (0, 2) for y in (0, 2)]
                 + [[9 * z + 3 * y + x for y in range(3)] for z in (0, 2) for x in (0, 2)]
                 + [[9 * z + 3 * y + x for z in range(3)] for y in (0, 2) for x in (0, 2)])

# Coefficient of t^i in the Lagrange polynomial of node a (nodes 0, 1/2, 1): row a, column i (jd_affine.py:53).
LAGRANGE_MONOMIAL = np.array([[1., -3., 2.], [0., 4., -4.], [0., -1., 2.]])
NINE = np.linspace(0., 1., 9)
STARTS, ITERATIONS, HALVINGS = 4, 30, 40
LOWERED_RELATIVE = 1e-3


def lagrange(t):
    """The three Lagrange polynomials on 0, 1/2, 1 at t (last axis)."""
    t = np.asarray(t, dtype=np.float64)
    return np.stack([2 * (t - .5) * (t - 1), -4 * t * (t - 1), 2 * t * (t - .5)], axis=-1)


def edge_minima(f0, f1, f2):
    """Interior minimum of the quadratic through (0, f0), (1/2, f1), (1, f2), inf when it has none in (0, 1)
    (jd_affine.py:24-36): q(t) = f0 + b t + a t^2 with a = 2(f0 - 2 f1 + f2), b = -3 f0 + 4 f1 - f2."""
    a = 2 * (f0 - 2 * f1 + f2)
    b = -3 * f0 + 4 * f1 - f2
    inside = (a > 0) & (b < 0) & (-b < 2 * a)
    out = np.full(np.shape(f0), np.inf)
    out[inside] = f0[inside] - b[inside] ** 2 / (4 * a[inside])
    return out


def cell_minima(J):
    """m_c per cell of (cells, 27) closed27 values (jd_affine.py:39-46)."""
    f = J[:, EDGES]
    return np.minimum(J.min(axis=1), edge_minima(f[..., 0], f[..., 1], f[..., 2]).min(axis=1))


def bernstein(J):
    """Bernstein coefficients of the triquadratic of each row (bernstein_barrier.py:26-38): per axis the middle
    entry becomes 2 b - (a + c)/2, x first, then y, then z, elementwise."""
    v = np.array(J, dtype=np.float64).reshape(-1, 3, 3, 3)
    v[..., 1] = 2. * v[..., 1] - .5 * (v[..., 0] + v[..., 2])
    v[..., 1, :] = 2. * v[..., 1, :] - .5 * (v[..., 0, :] + v[..., 2, :])
    v[..., 1, :, :] = 2. * v[..., 1, :, :] - .5 * (v[..., 0, :, :] + v[..., 2, :, :])
    return v.reshape(-1, 27)


def monomial(J):
    """C[n, k, j, i]: J = sum C z^k y^j x^i on the unit cell (jd_affine.py:70-71)."""
    v = np.asarray(J, dtype=np.float64).reshape(-1, 3, 3, 3)          # [n, z-node, y-node, x-node]
    M = LAGRANGE_MONOMIAL
    return np.einsum("nzyx,zk,yj,xi->nkji", v, M, M, M)


def grid_values(J):
    """J at the 9^3 points {0, 1/8, ..., 1}^3 in z-major order (index 81 iz + 9 iy + ix), (n, 729)."""
    L = lagrange(NINE)                                                 # [t, node]
    v = np.asarray(J, dtype=np.float64).reshape(-1, 3, 3, 3)
    return np.einsum("nzyx,Zz,Yy,Xx->nZYX", v, L, L, L).reshape(-1, 729)


def grid_point(index):
    """(x, y, z) of the 9^3 grid index."""
    index = np.asarray(index)
    return np.stack([NINE[index % 9], NINE[(index // 9) % 9], NINE[index // 81]], axis=-1)


def _bases(t):
    """Per-axis (value, first, second derivative) bases of 1, t, t^2."""
    one, zero = np.ones_like(t), np.zeros_like(t)
    return (np.stack([one, t, t * t], -1), np.stack([zero, one, 2 * t], -1), np.stack([zero, zero, 2 * one], -1))


def _contract(C, bz, by, bx):
    return np.einsum("nkji,nk,nj,ni->n", C, bz, by, bx)


def value(C, x):
    """J at points x (n, 3) = (x, y, z) of the cells C (n, 3, 3, 3)."""
    return _contract(C, _bases(x[:, 2])[0], _bases(x[:, 1])[0], _bases(x[:, 0])[0])


def derivatives(C, x):
    """Gradient (n, 3) and Hessian (n, 3, 3) in (x, y, z) order (jd_affine.py:82-92)."""
    (px, dx, hx), (py, dy, hy), (pz, dz, hz) = _bases(x[:, 0]), _bases(x[:, 1]), _bases(x[:, 2])
    g = np.stack([_contract(C, pz, py, dx), _contract(C, pz, dy, px), _contract(C, dz, py, px)], -1)
    H = np.empty((len(x), 3, 3))
    H[:, 0, 0], H[:, 1, 1], H[:, 2, 2] = _contract(C, pz, py, hx), _contract(C, pz, hy, px), _contract(C, hz, py, px)
    H[:, 0, 1] = H[:, 1, 0] = _contract(C, pz, dy, dx)
    H[:, 0, 2] = H[:, 2, 0] = _contract(C, dz, py, dx)
    H[:, 1, 2] = H[:, 2, 1] = _contract(C, dz, dy, px)
    return g, H


def bounded_newton(C, x, f, iterations=ITERATIONS, halvings=HALVINGS):
    """Monotone projected Newton on [0, 1]^3 from the points x with values f (jd_affine.py:95-127).

    Per iteration: a coordinate on a face whose gradient points out of the cell is fixed; the free block of the
    Hessian (identity on fixed coordinates) is diagonalised; if every eigenvalue exceeds 1e-12 max(1, |w|max) the
    step is the Newton step on the free coordinates, else the gradient divided by that scale. The step is halved
    (at most `halvings` times) until the clipped point strictly lowers the value; a row with no strict decrease
    stops. Returns the lowest value each row reached."""
    x, best = np.array(x, dtype=np.float64), np.array(f, dtype=np.float64)
    live = np.ones(len(x), dtype=bool)
    for _ in range(iterations):
        rows = np.flatnonzero(live)
        if rows.size == 0:
            break
        xr, Cr = x[rows], C[rows]
        g, H = derivatives(Cr, xr)
        fixed = ((xr <= 0.) & (g > 0.)) | ((xr >= 1.) & (g < 0.))
        gf = np.where(fixed, 0., g)
        keep = ~fixed
        Hf = H * keep[:, :, None] * keep[:, None, :] + np.eye(3) * fixed[:, :, None]
        w, Q = np.linalg.eigh(Hf)
        scale = np.maximum(1., np.abs(w).max(axis=1))
        definite = np.all(w > 1e-12 * scale[:, None], axis=1)
        inv = 1. / np.where(definite[:, None], w, 1.)
        newton = -np.einsum("nij,nj,nkj,nk->ni", Q, inv, Q, gf)
        d = np.where(definite[:, None], newton, -gf / scale[:, None])
        step = np.ones(rows.size)
        moved = np.zeros(rows.size, dtype=bool)
        for _ in range(halvings):
            xt = np.clip(xr + step[:, None] * d, 0., 1.)
            ft = value(Cr, xt)
            take = ~moved & (ft < best[rows])
            x[rows[take]], best[rows[take]] = xt[take], ft[take]
            moved |= take
            step = np.where(moved, step, .5 * step)
            if moved.all():
                break
        live[rows[~moved]] = False
    return best


def face_interior_minima(J, starts=STARTS, iterations=ITERATIONS):
    """mu_c per row of (n, 27) closed27 values (jd_affine.py:130-151): the least of the 9^3 grid values and of
    bounded Newton from the `starts` lowest grid points (ties: lower grid index first)."""
    J = np.asarray(J, dtype=np.float64).reshape(-1, 27)
    if len(J) == 0:
        return np.empty(0)
    C = monomial(J)
    grid = grid_values(J)
    first = np.argsort(grid, axis=1, kind="stable")[:, :starts]
    mu = grid.min(axis=1)
    for s in range(first.shape[1]):
        x0 = grid_point(first[:, s])
        mu = np.minimum(mu, bounded_newton(C, x0, value(C, x0), iterations))
    return mu


def candidates(J, minima, activation):
    """Cells whose Bernstein hull reaches below min(m_c, activation) (jd_affine.py:154-157)."""
    return np.flatnonzero(bernstein(J).min(axis=1) < np.minimum(minima, activation))


def extend(J, minima, activation):
    """m_c <- min(m_c, mu_c) on the candidate cells; returns (minima, stats) with the counts the reference
    records (jd_affine.py:176-202)."""
    minima = np.array(minima, dtype=np.float64)
    cand = candidates(J, minima, activation)
    stats = dict(candidates=int(cand.size), lowered_cells=0, lowered_cells_any=0, mu_min=float("inf"))
    if cand.size:
        mu = face_interior_minima(J[cand])
        old = minima[cand]
        stats.update(lowered_cells=int(np.sum(mu < old * (1 - LOWERED_RELATIVE))),
                     lowered_cells_any=int(np.sum(mu < old)), mu_min=float(mu.min()))
        minima[cand] = np.minimum(old, mu)
    return minima, stats


def node_min_jd(u, grid, closed_D, extension_activation=None):
    """Node JD measure over the full grid (+inf where no cell): the min over the cells around each node of m_c,
    extended on candidate cells when `extension_activation` is given (jd_affine.py:205-225). Returns
    (node, stats or None)."""
    _, _, J, _ = geometry(u, grid, closed_D)
    minima = cell_minima(J)
    stats = None
    if extension_activation is not None:
        minima, stats = extend(J, minima, extension_activation)
    node = np.full(grid.shape, np.inf).ravel()
    idx = cell_indices(grid)
    for a in range(8):
        np.minimum.at(node, idx[:, a], minima)
    return node.reshape(grid.shape), stats
