"""T66 global closed27 first JD root restated from jd_curl/three_d/solvers/global_jd_step.py
(oracle; test infrastructure only)."""
import numpy as np

from .q1 import cell_indices, cofactor, geometry

BISECTIONS = 50


def critical_points(b1, b2, b3):
    """global_jd_step.py:22-40."""
    a, b, c = 3 * b3, 2 * b2, b1
    one = np.ones_like(b1)
    with np.errstate(all="ignore"):
        linear = a == 0
        r_lin = np.where(linear & (b != 0), -c / np.where(b != 0, b, 1.), np.nan)
        disc = b * b - 4 * a * c
        sq = np.sqrt(np.where(disc >= 0, disc, 0.))
        q = -(b + np.where(b >= 0, sq, -sq)) / 2
        r1 = np.where(~linear & (disc >= 0) & (a != 0), q / np.where(a != 0, a, 1.), np.nan)
        r2 = np.where(~linear & (disc >= 0) & (q != 0), c / np.where(q != 0, q, 1.), np.nan)
        r1 = np.where(linear, r_lin, r1)
        r2 = np.where(linear, np.nan, r2)
    r1 = np.where((r1 > 0) & (r1 < 1), r1, one)
    r2 = np.where((r2 > 0) & (r2 < 1), r2, one)
    return np.minimum(r1, r2), np.maximum(r1, r2)


def _q(b0, b1, b2, b3, s):
    return b0 + s * (b1 + s * (b2 + s * b3))


def _touch(b0, b1, b2, b3, s):
    return 4 * np.finfo(float).eps * (np.abs(b0) + s * (np.abs(b1) + s * (np.abs(b2) + s * np.abs(b3))))


def first_roots(b0, b1, b2, b3):
    """global_jd_step.py:52-87: first root in (0, 1] (b0 > 0), inf if none."""
    b0, b1, b2, b3 = (np.asarray(x, dtype=np.float64) for x in (b0, b1, b2, b3))
    out = np.full(b0.shape, np.inf)
    with np.errstate(all="ignore"):
        maybe = ~(b0 > np.abs(b1) + np.abs(b2) + np.abs(b3))
    if not np.any(maybe):
        return out
    b0, b1, b2, b3 = b0[maybe], b1[maybe], b2[maybe], b3[maybe]
    t1, t2 = critical_points(b1, b2, b3)
    one = np.ones_like(b0)
    q1 = _q(b0, b1, b2, b3, t1) <= _touch(b0, b1, b2, b3, t1)
    q2 = _q(b0, b1, b2, b3, t2) <= _touch(b0, b1, b2, b3, t2)
    q3 = _q(b0, b1, b2, b3, one) <= 0
    hi = np.where(q1, t1, np.where(q2, t2, np.where(q3, 1., np.nan)))
    lo = np.where(q1, 0., np.where(q2, t1, np.where(q3, t2, np.nan)))
    has = np.isfinite(hi)
    lo, hi = lo[has], hi[has]
    c = [x[has] for x in (b0, b1, b2, b3)]
    for _ in range(BISECTIONS):
        mid = .5 * (lo + hi)
        neg = _q(*c, mid) <= 0
        hi = np.where(neg, mid, hi)
        lo = np.where(neg, lo, mid)
    roots = np.full(b0.shape, np.inf)
    roots[has] = lo
    out[maybe] = roots
    return out


def point_coefficients(F, V, delta):
    """global_jd_step.py:90-97."""
    AF, AV = cofactor(F), cofactor(V)
    c0 = np.sum(F[..., 0, :] * AF[..., 0, :], axis=-1) - delta
    c1 = np.sum(AF * V, axis=(-2, -1))
    c2 = np.sum(F * AV, axis=(-2, -1))
    c3 = np.sum(V[..., 0, :] * AV[..., 0, :], axis=-1)
    return c0, c1, c2, c3


def global_root(u, v, grid, D, delta, limit):
    """global_root_cpu (global_jd_step.py:107-129) on node fields u (base) and v (direction):
    (min root in s-units or inf, invalid base points, points)."""
    F, _, _, _ = geometry(u, grid, D)
    idx = cell_indices(grid)
    V = np.einsum("jpa,car->cprj", D, v.reshape(-1, 3)[idx])
    c0, c1, c2, c3 = point_coefficients(F, V, delta)
    coefficients = np.stack((c0, c1 * limit, c2 * limit ** 2, c3 * limit ** 3))
    if not np.all(np.isfinite(coefficients)):
        raise FloatingPointError("T66 nonfinite cubic coefficients")
    valid = c0 > 0
    best = np.inf
    if np.any(valid):
        r = first_roots(*(x[valid] for x in coefficients))
        best = float(r.min())
    return best, int(np.count_nonzero(~valid)), int(c0.size)
