"""Oracle: JD witnesses and the historical-witness step predictor -- TEST INFRASTRUCTURE ONLY.

Restates in NumPy:
- the point enclosure of J at a reference point (stationarity.py:39-56) that confirms a nominated witness;
- the witness nomination of the device screen (witness_screen.py:55-88): per tile of 8192 cells the least
  closed27 J (first occurrence), accepted when the enclosure's upper end is <= delta;
- the cubic det(A + s V) - delta and its first root in (0, limit] (jd_step.py:13-37);
- the witness memory: at most 128 recent (cell, reference point) witnesses, least recent dropped first, an
  undetermined-certificate count, and the prediction tau = min(limit, safety * least first root) over the closed27
  points and the witness points of every witness cell (jd_step.py:88-118, 151-239).

Pinned (tests/test_oracle_jd_step.py) by exact rational arithmetic (the enclosure contains the exact J), the cubic
identity det(A + s V) at four values of s, roots of cubics with known factors and a dense sign scan, and a
brute-force scan of det(A + s V) over s for the prediction.
"""
from collections import OrderedDict

import numpy as np

from .certificate import iadd, imul, isub, outward
from .q1 import cell_indices, geometry

CLOSED = np.array([(x, y, z) for z in (0., .5, 1.) for y in (0., .5, 1.) for x in (0., .5, 1.)])
TILE = 8192
CAPACITY = 128


def _idiv(a, b):
    with np.errstate(all="ignore"):
        e = (a[0] / b[0], a[0] / b[1], a[1] / b[0], a[1] / b[1])
    return outward(np.minimum.reduce(e), np.maximum.reduce(e))


def point_jd_interval(vertices, spacing, point):
    """(lo, hi, ok) enclosing J at `point` (x, y, z) of a cell with vertex values (8, 3), vertex a = x + 2y + 4z.
    ok is False where the reference raises (a nonfinite or inverted step)."""
    v = np.asarray(vertices, dtype=np.float64)
    ok = bool(np.all(np.isfinite(v)))
    P = lambda x: (np.float64(x), np.float64(x), True)

    def chk(r):
        nonlocal ok
        ok = ok and bool(r[2])
        return r

    F = [[None] * 3 for _ in range(3)]
    for j in range(3):
        d = [P(0.)] * 3
        for a in range(8):
            if a >> j & 1:
                continue
            b = a | 1 << j
            w = P(1.)
            for k in range(3):
                if k == j:
                    continue
                f = P(point[k]) if a >> k & 1 else chk(isub(P(1.), P(point[k])))
                w = chk(imul(w, f))
            for r in range(3):
                q = chk(_idiv(chk(isub(P(v[b, r]), P(v[a, r]))), P(spacing[j])))
                d[r] = chk(iadd(d[r], chk(imul(q, w))))
        for r in range(3):
            F[r][j] = chk(iadd(d[r], P(1. if r == j else 0.)))
    m = lambda x, y: chk(imul(x, y))
    s = lambda x, y: chk(isub(x, y))
    A = s(m(F[1][1], F[2][2]), m(F[1][2], F[2][1]))
    B = s(m(F[1][0], F[2][2]), m(F[1][2], F[2][0]))
    C = s(m(F[1][0], F[2][1]), m(F[1][1], F[2][0]))
    J = chk(iadd(s(m(F[0][0], A), m(F[0][1], B)), m(F[0][2], C)))
    return float(J[0]), float(J[1]), ok


def nominate_witness(u, grid, closed_D, delta, tile=TILE):
    """The first tile (in cell order) whose least closed27 J is <= delta and whose enclosure at that point
    confirms J <= delta: (cell, point) or None."""
    _, _, J, _ = geometry(u, grid, closed_D)
    vals = np.where(np.isfinite(J), J, np.inf)
    idx = cell_indices(grid)
    nodes = u.reshape(-1, 3)
    for start in range(0, len(vals), tile):
        block = vals[start:start + tile].ravel()
        at = int(np.argmin(block))
        if not block[at] <= delta:
            continue
        cell, p = start + at // 27, at % 27
        lo, hi, ok = point_jd_interval(nodes[idx[cell]], grid.spacing, CLOSED[p])
        if ok and hi <= delta:
            return cell, tuple(float(x) for x in CLOSED[p])
    return None


def cofactor(A):
    """Cofactor matrices of (..., 3, 3)."""
    C = np.empty_like(A)
    for r in range(3):
        for j in range(3):
            r1, r2, j1, j2 = (r + 1) % 3, (r + 2) % 3, (j + 1) % 3, (j + 2) % 3
            C[..., r, j] = A[..., r1, j1] * A[..., r2, j2] - A[..., r1, j2] * A[..., r2, j1]
    return C


def determinant_coefficients(A, V, delta=0.):
    """Ascending coefficients of det(A + s V) - delta (jd_step.py:13-17)."""
    return np.stack((np.linalg.det(A) - delta, np.sum(cofactor(A) * V, axis=(-2, -1)),
                     np.sum(A * cofactor(V), axis=(-2, -1)), np.linalg.det(V)), axis=-1)


def first_root_before(c, limit):
    """The least real root in (0, limit] of the cubic (ascending coefficients c), or None (jd_step.py:20-37).
    Raises FloatingPointError when the scaled coefficients are nonfinite or c0 <= 0."""
    c = np.asarray(c, dtype=float)
    with np.errstate(over="ignore", invalid="ignore"):
        c = c * np.array([1., limit, limit ** 2, limit ** 3])
    if not np.all(np.isfinite(c)) or c[0] <= 0:
        raise FloatingPointError("invalid cubic predictor base/coefficients")
    if c[0] > np.abs(c[1:]).sum():
        return None
    poly = np.trim_zeros(c[::-1], trim="f")
    if len(poly) == 1:
        return None
    roots = np.roots(poly / np.max(np.abs(poly)))
    real = [float(r.real) for r in roots if 0 < r.real <= 1 and abs(r.imag) <= 1e-8 * max(1., abs(r.real))]
    return None if not real else min(real) * limit


def point_gradients(u, grid, cell, points):
    """I + grad u at reference points (n, 3) of one cell: (n, 3, 3) [component, direction]."""
    nodes = u.reshape(-1, 3)[cell_indices(grid)[cell]]
    h = grid.spacing
    out = np.tile(np.eye(3), (len(points), 1, 1))
    for a in range(8):
        bits = [a >> j & 1 for j in range(3)]
        f = np.where(np.array(bits, dtype=bool), points, 1 - points)          # (n, 3)
        for j in range(3):
            o = [k for k in range(3) if k != j]
            w = (1. if bits[j] else -1.) / h[j] * f[:, o[0]] * f[:, o[1]]
            out[:, :, j] += w[:, None] * nodes[a][None, :]
    return out


class Memory:
    """Witness memory of one grid (jd_step.py:88-118, 151-239)."""

    def __init__(self):
        self.points = OrderedDict()
        self.uncertain = 0

    def witness(self, cell, point):
        key = (int(cell), tuple(float(x) for x in point))
        self.points[key] = None
        self.points.move_to_end(key)
        if len(self.points) > CAPACITY:
            self.points.popitem(last=False)

    def certificate(self, status, witness_cell=None, witness_point=None):
        if status == "uncertain":
            self.uncertain += 1
        if status == "infeasible" and witness_cell is not None and witness_point is not None:
            self.witness(witness_cell, witness_point)

    def predict(self, u, v, grid, delta, limit, safety=.9):
        """(tau, least root or None, fallback message or None) for base field u and direction field v."""
        if not self.points:
            return limit, None, None
        least = None
        try:
            for cell in sorted({c for c, _ in self.points}):
                pts = [tuple(p) for p in CLOSED]
                pts += [p for c, p in self.points if c == cell and p not in pts]
                pts = np.asarray(pts)
                A = point_gradients(u, grid, cell, pts)
                V = point_gradients(v, grid, cell, pts) - np.eye(3)
                for c in determinant_coefficients(A, V, delta):
                    r = first_root_before(c, limit)
                    if r is not None:
                        least = r if least is None else min(least, r)
        except (FloatingPointError, OverflowError) as exc:
            return limit, None, str(exc)
        if least is None:
            return limit, None, None
        return min(limit, safety * least), least, None
