"""JD witnesses and the historical-witness step predictor of the PG line search.

The reference keeps up to 128 recent physical witnesses -- (cell, reference point) pairs where a rejected
candidate had J <= delta -- and predicts the first trial step of each inner iteration from them
(jd_curl/three_d/solvers/jd_step.py:88-239): for every witness cell, at its 27 closed points and its witness
points, the first root in (0, limit] of det(A + s V) - delta (A = I + grad u of the accepted state, V = grad of
the direction) caps tau at safety * root. It also counts undetermined certificates (the trigger of the T69
metric extension).

Witnesses come from two places, as in the reference:
  - a candidate whose closed27 J reaches delta: the first tile of 8192 cells (in cell order) whose least closed27 J
    is <= delta nominates its first minimal point, and the point enclosure of J (interval arithmetic on the
    host) must confirm J <= delta (witness_screen.py:55-88);
  - an infeasible certificate's witness cell and reference point (jd_step.py:151-157).

The geometry runs on the device (seqvp_closed27_tile_minima, seqvp_witness_coefficients); the host confirms a
nominated point (seqvp_point_jd_interval_host) and finds the cubic roots with numpy.roots as the reference does
(first_root_before, jd_step.py:20-37) -- at most 128 cells of 27 or more points per inner step.
"""
from collections import OrderedDict

import numpy as np
import torch

from . import _lib

WITNESS_TILE = 8192        # the device screen's tile (problem.memory.cell_tile_size, problem.py:218-220)
CAPACITY = 128
CLOSED_POINTS = tuple((x, y, z) for z in (0., .5, 1.) for y in (0., .5, 1.) for x in (0., .5, 1.))


def first_root_before(c, limit):
    """Least real root in (0, limit] of c0 + c1 s + c2 s^2 + c3 s^3, or None (jd_step.py:20-37): the coefficients
    are scaled to s/limit; a nonfinite scaled cubic or c0 <= 0 raises FloatingPointError; c0 > |c1|+|c2|+|c3|
    has no root in (0, 1]; otherwise the companion roots (numpy.roots of the cubic over its largest coefficient)
    that are real to 1e-8 relative and lie in (0, 1]."""
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
    try:
        roots = np.roots(poly / np.max(np.abs(poly)))
    except np.linalg.LinAlgError as exc:
        raise FloatingPointError("cubic root failure") from exc
    real = [float(r.real) for r in roots if 0 < r.real <= 1 and abs(r.imag) <= 1e-8 * max(1., abs(r.real))]
    return None if not real else min(real) * limit


def nominate_witness(u, k, grid, delta, device, tile=WITNESS_TILE):
    """The confirmed witness of the candidate node field u (device), or None: (cell, (x, y, z))."""
    nz, ny, nx = grid.shape
    ncell = grid.n_cells
    ntile = -(-ncell // tile)
    out = torch.empty((ntile, 26), dtype=torch.float64, device=device)
    _lib.call("seqvp_closed27_tile_minima", _lib.ptr(u), _lib.ptr(k), nx, ny, nz, tile, _lib.ptr(out),
              _lib.stream_handle(device))
    host = out.cpu().numpy()
    spacing = np.ascontiguousarray(grid.spacing, dtype=np.float64)
    enclosure = np.empty(2)
    for t in np.flatnonzero(host[:, 0] <= delta):
        at = int(host[t, 1])
        cell, point = t * tile + at // 27, np.array(CLOSED_POINTS[at % 27])
        vertices = np.ascontiguousarray(host[t, 2:26])
        _lib.call("seqvp_point_jd_interval_host", _lib.hptr(vertices), _lib.hptr(spacing), _lib.hptr(point),
                  _lib.hptr(enclosure))
        if enclosure[1] <= delta:                      # NaN (a failed enclosure) never confirms
            return int(cell), tuple(float(x) for x in point)
    return None


class JDStepMemory:
    """Recent witnesses of one grid and the count of undetermined certificates (jd_step.py:88-118)."""

    def __init__(self):
        self.points = OrderedDict()
        self.uncertain = 0
        self.grid_shape = None
        self.predictions = 0

    def bind(self, grid):
        shape = (tuple(grid.shape), tuple(grid.spacing))
        if shape != self.grid_shape:
            self.grid_shape = shape
            self.points.clear()
            self.uncertain = 0
            self.ncell = grid.n_cells

    def witness(self, cell, point):
        point = tuple(float(x) for x in point)
        if (not 0 <= int(cell) < self.ncell or len(point) != 3
                or not all(np.isfinite(x) and 0 <= x <= 1 for x in point)):
            raise ValueError("invalid physical JD witness location")
        key = (int(cell), point)
        self.points[key] = None
        self.points.move_to_end(key)
        if len(self.points) > CAPACITY:
            self.points.popitem(last=False)

    def certificate(self, cert):
        """Record a certificate result (a FeasibilityResult or a dict of its fields; None: not reached)."""
        if cert is None:
            return
        get = cert.get if isinstance(cert, dict) else (lambda k: getattr(cert, k))
        if get("status") == "uncertain":
            self.uncertain += 1
        if (get("status") == "infeasible" and get("witness_cell") is not None
                and get("witness_reference_point") is not None):
            self.witness(get("witness_cell"), get("witness_reference_point"))

    def layout(self):
        """(cells (n,), points (n, 3)) of every witness cell's closed points and extra witness points."""
        cells, points = [], []
        for cell in sorted({c for c, _ in self.points}):
            extra = [p for c, p in self.points if c == cell and p not in CLOSED_POINTS]
            pts = list(CLOSED_POINTS) + extra
            cells += [cell] * len(pts)
            points += pts
        return np.array(cells, dtype=np.int64), np.array(points, dtype=np.float64)

    def predict(self, coefficients, limit, safety=.9):
        """(tau, info): `coefficients(cells, points)` returns the (n, 4) cubic coefficients of det(A + s V) - delta
        at the layout's points (host array); tau = min(limit, safety * least first root), or limit when there are
        no witnesses, no root or a failed cubic (jd_step.py:177-239)."""
        info = dict(policy="historical-cell-closed27-cubic-v1", witnesses=len(self.points), points=0,
                    initial_tau=float(limit), root=None, fallback=None)
        tau = limit
        if self.points:
            self.predictions += 1
            cells, points = self.layout()
            try:
                C = coefficients(cells, points)
                least = None
                # points are counted per witness cell once its cubics are done, so a failed cubic leaves the
                # count of the cells before it (jd_step.py:212-222)
                ends = list(np.flatnonzero(np.diff(cells)) + 1) + [len(cells)]
                start = 0
                for end in ends:
                    for c in C[start:end]:
                        r = first_root_before(c, limit)
                        if r is not None:
                            least = r if least is None else min(least, r)
                    info["points"] += end - start
                    start = end
                if least is not None:
                    info["root"] = least
                    tau = min(limit, safety * least)
            except (FloatingPointError, OverflowError) as exc:
                info["fallback"] = str(exc)
                tau = limit
        info["predicted_tau"] = float(tau)
        return tau, info
