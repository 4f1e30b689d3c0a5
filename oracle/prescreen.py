"""T75 prescreen oracle (test infrastructure only): ctypes binding of oracle/prescreen.c and a
pure-Python restatement of certificate_prescreen.py:60-121 for tiny inputs (Python floats are IEEE
binary64 with separately rounded operations, so both give the reference kernel's bits)."""
import ctypes
import math
from pathlib import Path

import numpy as np

from .q1 import VERTICES

LIB = Path(__file__).resolve().parent / "_build" / "libprescreen_oracle.so"
UNIT = 2. ** -53
ETA = 1e-290
SQRT27_UP = 5.196152422706633


def _gamma(n):
    return n * UNIT / (1 - n * UNIT)


G12, G6, G3 = _gamma(12), _gamma(6), _gamma(3)


def lower_bounds_c(u, grid, D, cells=None):
    lib = ctypes.CDLL(str(LIB))
    fn = lib.oracle_prescreen
    fn.restype = None
    vp, i64 = ctypes.c_void_p, ctypes.c_int64
    fn.argtypes = [vp, vp, vp, i64, i64, i64, i64, vp]
    u = np.ascontiguousarray(u, dtype=np.float64)
    D = np.ascontiguousarray(D, dtype=np.float64)
    nz, ny, nx = grid.shape
    if cells is None:
        n, cp = grid.n_cells, None
    else:
        cells = np.ascontiguousarray(cells, dtype=np.int64)
        n, cp = len(cells), cells.ctypes.data
    out = np.empty(n)
    fn(u.ctypes.data, D.ctypes.data, cp, n, nx, ny, nz, out.ctypes.data)
    return out


def lower_bound_py(u, grid, D, c):
    """One cell, pure Python, operation order of certificate_prescreen.py:66-121."""
    nz, ny, nx = grid.shape
    cz, cy, cx = grid.cell_shape
    x0, y0, z0 = c % cx, (c // cx) % cy, c // (cx * cy)
    base = (z0 * ny + y0) * nx + x0
    offsets = [(z * ny + y) * nx + x for x, y, z in VERTICES]
    flat = np.asarray(u).reshape(-1, 3)
    b = [0.] * 27
    sigma = phi = big = 0.
    for p in range(27):
        F = [[0.] * 3 for _ in range(3)]
        for r in range(3):
            for j in range(3):
                acc = mag = 0.
                for a in range(8):
                    term = float(flat[base + offsets[a], r]) * float(D[j, p, a])
                    acc += term
                    mag += abs(term)
                F[r][j] = acc
                if mag > sigma:
                    sigma = mag
            F[r][r] += 1.
        for r in range(3):
            for j in range(3):
                if abs(F[r][j]) > phi:
                    phi = abs(F[r][j])
        v = (F[0][0] * (F[1][1] * F[2][2] - F[1][2] * F[2][1]) + F[0][1] * (F[1][2] * F[2][0] - F[1][0] * F[2][2])
             + F[0][2] * (F[1][0] * F[2][1] - F[1][1] * F[2][0]))
        b[p] = v
        if abs(v) > big:
            big = abs(v)
    for stride in (1, 3, 9):
        for p in range(27):
            if (p // stride) % 3 == 1:
                b[p] = 2. * b[p] - .5 * (b[p - stride] + b[p + stride])
    e = G12 * (sigma + phi) + ETA
    P = phi + e
    ej = SQRT27_UP * e * (3. * P * P + 3. * P * e + e * e) + 6. * G6 * P * P * P
    bound = 1.01 * (27. * ej + 135. * G3 * big) + ETA
    if any(not math.isfinite(w) for w in b) or not math.isfinite(bound) or not math.isfinite(sigma):
        return float("nan")
    return min(b) - bound
