"""DST-I Poisson inverse restated from jd_curl/three_d/poisson.py (oracle; test infrastructure only).

apply() uses scipy.fft.dstn/idstn type 1, norm="ortho" -- the reference's own third-party routine
(poisson.py:99-106). apply_operator() restates the independent operator action (poisson.py:59-83)
that pins the inverse: M (M^-1 x) = x.
"""
import numpy as np
from scipy.fft import dstn, idstn


def spectrum(grid, metric="q1"):
    """poisson.py:16-30."""
    ell, mass = [], []
    for n, h in zip(grid.shape[::-1], grid.spacing):
        theta = np.pi * np.arange(1, n - 1) / (n - 1)
        ell.append((2 * np.sin(theta / 2) / h) ** 2)
        mass.append((2 + np.cos(theta)) / 3)
    ex, ey, ez = ell[0][None, None, :], ell[1][None, :, None], ell[2][:, None, None]
    mx, my, mz = mass[0][None, None, :], mass[1][None, :, None], mass[2][:, None, None]
    return ex + ey + ez if metric == "seven_point" else ex * my * mz + mx * ey * mz + mx * my * ez


def denominator(grid, a, b, metric="q1"):
    """poisson.py:118-131."""
    return grid.omega * (a + b * spectrum(grid, metric))


def apply(rhs, den):
    """PoissonInverse3D.apply (poisson.py:99-106) on a (3F,) vector."""
    fields = np.asarray(rhs, dtype=np.float64).reshape((3,) + den.shape)
    t = dstn(fields, type=1, norm="ortho", axes=(-3, -2, -1))
    return idstn(t / den, type=1, norm="ortho", axes=(-3, -2, -1)).ravel()


def _tridiagonal(field, axis, diagonal, off):
    result = diagonal * field
    left, right = [slice(None)] * field.ndim, [slice(None)] * field.ndim
    left[axis], right[axis] = slice(1, None), slice(None, -1)
    result[tuple(left)] += off * field[tuple(right)]
    result[tuple(right)] += off * field[tuple(left)]
    return result


def apply_operator(field, grid, metric="q1"):
    """Scalar KQ/omega action (poisson.py:59-83)."""
    out = np.zeros_like(field)
    for axis, h in enumerate(grid.spacing[::-1]):
        factor = (1 / np.float64(h)) ** 2
        term = _tridiagonal(field, axis, 2 * factor, -factor)
        if metric == "q1":
            for other in range(3):
                if other != axis:
                    term = _tridiagonal(term, other, 2 / 3, 1 / 6)
        out += term
    return out


def matvec(v, grid, a, b, metric="q1"):
    """PoissonInverse3D.matvec (poisson.py:108-115)."""
    fields = np.asarray(v).reshape((3,) + grid.interior_shape)
    return np.stack([grid.omega * (a * s + b * apply_operator(s, grid, metric)) for s in fields]).ravel()
