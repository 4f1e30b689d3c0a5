"""Not-a-knot cubic tensor spline restated from jd_curl/three_d/interpolation.py:29-40, 72-106
(oracle; test infrastructure only), using SciPy's make_interp_spline/NdBSpline as the reference does."""
import numpy as np
from scipy.interpolate import NdBSpline, make_interp_spline


def build(image, grid):
    coefficients, knots = np.array(image, dtype=np.float64), []
    for axis, coordinates in enumerate(grid.axes[::-1]):
        spline = make_interp_spline(coordinates, coefficients, k=3, axis=axis, bc_type="not-a-knot")
        coefficients = np.moveaxis(spline.c, 0, axis)
        knots.append(np.array(spline.t))
    return NdBSpline(tuple(knots), coefficients, (3, 3, 3), extrapolate=False)


def sample(spline, points_xyz, with_gradient=True):
    """Value and xyz gradient at (..., 3) points (interpolation.py:95-106)."""
    q = np.asarray(points_xyz).reshape(-1, 3)[:, ::-1]
    value = spline(q)
    if not with_gradient:
        return value.reshape(np.shape(points_xyz)[:-1]), None
    grad = np.stack([spline(q, nu=nu) for nu in ((0, 0, 1), (0, 1, 0), (1, 0, 0))], axis=-1)
    return value.reshape(np.shape(points_xyz)[:-1]), grad.reshape(np.shape(points_xyz))
