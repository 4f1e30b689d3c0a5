"""Moving-image sampler: not-a-knot cubic tensor spline, evaluated on the device.

build_image_sampler follows jd_curl/three_d/interpolation.py:29-40: the coefficients come from
SciPy's make_interp_spline (k=3, not-a-knot) along z, y, x in turn -- the reference's own
third-party routine, run once per moving image on the host. Sampling (value and physical
gradient at X + u or at arbitrary points) runs in libseqvp_hip.so, as TorchSpline.sample does
(backends/spline.py:81-127).
"""
from dataclasses import dataclass

import numpy as np
import torch
from scipy.interpolate import make_interp_spline

from . import _lib
from .problem import InputValidationError


@dataclass(frozen=True)
class SplineCoefficients:
    coefficients: np.ndarray     # (nz, ny, nx)
    knots: tuple                 # (tz, ty, tx)


def build_image_sampler(image, grid):
    image = np.asarray(image, dtype=np.float64)
    if image.shape != tuple(grid.shape):
        raise InputValidationError("image must have the grid shape")
    coefficients, knots = image.copy(), []
    for axis, coordinates in enumerate(grid.axes[::-1]):
        spline = make_interp_spline(coordinates, coefficients, k=3, axis=axis, bc_type="not-a-knot")
        coefficients = np.moveaxis(spline.c, 0, axis)
        knots.append(np.array(spline.t, dtype=np.float64))
    if not np.all(np.isfinite(coefficients)):
        raise FloatingPointError("nonfinite spline coefficients")
    return SplineCoefficients(np.ascontiguousarray(coefficients), tuple(knots))


class DeviceSpline:
    """Device-resident coefficients and knots of one moving image on one grid."""

    def __init__(self, sampler, grid, device="cuda"):
        self.grid = grid
        self.device = torch.device(device)
        self.c = torch.tensor(sampler.coefficients, dtype=torch.float64, device=self.device)
        self.t = tuple(torch.tensor(t, dtype=torch.float64, device=self.device) for t in sampler.knots)
        self.status = torch.zeros(1, dtype=torch.int64, device=self.device)
        x, y, z = grid.axes
        self.box = (float(x[0]), float(x[-1]), float(y[0]), float(y[-1]), float(z[0]), float(z[-1]))

    def _spline_args(self):
        cz, cy, cx = self.c.shape
        tz, ty, tx = self.t
        return (_lib.ptr(self.c), cz, cy, cx, _lib.ptr(tz), tz.numel(), _lib.ptr(ty), ty.numel(), _lib.ptr(tx),
                tx.numel())

    def warp(self, u, *, with_gradient=True, value=None, gradient=None):
        """W = M(X + u) and b = grad M(X + u) on the node grid. Returns (value, gradient, n_invalid);
        n_invalid counts nodes outside the support (strict on free nodes), a device int64 tensor."""
        nz, ny, nx = self.grid.shape
        if u.dtype != torch.float64 or tuple(u.shape) != (nz, ny, nx, 3) or not u.is_contiguous():
            raise InputValidationError("u must be a contiguous float64 (nz, ny, nx, 3) device field")
        value = torch.empty((nz, ny, nx), dtype=torch.float64, device=self.device) if value is None else value
        if with_gradient and gradient is None:
            gradient = torch.empty_like(u)
        hx, hy, hz = self.grid.spacing
        ox, oy, oz = self.grid.origin
        _lib.call("seqvp_spline_warp", _lib.ptr(u), *self._spline_args(), ox, oy, oz, hx, hy, hz, nx, ny, nz,
                  _lib.ptr(value), _lib.ptr(gradient if with_gradient else None), _lib.ptr(self.status),
                  _lib.stream_handle(self.device))
        return value, (gradient if with_gradient else None), self.status

    def sample(self, points, *, with_gradient=True):
        """Value and xyz gradient at (..., 3) points; raises if a point leaves the closed support."""
        if points.dtype != torch.float64 or points.shape[-1] != 3:
            raise InputValidationError("points must be float64 (..., 3)")
        flat = points.reshape(-1, 3).contiguous()
        n = flat.shape[0]
        value = torch.empty(n, dtype=torch.float64, device=self.device)
        gradient = torch.empty((n, 3), dtype=torch.float64, device=self.device) if with_gradient else None
        _lib.call("seqvp_spline_points", _lib.ptr(flat), n, *self._spline_args(), *self.box, _lib.ptr(value),
                  _lib.ptr(gradient), _lib.ptr(self.status), _lib.stream_handle(self.device))
        if int(self.status.item()):
            raise InputValidationError("spline query outside permitted support")
        return value.reshape(points.shape[:-1]), (gradient.reshape(points.shape) if with_gradient else None)
