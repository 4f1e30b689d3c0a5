"""Moving-image sampler: not-a-knot cubic tensor spline, evaluated on the device.

build_image_sampler follows jd_curl/three_d/interpolation.py:29-40: the coefficients come from
SciPy's make_interp_spline (k=3, not-a-knot) along z, y, x in turn -- the reference's own
third-party routine, on the host. DeviceSpline.from_image computes the same coefficients on the
device: the interpolant is linear in the image, so each axis pass is one dense matrix (the
inverse collocation matrix, make_interp_spline of the identity, n <= a few hundred) applied to
every line along that axis with the FP64 MFMA contraction of seqvp_axis_transform. Sampling
(value and physical gradient at X + u or at arbitrary points) runs in libseqvp_hip.so, as
TorchSpline.sample does (backends/spline.py:81-127).
"""
from functools import lru_cache
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


@lru_cache(maxsize=16)
def axis_operators(axes_key):
    """(matrices, knots) per axis z, y, x: the not-a-knot cubic coefficient map of one line (make_interp_spline of
    the identity: column j = the coefficients of unit data j) and the knot vector. axes_key = the axis
    coordinates as tuples (z, y, x)."""
    mats, knots = [], []
    for coords in axes_key:
        coords = np.asarray(coords, dtype=np.float64)
        spline = make_interp_spline(coords, np.eye(len(coords)), k=3, axis=0, bc_type="not-a-knot")
        mats.append(np.ascontiguousarray(spline.c, dtype=np.float64))
        knots.append(np.array(spline.t, dtype=np.float64))
    return tuple(mats), tuple(knots)


class DeviceSpline:
    """Device-resident coefficients and knots of one moving image on one grid."""

    def __init__(self, sampler, grid, device="cuda"):
        self.grid = grid
        self.device = torch.device(device)
        if sampler is None:          # filled by from_image
            return
        self.c = torch.tensor(sampler.coefficients, dtype=torch.float64, device=self.device)
        self.t = tuple(torch.tensor(t, dtype=torch.float64, device=self.device) for t in sampler.knots)
        self._finish()

    @classmethod
    def from_image(cls, image, grid, device=None):
        """The spline of a device image (nz, ny, nx) float64 with the coefficients computed on the device."""
        device = image.device if device is None else torch.device(device)
        if tuple(image.shape) != tuple(grid.shape) or image.dtype != torch.float64:
            raise InputValidationError("image must be a float64 device tensor of the grid shape")
        self = cls(None, grid, device)
        mats, knots = axis_operators(tuple(tuple(a) for a in grid.axes[::-1]))
        mz, my, mx = (torch.as_tensor(m, device=device) for m in mats)
        mx_t = mx.t().contiguous()
        nz, ny, nx = grid.shape
        self.c = torch.empty(grid.shape, dtype=torch.float64, device=device)
        ws = torch.empty(grid.shape, dtype=torch.float64, device=device)
        _lib.call("seqvp_axis_transform", _lib.ptr(image.contiguous()), _lib.ptr(self.c), _lib.ptr(mx_t),
                  _lib.ptr(my), _lib.ptr(mz), nx, ny, nz, _lib.ptr(ws), ws.numel() * 8, _lib.stream_handle(device))
        self.t = tuple(torch.as_tensor(t, device=device) for t in knots)
        self._finish()
        return self

    def _finish(self):
        grid = self.grid
        self.status = torch.zeros(1, dtype=torch.int64, device=self.device)
        x, y, z = grid.axes
        self.box = (float(x[0]), float(x[-1]), float(y[0]), float(y[-1]), float(z[0]), float(z[-1]))

    def _spline_args(self):
        cz, cy, cx = self.c.shape
        tz, ty, tx = self.t
        return (_lib.ptr(self.c), cz, cy, cx, _lib.ptr(tz), tz.numel(), _lib.ptr(ty), ty.numel(), _lib.ptr(tx),
                tx.numel())

    def warp(self, u, *, with_gradient=True, value=None, gradient=None, status=None):
        """W = M(X + u) and b = grad M(X + u) on the node grid. Returns (value, gradient, n_invalid);
        n_invalid counts nodes outside the support (strict on free nodes), a device int64 tensor (`status` when
        given, else the sampler's own)."""
        nz, ny, nx = self.grid.shape
        if u.dtype != torch.float64 or tuple(u.shape) != (nz, ny, nx, 3) or not u.is_contiguous():
            raise InputValidationError("u must be a contiguous float64 (nz, ny, nx, 3) device field")
        value = torch.empty((nz, ny, nx), dtype=torch.float64, device=self.device) if value is None else value
        if with_gradient and gradient is None:
            gradient = torch.empty_like(u)
        hx, hy, hz = self.grid.spacing
        ox, oy, oz = self.grid.origin
        status = self.status if status is None else status
        _lib.call("seqvp_spline_warp", _lib.ptr(u), *self._spline_args(), ox, oy, oz, hx, hy, hz, nx, ny, nz,
                  _lib.ptr(value), _lib.ptr(gradient if with_gradient else None), _lib.ptr(status),
                  _lib.stream_handle(self.device))
        return value, (gradient if with_gradient else None), status

    def sample(self, points, *, chunk_size=65536, strict=None, with_gradient=True):
        """Value and xyz gradient at (..., 3) device points (TorchSpline.sample, backends/spline.py:81-127): every
        point must lie in the closed support and, where the boolean mask `strict` (the query shape) is set, in the
        open support, else InputValidationError; the queries run in launches of at most `chunk_size` points."""
        if (not isinstance(points, torch.Tensor) or points.dtype != torch.float64 or points.shape[-1:] != (3,)
                or points.device != self.device):
            raise InputValidationError("spline requires explicit device float64 xyz points")
        if isinstance(chunk_size, bool) or not isinstance(chunk_size, int) or chunk_size < 1:
            raise InputValidationError("positive integer spline chunk required")
        flat = points.reshape(-1, 3).contiguous()
        n = flat.shape[0]
        mask = None
        if strict is not None:
            if (not isinstance(strict, torch.Tensor) or strict.dtype != torch.bool
                    or tuple(strict.shape) != tuple(points.shape[:-1]) or strict.device != self.device):
                raise InputValidationError("strict query mask mismatch")
            mask = strict.reshape(-1).contiguous().view(torch.uint8)
        value = torch.empty(n, dtype=torch.float64, device=self.device)
        gradient = torch.empty((n, 3), dtype=torch.float64, device=self.device) if with_gradient else None
        chunks = range(0, n, chunk_size)
        status = torch.zeros(max(1, len(chunks)), dtype=torch.int64, device=self.device)
        for k, start in enumerate(chunks):
            end = min(n, start + chunk_size)
            _lib.call("seqvp_spline_points", _lib.ptr(flat[start:end]), end - start,
                      None if mask is None else _lib.ptr(mask[start:end]), *self._spline_args(), *self.box,
                      _lib.ptr(value[start:end]), None if gradient is None else _lib.ptr(gradient[start:end]),
                      _lib.ptr(status[k:k + 1]), _lib.stream_handle(self.device))
        if int(status.sum().item()):
            raise InputValidationError("spline query outside permitted support")
        return value.reshape(points.shape[:-1]), (gradient.reshape(points.shape) if with_gradient else None)
