"""Device Q1 Poisson metric inverse via DST-I (the TorchPoisson / _HostPoisson interface).

poisson_spectrum / build_poisson_inverse follow jd_curl/three_d/poisson.py:16-30, 118-131 on the host
(for callers that hand over a PoissonInverse3D); the solver builds the same denominator on the device
from the separable 1-D spectrum factors (SpectralMetric), bitwise equal and without the 53 MB host
array per metric build. apply / apply_fp32 / apply_equilibrated follow torch_ops.py:98-136 and
hybrid.py:230-286, with every transform running in libseqvp_hip.so (seqvp_dst_poisson[_f32]).
"""
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .problem import InputValidationError

OP_DST = 6


def spectrum_factors(grid):
    """The 1-D factors of the spectrum per axis x, y, z: ell = (2 sin(theta/2)/h)^2 and m = (2 + cos theta)/3,
    theta = pi k/(n-1), k = 1..n-2 (poisson.py:19-23)."""
    ell, mass = [], []
    for n, h in zip(grid.shape[::-1], grid.spacing):
        theta = np.pi * np.arange(1, n - 1) / (n - 1)
        ell.append((2 * np.sin(theta / 2) / h) ** 2)
        mass.append((2 + np.cos(theta)) / 3)
    return ell, mass


def poisson_spectrum(grid, metric="q1"):
    """Q1 (or seven-point) spectrum on the interior DST-I modes (poisson.py:16-30)."""
    if metric not in ("seven_point", "q1"):
        raise InputValidationError("metric: expected seven_point or q1")
    ell, mass = [], []
    for n, h in zip(grid.shape[::-1], grid.spacing):
        theta = np.pi * np.arange(1, n - 1) / (n - 1)
        ell.append((2 * np.sin(theta / 2) / h) ** 2)
        mass.append((2 + np.cos(theta)) / 3)
    ex, ey, ez = ell[0][None, None, :], ell[1][None, :, None], ell[2][:, None, None]
    mx, my, mz = mass[0][None, None, :], mass[1][None, :, None], mass[2][:, None, None]
    result = ex + ey + ez if metric == "seven_point" else ex * my * mz + mx * ey * mz + mx * my * ez
    if np.any(result <= 0) or not np.all(np.isfinite(result)):
        raise FloatingPointError("nonpositive Poisson spectrum")
    return result


@dataclass(frozen=True)
class PoissonInverse3D:
    grid: object
    a: float
    b: float
    metric: str
    denominator: np.ndarray


def build_poisson_inverse(grid, *, a, b, metric="q1"):
    """omega (a + b spectrum) (poisson.py:118-131)."""
    if not (np.isfinite(a) and a >= 0 and np.isfinite(b) and b > 0):
        raise InputValidationError("a >= 0 and b > 0 required")
    den = grid.omega * (a + b * poisson_spectrum(grid, metric))
    if not np.all(np.isfinite(den)) or np.any(den <= 0):
        raise InputValidationError("Poisson coefficients/grid not representable in float64")
    return PoissonInverse3D(grid, float(a), float(b), metric, den)


@dataclass(frozen=True)
class SpectralMetric:
    """A Poisson metric given by (grid, a, b, metric) whose denominator is built on the device."""
    grid: object
    a: float
    b: float
    metric: str = "q1"


class DevicePoisson:
    """M^-1 on the device for one metric; same methods as TorchPoisson (torch_ops.py:107-136).
    `inverse` is a PoissonInverse3D (denominator uploaded) or a SpectralMetric (denominator built on the
    device from the separable 1-D spectrum factors, no host array)."""

    def __init__(self, inverse, device="cuda", ws=None, out=None):
        self.grid = inverse.grid
        self.device = torch.device(device)
        self.inverse = inverse
        if isinstance(inverse, SpectralMetric):
            if not (np.isfinite(inverse.a) and inverse.a >= 0 and np.isfinite(inverse.b) and inverse.b > 0):
                raise InputValidationError("a >= 0 and b > 0 required")
            ell, mass = spectrum_factors(self.grid)
            fac = [torch.tensor(v, dtype=torch.float64, device=self.device) for v in ell + mass]
            mz, my, mx = self.grid.interior_shape
            den = out if out is not None else torch.empty((mz, my, mx), dtype=torch.float64, device=self.device)
            _lib.call("seqvp_poisson_denominator", _lib.ptr(fac[0]), _lib.ptr(fac[3]), _lib.ptr(fac[1]),
                      _lib.ptr(fac[4]), _lib.ptr(fac[2]), _lib.ptr(fac[5]), mx, my, mz, int(inverse.metric == "q1"),
                      self.grid.omega, float(inverse.a), float(inverse.b), _lib.ptr(den), _lib.stream_handle(self.device))
            self._factors = fac      # kept alive until the launch has run
            self._denominator = den
        else:
            self._denominator = torch.tensor(np.ascontiguousarray(inverse.denominator), dtype=torch.float64,
                                             device=self.device)
        self.mz, self.my, self.mx = self.grid.interior_shape
        self.n = 3 * self.mz * self.my * self.mx
        self.ws = ws if ws is not None else _lib.Workspace(self.device)

    def _run(self, name, rhs, scale, negate, out):
        if rhs.dtype != torch.float64 or rhs.numel() != self.n or not rhs.is_contiguous():
            raise InputValidationError("rhs must be a contiguous float64 (3F,) device vector")
        if scale is not None and (scale.numel() != self.n or scale.dtype != torch.float64):
            raise InputValidationError("scale must be a float64 (3F,) device vector")
        out = torch.empty_like(rhs) if out is None else out
        buf, nbytes = self.ws.get(_lib.workspace_bytes(OP_DST, self.grid.shape))
        _lib.call(name, _lib.ptr(rhs), _lib.ptr(self._denominator), _lib.ptr(scale), int(negate), self.mx, self.my,
                  self.mz, 3, _lib.ptr(out), _lib.ptr(buf), nbytes, _lib.stream_handle(self.device))
        return out

    def apply(self, rhs, out=None):
        """M^-1 rhs, float64 (torch_ops.py:107-115)."""
        return self._run("seqvp_dst_poisson", rhs.reshape(-1), None, False, out)

    def apply_fp32(self, rhs, out=None):
        """The same solve with the transforms and the division in float32 (torch_ops.py:117-136)."""
        return self._run("seqvp_dst_poisson_f32", rhs.reshape(-1), None, False, out)

    def equilibrated_direction(self, scale, gradient, out=None, fp32=False):
        """-(s * M^-1 (s * g)) for device vectors, in one launch: the equilibrated PG direction of the device
        solver (the pre-scale, the solve and the negated post-scale of hybrid.py:230-286 fused)."""
        name = "seqvp_dst_poisson_f32" if fp32 else "seqvp_dst_poisson"
        return self._run(name, gradient.reshape(-1), scale.reshape(-1), True, out)

    # ---- the host-array protocol of the reference's device Poisson (_HostPoisson, hybrid.py:224-286) ----------
    def _device(self, a):
        if isinstance(a, torch.Tensor):
            return a.to(device=self.device, dtype=torch.float64).reshape(-1).contiguous()
        # a read-only host array (the reference freezes its arrays) is copied, not wrapped
        return torch.tensor(np.asarray(a, dtype=np.float64).reshape(-1), dtype=torch.float64, device=self.device)

    def prepare_equilibrated(self, scale, gradient):
        """The operands of apply_equilibrated: (scale, gradient) on the device (host arrays are uploaded)
        (hybrid.py:241-262). A nonfinite s * g shows as a nonfinite direction, which apply_equilibrated hands back
        as its ('x', ...) result."""
        s, g = self._device(scale), self._device(gradient)
        if s.numel() != self.n or g.numel() != self.n:
            raise InputValidationError("scale and gradient must be (3F,) vectors")
        return s, g

    def apply_equilibrated(self, prepared, between=None):
        """(direction, True) with direction = -(s * M^-1 (s * g)) as a host array when every value is finite, else
        ('x', M^-1 (s * g) as a host array) for the caller's own tail (hybrid.py:264-286). `between` is host work run
        once while the device computes."""
        s, g = prepared
        out = self.equilibrated_direction(s, g)
        if between is not None:
            between()
        host = out.cpu().numpy()
        if np.all(np.isfinite(host)):
            return host, True
        pre = (s.cpu().numpy() * g.cpu().numpy())
        return "x", self.apply(torch.as_tensor(pre, device=self.device)).cpu().numpy()


class TorchPoisson(DevicePoisson):
    """DevicePoisson with the reference TorchPoisson's argument order (torch_ops.py:107-111): (device, inverse)."""

    def __init__(self, device, inverse):
        super().__init__(inverse, getattr(device, "device", device))
