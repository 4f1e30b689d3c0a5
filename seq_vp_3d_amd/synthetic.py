"""Synthetic registration pairs (fixed image, known smooth displacement, moving image).

Fixed: K Gaussian blobs sum a_k exp(-|x - c_k|^2 / (2 sigma_k^2)) with centres uniform over the
interior node coordinates, sigma_k uniform in [sigma_lo, sigma_hi] voxels and a_k uniform in
[0.2, 1], min-max normalised to [0, 1]. Known displacement per component c:
A * sum_m s_cm sin(pi kx x^) sin(pi ky y^) sin(pi kz z^) with x^ = (x - x0)/(x_end - x0) (zero on every
face), integer k uniform in {1,2,3}, s uniform in [-1, 1]; A is bisected so that the minimum closed27
Jacobian determinant of id + v is at least J_min. Moving: the fixed image's cubic spline sampled at
X + v(X), plus N(0, noise^2). Draws come from numpy.random.default_rng(seed) in that order.
"""
from dataclasses import dataclass

import numpy as np
import torch

from .problem import make_grid_3d, q1_coefficients


@dataclass(frozen=True)
class PairSpec:
    shape: tuple            # (nz, ny, nx)
    blobs: int
    sigma: tuple            # (lo, hi) voxels
    modes: int
    jmin: float
    noise: float = .01
    seed: int = 0


CONFIG_SMALL = PairSpec((32, 32, 32), 12, (2., 6.), 3, .6, .01, 0)
CONFIG_NATIVE = PairSpec((160, 192, 224), 200, (3., 12.), 6, .5, .01, 1)


def scaled_spec(shape, seed, native=CONFIG_NATIVE):
    """Batch/serving pairs: native recipe with the blob count scaled by volume."""
    vol = float(np.prod(shape)) / float(np.prod(native.shape))
    return PairSpec(tuple(shape), max(1, int(round(native.blobs * vol))), native.sigma, native.modes, native.jmin,
                    native.noise, seed)


def _blobs(spec, grid, rng, device):
    nz, ny, nx = spec.shape
    x, y, z = (torch.tensor(a, dtype=torch.float64, device=device) for a in grid.axes)
    lo = [a[1] for a in grid.axes]
    hi = [a[-2] for a in grid.axes]
    centres = rng.uniform(lo, hi, size=(spec.blobs, 3))
    sigmas = rng.uniform(spec.sigma[0], spec.sigma[1], size=spec.blobs)
    amps = rng.uniform(.2, 1., size=spec.blobs)
    img = torch.zeros(spec.shape, dtype=torch.float64, device=device)
    h = grid.spacing
    for c, s, a in zip(centres, sigmas, amps):
        gx = torch.exp(-(x - c[0]) ** 2 / (2 * (s * h[0]) ** 2))
        gy = torch.exp(-(y - c[1]) ** 2 / (2 * (s * h[1]) ** 2))
        gz = torch.exp(-(z - c[2]) ** 2 / (2 * (s * h[2]) ** 2))
        img += a * gz[:, None, None] * gy[None, :, None] * gx[None, None, :]
    img = (img - img.min()) / (img.max() - img.min())
    return img


def _modes(spec, grid, rng, device):
    x, y, z = (torch.tensor(a, dtype=torch.float64, device=device) for a in grid.axes)
    xh = (x - x[0]) / (x[-1] - x[0])
    yh = (y - y[0]) / (y[-1] - y[0])
    zh = (z - z[0]) / (z[-1] - z[0])
    k = rng.integers(1, 4, size=(3, spec.modes, 3))
    s = rng.uniform(-1., 1., size=(3, spec.modes))
    w = torch.zeros(spec.shape + (3,), dtype=torch.float64, device=device)
    for c in range(3):
        for m in range(spec.modes):
            kx, ky, kz = (int(v) for v in k[c, m])
            w[..., c] += float(s[c, m]) * (torch.sin(np.pi * kz * zh)[:, None, None]
                                           * torch.sin(np.pi * ky * yh)[None, :, None]
                                           * torch.sin(np.pi * kx * xh)[None, None, :])
    return w


def closed27_min_jd(u, grid, chunk=1 << 18):
    """min over cells and closed27 points of det(I + grad u) (data-generation helper, any device)."""
    D, _ = q1_coefficients(grid.spacing, "closed27_uniform")
    D = torch.tensor(D, dtype=torch.float64, device=u.device)
    nz, ny, nx = grid.shape
    cz, cy, cx = grid.cell_shape
    flat = u.reshape(-1, 3)
    off = torch.tensor([(z * ny + y) * nx + x for z in (0, 1) for y in (0, 1) for x in (0, 1)], device=u.device)
    best = float("inf")
    for start in range(0, grid.n_cells, chunk):
        c = torch.arange(start, min(start + chunk, grid.n_cells), device=u.device)
        base = ((c // (cx * cy)) * ny + (c // cx) % cy) * nx + c % cx
        local = flat[base[:, None] + off[None, :]]                 # (C, 8, 3)
        F = torch.einsum("jpa,car->cprj", D, local) + torch.eye(3, dtype=u.dtype, device=u.device)
        J = (F[..., 0, 0] * (F[..., 1, 1] * F[..., 2, 2] - F[..., 1, 2] * F[..., 2, 1])
             + F[..., 0, 1] * (F[..., 1, 2] * F[..., 2, 0] - F[..., 1, 0] * F[..., 2, 2])
             + F[..., 0, 2] * (F[..., 1, 0] * F[..., 2, 1] - F[..., 1, 1] * F[..., 2, 0]))
        best = min(best, float(J.min()))
    return best


@dataclass
class Pair:
    grid: object
    fixed: torch.Tensor
    moving: torch.Tensor
    displacement: torch.Tensor
    amplitude: float
    spec: PairSpec


def device_min_jd(grid, device):
    """u -> min over cells and closed27 points of det(I + grad u), by the device barrier kernel (the min J output
    of FusedQ1.barrier_energy); the fast replacement of closed27_min_jd on a GPU."""
    from .problem import Q1Problem, Weights3D
    from .q1 import FusedQ1
    q = FusedQ1(Q1Problem(grid, Weights3D(1., 1., 1., 1.), 0., .05), device)
    return lambda u: float(q.barrier_energy(u.contiguous())[1].item())


def device_sampler(device):
    """sampler(image, grid, points) with the spline coefficients built on the device."""
    from .spline import DeviceSpline
    return lambda image, grid, points: DeviceSpline.from_image(image.contiguous(), grid, device).sample(
        points, with_gradient=False)[0]


def make_pair(spec, device="cpu", sampler=None, min_jd=None):
    """Builds the pair. `sampler(image, grid, points)` returns the spline value of `image` at
    `points` (the caller passes the device sampler on a GPU, or a host one on the CPU); `min_jd(u)`
    (default closed27_min_jd) is the closed27 min J used by the amplitude bisection."""
    nz, ny, nx = spec.shape
    grid = make_grid_3d(nx=nx, ny=ny, nz=nz, hx=1., hy=1., hz=1.)
    jd = (lambda u: closed27_min_jd(u, grid)) if min_jd is None else min_jd
    rng = np.random.default_rng(spec.seed)
    fixed = _blobs(spec, grid, rng, device)
    w = _modes(spec, grid, rng, device)
    lo, hi = 0., 1.
    while jd(hi * w) >= spec.jmin:
        hi *= 2.
        if hi > 1e6:
            break
    for _ in range(40):
        mid = .5 * (lo + hi)
        if jd(mid * w) >= spec.jmin:
            lo = mid
        else:
            hi = mid
    v = (lo * w).contiguous()
    ident = torch.tensor(grid.identity(), dtype=torch.float64, device=device)
    warped = sampler(fixed, grid, ident + v)
    noise = torch.tensor(rng.normal(0., spec.noise, size=spec.shape), dtype=torch.float64, device=device)
    moving = (warped + noise).contiguous()
    return Pair(grid, fixed.contiguous(), moving, v, lo, spec)
