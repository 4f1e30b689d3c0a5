"""Grid, weights and Q1 constant tables of the registration problem (host side).

Mirrors the parts of jd_curl.three_d that the device path reads: Grid3D (grid.py:11-70),
Weights3D (problem.py:189-198), the Q1 operator tables (cell_q1.py:101-129) and the constant
vector of the fused Q1 kernels (triton_q1.py:250-258). Every object here also accepts the
reference's own objects by duck typing (grid.shape/spacing/omega, weights.lam/mu/nu/kappa,
certificate_config.delta, barrier_activation), so a reference Problem3D can be handed over as is.
"""
from dataclasses import dataclass
from itertools import product

import math

import numpy as np

VERTICES = tuple((x, y, z) for z, y, x in product((0, 1), repeat=3))
GAUSS_NODES = (.5 - np.sqrt(15.) / 10, .5, .5 + np.sqrt(15.) / 10)
GAUSS_WEIGHTS = (5 / 18, 4 / 9, 5 / 18)
CLOSED_NODES = (0., .5, 1.)


class InputValidationError(ValueError):
    """Contract violation (the reference's validation.InputValidationError)."""


@dataclass(frozen=True)
class Grid3D:
    """Rectangular node grid: shape is (nz, ny, nx), spacing and origin are (x, y, z)."""
    shape: tuple
    spacing: tuple
    origin: tuple = (0., 0., 0.)

    @property
    def omega(self):
        return float(math.prod(self.spacing))        # ((h0 h1) h2), the same product as np.prod

    @property
    def interior_shape(self):
        return tuple(n - 2 for n in self.shape)

    @property
    def cell_shape(self):
        return tuple(n - 1 for n in self.shape)

    @property
    def n_cells(self):
        return math.prod(self.cell_shape)

    @property
    def n_nodes(self):
        return math.prod(self.shape)

    @property
    def n_free_scalar(self):
        return math.prod(self.interior_shape)

    @property
    def axes(self):
        nz, ny, nx = self.shape
        return tuple(o + h * np.arange(n, dtype=np.float64) for n, h, o in zip((nx, ny, nz), self.spacing, self.origin))

    def node_weights(self):
        """q = omega * tz * ty * tx with trapezoid weights 1/2 on the end nodes (grid.py:37-39)."""
        tr = []
        for n in self.shape[::-1]:
            w = np.ones(n)
            w[[0, -1]] = .5
            tr.append(w)
        tx, ty, tz = tr
        return self.omega * tz[:, None, None] * ty[None, :, None] * tx[None, None, :]

    def identity(self):
        x, y, z = self.axes
        out = np.empty(self.shape + (3,))
        out[..., 0], out[..., 1], out[..., 2] = x, y[:, None], z[:, None, None]
        return out


def make_grid_3d(*, nx, ny, nz, hx, hy, hz, x0=0., y0=0., z0=0.):
    """make_grid_3d (grid.py:49-70) without the unit/hash bookkeeping."""
    for n in (nx, ny, nz):
        if int(n) != n or n < 3:
            raise InputValidationError("grid sizes must be integers >= 3")
    for h in (hx, hy, hz):
        if not (np.isfinite(h) and h > 0):
            raise InputValidationError("spacing must be positive and finite")
    return Grid3D((int(nz), int(ny), int(nx)), (float(hx), float(hy), float(hz)), (float(x0), float(y0), float(z0)))


@dataclass(frozen=True)
class Weights3D:
    lam: float
    mu: float
    nu: float
    kappa: float


def q1_coefficients(spacing, location):
    """Q1 derivative table D (3, P, 8) and weights, as CellQ1Operators3D builds them (cell_q1.py:101-129):
    D[j, p, a] = ((+-1) / h_j) * f_o1 * f_o2 in that product order."""
    if location == "gauss27":
        nodes, weights = GAUSS_NODES, GAUSS_WEIGHTS
    elif location == "closed27_uniform":
        nodes, weights = CLOSED_NODES, (1 / 3, 1 / 3, 1 / 3)
    else:
        raise InputValidationError("location: gauss27 or closed27_uniform")
    omega = float(np.prod(spacing))
    points = np.array([(x, y, z) for z, y, x in product(nodes, repeat=3)])
    w = np.array([omega * x * y * z for z, y, x in product(weights, repeat=3)])
    if location == "closed27_uniform":
        w = np.full(27, omega / 27)
    D = np.empty((3, 27, 8))
    for a, vertex in enumerate(VERTICES):
        factors = [points[:, j] if vertex[j] else 1 - points[:, j] for j in range(3)]
        for j in range(3):
            o = [k for k in range(3) if k != j]
            D[j, :, a] = ((1 if vertex[j] else -1) / spacing[j] * factors[o[0]] * factors[o[1]])
    return D, w


@dataclass(frozen=True)
class Q1Problem:
    """The frozen constants of the hot path: grid, weights, certificate delta and barrier activation
    (closed27 localized barrier, the reference's MRI route, problem.py:283-311)."""
    grid: Grid3D
    weights: Weights3D
    delta: float = 0.
    activation: float = .05

    @property
    def barrier_weight(self):
        return self.weights.kappa * self.grid.omega / 27


def problem_constants(problem):
    """(grid, k[39], closed27 D (3,27,8), delta) from a Q1Problem or a reference Problem3D."""
    grid = problem.grid
    if hasattr(problem, "certificate_config"):          # a reference Problem3D
        delta = float(problem.certificate_config.delta)
        act = problem.barrier_activation
        if act is None or getattr(problem, "barrier_location", "closed27_uniform") != "closed27_uniform":
            raise InputValidationError("the device path needs the closed27 localized barrier")
        qw = np.asarray(problem._gauss.weights, dtype=np.float64)
        barrier_weight = float(problem.barrier_weight)
    else:
        delta, act = float(problem.delta), float(problem.activation)
        _, qw = q1_coefficients(grid.spacing, "gauss27")
        barrier_weight = float(problem.barrier_weight)
    if not (0 <= delta < act <= 1):
        raise InputValidationError("require 0 <= delta < activation <= 1")
    w = problem.weights
    k = np.concatenate([[1 / h for h in grid.spacing], GAUSS_NODES, qw,
                        [w.mu, w.nu, barrier_weight, delta, act, act - delta]]).astype(np.float64)
    D, _ = q1_coefficients(grid.spacing, "closed27_uniform")
    return grid, k, D, delta


def native_weights(shape_zyx, spacing_xyz=(1., 1., 1.), mu=.03, nu=.03, kappa=.002):
    """Default native-level weights: lam = 0.1 / L*^2 with L* the geometric-mean field of view
    (docs/method.md:252-255), mu = nu = 0.03, kappa = 0.002."""
    fov = [(n - 1) * h for n, h in zip(shape_zyx[::-1], spacing_xyz)]
    L = float(np.prod(fov)) ** (1 / 3)
    return Weights3D(lam=.1 / L ** 2, mu=mu, nu=nu, kappa=kappa)
