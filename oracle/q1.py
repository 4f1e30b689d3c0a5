"""NumPy restatement of the reference's Q1 cell model (oracle; test infrastructure only).

Layouts follow the reference: node fields (nz, ny, nx, 3) float64, components xyz, x fastest;
theta/gradient vectors component-major (3, nz-2, ny-2, nx-2) flattened; cell c = (cz*cy+cy_)*cx+cx_;
vertex a = x + 2y + 4z (jd_curl/three_d/cell_q1.py:10); point p = 9*iz + 3*iy + ix
(jd_curl/three_d/cell_q1.py:114). Sums that the reference takes in longdouble stay longdouble here.
"""
from itertools import product

import numpy as np

VERTICES = tuple((x, y, z) for z, y, x in product((0, 1), repeat=3))   # cell_q1.py:10
GAUSS_NODES = (.5 - np.sqrt(15.) / 10, .5, .5 + np.sqrt(15.) / 10)     # cell_q1.py:106
GAUSS_WEIGHTS = (5 / 18, 4 / 9, 5 / 18)                                 # cell_q1.py:107
CLOSED_NODES = (0., .5, 1.)                                             # cell_q1.py:111


class Grid:
    """Grid3D restatement (jd_curl/three_d/grid.py:11-70): shape zyx, spacing xyz, trapezoid mass."""

    def __init__(self, shape_zyx, spacing_xyz=(1., 1., 1.), origin_xyz=(0., 0., 0.)):
        self.shape = tuple(int(n) for n in shape_zyx)
        self.spacing = tuple(float(h) for h in spacing_xyz)
        self.origin = tuple(float(o) for o in origin_xyz)
        self.omega = float(np.prod(self.spacing))
        nz, ny, nx = self.shape
        self.axes = tuple(o + h * np.arange(n, dtype=np.float64)
                          for n, h, o in zip((nx, ny, nz), self.spacing, self.origin))
        self.trapezoid = []
        for n in (nx, ny, nz):
            w = np.ones(n)
            w[[0, -1]] = .5
            self.trapezoid.append(w)

    @property
    def cell_shape(self):
        return tuple(n - 1 for n in self.shape)

    @property
    def interior_shape(self):
        return tuple(n - 2 for n in self.shape)

    @property
    def n_cells(self):
        return int(np.prod(self.cell_shape))

    @property
    def n_free_scalar(self):
        return int(np.prod(self.interior_shape))

    def node_weights(self):
        """grid.py:37-39: q = omega * tz * ty * tx."""
        tx, ty, tz = self.trapezoid
        return self.omega * tz[:, None, None] * ty[None, :, None] * tx[None, None, :]

    def identity(self):
        """grid.py:41-46."""
        x, y, z = self.axes
        out = np.empty(self.shape + (3,))
        out[..., 0], out[..., 1], out[..., 2] = x, y[:, None], z[:, None, None]
        return out


def q1_tables(grid, location):
    """CellQ1Operators3D (cell_q1.py:101-129): points, weights, basis (P,8), derivative D (3,P,8)."""
    if location == "gauss27":
        nodes, weights = GAUSS_NODES, GAUSS_WEIGHTS
    elif location == "closed27_uniform":
        nodes, weights = CLOSED_NODES, (1 / 3, 1 / 3, 1 / 3)
    elif location == "corners8":
        nodes, weights = (0., 1.), (.5, .5)
    else:
        raise ValueError(location)
    points = np.array([(x, y, z) for z, y, x in product(nodes, repeat=3)])
    w = np.array([grid.omega * x * y * z for z, y, x in product(weights, repeat=3)])
    if location == "closed27_uniform":
        w = np.full(27, grid.omega / 27)
    basis = np.empty((len(points), 8))
    D = np.empty((3, len(points), 8))
    for a, vertex in enumerate(VERTICES):
        factors = [points[:, j] if vertex[j] else 1 - points[:, j] for j in range(3)]
        basis[:, a] = factors[0] * factors[1] * factors[2]
        for j in range(3):
            o = [k for k in range(3) if k != j]
            D[j, :, a] = ((1 if vertex[j] else -1) / grid.spacing[j] * factors[o[0]] * factors[o[1]])
    return points, w, basis, D


def cell_indices(grid, cells=None):
    """cell_node_indices (cell_q1.py:89-98) for all (or the listed) cells: (C, 8) flat node ids."""
    if cells is None:
        cells = np.arange(grid.n_cells)
    cz, cy, cx = np.unravel_index(np.asarray(cells), grid.cell_shape)
    _, ny, nx = grid.shape
    return np.stack([((cz + z) * ny + cy + y) * nx + cx + x for x, y, z in VERTICES], axis=-1)


def cofactor(F):
    """geometry.py:11-23: out[r,c] = F[j,l]F[k,m] - F[j,m]F[k,l] over cyclic (r,j,k), (c,l,m)."""
    out = np.empty(F.shape)
    for r, j, k in ((0, 1, 2), (1, 2, 0), (2, 0, 1)):
        for c, l, m in ((0, 1, 2), (1, 2, 0), (2, 0, 1)):
            out[..., r, c] = F[..., j, l] * F[..., k, m] - F[..., j, m] * F[..., k, l]
    return out


def geometry(u, grid, D, cells=None):
    """_compute_cell_geometry (geometry.py:60-89): F = I + grad u, cofactor A, J = F0.A0, curl."""
    idx = cell_indices(grid, cells)
    local = u.reshape(-1, 3)[idx]                         # (C, 8, 3)
    F = np.einsum("jpa,car->cprj", D, local)
    F = F + np.eye(3)
    A = cofactor(F)
    J = np.sum(F[..., 0, :] * A[..., 0, :], axis=-1)
    curl = np.stack((F[..., 2, 1] - F[..., 1, 2], F[..., 0, 2] - F[..., 2, 0],
                     F[..., 1, 0] - F[..., 0, 1]), axis=-1)
    return F, A, J, curl


def barrier_terms(J, delta, weight, activation):
    """Localized C2 barrier (barrier.py:7-36, activation branch): value, first, second."""
    J = np.asarray(J, dtype=np.float64)
    if np.any(J <= delta):
        raise FloatingPointError("barrier: J <= delta")
    value, first, second = (np.zeros_like(J) for _ in range(3))
    active = J < activation
    scale = activation - delta
    t = (J[active] - delta) / scale
    r = (activation - J[active]) / scale
    logt = np.log(t)
    near = r < .5
    logt[near] = np.log1p(-r[near])
    value[active] = -weight * r * r * logt
    first[active] = weight * (2 * r * logt - r * r / t) / scale
    second[active] = weight * (-2 * logt + 4 * r / t + (r / t) ** 2) / scale ** 2
    return value, first, second


def barrier_difference(base_j, trial_j, delta, activation):
    """Unweighted psi(J_trial) - psi(J_base) in longdouble (energy_comparison.py:14-46, localized)."""
    a = np.asarray(base_j, dtype=np.longdouble)
    b = np.asarray(trial_j, dtype=np.longdouble)
    d = np.longdouble(delta)
    act = np.longdouble(activation)
    a, b = np.minimum(a, act), np.minimum(b, act)
    scale = act - d
    ra, rb = (act - a) / scale, (act - b) / scale
    h = (b - a) / scale
    loga = np.log((a - d) / scale)
    near = ra < .5
    loga[near] = np.log1p(-ra[near])
    x = (b - a) / (a - d)
    ratio = np.empty_like(x)
    near = x >= -.5
    ratio[near] = np.log1p(x[near])
    ratio[~near] = np.log((b[~near] - d) / (a[~near] - d))
    return h * (ra + rb) * loga - rb * rb * ratio


class Model:
    """The frozen problem data the hot path reads (Problem3D, problem.py:283-311, closed27 localized barrier)."""

    def __init__(self, grid, *, lam, mu, nu, kappa, delta=0., activation=.05, fixed=None, data_weight=None):
        self.grid = grid
        self.lam, self.mu, self.nu, self.kappa = float(lam), float(mu), float(nu), float(kappa)
        self.delta, self.activation = float(delta), float(activation)
        self.gauss = q1_tables(grid, "gauss27")
        self.closed = q1_tables(grid, "closed27_uniform")
        self.fixed = fixed
        q = grid.node_weights()
        self.q = q
        self.qd = q if data_weight is None else q * data_weight   # data_weight.py:9-12

    @property
    def barrier_weight(self):
        """problem.py:309-311: kappa * omega / 27 (closed27 points)."""
        return self.kappa * self.grid.omega / 27

    def constants(self):
        """The 39-entry constant vector of the fused Q1 kernels (triton_q1_kernels.py header)."""
        g = self.grid
        k = np.zeros(39)
        k[0:3] = [1 / h for h in g.spacing]
        k[3:6] = GAUSS_NODES
        k[6:33] = self.gauss[1]
        k[33], k[34] = self.mu, self.nu
        k[35] = self.barrier_weight
        k[36], k[37] = self.delta, self.activation
        k[38] = self.activation - self.delta
        return k

    # ---- energies ------------------------------------------------------------------------
    def soft_energy(self, u, ref, base=None):
        """FusedQ1.soft_energy semantics (triton_q1.py:572-609; residuals.py:166-172;
        energy_comparison.py:196-200): [JD penalty, curl penalty, stable E difference vs base]."""
        _, qw, _, D = self.gauss
        _, _, J, C = geometry(u, self.grid, D)
        _, _, Jr, Cr = geometry(ref, self.grid, D)
        ld = np.longdouble
        jd = ld(.5) * ld(self.mu) * np.sum((qw * (J - Jr) ** 2).astype(ld), dtype=ld)
        curl = ld(.5) * ld(self.nu) * np.sum((qw[None, :, None] * (C - Cr) ** 2).astype(ld), dtype=ld)
        diff = ld(0)
        if base is not None:
            _, _, Ja, Ca = geometry(base, self.grid, D)
            sm, sn = np.sqrt(self.mu * qw), np.sqrt(self.nu * qw)
            for x, y in ((sm * (Ja - Jr), sm * (J - Jr)),
                         (sn[None, :, None] * (Ca - Cr), sn[None, :, None] * (C - Cr))):
                x, y = x.astype(ld), y.astype(ld)
                diff += np.sum((y - x) * (y + x), dtype=ld) / 2
        return float(jd), float(curl), float(diff)

    def barrier_energy(self, u, base=None):
        """FusedQ1.barrier_energy semantics (triton_q1.py:680-719; residuals.py:173-186;
        energy_comparison.py:234-240): [sum psi, min J, sum stable d psi, min base J]."""
        D = self.closed[3]
        _, _, J, _ = geometry(u, self.grid, D)
        w = self.barrier_weight
        value, _, _ = barrier_terms(J, self.delta, w, self.activation)
        out = [float(np.sum(value, dtype=np.longdouble)), float(J.min()), 0., np.inf]
        if base is not None:
            _, _, Ja, _ = geometry(base, self.grid, D)
            d = barrier_difference(Ja, J, self.delta, self.activation)
            out[2] = float(np.longdouble(w) * np.sum(d, dtype=np.longdouble))
            out[3] = float(Ja.min())
        return out

    def node_energies(self, u, ref, warped):
        """image and proximity terms (residuals.py:189-195)."""
        ld = np.longdouble
        image = .5 * np.sum(self.qd * (warped - self.fixed) ** 2, dtype=ld)
        prox = .5 * self.lam * np.sum(self.q[..., None] * (u - ref) ** 2, dtype=ld)
        return float(image), float(prox)

    # ---- derivatives ---------------------------------------------------------------------
    def _scatter(self, out, A, jd, curl, D):
        """_geometry_flux + _scatter_geometry (linearization.py:48-66): out (N,3) += sum_j D_j^T flux[...,r,j]."""
        flux = A * jd[..., None, None]
        if curl is not None:
            flux[..., 2, 1] += curl[..., 0]
            flux[..., 1, 2] -= curl[..., 0]
            flux[..., 0, 2] += curl[..., 1]
            flux[..., 2, 0] -= curl[..., 1]
            flux[..., 1, 0] += curl[..., 2]
            flux[..., 0, 1] -= curl[..., 2]
        local = np.einsum("cprj,jpa->car", flux, D)           # (C, 8, 3)
        idx = cell_indices(self.grid)
        flat = out.reshape(-1, 3)
        for a in range(8):
            np.add.at(flat, idx[:, a], local[:, a])

    def gradient_nodes(self, node_base, u, ref, soft=True, barrier=True):
        """FusedQ1.gradient_nodes (triton_q1.py:504-518) = base + Gauss soft VJP + closed27 barrier VJP
        (linearization.py:159-201), full node field (N,3)."""
        out = np.array(node_base, dtype=np.float64, copy=True)
        if soft:
            _, qw, _, D = self.gauss
            _, A, J, C = geometry(u, self.grid, D)
            _, _, Jr, Cr = geometry(ref, self.grid, D)
            self._scatter(out, A, self.mu * qw * (J - Jr), self.nu * qw[None, :, None] * (C - Cr), D)
        if barrier:
            D = self.closed[3]
            _, A, J, _ = geometry(u, self.grid, D)
            _, first, _ = barrier_terms(J, self.delta, self.barrier_weight, self.activation)
            self._scatter(out, A, first, None, D)
        return out

    def node_gradient_base(self, u, ref, warped, image_grad, total=True):
        """Node terms of the gradient (linearization.py:150-156; torch_ops.py:563-568)."""
        base = self.qd[..., None] * (warped - self.fixed)[..., None] * image_grad
        if total:
            base = base + self.lam * self.q[..., None] * (u - ref)
        return base

    def ggn_diagonal_nodes(self, node_base, u):
        """barrier_ggn_diagonal's cell part (linearization.py:357-380) added to node_base (N,3)."""
        out = np.array(node_base, dtype=np.float64, copy=True)
        _, qw, _, D = self.gauss
        Dc = self.closed[3]
        _, A, _, _ = geometry(u, self.grid, D)
        _, Ac, Jc, _ = geometry(u, self.grid, Dc)
        _, _, second = barrier_terms(Jc, self.delta, self.barrier_weight, self.activation)
        idx = cell_indices(self.grid)
        flat = out.reshape(-1, 3)
        for r in range(3):
            for a in range(8):
                lj = sum(A[..., r, j] * D[j, :, a] for j in range(3))
                lc = sum(Ac[..., r, j] * Dc[j, :, a] for j in range(3))
                cs = sum(D[j, :, a] ** 2 for j in range(3) if j != r)
                value = np.sum(qw * (self.mu * lj ** 2 + self.nu * cs), axis=-1)
                value += np.sum(second * lc ** 2, axis=-1)
                np.add.at(flat[:, r], idx[:, a], value)
        return out

    def ggn_diagonal_base(self, image_grad):
        """Node part of the GGN diagonal: q m b^2 + lam q (linearization.py:362)."""
        return self.qd[..., None] * image_grad ** 2 + self.lam * self.q[..., None]

    def hvp_nodes(self, u, v, image_grad):
        """barrier_ggn_hvp (linearization.py:327-354), full node field (N,3); used to pin the diagonal."""
        out = (self.qd[..., None] * image_grad * np.sum(image_grad * v, axis=-1)[..., None]
               + self.lam * self.q[..., None] * v)
        _, qw, _, D = self.gauss
        _, A, _, _ = geometry(u, self.grid, D)
        idx = cell_indices(self.grid)
        V = np.einsum("jpa,car->cprj", D, v.reshape(-1, 3)[idx])
        jd = np.sum(A * V, axis=(-2, -1))
        curl = np.stack((V[..., 2, 1] - V[..., 1, 2], V[..., 0, 2] - V[..., 2, 0],
                         V[..., 1, 0] - V[..., 0, 1]), axis=-1)
        self._scatter(out, A, self.mu * qw * jd, self.nu * qw[None, :, None] * curl, D)
        Dc = self.closed[3]
        _, Ac, Jc, _ = geometry(u, self.grid, Dc)
        _, _, second = barrier_terms(Jc, self.delta, self.barrier_weight, self.activation)
        Vc = np.einsum("jpa,car->cprj", Dc, v.reshape(-1, 3)[idx])
        self._scatter(out, Ac, second * np.sum(Ac * Vc, axis=(-2, -1)), None, Dc)
        return out

    # ---- T62 node JD measure ---------------------------------------------------------------
    def node_min_jd(self, u):
        """node_min_jd without the T69 extension (jd_affine.py:24-48, 205-225): node = min over its cells of
        min(closed27 J, exact edge-quadratic minima); +inf where no cell."""
        _, _, J, _ = geometry(u, self.grid, self.closed[3])
        edges = np.array([[9 * iz + 3 * iy + ix for ix in range(3)] for iz in (0, 2) for iy in (0, 2)]
                         + [[9 * iz + 3 * iy + ix for iy in range(3)] for iz in (0, 2) for ix in (0, 2)]
                         + [[9 * iz + 3 * iy + ix for iz in range(3)] for iy in (0, 2) for ix in (0, 2)])
        f = J[:, edges]
        f0, f1, f2 = f[..., 0], f[..., 1], f[..., 2]
        a = 2 * (f0 - 2 * f1 + f2)
        b = -3 * f0 + 4 * f1 - f2
        inside = (a > 0) & (b < 0) & (-b < 2 * a)
        vertex = np.full(a.shape, np.inf)
        vertex[inside] = f0[inside] - b[inside] ** 2 / (4 * a[inside])
        minima = np.minimum(J.min(axis=1), vertex.min(axis=1))
        node = np.full(self.grid.shape, np.inf).ravel()
        idx = cell_indices(self.grid)
        for a_ in range(8):
            np.minimum.at(node, idx[:, a_], minima)
        return node.reshape(self.grid.shape)


def inject_free(theta, grid):
    """boundary.py:8-12."""
    out = np.zeros(grid.shape + (3,))
    out[1:-1, 1:-1, 1:-1] = np.moveaxis(theta.reshape((3,) + grid.interior_shape), 0, -1)
    return out


def restrict_free(field, grid):
    """boundary.py:15-17 / linearization.py:84-88."""
    return np.moveaxis(field[1:-1, 1:-1, 1:-1], -1, 0).copy().ravel(order="C")
