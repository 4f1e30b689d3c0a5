"""CPU twin of seq_vp_3d_amd.hotpath.PairState built from the oracle (test infrastructure only).

It exposes the attributes and methods the solver (seq_vp_3d_amd/solver.py) calls, computed by the
NumPy oracle, so the driver's decisions can be tested on the CPU and compared step for step with the
device run.
"""
from dataclasses import dataclass

import numpy as np
import torch

from oracle import certificate as ocert, jd_affine as oja, jd_root as ojr, jd_step as ojs, poisson as opo, q1 as Q, spline as osp
from seq_vp_3d_amd.hotpath import Energies, TrialResult


class CpuPairState:
    def __init__(self, grid_shape, spacing, weights, fixed, moving, delta=0., activation=.05):
        self.grid = Q.Grid(grid_shape, spacing)
        w = weights
        self.m = Q.Model(self.grid, lam=w.lam, mu=w.mu, nu=w.nu, kappa=w.kappa, delta=delta, activation=activation,
                         fixed=np.asarray(fixed, dtype=np.float64))
        self.weights = w
        self.spline = osp.build(np.asarray(moving, dtype=np.float64), self.grid)
        self.shape = tuple(grid_shape)
        self.nfree = 3 * self.grid.n_free_scalar
        t = lambda a: torch.tensor(np.asarray(a, dtype=np.float64))
        self._t = t
        self.ref = t(np.zeros(self.shape + (3,)))
        self.theta = t(np.zeros(self.nfree))
        self.g = t(np.zeros(self.nfree))
        self.p = t(np.zeros(self.nfree))
        self.theta_t = t(np.zeros(self.nfree))
        self.D = self.m.closed[3]
        self.energy = None
        self.set_state(self.theta.numpy())

    # state ------------------------------------------------------------------------------------
    def _eval(self, theta):
        u = Q.inject_free(np.asarray(theta), self.grid)
        X = self.grid.identity()
        value, grad = osp.sample(self.spline, X + u)
        return u, value, grad

    def set_state(self, theta):
        self.certificate = None
        self.theta = self._t(theta)
        self.u, self.warped, self.bgrad = self._eval(self.theta.numpy())
        self.energy = self.evaluate_energy(self.u, self.warped)

    def evaluate_energy(self, u, warped):
        image, prox = self.m.node_energies(u, self.ref.numpy(), warped)
        jd, curl, _ = self.m.soft_energy(u, self.ref.numpy())
        bar = self.m.barrier_energy(u)
        return Energies(image, prox, jd, curl, bar[0])

    def dots(self, *pairs):
        return [float(np.dot(a.numpy().astype(np.longdouble), b.numpy())) for a, b in pairs]

    # metric / gradient / direction ---------------------------------------------------------------
    def metric(self, *, shift_factor=1., jd_affine=True, activation=.25, extension=True):
        m, w = self.m, self.weights
        base = m.ggn_diagonal_base(self.bgrad)
        diag = Q.restrict_free(m.ggn_diagonal_nodes(base, self.u), self.grid)
        b = Q.restrict_free(self.bgrad, self.grid)
        alpha = shift_factor * float(np.sum(b * b) / (self.nfree // 3)) / 3
        a = w.lam + alpha
        bp = w.mu / 2 + w.nu / 2
        self.den = opo.denominator(self.grid, a, bp)
        inv_h2 = sum(1 / h ** 2 for h in self.grid.spacing)
        mm = self.grid.omega * (a + bp * (8 / 9) * inv_h2)
        scale = np.sqrt(mm) / np.sqrt(diag)
        if jd_affine:
            node, self.extension_stats = oja.node_min_jd(self.u, self.grid, m.closed[3],
                                                         extension_activation=activation if extension else None)
            node = node[1:-1, 1:-1, 1:-1]
            r = np.minimum(1., np.maximum(node, 1e-12) / activation)
            scale = scale * np.broadcast_to(r, (3,) + r.shape).ravel()
        self.scale = scale
        return dict(shift_alpha=alpha)

    def gradient(self):
        m = self.m
        base = m.node_gradient_base(self.u, self.ref.numpy(), self.warped, self.bgrad)
        self.g = self._t(Q.restrict_free(m.gradient_nodes(base, self.u, self.ref.numpy()), self.grid))
        return self.g

    def gradient_norm(self):
        return float(np.linalg.norm(self.g.numpy())) / np.sqrt(self.grid.omega)

    def direction(self, fp32=False):
        s, g = self.scale, self.g.numpy()
        self.p = self._t(-(s * opo.apply(s * g, self.den)))
        return self.p

    def jd_bound(self, limit):
        v = Q.inject_free(self.p.numpy(), self.grid)
        return ojr.global_root(self.u, v, self.grid, self.D, self.m.delta, limit)

    def jd_bound_launch(self, limit):
        return limit

    def jd_bound_read(self, pending):
        return self.jd_bound(pending)

    def witness_coefficients(self, cells, points):
        v = Q.inject_free(self.p.numpy(), self.grid)
        out = []
        for cell in sorted(set(int(c) for c in cells)):
            pts = np.asarray(points)[np.asarray(cells) == cell]
            A = ojs.point_gradients(self.u, self.grid, cell, pts)
            V = ojs.point_gradients(v, self.grid, cell, pts) - np.eye(3)
            out.append(ojs.determinant_coefficients(A, V, self.m.delta))
        return np.concatenate(out)

    def direction_scalars(self, prev, norm=False, ratio=False, fp32=None, gradient=False):
        if gradient:
            self.gradient()
        if fp32 is not None:
            self.direction(fp32)
        (slope,) = self.dots((self.g, self.p))
        return (slope, (None if prev is None else self.bb2_products(prev)), self.step_bound(),
                self.gradient_norm() if norm else None, self._ratio if ratio else None)

    def direction_scalars_launch(self, prev, norm=False, ratio=False, fp32=None, gradient=False):
        return self.direction_scalars(prev, norm, ratio, fp32, gradient)

    def direction_scalars_read(self, pending):
        return pending

    def step_ratio_launch(self, a, b):
        self._ratio = self.step_ratio(a, b)

    def read_step_ratio(self):
        return self._ratio

    def step_bound(self):
        v = self.p.numpy().reshape((3,) + self.grid.interior_shape)
        pts = (self.grid.identity() + self.u)[1:-1, 1:-1, 1:-1]
        bound = np.inf
        for r, axis in enumerate(self.grid.axes):
            pos, neg = v[r] > 0, v[r] < 0
            if np.any(pos):
                bound = min(bound, float(np.min((axis[-1] - pts[..., r][pos]) / v[r][pos])))
            if np.any(neg):
                bound = min(bound, float(np.min((axis[0] - pts[..., r][neg]) / v[r][neg])))
        return bound

    def bb2_products(self, prev):
        s = self.theta.numpy() - prev["theta"].numpy()
        y = self.g.numpy() - prev["g"].numpy()
        iy = prev["p"].numpy() - self.p.numpy()
        return float(s @ y), float(np.linalg.norm(s)), float(np.linalg.norm(y)), float(y @ iy)

    def step_ratio(self, a, b):
        a, b = a.numpy(), b.numpy()
        length = float(np.linalg.norm([ax[-1] - ax[0] for ax in self.grid.axes]))
        self.step_max_diff = float(np.max(np.abs(a - b)))
        return float(np.max(np.abs(a - b)) / max(length, np.max(np.abs(b))))

    # trial -------------------------------------------------------------------------------------
    def trial(self, tau):
        theta_t = self.theta.numpy() + tau * self.p.numpy()
        self.theta_t = self._t(theta_t)
        u = Q.inject_free(theta_t, self.grid)
        X = self.grid.identity()
        pts = X + u
        lo = np.array([ax[0] for ax in self.grid.axes])
        hi = np.array([ax[-1] for ax in self.grid.axes])
        inner = pts[1:-1, 1:-1, 1:-1]
        if np.any(pts < lo) or np.any(pts > hi) or np.any(inner <= lo) or np.any(inner >= hi):
            return TrialResult(False, "INVALID_DOMAIN", None)
        J = Q.geometry(u, self.grid, self.D)[2]
        folded = not J.min() > self.m.delta
        if folded:
            witness = ojs.nominate_witness(u, self.grid, self.D, self.m.delta)
            if witness is not None:
                return TrialResult(False, "JACOBIAN_VIOLATION", None, witness=witness)
        cert = ocert.certify_map(u, self.grid.shape, self.grid.spacing, self.grid.axes, self.m.delta, D=self.D)
        if cert["status"] != "certified_feasible":
            reason = "FEASIBILITY_UNCERTAIN" if cert["status"] == "uncertain" else cert["reason"]
            return TrialResult(False, reason, None, unproven_cells=cert["n_uncertain"] + cert["n_infeasible"],
                               certificate=cert)
        if folded:
            return TrialResult(False, "JACOBIAN_VIOLATION", None, certificate=cert)
        value, grad = osp.sample(self.spline, pts)
        self._trial = (u, value, grad, cert)
        e = self.evaluate_energy(u, value)
        ref = self.ref.numpy()
        _, _, sdiff = self.m.soft_energy(u, ref, self.u)
        bar = self.m.barrier_energy(u, self.u)
        qd = self.m.qd
        x, y = np.sqrt(qd) * (self.warped - self.m.fixed), np.sqrt(qd) * (value - self.m.fixed)
        img = float(np.sum(((y - x) * (y + x)).astype(np.longdouble)) / 2)
        sl = np.sqrt(self.weights.lam * self.m.q)[..., None]
        x, y = sl * (self.u - ref), sl * (u - ref)
        prox = float(np.sum(((y - x) * (y + x)).astype(np.longdouble)) / 2)
        total = img + prox + sdiff + bar[2]
        outer = img + bar[2]
        return TrialResult(True, "", e, total, outer, min(bar[1], bar[3]) - self.m.delta, 0, None, cert)

    def promote(self, energy, keep_history=False):
        prev = dict(theta=self.theta.clone(), g=self.g.clone(), p=self.p.clone()) if keep_history else None
        self.theta = self.theta_t.clone()
        self.u, self.warped, self.bgrad, self.certificate = self._trial
        self.energy = energy
        return prev

    # outer -------------------------------------------------------------------------------------
    def snapshot(self):
        return dict(theta=self.theta.clone(), u=self.u.copy(), warped=self.warped.copy(), bgrad=self.bgrad.copy(),
                    energy=self.energy, certificate=self.certificate)

    def restore(self, snap):
        self.theta = snap["theta"].clone()
        self.u, self.warped, self.bgrad = snap["u"].copy(), snap["warped"].copy(), snap["bgrad"].copy()
        self.energy = snap["energy"]
        self.certificate = snap["certificate"]

    def compare_to(self, snap):
        e0, e1 = snap["energy"], self.energy
        ref = self.ref.numpy()
        _, _, sdiff = self.m.soft_energy(self.u, ref, snap["u"])
        bar = self.m.barrier_energy(self.u, snap["u"])
        qd = self.m.qd
        x, y = np.sqrt(qd) * (snap["warped"] - self.m.fixed), np.sqrt(qd) * (self.warped - self.m.fixed)
        img = float(np.sum(((y - x) * (y + x)).astype(np.longdouble)) / 2)
        sl = np.sqrt(self.weights.lam * self.m.q)[..., None]
        x, y = sl * (snap["u"] - ref), sl * (self.u - ref)
        prox = float(np.sum(((y - x) * (y + x)).astype(np.longdouble)) / 2)
        del e0, e1
        return img + prox + sdiff + bar[2], img + bar[2], min(bar[1], bar[3]) - self.m.delta

    def outer_norms(self):
        m = self.m
        ref = self.ref.numpy()
        image = m.node_gradient_base(self.u, ref, self.warped, self.bgrad, total=False)
        total = m.node_gradient_base(self.u, ref, self.warped, self.bgrad)
        outer = m.gradient_nodes(image, self.u, ref, soft=False, barrier=True)
        tot = m.gradient_nodes(total, self.u, ref, soft=True, barrier=True)
        r = np.sqrt(self.grid.omega)
        return tuple(float(np.linalg.norm(Q.restrict_free(f, self.grid))) / r for f in (outer, image, tot))

    def refresh_reference(self):
        self.ref = self._t(self.u.copy())
        self.energy = self.evaluate_energy(self.u, self.warped)
        assert (self.energy.proximity, self.energy.jd, self.energy.curl) == (0., 0., 0.)
