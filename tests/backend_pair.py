"""The oracle driver's numerics served by AmdBackend's `backend=` methods (test infrastructure only).

oracle/driver.py restates the reference's solve_poisson_gradient / solve_sequential on a Pair object. BackendPair
answers the same questions the way the reference's driver asks its backend (poisson_gradient.py:387-1646,
sequential.py:288-689): evaluate_trial + promote_trial, gradient, outer_norms, diagonal + node_min_jd +
poisson(...).prepare_equilibrated / apply_equilibrated for the metric and direction, global_jd_root and
stable_energy_reduction. The host arithmetic that the reference driver itself does (the image shift, the metric
diagonal, the affine factors, the step bound, the step ratio and the witness predictions) stays on the host, as there.
"""
import numpy as np

from oracle import driver as od


class BackendPair(od.Pair):
    def __init__(self, backend, shape, spacing, weights, fixed, moving, delta=0., activation=.05):
        from seq_vp_3d_amd import backend as B
        from seq_vp_3d_amd.certificate import CertificateConfig
        from seq_vp_3d_amd.problem import Weights3D, make_grid_3d
        super().__init__(shape, spacing, weights, fixed, moving, delta, activation)
        nz, ny, nx = shape
        self.B = B
        self.backend = backend
        self.rgrid = make_grid_3d(nx=nx, ny=ny, nz=nz, hx=spacing[0], hy=spacing[1], hz=spacing[2])
        self.base_problem = B.make_problem(grid=self.rgrid, fixed_image=fixed, moving_image=moving,
                                           u_ref=np.zeros(tuple(shape) + (3,)),
                                           weights=Weights3D(weights.lam, weights.mu, weights.nu, weights.kappa),
                                           certificate_config=CertificateConfig(delta=delta),
                                           barrier_activation=activation)
        self._problems = {}
        self.calls = {}

    def _count(self, name):
        self.calls[name] = self.calls.get(name, 0) + 1

    def problem(self, ref):
        key = np.ascontiguousarray(ref).tobytes()
        prob = self._problems.get(key)
        if prob is None:
            prob = self._problems[key] = self.B.with_reference(self.base_problem, ref)
        return prob

    def _lin(self, state):
        """Linearization3D(problem, state): the problem the state was evaluated with."""
        return self.B.build_linearization(state.problem, state.accepted)

    # ---- evaluation ----------------------------------------------------------------------------------
    def evaluate(self, theta, ref):
        prob = self.problem(ref)
        seen = {}
        self._count("evaluate_trial")
        trial = self.backend.evaluate_trial(prob, theta, certificate_observer=lambda c: seen.setdefault("cert", c),
                                            witness_observer=lambda cell, point: seen.setdefault("w", (cell, point)))
        cert = seen.get("cert")
        record = None if cert is None else dict(status=cert.status, reason=cert.reason, witness_cell=cert.witness_cell,
                                                witness_reference_point=cert.witness_reference_point)
        if isinstance(trial, self.B.InvalidTrial):
            return od.Invalid(trial.reason, certificate=record, witness=seen.get("w"))
        self._count("promote_trial")
        acc = self.backend.promote_trial(prob, trial)
        e = acc.energy
        state = od.State(np.array(acc.theta), np.array(acc.u), np.array(acc.warped), np.array(acc.image_grad),
                         od.Energy(e.image, e.proximity, e.jd, e.curl, e.barrier), record)
        state.accepted, state.problem = acc, prob
        return state

    def rereference(self, state, ref):
        return self.evaluate(state.theta, ref)

    def reduction(self, base, trial, ref, outer_only=False):
        self._count("stable_energy_reduction")
        assert trial.problem is self.problem(ref) and base.problem is trial.problem
        r = self.backend.stable_energy_reduction(trial.problem, base.accepted, trial.accepted,
                                                 outer_only=outer_only)
        return r.actual, r.resolution, r.certified

    # ---- derivatives, metric, direction ----------------------------------------------------------------
    def gradient(self, state, ref, term="total"):
        assert state.problem is self.problem(ref)
        lin = self._lin(state)
        if term == "total":
            self._count("gradient")
            return self.backend.gradient(lin, 0.)
        self._count("outer_gradient_vectors")
        outer, image, _ = self.backend.outer_gradient_vectors(lin)
        return outer if term == "outer" else image

    def outer_norms(self, state, ref):
        self._count("outer_norms")
        assert state.problem is self.problem(ref)
        return self.backend.outer_norms(self._lin(state), self.rgrid.omega)

    def metric(self, state, config, uncertain):
        """_pg_metric_inverse's T62-family branch with a backend (poisson_gradient.py:400-576, metric_host off)."""
        from seq_vp_3d_amd.poisson import build_poisson_inverse
        g = self.rgrid
        lin = self._lin(state)
        b = state.image_grad[1:-1, 1:-1, 1:-1]
        q = g.node_weights()[1:-1, 1:-1, 1:-1] / g.omega
        alpha = config.shift_factor * float(np.mean(q * np.sum(b * b, axis=-1))) / 3
        a, bp = self.lam + alpha, self.model.mu / 2 + self.model.nu / 2
        self._count("poisson")
        inverse = self.backend.poisson(build_poisson_inverse(g, a=a, b=bp))
        self._count("diagonal")
        diag = self.backend.device.download(self.backend.diagonal(lin))
        if not np.all(np.isfinite(diag)) or np.any(diag <= 0):
            raise FloatingPointError("T59 GGN diagonal must be positive")
        m = g.omega * (a + bp * (8 / 9) * sum(1 / h ** 2 for h in g.spacing))
        scale = np.sqrt(m) / np.sqrt(diag)
        policy = config.metric_policy
        extend = policy.endswith("_v3") or (policy.endswith("_v2") and uncertain > 0)
        self._count("node_min_jd")
        stats = {}
        kwargs = dict(extension_activation=config.jd_affine_activation, stats=stats) if extend else {}
        node = self.backend.node_min_jd(lin, **kwargs)
        self.last_extension = stats
        inner = node[1:-1, 1:-1, 1:-1]
        if not np.all(np.isfinite(inner)):
            raise FloatingPointError("T62 node JD not finite")
        r = np.minimum(1., np.maximum(inner, 1e-12) / config.jd_affine_activation)
        return scale * np.broadcast_to(r, (3,) + r.shape).ravel(), inverse

    def direction(self, g, scale, inverse):
        """fused_direction_checked's device branch (poisson_gradient.py:338-362)."""
        self._count("apply_equilibrated")
        d, ok = inverse.apply_equilibrated(inverse.prepare_equilibrated(scale, g))
        if isinstance(d, str):
            return -(scale * ok)
        return np.array(d)

    def global_root(self, state, p, limit):
        self._count("global_jd_root")
        return self.backend.global_jd_root(self._lin(state), p, limit)
