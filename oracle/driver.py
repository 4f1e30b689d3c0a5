"""The certified PG inner solve and the proximal outer loop, restated from the reference (oracle; test
infrastructure only).

An independent restatement of the decisions of
  solve_poisson_gradient (jd_curl/three_d/solvers/poisson_gradient.py:1090-1646) on the native route:
    ConservativePGConfig3D with the T62/T69/T73 metric (_pg_metric_inverse, :400-576), the historical-witness
    prediction (prediction=True), the T66 global root (:708-735) when jd_step_global, the T63 resolution exit;
  _solve_sequential (jd_curl/three_d/sequential.py:165-689) with SequentialReuseConfig3D
    (reuse_initial_reference, joint outer gradients);
and of the CPU trial evaluation of the device route (accelerated.py:1666-1730: domain check, closed27 witness
screen, whole-cell certificate, sampling, energies), the stable comparisons (energy_comparison.py:64-138) and the
step algebra (_step_proposal :955-990, sampling_step_bound :993-1011, step_ratio common.py:70-81).

The arithmetic comes from the other oracle modules (q1, spline, poisson, jd_affine, jd_root, jd_step, certificate);
this module only orders it the way the reference does. It is written from the reference's source, not from the
library's driver, so the library's stepwise decisions can be checked against it.
"""
from collections import namedtuple
from dataclasses import dataclass, field

import numpy as np

from . import certificate as ocert, jd_affine as oja, jd_root as ojr, jd_step as ojs, poisson as opo, q1 as Q
from . import spline as osp

EPS = float(np.finfo(float).eps)

Energy = namedtuple("Energy", "image proximity jd curl barrier")


def total(e):
    """EnergyTerms3D.total (residuals.py:31-33), curvature 0 on the native level."""
    return e.image + e.proximity + e.jd + e.curl + e.barrier


def outer(e):
    """EnergyTerms3D.outer (residuals.py:35-37)."""
    return e.image + e.barrier


@dataclass
class State:
    theta: np.ndarray
    u: np.ndarray
    warped: np.ndarray
    image_grad: np.ndarray
    energy: Energy
    certificate: dict


@dataclass
class Invalid:
    reason: str
    certificate: dict | None = None
    witness: tuple | None = None


@dataclass(frozen=True)
class PGConfig:
    """ConservativePGConfig3D fields of the native route (poisson_gradient.py:24-235)."""
    max_iterations: int = 5
    initial_step: float = 1.
    step_min: float = 1e-8
    step_max: float = 10.
    max_backtracks: int = 25
    c1: float = 1e-4
    step_tol: float = 1e-12
    stall_steps: int = 5
    jd_step_safety: float = .9
    metric_policy: str = "image_shift_ggn_equilibrated_jd_affine_v3"
    shift_factor: float = 1.
    jd_affine_activation: float = .25
    resolution_exit: bool = True
    jd_step_global: bool = False


@dataclass(frozen=True)
class SequentialConfig:
    """SequentialReuseConfig3D (sequential.py:28-56)."""
    max_iterations: int = 3
    eta0: float = .01
    step_tol: float = 1e-12
    stall_steps: int = 5


@dataclass
class Inner:
    state: State
    status: str
    reason: str
    gradient_norm: float | None
    accepted_steps: int
    history: list = field(default_factory=list)


class Pair:
    """The frozen problem of one registration pair: grid, weights, images, certificate delta, barrier activation."""

    def __init__(self, shape, spacing, weights, fixed, moving, delta=0., activation=.05):
        self.grid = Q.Grid(shape, spacing)
        self.model = Q.Model(self.grid, lam=weights.lam, mu=weights.mu, nu=weights.nu, kappa=weights.kappa,
                             delta=delta, activation=activation, fixed=np.asarray(fixed, dtype=np.float64))
        self.lam = float(weights.lam)
        self.spline = osp.build(np.asarray(moving, dtype=np.float64), self.grid)
        self.D = self.model.closed[3]
        self.delta = float(delta)

    # ---- evaluation ----------------------------------------------------------------------------------
    def domain_invalid(self, u):
        """check_map_domain (interpolation.py:43-69)."""
        bad = np.zeros(self.grid.shape, dtype=bool)
        with np.errstate(over="ignore", invalid="ignore"):
            for j, coords in enumerate(self.grid.axes):
                sh = [1, 1, 1]
                sh[2 - j] = len(coords)
                value = u[..., j] + coords.reshape(sh)
                bad |= ~np.isfinite(value) | (value < coords[0]) | (value > coords[-1])
                free = value[1:-1, 1:-1, 1:-1]
                bad[1:-1, 1:-1, 1:-1] |= (free <= coords[0]) | (free >= coords[-1])
        return bool(np.any(bad))

    def energy(self, u, warped, ref):
        """_energy (residuals.py:189-206): image, proximity vs ref, Gauss JD/curl change vs ref, closed27 barrier."""
        m = self.model
        image, prox = m.node_energies(u, ref, warped)
        jd, curl, _ = m.soft_energy(u, ref)
        return Energy(image, prox, jd, curl, m.barrier_energy(u)[0])

    def evaluate(self, theta, ref):
        """evaluate_trial + promote_trial of the device route (accelerated.py:1666-1730, 2850-2874)."""
        theta = np.asarray(theta, dtype=np.float64)
        if not np.all(np.isfinite(theta)):
            return Invalid("NONFINITE_STATE")
        u = Q.inject_free(theta, self.grid)
        if self.domain_invalid(u):
            return Invalid("INVALID_DOMAIN")
        J = Q.geometry(u, self.grid, self.D)[2]
        if not J.min() > self.delta:
            # witness_screen.screen (witness_screen.py:55-88): a confirmed closed27 point with J <= delta
            witness = ojs.nominate_witness(u, self.grid, self.D, self.delta)
            if witness is not None:
                return Invalid("JACOBIAN_VIOLATION", witness=witness)
        cert = ocert.certify_map(u, self.grid.shape, self.grid.spacing, self.grid.axes, self.delta, D=self.D)
        if cert["status"] != "certified_feasible":
            reason = "FEASIBILITY_UNCERTAIN" if cert["status"] == "uncertain" else cert["reason"]
            return Invalid(reason, certificate=cert)
        warped, grad = osp.sample(self.spline, self.grid.identity() + u)
        return State(theta.copy(), u, warped, grad, self.energy(u, warped, ref), cert)

    def rereference(self, state, ref):
        """The same map's energy against another reference (make_self_reference + prepare_base)."""
        return State(state.theta, state.u, state.warped, state.image_grad, self.energy(state.u, state.warped, ref),
                     state.certificate)

    # ---- comparisons ---------------------------------------------------------------------------------
    def differences(self, base, trial, ref):
        """(E difference, F difference) in stable form (energy_comparison.py:80-114): node residual differences,
        Gauss change differences and barrier differences, summed in long double."""
        m = self.model
        ld = np.longdouble
        x, y = np.sqrt(m.qd) * (base.warped - m.fixed), np.sqrt(m.qd) * (trial.warped - m.fixed)
        image = np.sum(((y - x) * (y + x)).astype(ld)) / 2
        sl = np.sqrt(self.lam * m.q)[..., None]
        x, y = sl * (base.u - ref), sl * (trial.u - ref)
        prox = np.sum(((y - x) * (y + x)).astype(ld)) / 2
        soft = ld(m.soft_energy(trial.u, ref, base.u)[2])
        barrier = ld(m.barrier_energy(trial.u, base.u)[2])
        return float(image + prox + soft + barrier), float(image + barrier)

    def reduction(self, base, trial, ref, outer_only=False):
        """_reduction (energy_comparison.py:117-124): actual = -difference, resolution = 16 eps max(E)."""
        dE, dF = self.differences(base, trial, ref)
        f = outer if outer_only else total
        actual = -(dF if outer_only else dE)
        resolution = 16 * EPS * max(f(base.energy), f(trial.energy))
        return actual, resolution, actual - resolution

    # ---- derivatives, metric, direction ----------------------------------------------------------------
    def gradient(self, state, ref, term="total"):
        """energy_gradient (linearization.py:150-201) on the free nodes: total (E), image (D) or outer (F)."""
        m = self.model
        if term == "outer":
            base = m.node_gradient_base(state.u, ref, state.warped, state.image_grad, total=False)
            return Q.restrict_free(m.gradient_nodes(base, state.u, ref, soft=False, barrier=True), self.grid)
        if term == "image":
            return Q.restrict_free(m.node_gradient_base(state.u, ref, state.warped, state.image_grad, total=False),
                                   self.grid)
        base = m.node_gradient_base(state.u, ref, state.warped, state.image_grad)
        return Q.restrict_free(m.gradient_nodes(base, state.u, ref), self.grid)

    def norm(self, g):
        """gradient_norm_mass (solvers/common.py:182-191)."""
        return float(np.linalg.norm(g)) / np.sqrt(self.grid.omega)

    def outer_norms(self, state, ref):
        """(GF, GD, GE): outer F, image D and inner E gradient norms (sequential.py:288-338)."""
        return tuple(self.norm(self.gradient(state, ref, t)) for t in ("outer", "image", "total"))

    def metric(self, state, config, uncertain):
        """_pg_metric_inverse for the T62/T69/T73 policies (poisson_gradient.py:400-576): (scale, denominator)."""
        m, g = self.model, self.grid
        b = state.image_grad[1:-1, 1:-1, 1:-1]
        q = g.node_weights()[1:-1, 1:-1, 1:-1] / g.omega
        alpha = config.shift_factor * float(np.mean(q * np.sum(b * b, axis=-1))) / 3      # :269-278
        a, bp = self.lam + alpha, m.mu / 2 + m.nu / 2
        den = opo.denominator(g, a, bp)
        diag = Q.restrict_free(m.ggn_diagonal_nodes(m.ggn_diagonal_base(state.image_grad), state.u), g)
        if not np.all(np.isfinite(diag)) or np.any(diag <= 0):
            raise FloatingPointError("T59 GGN diagonal must be positive")
        mm = g.omega * (a + bp * (8 / 9) * sum(1 / h ** 2 for h in g.spacing))       # metric_diagonal :281-284
        scale = np.sqrt(mm) / np.sqrt(diag)
        policy = config.metric_policy
        extend = policy.endswith("_v3") or (policy.endswith("_v2") and uncertain > 0)
        node, _ = oja.node_min_jd(state.u, g, self.D,
                                  extension_activation=config.jd_affine_activation if extend else None)
        inner = node[1:-1, 1:-1, 1:-1]
        if not np.all(np.isfinite(inner)):
            raise FloatingPointError("T62 node JD not finite")
        r = np.minimum(1., np.maximum(inner, 1e-12) / config.jd_affine_activation)      # affine_factors
        return scale * np.broadcast_to(r, (3,) + r.shape).ravel(), den

    def direction(self, g, scale, den):
        """EquilibratedInverse.apply (poisson_gradient.py:287-293): p = -S M^-1 (S g)."""
        return -(scale * opo.apply(scale * g, den))

    def step_bound(self, state, p):
        """sampling_step_bound (poisson_gradient.py:993-1011)."""
        v = p.reshape((3,) + self.grid.interior_shape)
        pts = (self.grid.identity() + state.u)[1:-1, 1:-1, 1:-1]
        bound = np.inf
        for r, axis in enumerate(self.grid.axes):
            pos, neg = v[r] > 0, v[r] < 0
            if np.any(pos):
                bound = min(bound, float(np.min((axis[-1] - pts[..., r][pos]) / v[r][pos])))
            if np.any(neg):
                bound = min(bound, float(np.min((axis[0] - pts[..., r][neg]) / v[r][neg])))
        return bound

    def global_root(self, state, p, limit):
        """global_root_cpu (global_jd_step.py:107-129): (first root in s-units or inf, invalid base points, points)."""
        return ojr.global_root(state.u, Q.inject_free(p, self.grid), self.grid, self.D, self.delta, limit)

    def step_ratio(self, delta, theta):
        """step_ratio (solvers/common.py:70-81)."""
        length = float(np.linalg.norm([ax[-1] - ax[0] for ax in self.grid.axes]))
        return float(np.max(np.abs(delta)) / max(length, np.max(np.abs(theta))))


def step_proposal(config, theta, g, p, previous):
    """_step_proposal (poisson_gradient.py:955-990): accepted-only BB2 with its fallbacks."""
    protect = lambda v: min(config.step_max, max(config.step_min, v))
    if previous is None:
        return protect(config.initial_step), "initial"
    old_theta, old_g, old_p, old_tau = previous
    fallback = protect(2 * min(old_tau, config.step_max / 2))
    s, y, iy = theta - old_theta, g - old_g, old_p - p
    if not all(np.all(np.isfinite(a)) for a in (s, y, iy)):
        return fallback, "fallback_nonfinite_secant"
    sy = float(s @ y)
    if sy <= 0:
        return fallback, "fallback_nonpositive_curvature"
    cosine = float((s / np.linalg.norm(s)) @ (y / np.linalg.norm(y)))
    if not np.isfinite(cosine):
        return fallback, "fallback_nonfinite_secant"
    if cosine <= 1e-12:
        return fallback, "fallback_small_curvature"
    yy = float(y @ iy)
    if yy <= 0:
        return fallback, "fallback_nonpositive_metric_curvature"
    proposal = sy / yy
    if not np.isfinite(proposal) or proposal <= 0:
        return fallback, "fallback_nonpositive_step"
    return protect(proposal), "bb2"


def precision_stagnation_evidence(rows, untruncated):
    """energy_comparison.py:64-77: every attempt of an untruncated search a valid, unresolved comparison."""
    if not rows or not untruncated:
        return False
    if any(r["reason"] != "ENERGY_RESOLUTION" or not r["trial_valid"] or r["accepted"]
           or r["energy_resolution"] is None for r in rows):
        return False
    return all(r["actual_reduction"] is not None and np.isfinite(r["actual_reduction"])
               and np.isfinite(r["energy_resolution"]) and r["energy_resolution"] >= 0
               and abs(r["actual_reduction"]) <= r["energy_resolution"] for r in rows)


def failed_search_status(rows):
    """_failed_search_status (poisson_gradient.py:1014-1025)."""
    reasons = {r["reason"] for r in rows}
    if reasons == {"INVALID_DOMAIN"}:
        return "domain_blocked"
    if reasons == {"FEASIBILITY_UNCERTAIN"}:
        return "feasibility_uncertain"
    if reasons == {"JACOBIAN_VIOLATION"}:
        return "feasibility_blocked"
    if reasons and all(r.startswith("NONFINITE") for r in reasons):
        return "nonfinite_evaluation"
    return "line_search_failed"


def _observe(memory, trial):
    """certificate_seen / witness_seen (poisson_gradient.py:1137-1150) for an evaluated trial."""
    if isinstance(trial, Invalid):
        if trial.witness is not None:
            memory.witness(*trial.witness)
        cert = trial.certificate
    else:
        cert = trial.certificate
    if cert is not None:
        memory.certificate(cert["status"], cert["witness_cell"], cert["witness_reference_point"])


def solve_poisson_gradient(pair, base, ref, target, config, memory, observer=None):
    """One frozen-reference inner solve from the accepted state `base` (poisson_gradient.py:1090-1646)."""
    grid = pair.grid
    history, accepted_steps, small_steps = [], 0, 0
    previous = None
    inverse = None
    G = None

    def finish(status, reason):
        return Inner(base, status, reason, G, accepted_steps, history)

    while True:
        try:
            g = pair.gradient(base, ref)
            if not np.all(np.isfinite(g)):
                raise FloatingPointError("gradient")
            G = pair.norm(g)
        except FloatingPointError as exc:
            G = None
            return finish("nonfinite_evaluation", str(exc))
        if G <= target:
            return finish("converged_gradient", "GRADIENT_TARGET")
        if small_steps >= config.stall_steps:
            return finish("stagnated", "SMALL_ACCEPTED_STEPS")
        if accepted_steps >= config.max_iterations:
            return finish("max_iterations", "ACCEPTED_STEP_BUDGET")
        if inverse is None:
            try:
                inverse = pair.metric(base, config, memory.uncertain)
            except FloatingPointError as exc:
                return finish("nonfinite_evaluation", str(exc))
        p = pair.direction(g, *inverse)
        slope = float(g @ p)
        if not np.all(np.isfinite(p)) or not np.isfinite(slope):
            return finish("nonfinite_evaluation", "PG direction")
        if slope >= 0:
            return finish("non_descent_direction", "NON_DESCENT_DIRECTION")
        proposal, rule = step_proposal(config, base.theta, g, p, previous)
        bound = pair.step_bound(base, p)
        if bound <= 0:
            return finish("domain_blocked", "NO_POSITIVE_DOMAIN_STEP")
        tau = min(proposal, .9 * bound)
        v = Q.inject_free(p, grid)
        predicted, root, fallback = memory.predict(base.u, v, grid, pair.delta, tau, config.jd_step_safety)
        prediction = dict(predicted_tau=predicted, root=root, fallback=fallback)
        if config.jd_step_global:
            # _global_prediction (:708-735)
            try:
                s, invalid, points = pair.global_root(base, p, tau)
                found = None if not np.isfinite(s) else s * tau
                prediction["global"] = dict(root=found, invalid_base_points=invalid, points=points)
                if found is not None and found != 0:
                    predicted = min(predicted, config.jd_step_safety * found)
            except FloatingPointError as exc:
                prediction["global"] = dict(fallback=str(exc))
        tau = predicted
        untruncated = tau == proposal
        attempts = []
        global_attempts = 0
        exit_floor = 15 * EPS * total(base.energy) if config.resolution_exit else None
        accepted = False
        for attempt in range(config.max_backtracks + 1):
            candidate = base.theta + tau * p
            required = -config.c1 * tau * slope
            trial = pair.evaluate(candidate, ref)
            _observe(memory, trial)
            valid = not isinstance(trial, Invalid)
            actual = resolution = certified = relative = None
            accepted = False
            if valid:
                relative = pair.step_ratio(candidate - base.theta, base.theta)
                actual, resolution, certified = pair.reduction(base, trial, ref)
                if not np.isfinite(actual):
                    reason = "NONFINITE_COMPARISON"
                else:
                    accepted = certified > 0 and certified >= required
                    reason = ("ACCEPTED" if accepted else "ENERGY_RESOLUTION" if abs(actual) <= resolution
                              else "NO_STRICT_DECREASE" if actual < 0 else "ARMIJO_REJECTED")
            else:
                reason = trial.reason
            exit_info = next_prediction = None
            if (config.resolution_exit and not accepted and not reason.startswith("NONFINITE")
                    and global_attempts + 1 <= config.max_backtracks):
                if reason == "JACOBIAN_VIOLATION":
                    pred, r, fb = memory.predict(base.u, v, grid, pair.delta, tau, config.jd_step_safety)
                    next_tau, next_prediction = min(.5 * tau, pred), dict(predicted_tau=pred, root=r, fallback=fb)
                else:
                    next_tau, next_prediction = .5 * tau, prediction
                decrease = next_tau * abs(slope)
                exit_info = dict(next_tau=next_tau, predicted_decrease=decrease, resolution_floor=exit_floor,
                                 stop=bool(decrease <= exit_floor))
            row = dict(iteration=accepted_steps, attempt=attempt + 1, accepted=bool(accepted), reason=reason,
                       trial_valid=valid, energy_resolution=resolution, actual_reduction=actual,
                       certified_reduction=certified, energy_base=total(base.energy),
                       energy_trial=total(trial.energy) if valid else None, tau=tau, step_proposal=proposal,
                       step_rule=rule, untruncated=untruncated, domain_step_bound=bound, gradient_norm=G,
                       step_ratio=relative, gradient_dot_direction=slope, jd_prediction=prediction,
                       resolution_exit=exit_info, witnesses_after_trial=len(memory.points))
            history.append(row)
            attempts.append(row)
            if observer is not None:
                observer(row)
            if reason.startswith("NONFINITE"):
                return finish("nonfinite_evaluation", reason)
            if accepted:
                small_steps = small_steps + 1 if relative <= config.step_tol else 0
                previous = (base.theta, g, p, tau)
                base = trial
                accepted_steps += 1
                break
            global_attempts += 1
            if global_attempts > config.max_backtracks:
                break
            if exit_info is not None:
                tau, prediction = exit_info["next_tau"], next_prediction
                if exit_info["stop"]:
                    break
            elif reason == "JACOBIAN_VIOLATION":
                pred, r, fb = memory.predict(base.u, v, grid, pair.delta, tau, config.jd_step_safety)
                prediction = dict(predicted_tau=pred, root=r, fallback=fb)
                tau = min(.5 * tau, pred)
            else:
                tau *= .5
        if not accepted:
            evidence = precision_stagnation_evidence(attempts, untruncated)
            status = "precision_stagnated" if evidence else failed_search_status(attempts)
            return finish(status, attempts[-1]["reason"])


@dataclass
class Outcome:
    status: str
    reason: str
    state: State
    ref: np.ndarray
    G0: float
    GF: float
    threshold: float
    history: list
    inner_history: list


def solve_sequential(pair, theta, config=SequentialConfig(), inner_config=PGConfig(), observer=None,
                     inner_observer=None):
    """The proximal outer loop (sequential.py:165-689) from theta, the start state being its own reference
    (reuse_initial_reference / refresh)."""
    memory = ojs.Memory()
    history, inner_rows = [], []
    u0 = Q.inject_free(np.asarray(theta, dtype=np.float64), pair.grid)
    raw = pair.evaluate(theta, u0)
    if isinstance(raw, Invalid):
        raise ValueError(f"invalid start state: {raw.reason}")
    ref, current = raw.u.copy(), raw
    GF, GD, GE = pair.outer_norms(current, ref)
    G0 = GF
    threshold = 1e-8 + 1e-6 * G0
    accepted_count, small_steps = 0, 0

    def finish(status, reason):
        return Outcome(status, reason, current, ref, G0, GF, threshold, history, inner_rows)

    while True:
        if GF <= threshold:
            return finish("converged_regularized_gradient", "Feasible F=D+B gradient criterion")
        if small_steps >= config.stall_steps:
            return finish("outer_stagnated", "Small accepted steps with nonzero outer gradient")
        if accepted_count >= config.max_iterations:
            return finish("outer_max_iterations", "Outer budget exhausted; not convergence")
        start = current
        eta = config.eta0 if G0 == 0 else config.eta0 * np.sqrt(min(GF, G0) / G0)
        target = float(eta * GF)
        inner = solve_poisson_gradient(pair, start, ref, target, inner_config, memory,
                                       observer=lambda row: inner_rows.append(dict(row, outer=len(history))))
        candidate = inner.state
        GFa, GDa, GEa = pair.outer_norms(candidate, ref)
        relative = pair.step_ratio(candidate.theta - start.theta, start.theta)
        totalE = pair.reduction(start, candidate, ref)
        fixedF = pair.reduction(start, candidate, ref, outer_only=True)
        cert = totalE[2] > 0 and fixedF[2] > 0
        eligible = inner.status in ("converged_gradient", "max_iterations", "stagnated")
        same = bool(np.array_equal(start.theta, candidate.theta))
        accepted = bool(eligible and cert and not same)
        reason = ("ACCEPTED" if accepted else "INNER_FAILED_NOT_COMMITTED" if not eligible
                  else "NO_PROGRESS" if same else "DESCENT_CERTIFICATE_FAILED")
        row = dict(outer_i=len(history), E_start=total(start.energy), E_candidate=total(candidate.energy),
                   F_start=outer(start.energy), F_candidate=outer(candidate.energy),
                   actual_reduction_E=totalE[0], resolution_E=totalE[1], actual_reduction_F=fixedF[0],
                   resolution_F=fixedF[1], GF_initial=G0, GF_threshold=threshold, GF_before=GF, GF_candidate=GFa,
                   GD_candidate=GDa, GE_candidate=GEa, forcing_eta=float(eta), forcing_target=target,
                   inner_status=inner.status, inner_reason=inner.reason, inner_accepted_steps=inner.accepted_steps,
                   outer_accepted=accepted, reason=reason, step_ratio=relative)
        history.append(row)
        if observer is not None:
            observer(row)
        if not accepted:
            if inner.status == "precision_stagnated":
                return finish("outer_precision_stagnated", reason)
            if inner.status in ("domain_blocked", "feasibility_blocked", "feasibility_uncertain"):
                return finish("outer_" + inner.status, inner.reason)
            if not eligible:
                return finish("outer_inner_failed", inner.status + ": " + inner.reason)
            return finish("outer_no_progress" if reason == "NO_PROGRESS" else "outer_candidate_rejected", reason)
        accepted_count += 1
        current, GF, GD, GE = candidate, GFa, GDa, GEa
        small_steps = small_steps + 1 if relative <= config.step_tol else 0
        if GF <= threshold or accepted_count >= config.max_iterations or small_steps >= config.stall_steps:
            continue
        ref = current.u.copy()
        current = pair.rereference(current, ref)
        GE = GF
