"""Device-resident certified PG solver: the inner projected-gradient loop and the outer proximal loop.

Restates the decisions of the reference's native route -- solve_poisson_gradient with
ConservativePGConfig3D(resolution_exit=True) and the image_shift_ggn_equilibrated_jd_affine_v3 metric (T62 node
measure with the T69 face/interior extension always on; jd_curl/three_d/solvers/poisson_gradient.py:400-576,
1090-1646; docs/usage.md:84-93), with the T66 global root bound as an option (jd_step_global, off by default as in
ConservativePGConfig3D, poisson_gradient.py:143), and
solve_sequential (jd_curl/three_d/sequential.py:151-689) -- on a PairState whose arrays never leave HBM.
Only scalars cross PCIe. The first trial step of an inner iteration is the historical-witness prediction
(jd_step.py:177-239), capped by the T66 global closed27 root when jd_step_global is set. A trial is valid only
when its closed27 samples are above delta and the whole-cell certificate proves J > delta on every cell (GPU
edge hull and prescreen of every cell, the CPU interval Bernstein fallback for the cells they leave undecided;
an uncertain certificate makes the trial FEASIBILITY_UNCERTAIN, a witness JACOBIAN_VIOLATION, as
accelerated.py:1692-1729). Witnesses of rejected candidates feed the predictor.
"""
from dataclasses import dataclass, field
from time import perf_counter

import numpy as np
import torch

from .jd_step import JDStepMemory

EPS = float(np.finfo(float).eps)

# JD-affine metric policies (poisson_gradient.py:116-119, 480-490): the T62 node measure, the T69 measure with the
# face/interior extension once the pair has met an undetermined certificate, and T73 with the extension always on
# (the native default, usage.md:85).
T62_POLICY = "image_shift_ggn_equilibrated_jd_affine_v1"
T69_POLICY = "image_shift_ggn_equilibrated_jd_affine_v2"
T73_POLICY = "image_shift_ggn_equilibrated_jd_affine_v3"


@dataclass(frozen=True)
class PGConfig:
    """ConservativePGConfig3D defaults of the native level (poisson_gradient.py:23-235; usage.md:85-93)."""
    max_iterations: int = 5
    initial_step: float = 1.
    step_min: float = 1e-8
    step_max: float = 10.
    max_backtracks: int = 25
    c1: float = 1e-4
    step_tol: float = 1e-12
    stall_steps: int = 5
    jd_step_safety: float = .9
    resolution_exit: bool = True
    jd_step_global: bool = False         # T66 global closed27 root cap (poisson_gradient.py:143 default)
    shift_factor: float = 1.
    jd_affine_activation: float = .25
    metric_policy: str = T73_POLICY
    fp32_direction: bool = False
    fp32_screen: bool = False            # precision form: FP32 screening energies with FP64 rechecks


FP32_SCREEN_CHECKS = 3       # first screened trials per state also evaluated in FP64 and compared
FP32_SCREEN_RECHECK = 50     # then every 50th (accelerated.py:67-76, 2690-2728)
CRITICAL_GUARDS = 8.         # a decision within 8x the band is rechecked in FP64 (accelerated.py:3084-3095)


def screen_record(st):
    rec = getattr(st, "screen_record", None)
    if rec is None:
        rec = st.screen_record = dict(uses=0, checks=0, rechecks=0, disabled=False, mismatches=0, max_rel=0.)
    return rec


def screened_trial(st, tau, required):
    """One trial with the FP32 screen: band-checked decisions, FP64 recheck near the decision boundary,
    first-use and periodic FP64 comparisons that switch the screen off on a violation."""
    rec = screen_record(st)
    tr = st.trial(tau, screen=True)
    if not tr.ok or tr.unproven_cells:
        return tr
    n = rec["uses"]
    rec["uses"] += 1
    if n < FP32_SCREEN_CHECKS or (n + 1 - FP32_SCREEN_CHECKS) % FP32_SCREEN_RECHECK == 0:
        exact = st.trial_energies(screen=False)
        rec["checks"] += 1
        e, x = tr.energy, exact.energy
        gaps = [abs(e.jd - x.jd) / max(2. ** -24 * 64 * x.jd, 1e-300),
                abs(e.curl - x.curl) / max(2. ** -24 * 64 * x.curl, 1e-300),
                abs(e.barrier - x.barrier) / max(2. ** -24 * 4096 * x.barrier, 1e-300),
                abs(tr.total_difference - exact.total_difference) / max(tr.band, 1e-300),
                abs(tr.outer_difference - exact.outer_difference) / max(tr.band, 1e-300)]
        rel = max(0. if g == 0 else g for g in gaps)
        rec["max_rel"] = max(rec["max_rel"], min(rel, 1e300))
        if not rel <= 1.:
            rec["mismatches"] += 1
            rec["disabled"] = True
        return exact
    resolution = max(16 * EPS * max(st.energy.total, tr.energy.total), tr.band)
    actual = -tr.total_difference
    if abs(actual) <= CRITICAL_GUARDS * resolution or abs(actual - resolution - required) <= CRITICAL_GUARDS * resolution:
        rec["rechecks"] += 1
        return st.trial_energies(screen=False)
    return tr


@dataclass(frozen=True)
class SequentialConfig:
    """SequentialReuseConfig3D (sequential.py:28-56) with the native run's max_iterations."""
    max_iterations: int = 1000
    eta0: float = .01
    step_tol: float = 1e-12
    stall_steps: int = 5


@dataclass
class InnerResult:
    status: str
    reason: str
    accepted_steps: int
    gradient_norm: float | None
    trials: int
    history: list = field(default_factory=list)


@dataclass
class SequentialResult:
    status: str
    reason: str
    outer_accepted: int
    G0: float
    GF: float
    threshold: float
    history: list
    seconds: float
    energy: object = None
    GD: float | None = None
    GE: float | None = None
    native_mse: float | None = None     # MSE of the baseline (u = 0 with baseline_at_zero, else the start)
    final_mse: float | None = None
    mse_ratio: float | None = None

    @property
    def success(self):
        """Only the regularized gradient criterion is success (sequential.py:281)."""
        return self.status == "converged_regularized_gradient"


_WEIGHT_SUMS = {}


def reported_mse(grid, image_energy):
    """The reported MSE 2 D / sum q, the node-weight sum in long double (sequential.py:95-118), computed once per
    grid shape and spacing (the sum over the full node array is ~0.1 s at the native size)."""
    key = (tuple(grid.shape), tuple(float(h) for h in grid.spacing))
    total = _WEIGHT_SUMS.get(key)
    if total is None:
        total = _WEIGHT_SUMS[key] = np.sum(grid.node_weights(), dtype=np.longdouble)
    return float(2 * np.longdouble(image_energy) / total)


CERTIFICATE_FIELDS = ("status", "reason", "min_jd_lower_bound", "n_fast_cells", "n_prescreen_cells",
                      "n_fallback_cells", "n_boxes", "n_uncertain", "n_infeasible")


def _field(cert, key):
    """One field of a FeasibilityResult or of a dict of its fields (None for no certificate)."""
    if cert is None:
        return None
    return cert.get(key) if isinstance(cert, dict) else getattr(cert, key)


def certificate_record(cert):
    """The certificate fields a history row keeps (a FeasibilityResult or a dict of its fields; None)."""
    if cert is None:
        return None
    return {k: _field(cert, k) for k in CERTIFICATE_FIELDS}


def _protect(config, value):
    return min(config.step_max, max(config.step_min, value))


def step_proposal(st, config, previous, products=None):
    """Accepted-only BB2 with the reference's fallbacks (poisson_gradient.py:955-990). `products` are the
    (s.y, |s|, |y|, y.iy) of `previous` when already read."""
    if previous is None:
        return _protect(config, config.initial_step), "initial"
    fallback = _protect(config, 2 * min(previous["tau"], config.step_max / 2))
    sy, ns, ny, yiy = st.bb2_products(previous) if products is None else products
    if not all(np.isfinite(v) for v in (sy, ns, ny, yiy)):
        return fallback, "fallback_nonfinite_secant"
    if sy <= 0:
        return fallback, "fallback_nonpositive_curvature"
    cosine = sy / (ns * ny) if ns > 0 and ny > 0 else 0.
    if not np.isfinite(cosine) or cosine <= 1e-12:
        return fallback, "fallback_small_curvature"
    if yiy <= 0:
        return fallback, "fallback_nonpositive_metric_curvature"
    proposal = sy / yiy
    if not np.isfinite(proposal) or proposal <= 0:
        return fallback, "fallback_nonpositive_step"
    return _protect(config, proposal), "bb2"


def precision_stagnation_evidence(attempts, untruncated):
    """Every attempt of a search whose first step was the untruncated proposal is a valid comparison that the
    resolution cannot decide (energy_comparison.py:64-77)."""
    if not attempts or not untruncated:
        return False
    if any(r["reason"] != "ENERGY_RESOLUTION" or not r["trial_valid"] or r["accepted"]
           or r["energy_resolution"] is None for r in attempts):
        return False
    return all(r["actual_reduction"] is not None and np.isfinite(r["actual_reduction"])
               and np.isfinite(r["energy_resolution"]) and r["energy_resolution"] >= 0
               and abs(r["actual_reduction"]) <= r["energy_resolution"] for r in attempts)


def failed_search_status(attempts):
    """Status of a line search that accepted nothing, from the set of its reasons (poisson_gradient.py:1014-1025)."""
    reasons = {r["reason"] for r in attempts}
    if reasons == {"INVALID_DOMAIN"}:
        return "domain_blocked"
    if reasons == {"FEASIBILITY_UNCERTAIN"}:
        return "feasibility_uncertain"
    if reasons == {"JACOBIAN_VIOLATION"}:
        return "feasibility_blocked"
    if reasons and all(r.startswith("NONFINITE") for r in reasons):
        return "nonfinite_evaluation"
    return "line_search_failed"


def solve_poisson_gradient(st, target, config=PGConfig(), budget=None, observer=None, memory=None):
    """One frozen-reference inner solve on the accepted state of `st` (poisson_gradient.py:1090-1646). `memory` is
    the JD witness memory of the level (a fresh one when None, poisson_gradient.py:1130-1136)."""
    if memory is None:
        memory = JDStepMemory()
    memory.bind(st.grid)
    history = []
    accepted_steps, small_steps, trials = 0, 0, 0
    previous = None
    metric_ready = False
    G = None

    def finish(status, reason):
        return InnerResult(status, reason, accepted_steps, G, trials, history)

    ratio_pending = False       # the last accepted step's ratio is on the device, read with the next scalars
    prefetched = None           # the next iteration's direction scalars, enqueued at the acceptance
    chained = None              # the next iteration's first attempt, enqueued behind them (PairState.chain_launch)
    while True:
        scalars = None
        try:
            if accepted_steps >= config.max_iterations:
                # this gradient ends the solve (every exit below returns): take it from the joint outer
                # triple, which the outer loop then reuses (the reference's final_gradient,
                # accelerated.py:1451; sequential.py:288-338)
                st.outer_cache = st.outer_norms()
                G = st.outer_cache[2]
            elif metric_ready:
                # the gradient, the direction and its scalars are enqueued before the gradient checks (at the
                # acceptance that ended the last search, when there was one) and read with the gradient norm (and
                # the last step's ratio) in one synchronisation; when a check below ends the solve they are simply
                # not used
                if prefetched is not None:
                    scalars, prefetched = st.direction_scalars_read(prefetched), None
                else:
                    scalars = st.direction_scalars(previous, norm=True, ratio=ratio_pending,
                                                   fp32=config.fp32_direction, gradient=True)
                G = scalars[3]
            else:
                st.gradient()
                G = st.gradient_norm()
        except FloatingPointError as exc:
            return finish("nonfinite_evaluation", str(exc))
        if ratio_pending:
            relative = scalars[4] if scalars is not None else st.read_step_ratio()
            small_steps = small_steps + 1 if relative <= config.step_tol else 0
            ratio_pending = False
        if not np.isfinite(G):
            return finish("nonfinite_evaluation", "nonfinite gradient")
        if G <= target:
            return finish("converged_gradient", "GRADIENT_TARGET")
        if budget is not None and budget():
            return finish("work_budget_exhausted", "WALL_BUDGET")
        if small_steps >= config.stall_steps:
            return finish("stagnated", "SMALL_ACCEPTED_STEPS")
        if accepted_steps >= config.max_iterations:
            return finish("max_iterations", "ACCEPTED_STEP_BUDGET")
        if not metric_ready:
            if config.metric_policy not in (T62_POLICY, T69_POLICY, T73_POLICY):
                raise ValueError(f"unknown metric policy {config.metric_policy!r}")
            extend = config.metric_policy == T73_POLICY or (config.metric_policy == T69_POLICY
                                                             and memory.uncertain > 0)
            try:
                st.metric(shift_factor=config.shift_factor, activation=config.jd_affine_activation,
                          extension=extend)
            except FloatingPointError as exc:
                return finish("nonfinite_evaluation", str(exc))
            metric_ready = True
        if scalars is None:
            # the direction, the slope, the BB2 products and the domain bound come back together (the bound and
            # the products are not used when the slope ends the solve)
            scalars = st.direction_scalars(previous, fp32=config.fp32_direction)
        slope, products, bound = scalars[:3]
        if not np.isfinite(slope):
            return finish("nonfinite_evaluation", "PG directional derivative")
        if slope >= 0:
            return finish("non_descent_direction", "NON_DESCENT_DIRECTION")
        proposal, rule = step_proposal(st, config, previous, products)
        if bound <= 0:
            return finish("domain_blocked", "NO_POSITIVE_DOMAIN_STEP")
        tau = min(proposal, .9 * bound)
        # a chained first attempt is used only when its device step is this step, bit for bit
        chain, chained = chained, None
        if chain is not None and not st.chain_matches(chain, tau):
            chain = None
        # JD step prediction: the historical witnesses, then (T66) the global closed27 root on the same limit
        # (poisson_gradient.py:1356-1366, 708-735); the root's sweep is enqueued first and runs while the host
        # evaluates the witnesses (or already ran in the chain)
        pending_root = st.jd_bound_launch(tau) if config.jd_step_global and chain is None else None
        predicted, prediction = memory.predict(st.witness_coefficients, tau, config.jd_step_safety)
        root_info = None
        if config.jd_step_global:
            try:
                root_s, invalid, points = st.chain_root(chain) if chain is not None else st.jd_bound_read(pending_root)
            except FloatingPointError as exc:
                root_s, invalid, points = float("nan"), 0, 0
                root_info = dict(fallback=str(exc))
            if root_info is None:
                root = None if not np.isfinite(root_s) else root_s * tau
                root_info = dict(limit=tau, root=root, invalid_base_points=invalid, points=points)
                if root == 0:
                    root_info["fallback"] = "zero_root_below_resolution"
                elif root is not None:
                    predicted = min(predicted, config.jd_step_safety * root)
            prediction["global"] = root_info
        tau = predicted
        untruncated = tau == proposal
        exit_floor = 15 * EPS * st.energy.total if config.resolution_exit else None
        global_attempts = 0
        first_attempt = len(history)
        accepted = False
        for attempt in range(config.max_backtracks + 1):
            if budget is not None and budget():
                return finish("work_budget_exhausted", "WALL_BUDGET")
            trials += 1
            required = -config.c1 * tau * slope
            if config.fp32_screen and not screen_record(st)["disabled"]:
                tr = screened_trial(st, tau, required)
            elif attempt == 0 and chain is not None and st.chain_step(chain) == tau:
                tr = st.chain_trial(chain)
            else:
                tr = st.trial(tau)
            if tr.witness is not None:
                memory.witness(*tr.witness)
            memory.certificate(tr.certificate)
            actual = resolution = certified = None
            if not tr.ok:
                reason = tr.reason
            elif tr.unproven_cells:
                reason = "FEASIBILITY_UNCERTAIN"
            else:
                actual = -tr.total_difference
                resolution = 16 * EPS * max(st.energy.total, tr.energy.total)
                if tr.band is not None:
                    resolution = max(resolution, tr.band)
                certified = actual - resolution
                if not np.isfinite(actual):
                    reason = "NONFINITE_COMPARISON"
                else:
                    accepted = certified > 0 and certified >= required
                    reason = ("ACCEPTED" if accepted else "ENERGY_RESOLUTION" if abs(actual) <= resolution
                              else "NO_STRICT_DECREASE" if actual < 0 else "ARMIJO_REJECTED")
            exit_info = next_prediction = None
            if (config.resolution_exit and not accepted and not reason.startswith("NONFINITE")
                    and global_attempts + 1 <= config.max_backtracks):
                if reason == "JACOBIAN_VIOLATION":
                    # the memory now holds this trial's witness: the next step is also capped by its prediction
                    # (poisson_gradient.py:1503-1508)
                    repredicted, next_prediction = memory.predict(st.witness_coefficients, tau, config.jd_step_safety)
                    next_tau = min(.5 * tau, repredicted)
                else:
                    next_tau, next_prediction = .5 * tau, prediction
                decrease = next_tau * abs(slope)
                exit_info = dict(next_tau=next_tau, predicted_decrease=decrease, resolution_floor=exit_floor,
                                 stop=bool(decrease <= exit_floor))
            energy_base = st.energy.total
            if accepted:
                # the step ratio (small-step count) is read at the next synchronisation, before any check uses it;
                # the next iteration's gradient, direction and scalars are enqueued before the host records this
                # attempt, so that the device works while it does
                st.step_ratio_launch(st.theta_t, st.theta)
                ratio_pending = True
                previous = st.promote(tr.energy, keep_history=True)
                previous["tau"] = tau
                if metric_ready and accepted_steps + 1 < config.max_iterations:
                    prefetched = st.direction_scalars_launch(previous, norm=True, ratio=True,
                                                             fp32=config.fp32_direction, gradient=True)
                    # with an empty witness memory the next first attempt's step depends only on device values:
                    # it is enqueued too (the host checks it before use)
                    launch = getattr(st, "chain_launch", None)
                    if launch is not None and not config.fp32_screen and not memory.points:
                        chained = launch(tau, config.step_min, config.step_max, config.initial_step,
                                         config.jd_step_safety, root=config.jd_step_global)
            row = dict(iteration=accepted_steps, attempt=attempt + 1, accepted=bool(accepted), reason=reason, tau=tau,
                       step_proposal=proposal, step_rule=rule, untruncated=untruncated, domain_step_bound=bound,
                       gradient_norm=G, gradient_dot_direction=slope, trial_valid=bool(tr.ok and not tr.unproven_cells),
                       actual_reduction=actual,
                       energy_resolution=resolution, certified_reduction=certified, energy_base=energy_base,
                       energy_trial=tr.energy.total if tr.energy is not None else None,
                       unproven_cells=tr.unproven_cells, certificate=certificate_record(tr.certificate),
                       jd_prediction=prediction, resolution_exit=exit_info)
            history.append(row)
            if observer is not None:
                observer(row)
            if reason.startswith("NONFINITE"):
                return finish("nonfinite_evaluation", reason)
            if accepted:
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
                repredicted, prediction = memory.predict(st.witness_coefficients, tau, config.jd_step_safety)
                tau = min(.5 * tau, repredicted)
            else:
                tau *= .5
        if not accepted:
            attempts = history[first_attempt:]
            status = ("precision_stagnated" if precision_stagnation_evidence(attempts, untruncated)
                      else failed_search_status(attempts))
            return finish(status, attempts[-1]["reason"])


def solve_sequential(st, config=SequentialConfig(), inner_config=PGConfig(), budget=None, observer=None,
                     max_outer=None, baseline_at_zero=False, inner_observer=None):
    """Outer proximal loop (sequential.py:441-689): inner solve from the committed state, stable E and F
    comparisons against the start, commit, self-reference refresh. With baseline_at_zero (the native level of
    the four-level route, multiresolution.py:496-502) G0, its threshold and the baseline MSE are those of u = 0
    (sequential.py:396-419), whatever the start."""
    started = perf_counter()
    history = []
    memory = JDStepMemory()             # the level's witness memory, shared by its inner solves (sequential.py:214-215)
    memory.bind(st.grid)
    st.refresh_reference()
    GF, GD, GE = st.outer_norms()
    G0 = GF
    native = reported_mse(st.grid, st.energy.image)
    if baseline_at_zero and bool(torch.any(st.theta != 0)):
        # F = D + B does not depend on the reference, so the zero map's GF needs only its image and barrier terms
        start = st.snapshot()
        st.set_state(theta=torch.zeros_like(st.theta))
        G0 = st.outer_norms()[0]
        native = reported_mse(st.grid, st.energy.image)
        st.restore(start)
    threshold = 1e-8 + 1e-6 * G0
    accepted_count, small_steps = 0, 0
    limit = config.max_iterations if max_outer is None else min(config.max_iterations, max_outer)

    def finish(status, reason):
        final = reported_mse(st.grid, st.energy.image)
        return SequentialResult(status, reason, accepted_count, G0, GF, threshold, history,
                                perf_counter() - started, st.energy, GD, GE, native, final,
                                None if native == 0 else final / native)

    while True:
        if budget is not None and budget():
            return finish("outer_time_budget", "WALL_BUDGET")
        if GF <= threshold:
            return finish("converged_regularized_gradient", "Feasible F=D+B gradient criterion")
        if small_steps >= config.stall_steps:
            return finish("outer_stagnated", "Small accepted steps with nonzero outer gradient")
        if accepted_count >= limit:
            return finish("outer_max_iterations", "Outer budget exhausted; not convergence")
        tick = perf_counter()
        start = st.snapshot()
        st.outer_cache = None
        eta = config.eta0 if G0 == 0 else config.eta0 * np.sqrt(min(GF, G0) / G0)
        target = float(eta * GF)
        inner = solve_poisson_gradient(st, target, inner_config, budget=budget, memory=memory,
                                       observer=inner_observer)
        eligible = inner.status in ("converged_gradient", "max_iterations", "stagnated")
        total = fixed = None
        GFa = GDa = GEa = relative = None
        # an inner solve without an accepted step hands back the start itself: no progress, nothing to compare
        cert, same = False, True
        if inner.accepted_steps > 0:
            cached = getattr(st, "outer_cache", None)
            GFa, GDa, GEa = cached if cached is not None else st.outer_norms()
            relative = st.step_ratio(st.theta, start["theta"])
            dE, dF, gap = st.compare_to(start)
            resE = 16 * EPS * max(start["energy"].total, st.energy.total)
            resF = 16 * EPS * max(start["energy"].outer, st.energy.outer)
            total, fixed = (-dE, resE), (-dF, resF)
            cert = (-dE - resE) > 0 and (-dF - resF) > 0
            # theta equals the start's iff max|theta - start| (left by step_ratio) is 0: theta is an accepted
            # trial, whose energy (and so every entry) is finite, and -0.0 meets 0.0 in both tests
            same = st.step_max_diff == 0.0
        accepted = bool(eligible and cert and not same)
        reason = ("ACCEPTED" if accepted else "INNER_FAILED_NOT_COMMITTED" if not eligible
                  else "NO_PROGRESS" if same else "DESCENT_CERTIFICATE_FAILED")
        valid = inner.accepted_steps > 0
        se, ce = start["energy"], st.energy
        cc = st.certificate if valid else None
        candidate_mse = reported_mse(st.grid, ce.image) if valid else None
        row = dict(outer_i=len(history), E_start=se.total, E_candidate=ce.total if valid else None,
                   F_start=se.outer, F_candidate=ce.outer if valid else None,
                   D_start=se.image, D_candidate=ce.image if valid else None,
                   B_start=se.barrier, B_candidate=ce.barrier if valid else None,
                   penalty_proximity=ce.proximity if valid else None, penalty_jd=ce.jd if valid else None,
                   penalty_curl=ce.curl if valid else None,
                   actual_reduction_E=None if total is None else total[0],
                   resolution_E=None if total is None else total[1],
                   actual_reduction_F=None if fixed is None else fixed[0],
                   resolution_F=None if fixed is None else fixed[1],
                   GF_initial=G0, GF_threshold=threshold,
                   GF_before=GF, GF_candidate=GFa, GD_candidate=GDa, GE_candidate=GEa, forcing_eta=float(eta),
                   forcing_target=target, forcing_passed=bool(valid and GEa is not None and GEa <= target),
                   inner_status=inner.status, inner_reason=inner.reason,
                   inner_accepted_steps=inner.accepted_steps, inner_trials=inner.trials, outer_accepted=accepted,
                   reason=reason, step_ratio=relative,
                   candidate_min_jd_lower_bound=_field(cc, "min_jd_lower_bound"),
                   candidate_sampled_min_jd=_field(cc, "sampled_min_jd"),
                   candidate_mse_ratio=None if candidate_mse is None or native == 0 else candidate_mse / native,
                   seconds=perf_counter() - tick, elapsed_seconds=perf_counter() - started)
        time_limited = False
        if accepted and budget is not None and budget():
            # no commit once the budget has expired, checked right before the state transition (sequential.py:627-637)
            accepted, time_limited, reason = False, True, "WALL_BUDGET_NOT_COMMITTED"
            row.update(outer_accepted=False, reason=reason)
        committed = ce if accepted else se
        committed_mse = candidate_mse if accepted else reported_mse(st.grid, se.image)
        row.update(committed_mse_ratio=None if native == 0 else committed_mse / native,
                   committed_GF=GFa if accepted else GF, committed_E=committed.total, committed_F=committed.outer)
        history.append(row)
        if observer is not None:
            observer(row)
        if not accepted:
            st.restore(start)
            if time_limited:
                return finish("outer_time_budget", reason)
            if inner.status == "precision_stagnated":
                return finish("outer_precision_stagnated", reason)
            if inner.status in ("domain_blocked", "feasibility_blocked", "feasibility_uncertain"):
                return finish("outer_" + inner.status, inner.reason)
            if not eligible:
                return finish("outer_inner_failed", inner.status + ": " + inner.reason)
            return finish("outer_no_progress" if reason == "NO_PROGRESS" else "outer_candidate_rejected", reason)
        accepted_count += 1
        GF, GD, GE = GFa, GDa, GEa
        small_steps = small_steps + 1 if relative <= config.step_tol else 0
        if (GF <= threshold or accepted_count >= limit or small_steps >= config.stall_steps
                or (budget is not None and budget())):
            continue
        st.refresh_reference()
