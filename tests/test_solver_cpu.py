"""The device solver's decision logic (seq_vp_3d_amd/solver.py) run on the oracle-backed CPU twin."""
import numpy as np
import pytest
import torch

from cpu_pair import CpuPairState
from oracle import q1 as Q, spline as osp
from seq_vp_3d_amd.problem import native_weights
from seq_vp_3d_amd.solver import PGConfig, SequentialConfig, solve_poisson_gradient, solve_sequential, step_proposal
from seq_vp_3d_amd.synthetic import PairSpec, make_pair


def _sampler(img, grid, pts):
    s = osp.build(img.numpy(), Q.Grid(grid.shape, grid.spacing))
    return torch.tensor(osp.sample(s, pts.numpy(), with_gradient=False)[0])


@pytest.fixture(scope="module")
def pair():
    spec = PairSpec((14, 15, 16), 6, (2., 4.), 3, .6, .01, 3)
    return spec, make_pair(spec, "cpu", _sampler)


def _state(spec, pair):
    st = CpuPairState(spec.shape, (1., 1., 1.), native_weights(spec.shape), pair.fixed.numpy(), pair.moving.numpy())
    st.set_state(Q.restrict_free(.5 * pair.displacement.numpy(), st.grid))
    return st


def test_outer_loop_commits_and_decreases_energy(pair):
    spec, p = pair
    st = _state(spec, p)
    rows = []
    res = solve_sequential(st, SequentialConfig(), PGConfig(), max_outer=2, observer=rows.append)
    assert res.outer_accepted == 2 and res.status == "outer_max_iterations"
    for r in rows:
        assert r["reason"] == "ACCEPTED" and r["actual_reduction_E"] > r["resolution_E"]
        assert r["actual_reduction_F"] > r["resolution_F"] and r["inner_accepted_steps"] >= 1
    assert res.GF < res.G0


def test_inner_attempt_rows_follow_the_acceptance_rule(pair):
    spec, p = pair
    st = _state(spec, p)
    st.refresh_reference()
    rows = []
    inner = solve_poisson_gradient(st, 0., PGConfig(max_iterations=3), observer=rows.append)
    assert inner.accepted_steps == 3 and inner.status == "max_iterations"
    for r in rows:
        if r["accepted"]:
            assert r["certified_reduction"] > 0 and r["certified_reduction"] >= -1e-4 * r["tau"] * r["gradient_dot_direction"]
        assert r["tau"] <= .9 * r["domain_step_bound"] + 1e-300
        root = r["jd_prediction"].get("global")
        if root and root["root"] is not None:
            assert r["tau"] <= .9 * root["root"] * (1 + 1e-15) or r["tau"] < root["limit"]
    assert rows[0]["step_rule"] == "initial" and all(r["step_rule"] != "initial" for r in rows[1:])


def test_bb2_fallbacks():
    class S:
        def __init__(self, vals):
            self.vals = vals

        def bb2_products(self, prev):
            return self.vals

    cfg = PGConfig()
    prev = dict(tau=.3)
    assert step_proposal(S((1., 1., 1., 2.)), cfg, prev) == (.5, "bb2")
    assert step_proposal(S((-1., 1., 1., 2.)), cfg, prev) == (.6, "fallback_nonpositive_curvature")
    assert step_proposal(S((1e-20, 1., 1., 2.)), cfg, prev) == (.6, "fallback_small_curvature")
    assert step_proposal(S((1., 1., 1., -2.)), cfg, prev) == (.6, "fallback_nonpositive_metric_curvature")
    assert step_proposal(S((1., 1., 1., 1e-12)), cfg, prev) == (10., "bb2")
    assert step_proposal(S((1., 1., 1., 1.)), cfg, None) == (1., "initial")


def test_zero_baseline_and_reported_mse(pair):
    """baseline_at_zero: G0 and the baseline MSE are the zero map's whatever the start (sequential.py:396-419);
    the MSE is 2 D / sum q, here checked against sum q (W - F)^2 / sum q of the zero map, where W is the spline
    of the moving image at the nodes (= the moving image)."""
    from seq_vp_3d_amd.problem import make_grid_3d
    spec, p = pair
    zero = CpuPairState(spec.shape, (1., 1., 1.), native_weights(spec.shape), p.fixed.numpy(), p.moving.numpy())
    zero.refresh_reference()
    g0_zero = zero.outer_norms()[0]
    q = make_grid_3d(nx=spec.shape[2], ny=spec.shape[1], nz=spec.shape[0], hx=1., hy=1., hz=1.).node_weights()
    mse_zero = float(np.sum(q * (p.moving.numpy() - p.fixed.numpy()) ** 2) / np.sum(q))
    st = _state(spec, p)
    res_start = solve_sequential(st, SequentialConfig(), PGConfig(), max_outer=0)
    st = _state(spec, p)
    res_zero = solve_sequential(st, SequentialConfig(), PGConfig(), max_outer=0, baseline_at_zero=True)
    assert res_zero.G0 == g0_zero and res_start.G0 != g0_zero
    assert res_zero.threshold == 1e-8 + 1e-6 * g0_zero
    assert abs(res_zero.native_mse - mse_zero) <= 1e-9 * mse_zero
    assert res_start.native_mse == res_start.final_mse != res_zero.native_mse      # the start is not u = 0
    assert res_zero.mse_ratio == res_zero.final_mse / res_zero.native_mse
    # the state is back where it started
    np.testing.assert_array_equal(st.theta.numpy(), _state(spec, p).theta.numpy())


class _Scripted:
    """A state whose trials are scripted: the line search meets a Jacobian violation at the first step, and the
    witness cubic of the memory has its first root at `root` (c0 + c1 s with c0 = 1, c1 = -1/root)."""

    def __init__(self, root, outcomes):
        from seq_vp_3d_amd.hotpath import Energies
        self.grid = type("G", (), dict(shape=(4, 4, 4), spacing=(1., 1., 1.), n_cells=27))()
        self.energy = Energies(1., 0., 0., 0., 0.)
        self.root, self.outcomes, self.taus = root, list(outcomes), []
        self.theta = self.theta_t = torch.zeros(24, dtype=torch.float64)

    def gradient(self):
        pass

    def gradient_norm(self):
        return 1.

    def outer_norms(self):
        return 1., 1., 1.

    def metric(self, **kw):
        pass

    def direction(self, fp32=False):
        pass

    def direction_scalars(self, previous, norm=False, ratio=False, fp32=None, gradient=False):
        return -1., None, 1e300, 1., None

    def witness_coefficients(self, cells, points):
        return np.tile([1., -1. / self.root, 0., 0.], (len(cells), 1))

    def trial(self, tau):
        from seq_vp_3d_amd.hotpath import Energies, TrialResult
        self.taus.append(tau)
        kind = self.outcomes.pop(0)
        if kind == "fold":
            return TrialResult(False, "JACOBIAN_VIOLATION", None, witness=(5, (.5, .5, .5)))
        return TrialResult(True, "", Energies(1. - tau, 0., 0., 0., 0.), -tau, -tau)

    def step_ratio_launch(self, a, b):
        pass

    def read_step_ratio(self):
        return 1.

    def promote(self, energy, keep_history=False):
        self.energy = energy
        return {}


@pytest.mark.parametrize("resolution_exit", [True, False])
def test_violation_repredicts_the_next_step(resolution_exit):
    """After a JACOBIAN_VIOLATION the memory holds the trial's witness and the next step is
    min(tau / 2, safety * its first root) (poisson_gradient.py:1503-1508, 1608-1610), not a plain halving."""
    from seq_vp_3d_amd.jd_step import JDStepMemory
    st = _Scripted(root=.1, outcomes=["fold", "ok"])
    mem = JDStepMemory()
    mem.bind(st.grid)
    cfg = PGConfig(max_iterations=1, initial_step=1., resolution_exit=resolution_exit)
    inner = solve_poisson_gradient(st, 0., cfg, memory=mem)
    assert inner.accepted_steps == 1
    assert st.taus[0] == 1. and st.taus[1] == pytest.approx(.9 * .1, rel=1e-15)
    # with the root beyond tau / 2 the halving binds
    st = _Scripted(root=.9, outcomes=["fold", "ok"])
    mem = JDStepMemory()
    mem.bind(st.grid)
    solve_poisson_gradient(st, 0., cfg, memory=mem)
    assert st.taus[1] == .5


def test_failed_search_status_follows_the_reasons():
    from seq_vp_3d_amd.solver import failed_search_status, precision_stagnation_evidence
    row = lambda reason, **kw: dict(dict(reason=reason, trial_valid=True, accepted=False, energy_resolution=1e-12,
                                         actual_reduction=5e-13), **kw)
    assert failed_search_status([row("JACOBIAN_VIOLATION")] * 3) == "feasibility_blocked"
    assert failed_search_status([row("INVALID_DOMAIN")]) == "domain_blocked"
    assert failed_search_status([row("FEASIBILITY_UNCERTAIN")] * 2) == "feasibility_uncertain"
    assert failed_search_status([row("ARMIJO_REJECTED"), row("JACOBIAN_VIOLATION")]) == "line_search_failed"
    rows = [row("ENERGY_RESOLUTION")] * 3
    assert precision_stagnation_evidence(rows, True)
    assert not precision_stagnation_evidence(rows, False)             # a truncated first step is no evidence
    assert not precision_stagnation_evidence(rows + [row("NO_STRICT_DECREASE")], True)
    assert not precision_stagnation_evidence([row("ENERGY_RESOLUTION", actual_reduction=2e-12)], True)
