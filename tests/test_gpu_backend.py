"""AmdBackend, the reference's `backend=` object (backends/accelerated.py:363; hybrid.py:51-286), on the device:
its methods against the oracle, then the oracle's restatement of the reference's solve_sequential /
solve_poisson_gradient (oracle/driver.py) driven only through those methods, step for step against the same flow on
the oracle's own arithmetic."""
import numpy as np
import pytest

from oracle import driver as od, q1 as Q, spline as osp

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

REL = 1e-12


def _pair():
    from seq_vp_3d_amd.problem import native_weights
    from seq_vp_3d_amd.synthetic import PairSpec, make_pair

    def sampler(img, grid, pts):
        s = osp.build(img.cpu().numpy(), Q.Grid(grid.shape, grid.spacing))
        return torch.tensor(osp.sample(s, pts.cpu().numpy(), with_gradient=False)[0], device=img.device)

    spec = PairSpec((14, 15, 16), 6, (2., 4.), 3, .6, .01, 3)
    pair = make_pair(spec, "cpu", sampler)
    return spec, pair, native_weights(spec.shape)


def close(a, b, rel=REL):
    return abs(a - b) <= rel * max(abs(a), abs(b), 1e-300)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def test_backend_methods_against_the_oracle(gpu_device):
    from backend_pair import BackendPair
    from seq_vp_3d_amd.backend import AmdBackend, EnergyReduction, InvalidTrial, build_linearization
    print("case backend methods 14x15x16", flush=True)
    spec, pair, w = _pair()
    shape, spacing = spec.shape, (1., 1., 1.)
    backend = AmdBackend(gpu_device)
    bp = BackendPair(backend, shape, spacing, w, pair.fixed.numpy(), pair.moving.numpy())
    op = od.Pair(shape, spacing, w, pair.fixed.numpy(), pair.moving.numpy())
    g = op.grid
    theta = Q.restrict_free(.5 * pair.displacement.numpy(), g)
    ref = Q.inject_free(.8 * theta, g)
    want = op.evaluate(theta, ref)
    prob = bp.problem(ref)
    trial = backend.evaluate_trial(prob, theta)
    for key in want.energy._fields:
        assert close(getattr(trial.energy, key), getattr(want.energy, key)), key
    assert trial.certificate.status == "certified_feasible" and np.array_equal(trial.u, want.u)
    acc = backend.promote_trial(prob, trial)
    assert rel(acc.image_grad, want.image_grad) <= REL and rel(acc.warped, want.warped) <= REL
    lin = build_linearization(prob, acc)
    assert rel(backend.gradient(lin, 0.), op.gradient(want, ref)) <= REL
    for a, b in zip(backend.outer_norms(lin, g.omega), op.outer_norms(want, ref)):
        assert close(a, b)
    diag = backend.diagonal_launch(lin).result()
    m = op.model
    wdiag = Q.restrict_free(m.ggn_diagonal_nodes(m.ggn_diagonal_base(want.image_grad), want.u), g)
    assert rel(diag, wdiag) <= REL
    stats = {}
    node = backend.node_min_jd(lin, extension_activation=.25, stats=stats)
    from oracle import jd_affine as oja
    wnode, wstats = oja.node_min_jd(want.u, g, op.D, extension_activation=.25)
    np.testing.assert_allclose(node, wnode, rtol=1e-12, atol=0)
    if wstats["candidates"]:
        assert stats["candidates"] == wstats["candidates"]
    p = np.random.default_rng(5).normal(size=theta.size) * .3
    for limit in (.5, 4.):
        s, inv, pts = backend.global_jd_root(lin, p, limit)
        ws, winv, wpts = op.global_root(want, p, limit)
        assert (inv, pts) == (winv, wpts) and (np.isinf(ws) and np.isinf(s) or close(s, ws))
    # the Poisson protocol and a stable comparison
    from seq_vp_3d_amd.poisson import build_poisson_inverse
    from oracle import poisson as opo
    P = backend.poisson(build_poisson_inverse(bp.rgrid, a=.2, b=.03))
    scale = np.random.default_rng(6).uniform(.5, 2., size=theta.size)
    d, ok = P.apply_equilibrated(P.prepare_equilibrated(scale, p))
    assert ok is True and rel(d, -(scale * opo.apply(scale * p, opo.denominator(g, .2, .03)))) <= REL
    assert rel(P.apply(p), opo.apply(p, opo.denominator(g, .2, .03))) <= REL
    other = backend.promote_trial(prob, backend.evaluate_trial(prob, .9 * theta))
    wother = op.evaluate(.9 * theta, ref)
    for outer_only in (False, True):
        r = backend.stable_energy_reduction(prob, acc, other, outer_only=outer_only)
        actual, resolution, certified = op.reduction(want, wother, ref, outer_only=outer_only)
        assert isinstance(r, EnergyReduction) and close(r.actual, actual, 1e-10) and close(r.resolution, resolution)
    E, F = backend.stable_energy_reductions(prob, acc, other)
    assert E == backend.stable_energy_reduction(prob, acc, other)
    assert F == backend.stable_energy_reduction(prob, acc, other, outer_only=True)
    # refusals come back as the reference's InvalidTrial
    bad = theta.copy()
    bad[3] = np.nan
    assert backend.evaluate_trial(prob, bad).reason == "NONFINITE_STATE"
    far = theta + 50.
    out = backend.evaluate_trial(prob, far)
    assert isinstance(out, InvalidTrial) and out.reason == "INVALID_DOMAIN"


def _same_flow(a, b):
    assert a.status == b.status and len(a.history) == len(b.history)
    for x, y in zip(a.history, b.history):
        assert (x["reason"], x["inner_status"], x["inner_accepted_steps"]) == \
               (y["reason"], y["inner_status"], y["inner_accepted_steps"])
        for key in ("E_start", "E_candidate", "GF_candidate", "GE_candidate", "actual_reduction_E"):
            assert close(x[key], y[key], 1e-10 if key == "actual_reduction_E" else REL), (x["outer_i"], key)
    assert len(a.inner_history) == len(b.inner_history)
    for x, y in zip(a.inner_history, b.inner_history):
        where = (x["outer"], x["iteration"], x["attempt"])
        assert (x["reason"], x["accepted"], x["step_rule"]) == (y["reason"], y["accepted"], y["step_rule"]), where
        for key in ("tau", "energy_base", "gradient_norm", "gradient_dot_direction"):
            assert close(x[key], y[key]), (where, key, x[key], y[key])


@pytest.mark.parametrize("jd_step_global,initial_step", [(True, 1.), (False, 10.)])
def test_reference_flow_on_the_backend(gpu_device, jd_step_global, initial_step):
    from backend_pair import BackendPair
    from seq_vp_3d_amd.backend import AmdBackend
    print("case reference flow on AmdBackend", jd_step_global, initial_step, flush=True)
    spec, pair, w = _pair()
    shape, spacing = spec.shape, (1., 1., 1.)
    theta0 = Q.restrict_free(.5 * pair.displacement.numpy(), Q.Grid(shape, spacing))
    cfg = od.PGConfig(jd_step_global=jd_step_global, initial_step=initial_step)
    outer = od.SequentialConfig(max_iterations=3)
    want = od.solve_sequential(od.Pair(shape, spacing, w, pair.fixed.numpy(), pair.moving.numpy()), theta0, outer, cfg)
    bp = BackendPair(AmdBackend(gpu_device), shape, spacing, w, pair.fixed.numpy(), pair.moving.numpy())
    got = od.solve_sequential(bp, theta0, outer, cfg)
    _same_flow(want, got)
    np.testing.assert_allclose(got.state.theta, want.state.theta, rtol=0, atol=REL * np.abs(want.state.theta).max())
    used = {"evaluate_trial", "promote_trial", "gradient", "outer_norms", "poisson", "diagonal", "node_min_jd",
            "apply_equilibrated", "stable_energy_reduction"} | ({"global_jd_root"} if jd_step_global else set())
    assert used <= set(bp.calls), sorted(used - set(bp.calls))


def _classify(actual, resolution):
    """accept / within the resolution / increase, as the inner loop reads a reduction (energy_comparison.py:117-124)."""
    return "accept" if actual - resolution >= 0 else "resolution" if abs(actual) <= resolution else "increase"


def test_near_margin_decisions_match_the_longdouble_comparison(gpu_device):
    """The reference sends a device energy difference within 8 x 4096 eps of the decision margin to its CPU long
    double comparison (accelerated.py:3084-3111). The device path here decides every trial on its double-double
    sums; on trials placed around both margins (actual = resolution and actual = 0) its reductions equal the
    oracle's long double reductions of the same states to a thousandth of the resolution, and every decision is
    the same."""
    from backend_pair import BackendPair
    from seq_vp_3d_amd.backend import AmdBackend
    print("case near-margin reductions 14x15x16", flush=True)
    spec, pair, w = _pair()
    shape, spacing = spec.shape, (1., 1., 1.)
    bp = BackendPair(AmdBackend(gpu_device), shape, spacing, w, pair.fixed.numpy(), pair.moving.numpy())
    op = od.Pair(shape, spacing, w, pair.fixed.numpy(), pair.moving.numpy())
    g = op.grid
    theta = Q.restrict_free(.5 * pair.displacement.numpy(), g)
    ref = Q.inject_free(.8 * theta, g)
    base = bp.evaluate(theta, ref)
    grad = bp.gradient(base, ref)
    p = -grad / np.linalg.norm(grad)
    slope = float(np.linalg.norm(grad))                    # actual(h) = h * slope + O(h^2) along p
    h0 = 16 * od.EPS * od.total(base.energy) / slope       # the step where actual reaches the resolution
    near = {"accept": 0, "resolution": 0, "increase": 0}
    for f in (-4., -1., -.5, 1e-3, .25, .5, .9, .97, 1.03, 1.1, 2., 4., 16.):
        trial = bp.evaluate(theta + f * h0 * p, ref)
        assert isinstance(trial, od.State)
        for outer_only in (False, True):
            actual, resolution, certified = bp.reduction(base, trial, ref, outer_only=outer_only)
            want = od.Pair.reduction(op, base, trial, ref, outer_only=outer_only)
            assert resolution == want[1]
            assert abs(actual - want[0]) <= 1e-3 * want[1], (f, outer_only, actual, want[0], want[1])
            kind = _classify(want[0], want[1])
            assert _classify(actual, resolution) == kind, (f, outer_only, actual, want)
            if not outer_only and min(abs(want[0] - want[1]), abs(want[0])) <= 8 * want[1]:
                near[kind] += 1
    assert all(v >= 1 for v in near.values()), near
