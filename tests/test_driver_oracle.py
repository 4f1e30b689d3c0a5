"""The library's driver (seq_vp_3d_amd/solver.py) against the independent restatement of the reference's inner PG
solve and outer loop (oracle/driver.py), both on the CPU: every attempt's decision, step and trial energy, and every
outer row, on a pair whose line searches accept, backtrack and meet Jacobian violations."""
import numpy as np
import pytest
import torch

from cpu_pair import CpuPairState
from oracle import driver as od, q1 as Q, spline as osp
from seq_vp_3d_amd.problem import native_weights
from seq_vp_3d_amd.solver import PGConfig, SequentialConfig, solve_sequential
from seq_vp_3d_amd.synthetic import PairSpec, make_pair

REL = 1e-12


def _sampler(img, grid, pts):
    s = osp.build(img.numpy(), Q.Grid(grid.shape, grid.spacing))
    return torch.tensor(osp.sample(s, pts.numpy(), with_gradient=False)[0])


@pytest.fixture(scope="module")
def pair():
    spec = PairSpec((14, 15, 16), 6, (2., 4.), 3, .6, .01, 3)
    return spec, make_pair(spec, "cpu", _sampler)


def close(a, b, rel=REL):
    if a is None or b is None:
        return a is None and b is None
    return abs(a - b) <= rel * max(abs(a), abs(b), 1e-300)


def _deviation(a, b, scale):
    if a is None or b is None:
        return 0. if a is None and b is None else float("inf")
    return abs(a - b) / max(scale, 1e-300)


def compare_runs(oracle_out, oracle_inner, res, rows, inner_rows, rel=REL, late=None):
    """Every decision equal, every value within `rel` of its scale: the value's own magnitude, except for the
    outer E-gradient norm GE, a difference of image, proximity, soft and barrier terms of the size of the image
    norms GF and GD (a small GE carries their rounding), whose scale is the largest of the three norms. late =
    [(k, rel_k), ...] in increasing k: the values of outer steps k and later within rel_k instead. All deviations
    are gathered first, so that a failure names the worst entry of every key."""
    assert res.status == oracle_out.status
    assert len(rows) == len(oracle_out.history)
    worst, by_outer = {}, {}

    def check(key, where, a, b, scale=None):
        d = _deviation(a, b, max(abs(a), abs(b)) if scale is None and a is not None and b is not None else scale)
        if d > worst.get(key, (-1., None))[0]:
            worst[key] = (d, where, a, b)
        outer = where[0] if isinstance(where, tuple) else where
        by_outer[outer] = max(by_outer.get(outer, 0.), d)

    for a, b in zip(oracle_out.history, rows):
        assert (a["reason"], a["inner_status"], a["inner_accepted_steps"]) == \
               (b["reason"], b["inner_status"], b["inner_accepted_steps"]), (a["outer_i"], a["reason"], b["reason"])
        for key in ("E_start", "F_start", "GF_before", "forcing_target"):
            check(key, a["outer_i"], a[key], b[key])
        if b["E_candidate"] is not None:
            for key in ("E_candidate", "F_candidate", "GF_candidate", "GD_candidate"):
                check(key, a["outer_i"], a[key], b[key])
            norms = max(abs(a[k]) for k in ("GE_candidate", "GF_candidate", "GD_candidate"))
            check("GE_candidate", a["outer_i"], a["GE_candidate"], b["GE_candidate"], norms)
    assert len(inner_rows) == len(oracle_inner)
    for a, b in zip(oracle_inner, inner_rows):
        where = (a["outer"], a["iteration"], a["attempt"])
        assert (a["iteration"], a["attempt"], a["reason"], a["accepted"], a["step_rule"]) == \
               (b["iteration"], b["attempt"], b["reason"], b["accepted"], b["step_rule"]), (where, a["reason"], b["reason"])
        for key in ("tau", "step_proposal", "energy_base", "energy_trial", "gradient_norm", "gradient_dot_direction"):
            check(key, where, a[key], b[key])
    report = {k: (f"{d:.2e}", *rest) for k, (d, *rest) in sorted(worst.items())}
    print("worst relative deviations:", report, flush=True)
    print("worst deviation per outer step:", " ".join(f"{k}:{v:.1e}" for k, v in sorted(by_outer.items())),
          flush=True)
    if late is None:
        assert all(d <= rel for d, *_ in worst.values()), report
    else:
        def bound(o):
            return ([rel] + [r for k, r in late if o >= k])[-1]
        bad = {o: d for o, d in by_outer.items() if d > bound(o)}
        assert not bad, (bad, report)


@pytest.mark.parametrize("jd_step_global,initial_step", [(False, 1.), (True, 1.), (False, 10.)])
def test_library_driver_is_the_reference_flow(pair, jd_step_global, initial_step):
    spec, p = pair
    w = native_weights(spec.shape)
    theta0 = Q.restrict_free(.5 * p.displacement.numpy(), Q.Grid(spec.shape, (1., 1., 1.)))
    outer = 3
    # the oracle's restatement of the reference
    opair = od.Pair(spec.shape, (1., 1., 1.), w, p.fixed.numpy(), p.moving.numpy())
    ocfg = od.PGConfig(jd_step_global=jd_step_global, initial_step=initial_step)
    oracle_out = od.solve_sequential(opair, theta0, od.SequentialConfig(max_iterations=outer), ocfg)
    oracle_inner = oracle_out.inner_history
    # the library's driver on the oracle-backed CPU twin
    st = CpuPairState(spec.shape, (1., 1., 1.), w, p.fixed.numpy(), p.moving.numpy())
    st.set_state(theta0)
    rows, inner_rows = [], []
    res = solve_sequential(st, SequentialConfig(), PGConfig(jd_step_global=jd_step_global, initial_step=initial_step),
                           max_outer=outer, observer=rows.append, inner_observer=inner_rows.append)
    compare_runs(oracle_out, oracle_inner, res, rows, inner_rows)
    reasons = {r["reason"] for r in oracle_inner}
    if initial_step == 10.:
        # the long first steps fold the map: the witness path (re-prediction after a violation) is exercised
        assert "JACOBIAN_VIOLATION" in reasons
    np.testing.assert_allclose(st.theta.numpy(), oracle_out.state.theta, rtol=0,
                               atol=REL * np.abs(oracle_out.state.theta).max())
