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


def compare_runs(oracle_out, oracle_inner, res, rows, inner_rows):
    assert res.status == oracle_out.status
    assert len(rows) == len(oracle_out.history)
    for a, b in zip(oracle_out.history, rows):
        assert (a["reason"], a["inner_status"], a["inner_accepted_steps"]) == \
               (b["reason"], b["inner_status"], b["inner_accepted_steps"]), (a["outer_i"], a["reason"], b["reason"])
        for key in ("E_start", "F_start", "GF_before", "forcing_target"):
            assert close(a[key], b[key]), (a["outer_i"], key, a[key], b[key])
        if b["E_candidate"] is not None:
            for key in ("E_candidate", "F_candidate", "GF_candidate", "GD_candidate", "GE_candidate"):
                assert close(a[key], b[key]), (a["outer_i"], key, a[key], b[key])
    assert len(inner_rows) == len(oracle_inner)
    for a, b in zip(oracle_inner, inner_rows):
        where = (a["outer"], a["iteration"], a["attempt"])
        assert (a["iteration"], a["attempt"], a["reason"], a["accepted"], a["step_rule"]) == \
               (b["iteration"], b["attempt"], b["reason"], b["accepted"], b["step_rule"]), (where, a["reason"], b["reason"])
        for key in ("tau", "step_proposal", "energy_base", "energy_trial", "gradient_norm", "gradient_dot_direction"):
            assert close(a[key], b[key]), (where, key, a[key], b[key])


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
