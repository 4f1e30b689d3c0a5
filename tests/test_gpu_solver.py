"""The device solver against the oracle's restatement of the reference's driver and against the oracle-backed CPU
twin, step by step (same decisions, same tau)."""
import numpy as np
import pytest

from cpu_pair import CpuPairState
from oracle import q1 as Q, spline as osp

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.mark.parametrize("jd_step_global,initial_step", [(False, 1.), (True, 1.), (False, 10.)])
def test_device_solver_matches_the_reference_flow(gpu_device, jd_step_global, initial_step):
    """The device driver on PairState against the oracle's independent restatement of the reference's inner PG
    solve and outer loop (oracle/driver.py): every attempt's decision, tau and trial energy, and every outer row, to
    1e-12 relative, over 4 outer steps."""
    from test_driver_oracle import compare_runs
    from oracle import driver as od
    from seq_vp_3d_amd.hotpath import PairState
    from seq_vp_3d_amd.problem import Q1Problem, native_weights
    from seq_vp_3d_amd.solver import PGConfig, SequentialConfig, solve_sequential
    from seq_vp_3d_amd.synthetic import PairSpec, make_pair
    print("case solver vs driver oracle 14x15x16", jd_step_global, initial_step, flush=True)

    def sampler(img, grid, pts):
        s = osp.build(img.cpu().numpy(), Q.Grid(grid.shape, grid.spacing))
        return torch.tensor(osp.sample(s, pts.cpu().numpy(), with_gradient=False)[0], device=img.device)

    spec = PairSpec((14, 15, 16), 6, (2., 4.), 3, .6, .01, 3)
    pair = make_pair(spec, "cpu", sampler)
    w = native_weights(spec.shape)
    theta0 = Q.restrict_free(.5 * pair.displacement.numpy(), Q.Grid(spec.shape, (1., 1., 1.)))
    outer = 4
    opair = od.Pair(spec.shape, (1., 1., 1.), w, pair.fixed.numpy(), pair.moving.numpy())
    want = od.solve_sequential(opair, theta0, od.SequentialConfig(max_iterations=outer),
                               od.PGConfig(jd_step_global=jd_step_global, initial_step=initial_step))
    gpu = PairState(Q1Problem(pair.grid, w, 0., .05), pair.fixed, pair.moving, gpu_device)
    gpu.set_state(theta=torch.tensor(theta0, device=gpu_device))
    rows, inner_rows = [], []
    res = solve_sequential(gpu, SequentialConfig(), PGConfig(jd_step_global=jd_step_global, initial_step=initial_step),
                           max_outer=outer, observer=rows.append, inner_observer=inner_rows.append)
    compare_runs(want, want.inner_history, res, rows, inner_rows)
    np.testing.assert_allclose(gpu.theta.cpu().numpy(), want.state.theta, rtol=0,
                               atol=1e-12 * np.abs(want.state.theta).max())


def test_fp32_screen_keeps_the_fp64_decisions(gpu_device):
    from seq_vp_3d_amd.hotpath import PairState
    from seq_vp_3d_amd.problem import Q1Problem, native_weights
    from seq_vp_3d_amd.solver import PGConfig, SequentialConfig, solve_sequential
    from seq_vp_3d_amd.synthetic import PairSpec, make_pair
    print("case fp32 screen driver 20x22x24", flush=True)

    def sampler(img, grid, pts):
        s = osp.build(img.cpu().numpy(), Q.Grid(grid.shape, grid.spacing))
        return torch.tensor(osp.sample(s, pts.cpu().numpy(), with_gradient=False)[0], device=img.device)

    spec = PairSpec((20, 22, 24), 8, (2., 5.), 3, .6, .01, 5)
    pair = make_pair(spec, "cpu", sampler)
    w = native_weights(spec.shape)
    runs = []
    for screen in (False, True):
        st = PairState(Q1Problem(pair.grid, w, 0., .05), pair.fixed, pair.moving, gpu_device)
        st.set_state(theta=torch.zeros(st.nfree, dtype=torch.float64, device=gpu_device))
        rows = []
        res = solve_sequential(st, SequentialConfig(), PGConfig(fp32_screen=screen), max_outer=3, observer=rows.append)
        runs.append((res, rows, st))
    (r64, rows64, st64), (r32, rows32, st32) = runs
    assert r64.status == r32.status and r64.outer_accepted == r32.outer_accepted
    assert [(a["reason"], a["inner_accepted_steps"], a["inner_trials"]) for a in rows64] == \
           [(a["reason"], a["inner_accepted_steps"], a["inner_trials"]) for a in rows32]
    assert torch.equal(st64.theta, st32.theta)
    rec = st32.screen_record
    assert rec["uses"] > 0 and rec["checks"] >= 1 and not rec["disabled"] and rec["max_rel"] <= 1.


@pytest.mark.parametrize("initial_step,jd_step_global", [(1., True), (10., True), (1., False)])
def test_chained_first_attempts_are_the_host_path(gpu_device, initial_step, jd_step_global):
    """PairState enqueues each next line search's first attempt on the device at the acceptance (step from the
    direction scalars, with the T66 bound on also the root and the capped step, then the trial); the driver uses it
    only when the device step is its own step bit for bit. Every history row (inner and outer) and the final state
    equal those of the host path (chain off); with initial step 10 the BB2 fallbacks and the root cap bind more
    often."""
    from seq_vp_3d_amd.hotpath import PairState
    from seq_vp_3d_amd.problem import Q1Problem, native_weights
    from seq_vp_3d_amd.solver import PGConfig, SequentialConfig, solve_sequential
    from seq_vp_3d_amd.synthetic import PairSpec, make_pair
    print("case chained first attempts 20x22x24", initial_step, flush=True)

    def sampler(img, grid, pts):
        s = osp.build(img.cpu().numpy(), Q.Grid(grid.shape, grid.spacing))
        return torch.tensor(osp.sample(s, pts.cpu().numpy(), with_gradient=False)[0], device=img.device)

    spec = PairSpec((20, 22, 24), 8, (2., 5.), 3, .6, .01, 5)
    pair = make_pair(spec, "cpu", sampler)
    w = native_weights(spec.shape)
    runs = []
    for chain in (False, True):
        st = PairState(Q1Problem(pair.grid, w, 0., .05), pair.fixed, pair.moving, gpu_device)
        st.chain = chain
        st.set_state(theta=torch.zeros(st.nfree, dtype=torch.float64, device=gpu_device))
        rows, inner = [], []
        res = solve_sequential(st, SequentialConfig(),
                               PGConfig(jd_step_global=jd_step_global, initial_step=initial_step),
                               max_outer=3, observer=rows.append, inner_observer=inner.append)
        runs.append((res, rows, inner, st))
    (r0, rows0, inner0, st0), (r1, rows1, inner1, st1) = runs
    assert st0.chained_trials == 0 and st1.chained_trials > 0
    assert r0.status == r1.status and r0.outer_accepted == r1.outer_accepted
    skip = {"seconds", "elapsed_seconds"}
    assert len(inner0) == len(inner1)
    for a, b in zip(inner0, inner1):
        assert {k: v for k, v in a.items() if k not in skip} == {k: v for k, v in b.items() if k not in skip}
    for a, b in zip(rows0, rows1):
        assert {k: v for k, v in a.items() if k not in skip} == {k: v for k, v in b.items() if k not in skip}
    assert torch.equal(st0.theta, st1.theta) and torch.equal(st0.u, st1.u)


@pytest.mark.parametrize("jd_step_global", [False, True])
def test_witness_memory_and_prediction_match_cpu_twin(gpu_device, jd_step_global):
    """A first step far too long folds the map: the device nominates the same confirmed witness, the memory
    predicts the same steps, and every attempt makes the same decision with the same tau."""
    from seq_vp_3d_amd.hotpath import PairState
    from seq_vp_3d_amd.jd_step import JDStepMemory
    from seq_vp_3d_amd.problem import Q1Problem, native_weights
    from seq_vp_3d_amd.solver import PGConfig, solve_poisson_gradient
    from seq_vp_3d_amd.synthetic import PairSpec, make_pair
    print(f"case witness memory jd_step_global={jd_step_global}", flush=True)

    def sampler(img, grid, pts):
        s = osp.build(img.cpu().numpy(), Q.Grid(grid.shape, grid.spacing))
        return torch.tensor(osp.sample(s, pts.cpu().numpy(), with_gradient=False)[0], device=img.device)

    spec = PairSpec((14, 15, 16), 6, (2., 4.), 3, .6, .01, 3)
    pair = make_pair(spec, "cpu", sampler)
    w = native_weights(spec.shape)
    config = PGConfig(max_iterations=5, jd_step_global=jd_step_global, initial_step=10.)
    cpu = CpuPairState(spec.shape, (1., 1., 1.), w, pair.fixed.numpy(), pair.moving.numpy())
    theta0 = Q.restrict_free(.5 * pair.displacement.numpy(), cpu.grid)
    cpu.set_state(theta0)
    gpu = PairState(Q1Problem(pair.grid, w, 0., .05), pair.fixed, pair.moving, gpu_device)
    gpu.set_state(theta=torch.tensor(theta0, device=gpu_device))
    runs = []
    for st in (cpu, gpu):
        st.refresh_reference()
        rows, mem = [], JDStepMemory()
        res = solve_poisson_gradient(st, 0., config, observer=rows.append, memory=mem)
        runs.append((res, rows, mem))
    (rc, rows_c, mem_c), (rg, rows_g, mem_g) = runs
    assert rc.status == rg.status and rc.accepted_steps == rg.accepted_steps
    assert len(rows_c) == len(rows_g)
    for a, b in zip(rows_c, rows_g):
        assert a["reason"] == b["reason"]
        assert abs(a["tau"] - b["tau"]) <= 1e-12 * a["tau"]
        pa, pb = a["jd_prediction"], b["jd_prediction"]
        assert pa["witnesses"] == pb["witnesses"] and pa["points"] == pb["points"]
        assert (pa["root"] is None) == (pb["root"] is None)
        if pa["root"] is not None:
            assert abs(pa["root"] - pb["root"]) <= 1e-12 * pa["root"]
    if not jd_step_global:
        assert any(r["reason"] == "JACOBIAN_VIOLATION" for r in rows_g)
        assert len(mem_g.points) >= 1
    assert list(mem_c.points) == list(mem_g.points)


def test_root_launch_then_witness_cubics_then_read(gpu_device):
    """The solver enqueues the T66 root sweep, evaluates the witness cubics (which write the direction's node field
    on the same stream) and only then reads the root: the value is the synchronous root's, bit for bit."""
    from seq_vp_3d_amd.hotpath import PairState
    from seq_vp_3d_amd.jd_step import JDStepMemory
    from seq_vp_3d_amd.problem import Q1Problem, native_weights
    from seq_vp_3d_amd.synthetic import PairSpec, make_pair
    print("case root launch ordering", flush=True)

    def sampler(img, grid, pts):
        s = osp.build(img.cpu().numpy(), Q.Grid(grid.shape, grid.spacing))
        return torch.tensor(osp.sample(s, pts.cpu().numpy(), with_gradient=False)[0], device=img.device)

    spec = PairSpec((14, 15, 16), 6, (2., 4.), 3, .6, .01, 3)
    pair = make_pair(spec, "cpu", sampler)
    w = native_weights(spec.shape)
    st = PairState(Q1Problem(pair.grid, w, 0., .05), pair.fixed, pair.moving, gpu_device)
    st.set_state(theta=torch.tensor(Q.restrict_free(.5 * pair.displacement.numpy(), Q.Grid(spec.shape, (1., 1., 1.))),
                                    device=gpu_device))
    st.refresh_reference()
    st.gradient()
    st.metric(shift_factor=1., activation=.25, extension=True)
    st.direction()
    mem = JDStepMemory()
    mem.bind(pair.grid)
    mem.witness(40, (.3, .6, .2))
    mem.witness(700, (.5, .5, .5))
    for limit in (.5, 3., 20.):
        want = st.jd_bound(limit)
        pending = st.jd_bound_launch(limit)
        mem.predict(st.witness_coefficients, limit, .9)
        assert st.jd_bound_read(pending) == want, limit


def _decode(v):
    return float.fromhex(v["hex"]) if isinstance(v, dict) else v


def _follow_stored_trajectory(gpu_device, name, late=None, theta_rel=1e-10):
    import json
    from pathlib import Path
    from types import SimpleNamespace

    from test_driver_oracle import compare_runs
    from seq_vp_3d_amd.hotpath import PairState
    from seq_vp_3d_amd.problem import Q1Problem, native_weights
    from seq_vp_3d_amd.solver import PGConfig, SequentialConfig, solve_sequential
    from seq_vp_3d_amd.synthetic import CONFIG_SMALL, make_pair

    path = Path(__file__).resolve().parent / "golden" / name
    doc = json.loads(path.read_text())
    print("case solver vs stored oracle trajectory 32^3,", doc["config"]["outer"], "outer steps", flush=True)

    def sampler(img, grid, pts):
        s = osp.build(img.cpu().numpy(), Q.Grid(grid.shape, grid.spacing))
        return torch.tensor(osp.sample(s, pts.cpu().numpy(), with_gradient=False)[0], device=img.device)

    pair = make_pair(CONFIG_SMALL, "cpu", sampler)
    assert float(pair.fixed.sum()) == _decode(doc["pair"]["fixed_sum"])
    assert float(pair.moving.sum()) == _decode(doc["pair"]["moving_sum"])
    want = SimpleNamespace(status=doc["status"],
                           history=[{k: _decode(v) for k, v in r.items()} for r in doc["history"]])
    want_inner = [{k: _decode(v) for k, v in r.items()} for r in doc["inner_history"]]
    w = native_weights(CONFIG_SMALL.shape)
    gpu = PairState(Q1Problem(pair.grid, w, 0., .05), pair.fixed, pair.moving, gpu_device)
    gpu.set_state(theta=torch.zeros(gpu.nfree, dtype=torch.float64, device=gpu_device))
    rows, inner_rows = [], []
    res = solve_sequential(gpu, SequentialConfig(), PGConfig(), max_outer=doc["config"]["outer"],
                           observer=rows.append, inner_observer=inner_rows.append)
    compare_runs(want, want_inner, res, rows, inner_rows, late=late)
    theta = gpu.theta.cpu().numpy()
    got = theta[np.asarray(doc["theta"]["index"])]
    ref = np.array([_decode(v) for v in doc["theta"]["values"]])
    scale = np.abs(ref).max()
    print("final theta: max deviation / max |theta| =", float(np.abs(got - ref).max() / scale), flush=True)
    np.testing.assert_allclose(got, ref, rtol=0, atol=theta_rel * scale)
    assert abs(np.linalg.norm(theta) - _decode(doc["theta"]["norm"])) <= theta_rel * _decode(doc["theta"]["norm"])


def test_device_solver_follows_the_stored_oracle_trajectory(gpu_device):
    """The 32^3 CPU-sized pair (12 blobs, true JD >= 0.6), 10 certified outer steps from u = 0 with the reference's
    defaults: the device driver on PairState against the oracle driver's trajectory stored by
    tests/golden/make_driver_trajectory.py (every attempt's decision, tau and trial energy and every outer row to
    1e-12 relative; 97 entries of the final theta to 1e-10 of its largest)."""
    _follow_stored_trajectory(gpu_device, "driver_trajectory_32.json")


def test_device_solver_follows_50_stored_oracle_outer_steps(gpu_device):
    """The same pair for 50 outer steps (250 attempts): every decision (reason, acceptance, step rule, inner
    status and step count) identical throughout. The values stay at the rounding floor (at most 1.2e-12 relative)
    through outer step 37; from step 38 the solver amplifies the last-bit differences of the two implementations
    (FMA contraction, summation order) by about 1.6x per step, to 2.1e-9 at step 49 (the BB2 step proposal, a
    ratio of differences, is the most sensitive). Bars: 1e-12 for steps 0-9, 1e-11 for 10-37, 1e-8 from 38 on; the
    final theta to 1e-10 of its largest entry (measured 3.3e-12)."""
    _follow_stored_trajectory(gpu_device, "driver_trajectory_32_50.json", late=[(10, 1e-11), (38, 1e-8)])


def test_first_inner_gradient_after_refresh_is_the_outer_gradient(gpu_device):
    """After a self-reference refresh the total gradient equals the outer (image + barrier) gradient of the joint
    pass bit for bit, so PairState.gradient takes that vector; after any state change it sweeps again."""
    from seq_vp_3d_amd.hotpath import PairState
    from seq_vp_3d_amd.problem import Q1Problem, native_weights
    from seq_vp_3d_amd.synthetic import PairSpec, make_pair
    print("case gradient reuse after refresh 18x19x20", flush=True)
    spec = PairSpec((18, 19, 20), 8, (2., 4.), 3, .6, .01, 5)

    def sampler(img, grid, pts):
        s = osp.build(img.cpu().numpy(), Q.Grid(grid.shape, grid.spacing))
        return torch.tensor(osp.sample(s, pts.cpu().numpy(), with_gradient=False)[0], device=img.device)

    pair = make_pair(spec, "cpu", sampler)
    w = native_weights(spec.shape)
    st = PairState(Q1Problem(pair.grid, w, 0., .05), pair.fixed, pair.moving, gpu_device)
    theta = Q.restrict_free(.7 * pair.displacement.numpy(), Q.Grid(spec.shape, (1., 1., 1.)))
    st.set_state(theta=torch.tensor(theta, device=gpu_device))
    st.refresh_reference()
    st.outer_norms()
    gf = st.diag.clone()
    swept = st.q1.gradient_free(st.warped, st.fixed, st.bgrad, st.u, st.ref, w.lam, data_weight=st.data_weight).clone()
    g = st.gradient()
    assert torch.equal(g, gf) and torch.equal(g, swept)
    assert float(torch.linalg.vector_norm(g)) > 0
    # a new state (and then a stale joint vector) is swept again
    st.set_state(theta=torch.tensor(.9 * theta, device=gpu_device))
    g2 = st.gradient().clone()
    assert not torch.equal(g2, gf)
    want = st.q1.gradient_free(st.warped, st.fixed, st.bgrad, st.u, st.ref, w.lam, data_weight=st.data_weight)
    assert torch.equal(g2, want)
