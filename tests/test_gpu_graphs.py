"""Captured solver segments (seq_vp_3d_amd/graphs.py): the trial and direction graphs give the bits of the same
calls issued one by one, on every replay, with eager work and synchronisations between the replays."""
import numpy as np
import pytest

from oracle import q1 as Q, spline as osp

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _pair(spec):
    from seq_vp_3d_amd.synthetic import make_pair

    def sampler(img, grid, pts):
        s = osp.build(img.cpu().numpy(), Q.Grid(grid.shape, grid.spacing))
        return torch.tensor(osp.sample(s, pts.cpu().numpy(), with_gradient=False)[0], device=img.device)

    return make_pair(spec, "cpu", sampler)


def _state(pair, spec, device, graphs):
    from seq_vp_3d_amd.hotpath import PairState
    from seq_vp_3d_amd.problem import Q1Problem, native_weights
    st = PairState(Q1Problem(pair.grid, native_weights(spec.shape), 0., .05), pair.fixed, pair.moving, device,
                   graphs=graphs)
    theta = Q.restrict_free(.5 * pair.displacement.numpy(), Q.Grid(spec.shape, (1., 1., 1.)))
    st.set_state(theta=torch.tensor(theta, device=device))
    return st


def test_trial_and_direction_replays_equal_eager_calls(gpu_device):
    """Each replay of the trial and direction segments (several taus, both buffer parities of the promotion swap)
    against the eager calls on a twin state, bit for bit; eager library work and a synchronisation run between the
    replays."""
    from seq_vp_3d_amd.synthetic import PairSpec
    print("case graph replays vs eager 18x19x20", flush=True)
    spec = PairSpec((18, 19, 20), 8, (2., 4.), 3, .6, .01, 5)
    pair = _pair(spec)
    states = [_state(pair, spec, gpu_device, g) for g in (True, False)]
    for st in states:
        st.refresh_reference()
        st.gradient()
        st.metric(shift_factor=1., activation=.25, extension=True)
    prev = [None, None]
    for it in range(4):
        scal = [st.direction_scalars(p, norm=True, fp32=False, gradient=True) for st, p in zip(states, prev)]
        assert scal[0] == scal[1], it
        assert torch.equal(states[0].p, states[1].p) and torch.equal(states[0].g, states[1].g)
        for tau in (.5 * scal[0][2], .25 * scal[0][2], .1 * scal[0][2]):
            res = [st.trial(tau) for st in states]
            a, b = res
            assert (a.ok, a.reason, a.total_difference, a.outer_difference, a.min_gap) == \
                   (b.ok, b.reason, b.total_difference, b.outer_difference, b.min_gap), (it, tau)
            assert a.energy == b.energy
            for name in ("u_t", "theta_t", "warped_t", "bgrad_t"):
                assert torch.equal(getattr(states[0], name), getattr(states[1], name)), name
            # eager work on the same stream between replays
            for st in states:
                st.q1.soft_energy(st.u, st.ref).tolist()
        # promote the last trial (swaps the buffers: the next iteration replays the other parity's graphs)
        if res[0].ok:
            prev = [st.promote(r.energy, keep_history=True) for st, r in zip(states, res)]
            for p in prev:
                p["tau"] = .1
    g = states[0].graphs
    assert g.captures >= 2 and g.launches > g.captures


def test_solver_with_graphs_is_the_solver_without(gpu_device):
    """Three certified outer steps of solve_sequential with and without captured segments: every row and attempt,
    and the final theta, bit for bit."""
    from seq_vp_3d_amd.solver import PGConfig, SequentialConfig, solve_sequential
    from seq_vp_3d_amd.synthetic import PairSpec
    print("case solver with vs without graphs 20x22x24", flush=True)
    spec = PairSpec((20, 22, 24), 8, (2., 4.), 3, .6, .01, 7)
    pair = _pair(spec)
    runs = []
    for graphs in (True, False):
        st = _state(pair, spec, gpu_device, graphs)
        rows, inner = [], []
        res = solve_sequential(st, SequentialConfig(), PGConfig(jd_step_global=True), max_outer=3,
                               observer=rows.append, inner_observer=inner.append)
        runs.append((res, rows, inner, st.theta.cpu().numpy()))
    (ra, rows_a, inner_a, ta), (rb, rows_b, inner_b, tb) = runs
    assert ra.status == rb.status and ra.outer_accepted == rb.outer_accepted
    skip = {"seconds", "elapsed_seconds"}
    assert [{k: v for k, v in r.items() if k not in skip} for r in rows_a] == \
           [{k: v for k, v in r.items() if k not in skip} for r in rows_b]
    assert len(inner_a) == len(inner_b)
    for a, b in zip(inner_a, inner_b):
        assert (a["reason"], a["tau"], a["energy_trial"], a["gradient_norm"]) == \
               (b["reason"], b["tau"], b["energy_trial"], b["gradient_norm"])
    assert np.array_equal(ta, tb)
