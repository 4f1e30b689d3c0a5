"""Device pre-optimiser (seqvp_preopt_*) and pyramid resampling against the oracle."""
import numpy as np
import pytest

from oracle import preopt as op

pytestmark = pytest.mark.gpu

H = (1.1, .9, 1.3)
SHAPE = (9, 10, 11)


def _inputs(seed=0, amp=.35):
    rng = np.random.default_rng(seed)
    D, Hh, W = SHAPE
    z, y, x = np.meshgrid(*(np.linspace(0, 1, n) for n in SHAPE), indexing="ij")
    F = np.sin(3 * x) * np.cos(2 * y) + .5 * z * z
    M = np.sin(3 * x + .2) * np.cos(2 * y - .1) + .5 * (z + .05) ** 2
    u = rng.normal(size=(3, D - 2, Hh - 2, W - 2)) * amp
    return u, F, M


def _grid():
    from seq_vp_3d_amd.problem import make_grid_3d
    return make_grid_3d(nx=SHAPE[2], ny=SHAPE[1], nz=SHAPE[0], hx=H[0], hy=H[1], hz=H[2])


@pytest.mark.parametrize("fp32", [False, True])
def test_energy_and_gradient_match_the_oracle(gpu_device, fp32):
    import torch
    from seq_vp_3d_amd.preopt import LevelState, energy_grad
    u, F, M = _inputs()
    args = (1.3, 2e-2, 4., .9)           # mu, alpha, kappa, j0 (the floor is active on part of the cells)
    want, g = op.energy_grad(u, F, M, H, *args)
    assert want["floor"] > 0
    dt = torch.float32 if fp32 else torch.float64
    st = LevelState(_grid(), gpu_device, fp32=fp32, light=True)
    T = lambda a: torch.tensor(a, dtype=dt, device=gpu_device).contiguous()
    print("case preopt energy_grad fp32" if fp32 else "case preopt energy_grad fp64", flush=True)
    terms, grad = energy_grad(st, T(u.reshape(-1)), T(F), T(M), *args)
    got = dict(zip(("total", "data", "geometry", "floor", "curvature", "min_det"), terms.tolist()))
    tol_e, tol_g = (1e-5, 1e-4) if fp32 else (1e-12, 1e-12)
    for k, v in want.items():
        assert abs(got[k] - v) <= tol_e * max(abs(v), 1e-3), k
    gg = grad.double().cpu().numpy().reshape(g.shape)
    assert np.linalg.norm(gg - g) <= tol_g * np.linalg.norm(g)


def test_step_matches_the_oracle_update(gpu_device):
    import torch
    from seq_vp_3d_amd import _lib
    from seq_vp_3d_amd.preopt import LevelState
    u, F, M = _inputs(amp=.05)
    _, g = op.energy_grad(u, F, M, H, 1., 1e-3, 10., .2)
    wg, rg = op.gauss_kernel(1.)
    wf, rf = op.gauss_kernel(.7)
    st = LevelState(_grid(), gpu_device, fp32=False)
    T = lambda a: torch.tensor(np.ascontiguousarray(a).reshape(-1), dtype=torch.float64, device=gpu_device)
    st.u.copy_(T(u))
    m, v = np.zeros_like(u), np.zeros_like(u)
    um = u.copy()
    W1, W2 = T(wg), T(wf)
    nz, ny, nx = SHAPE
    print("case preopt step", flush=True)
    for it in range(3):
        grad_it = g * (1 + .3 * it)
        st.g.copy_(T(grad_it))
        _lib.call("seqvp_preopt_step", _lib.ptr(st.u), _lib.ptr(st.m), _lib.ptr(st.v), _lib.ptr(st.g),
                  _lib.ptr(st.tmp), nx, ny, nz, _lib.ptr(W1), rg, _lib.ptr(W2), rf, 1., .5, .9, .99, 1e-8, 0,
                  _lib.ptr(st.state), _lib.ptr(st.pmax), _lib.stream_handle(gpu_device))
        um, m, v, capped = op.adam_step(um, m, v, grad_it, it + 1, 1., .5, .9, .99, 1e-8, w_grad=wg, w_field=wf)
    for got, want in ((st.u, um), (st.m, m), (st.v, v)):
        w = want.reshape(-1)
        assert np.max(np.abs(got.cpu().numpy() - w)) <= 1e-13 * max(1., np.max(np.abs(w)))
    t, hits = st.state.tolist()
    assert t == 3 and hits >= 1


def test_pyramid_resampling(gpu_device):
    import torch
    from seq_vp_3d_amd.preopt import coarse_grid, downsample, prolong, taper
    from seq_vp_3d_amd.problem import make_grid_3d
    rng = np.random.default_rng(5)
    grid = make_grid_3d(nx=29, ny=26, nz=23, hx=1.1, hy=.9, hz=1.3)
    img = rng.normal(size=grid.shape).cumsum(axis=0)
    print("case downsample", flush=True)
    got, cg = downsample(torch.tensor(img, device=gpu_device), grid, (7, 8, 9))
    want = op.downsample(img, (7, 8, 9))
    assert np.max(np.abs(got.cpu().numpy() - want)) <= 1e-13 * np.max(np.abs(want))
    # prolongation: bitwise the host arithmetic; affine fields reproduced
    u = np.zeros(cg.shape + (3,))
    u[1:-1, 1:-1, 1:-1] = rng.normal(size=(5, 6, 7, 3))
    fine = coarse_grid(grid, (12, 14, 17))
    print("case prolong", flush=True)
    p = prolong(torch.tensor(u, device=gpu_device), cg, fine).cpu().numpy()
    want = op.prolong(u, cg.axes[::-1], fine.axes[::-1], cg.spacing[::-1])
    assert np.array_equal(p, want)
    t = taper(torch.tensor(p, device=gpu_device), fine, 3).cpu().numpy()
    rs = [np.minimum(np.minimum(np.arange(n), n - 1 - np.arange(n)) / 3., 1.) for n in fine.shape]
    w = rs[0][:, None, None] * rs[1][None, :, None] * rs[2][None, None, :]
    assert np.array_equal(t, w[..., None] * p)


def test_graph_replays_equal_eager_iterations(gpu_device):
    import torch
    from seq_vp_3d_amd.preopt import LevelState, PreOptimizer, capture
    u, F, M = _inputs(amp=.1)
    pre = PreOptimizer(gpu_device)
    T = lambda a, dt=torch.float32: torch.tensor(np.ascontiguousarray(a), dtype=dt, device=gpu_device).contiguous()
    Fd, Md = T(F), T(M)
    wts = pre._weights(True)
    runs = []
    print("case preopt graph", flush=True)
    for mode in ("eager", "graph"):
        st = LevelState(_grid(), gpu_device, fp32=True)
        st.u.copy_(T(u.reshape(-1)))
        pre._iteration(st, Fd, Md, 2, wts)
        if mode == "eager":
            for _ in range(6):
                pre._iteration(st, Fd, Md, 2, wts)
        else:
            g = capture(lambda: pre._iteration(st, Fd, Md, 2, wts), gpu_device)
            for k in range(6):
                g.replay()
                if k == 2:                        # eager work and a synchronisation between replays
                    torch.cuda.synchronize(gpu_device)
                    _ = (Fd * 2).sum().item()
        torch.cuda.synchronize(gpu_device)
        runs.append([x.cpu().numpy().copy() for x in (st.u, st.m, st.v)] + [st.state.cpu().numpy().copy()])
    for a, b in zip(*runs):
        assert np.array_equal(a, b)
    assert runs[1][3][0] == 7


def test_preoptimize_small_pair_and_certified_transfer(gpu_device):
    import torch
    from seq_vp_3d_amd.hotpath import PairState
    from seq_vp_3d_amd.preopt import PreOptimizer, PreoptConfig, certify_transfer, taper
    from seq_vp_3d_amd.problem import Q1Problem, native_weights
    from seq_vp_3d_amd.spline import DeviceSpline, build_image_sampler
    from seq_vp_3d_amd.synthetic import CONFIG_SMALL, make_pair

    def sampler(image, grid, points):
        return DeviceSpline(build_image_sampler(image.cpu().numpy(), grid), grid, gpu_device).sample(
            points, with_gradient=False)[0]
    print("case preoptimize 32^3", flush=True)
    pair = make_pair(CONFIG_SMALL, gpu_device, sampler)
    cfg = PreoptConfig(iterations=(20, 20, 20, 10), lr=.5, sigma_field=1., mu=(1., 1., .01, .003))
    u, rec = PreOptimizer(gpu_device, cfg)(pair.fixed, pair.moving, pair.grid)
    assert u.shape == pair.grid.shape + (3,) and bool(torch.isfinite(u).all())
    for lv in rec["levels"]:
        e = lv["end_errors"]
        print(lv["level"], lv["first"], lv["last"], e, flush=True)
        assert e["energy_rel_error"] <= 1e-5 and e["grad_rel_l2_error"] <= 1e-3, lv
        assert np.isfinite(lv["last"]["total"])
    # the pre-optimised field matches the images better than the identity at the native level
    from seq_vp_3d_amd.preopt import LevelState, energy_grad
    st0 = LevelState(pair.grid, gpu_device, fp32=False, light=True)
    zero = torch.zeros(3 * pair.grid.n_free_scalar, dtype=torch.float64, device=gpu_device)
    data0 = energy_grad(st0, zero, pair.fixed, pair.moving, cfg.mu[3], cfg.alpha[3], cfg.kappa_pre, cfg.j0)[0][1].item()
    assert rec["levels"][3]["end_errors"]["energy_fp64"] < data0
    seed = taper(u, pair.grid, cfg.taper)
    st = PairState(Q1Problem(pair.grid, native_weights(pair.grid.shape), 0., .05), pair.fixed, pair.moving,
                   gpu_device)
    alpha, attempts = certify_transfer(st, seed)
    assert alpha is not None, attempts
    assert st.energy.total > 0
