"""Parity of every device entry point with the CPU oracle on small grids (runs through the C ABI)."""
import numpy as np
import pytest

from conftest import counterexample_cell, rough_field, smooth_field
from oracle import jd_root as ojr, poisson as opo, prescreen as ops, q1 as Q, spline as osp

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

SHAPE = (9, 10, 11)
SPACING = (1., 1.3, .8)


def case(activation=.05, shape=SHAPE, spacing=SPACING):
    from seq_vp_3d_amd.problem import Q1Problem, Weights3D, make_grid_3d
    nz, ny, nx = shape
    grid = make_grid_3d(nx=nx, ny=ny, nz=nz, hx=spacing[0], hy=spacing[1], hz=spacing[2])
    w = Weights3D(lam=2e-3, mu=.03, nu=.05, kappa=.002)
    prob = Q1Problem(grid, w, 0., activation)
    og = Q.Grid(shape, spacing)
    model = Q.Model(og, lam=w.lam, mu=w.mu, nu=w.nu, kappa=w.kappa, delta=0., activation=activation)
    return prob, og, model


def fields(shape, seed):
    ref = smooth_field(shape, .15, seed)
    base = ref + smooth_field(shape, .3, seed + 1) + rough_field(shape, .03, seed + 2)
    u = base + smooth_field(shape, .1, seed + 3) + rough_field(shape, .02, seed + 4)
    return ref, base, u


def dev(a, device):
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device=device)


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("activation", [.05, .9])
def test_soft_and_barrier_energy(gpu_device, activation):
    from seq_vp_3d_amd.q1 import FusedQ1
    print("case soft/barrier energy", activation, flush=True)
    prob, og, m = case(activation)
    ref, base, u = fields(og.shape, 20)
    q = FusedQ1(prob, gpu_device)
    U, R, B = dev(u, gpu_device), dev(ref, gpu_device), dev(base, gpu_device)
    got = q.soft_energy(U, R, B).cpu().numpy()
    jd, curl, diff = m.soft_energy(u, ref, base)
    assert abs(got[0] - jd) <= 1e-12 * jd and abs(got[1] - curl) <= 1e-12 * curl
    assert abs(got[2] - diff) <= 1e-12 * (jd + curl)
    got0 = q.soft_energy(U, R).cpu().numpy()
    assert got0[2] == 0. and abs(got0[0] - jd) <= 1e-12 * jd
    bar = q.barrier_energy(U, B).cpu().numpy()
    val, jmin, chg, amin = m.barrier_energy(u, base)
    if activation == .9:
        assert val > 0 and chg != 0
    assert abs(bar[0] - val) <= 1e-12 * max(val, 1e-300) + 1e-300
    # minima of J: the device forms J with fused multiply-adds, so they agree to a few ulps
    assert abs(bar[1] - jmin) <= 1e-13 * abs(jmin) and abs(bar[3] - amin) <= 1e-13 * abs(amin)
    w = m.barrier_weight
    assert abs(w * bar[2] - chg) <= 1e-12 * max(abs(val), abs(chg), 1e-300)
    zero = q.soft_energy(R, R, R).cpu().numpy()
    assert np.array_equal(zero, np.zeros(3))
    # the fused trial sweep: the same bits as the two separate sweeps, with and without a base
    for base in (B, None):
        fused = q.trial_energy(U, R, base)
        assert torch.equal(fused, torch.cat((q.soft_energy(U, R, base), q.barrier_energy(U, base)))), base is None
    # the reference passed as the base (its own buffer): the base's point terms are taken from the reference's, the
    # same bits as a separate copy of it; against the oracle the difference is the soft energy itself
    same = q.trial_energy(U, R, R)
    assert torch.equal(same, q.trial_energy(U, R, R.clone()))
    jd, curl, diff = m.soft_energy(u, ref, ref)
    got = same.cpu().numpy()
    assert abs(got[2] - diff) <= 1e-12 * (jd + curl) and abs(got[2] - (jd + curl)) <= 1e-12 * (jd + curl)


@pytest.mark.parametrize("activation", [.05, .7])
def test_trial_energy_with_a_quiet_base(gpu_device, activation):
    """A base with no active closed27 point, its least J passed in: the sweep skips the base's closed27 J and gives
    the same bits (activation .05: no active point anywhere; .7: the base, a copy of the reference, has none while
    the trial has some). A passed minimum below the activation gives a NaN barrier value."""
    from seq_vp_3d_amd.q1 import FusedQ1
    print("case quiet base", activation, flush=True)
    prob, og, m = case(activation)
    ref, base, u = fields(og.shape, 20)
    q = FusedQ1(prob, gpu_device)
    U, R = dev(u, gpu_device), dev(ref, gpu_device)
    B = R.clone() if activation == .7 else dev(base, gpu_device)
    full = q.trial_energy(U, R, B)
    bmin = q.trial_energy(B, R)[4:5].clone()
    assert float(bmin) >= q.activation and float(bmin) == float(full[6])
    if activation == .7:
        assert float(full[3]) > 0 and float(full[4]) < q.activation
    for _ in range(2):
        assert torch.equal(q.trial_energy(U, R, B, base_min=bmin, quiet=True), full)
        assert torch.equal(q.trial_energy(U, R, B, base_min=bmin), full)      # known, not declared quiet
    low = torch.full((1,), q.activation / 2, dtype=torch.float64, device=gpu_device)
    assert torch.isnan(q.trial_energy(U, R, B, base_min=low, quiet=True)[3])


@pytest.mark.parametrize("activation", [.05, .5])
def test_trial_energy_skips_are_the_full_sweep(gpu_device, activation):
    """On a grid of 1,400 workgroups (more than the GPU holds at once, so later waves meet the running least J of
    earlier ones) with a fold-like dip placed late in the cell order: the trial sweep, which skips the closed-point
    J of cells whose edges bound it above the activation and above the least J met so far (and of base cells
    above the base's known least J), gives the separate soft and barrier sweeps' bits, for a general base, a quiet
    one, the reference as base, and no base."""
    from seq_vp_3d_amd.q1 import FusedQ1
    print("case trial sweep skips 64x72x80", activation, flush=True)
    prob, og, m = case(activation, shape=(64, 72, 80), spacing=(1., 1., 1.))
    shape = og.shape
    ref = smooth_field(shape, 1., 5)
    base = ref + smooth_field(shape, 2., 6)
    nz, ny, nx = shape
    z, y, x = np.meshgrid(*(np.arange(n) for n in shape), indexing="ij")
    dip = np.exp(-((z - .8 * nz) ** 2 + (y - ny / 2) ** 2 + (x - nx / 2) ** 2) / 6.)
    u = base + smooth_field(shape, 1., 7) + 2.2 * dip[..., None] * np.array([0., 0., -1.]) * np.sign(x - nx / 2)[..., None]
    for f in (ref, base, u):
        f[[0, -1]] = 0.
        f[:, [0, -1]] = 0.
        f[:, :, [0, -1]] = 0.
    q = FusedQ1(prob, gpu_device)
    U, R, B = dev(u, gpu_device), dev(ref, gpu_device), dev(base, gpu_device)
    want = torch.cat((q.soft_energy(U, R, B), q.barrier_energy(U, B)))
    assert float(want[4]) < float(want[6])            # the dip lowers the trial's least J
    for _ in range(2):
        assert torch.equal(q.trial_energy(U, R, B), want)
    bmin = q.trial_energy(B, R)[4:5].clone()
    assert torch.equal(q.trial_energy(U, R, B, base_min=bmin), want)
    if float(bmin) >= q.activation:
        assert torch.equal(q.trial_energy(U, R, B, base_min=bmin, quiet=True), want)
    want_ref = torch.cat((q.soft_energy(U, R, R), q.barrier_energy(U, R)))
    rmin = q.trial_energy(R, R)[4:5].clone()
    assert torch.equal(q.trial_energy(U, R, R), want_ref)
    assert torch.equal(q.trial_energy(U, R, R, base_min=rmin), want_ref)
    want0 = torch.cat((q.soft_energy(U, R), q.barrier_energy(U)))
    assert torch.equal(q.trial_energy(U, R), want0)


def test_energy_nan_propagates_to_min_j(gpu_device):
    from seq_vp_3d_amd.q1 import FusedQ1
    prob, og, m = case()
    _, _, u = fields(og.shape, 30)
    u[4, 5, 5, 0] = np.nan
    q = FusedQ1(prob, gpu_device)
    out = q.barrier_energy(dev(u, gpu_device)).cpu().numpy()
    assert np.isnan(out[1])


@pytest.mark.parametrize("activation", [.05, .9])
def test_gradient_and_joint(gpu_device, activation):
    from seq_vp_3d_amd.q1 import FusedQ1
    print("case gradient/joint", activation, flush=True)
    prob, og, m = case(activation)
    ref, _, u = fields(og.shape, 40)
    nb = np.random.default_rng(41).normal(size=u.shape)
    q = FusedQ1(prob, gpu_device)
    U, R, NB = dev(u, gpu_device), dev(ref, gpu_device), dev(nb, gpu_device)
    for soft, barrier in ((True, True), (True, False), (False, True)):
        got = q.gradient_nodes(NB, U, R, soft=soft, barrier=barrier).cpu().numpy()
        want = m.gradient_nodes(nb, u, ref, soft=soft, barrier=barrier)
        assert rel(got - nb, want - nb) <= 1e-12, (soft, barrier)
    img = np.random.default_rng(42).normal(size=u.shape)
    outer, total = q.joint(dev(img, gpu_device), NB, U, R)
    wo = m.gradient_nodes(img, u, ref, soft=False, barrier=True)
    wt = m.gradient_nodes(nb, u, ref, soft=True, barrier=True)
    assert rel(outer.cpu().numpy() - img, wo - img) <= 1e-12 or activation == .05 and np.allclose(outer.cpu().numpy(), wo, rtol=0, atol=1e-15)
    assert rel(total.cpu().numpy() - nb, wt - nb) <= 1e-12


@pytest.mark.parametrize("activation", [.05, .9])
def test_ggn_diagonal(gpu_device, activation):
    from seq_vp_3d_amd.q1 import FusedQ1
    print("case ggn", activation, flush=True)
    prob, og, m = case(activation)
    _, _, u = fields(og.shape, 50)
    b = np.random.default_rng(51).normal(size=u.shape)
    base = m.ggn_diagonal_base(b)
    q = FusedQ1(prob, gpu_device)
    got = q.diagonal_nodes(dev(base, gpu_device), dev(u, gpu_device)).cpu().numpy()
    want = m.ggn_diagonal_nodes(base, u)
    assert rel(got - base, want - base) <= 1e-12


def test_node_min_jd(gpu_device):
    from seq_vp_3d_amd.q1 import FusedQ1
    prob, og, m = case()
    _, _, u = fields(og.shape, 60)
    q = FusedQ1(prob, gpu_device)
    got = q.node_min_jd(dev(u, gpu_device)).cpu().numpy()
    want = m.node_min_jd(u)
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=0)


def test_jd_root(gpu_device):
    from seq_vp_3d_amd.q1 import FusedJDRoot, FusedQ1
    print("case jd root", flush=True)
    prob, og, m = case()
    _, base, _ = fields(og.shape, 70)
    v = smooth_field(og.shape, 1.5, 71) + rough_field(og.shape, .3, 72)
    q = FusedQ1(prob, gpu_device)
    root = FusedJDRoot(prob, q1=q)
    for limit in (.5, 2., 10.):
        got = root(dev(base, gpu_device), dev(v, gpu_device), limit)
        want = ojr.global_root(base, v, og, m.closed[3], 0., limit)
        assert got[1:] == want[1:]
        if np.isinf(want[0]):
            assert np.isinf(got[0])
        else:
            assert abs(got[0] - want[0]) <= 1e-12 * want[0], (limit, got, want)
    # the direction as a free-node vector: bitwise the root of its node field
    pf = Q.restrict_free(v, og)
    vn = Q.inject_free(pf, og)
    for limit in (.5, 2., 10.):
        assert root.from_free(dev(base, gpu_device), dev(pf, gpu_device), limit) == \
            root(dev(base, gpu_device), dev(vn, gpu_device), limit), limit


@pytest.mark.parametrize("amp", [.02, .3, 1.5])
def test_jd_root_where_most_cells_are_skipped(gpu_device, amp):
    """The root sweep skips the cells its edge bound shows root-free: directions small everywhere but for a local
    bump (most waves skipped, the bump's cells evaluated) give the oracle's root (1e-12) and counts exactly; the free
    vector gives the node field's results bit for bit; a NaN in the direction sets the nonfinite flag."""
    from seq_vp_3d_amd.q1 import FusedJDRoot, FusedQ1
    print("case jd root with skipped cells", amp, flush=True)
    prob, og, m = case(shape=(17, 18, 19))
    _, base, _ = fields(og.shape, 70)
    v = smooth_field(og.shape, .01, 73)
    nz, ny, nx = og.shape
    z, y, x = np.meshgrid(*(np.arange(n) for n in og.shape), indexing="ij")
    bump = np.exp(-((z - nz / 3) ** 2 + (y - ny / 2) ** 2 + (x - 2 * nx / 3) ** 2) / 4.)
    v += amp * bump[..., None] * np.array([1., -.6, .4])
    v[[0, -1]] = 0.
    v[:, [0, -1]] = 0.
    v[:, :, [0, -1]] = 0.
    q = FusedQ1(prob, gpu_device)
    root = FusedJDRoot(prob, q1=q)
    B = dev(base, gpu_device)
    pf = Q.restrict_free(v, og)
    vn = Q.inject_free(pf, og)
    found = 0
    for limit in (.05, .5, 2., 10.):
        got = root(B, dev(vn, gpu_device), limit)
        want = ojr.global_root(base, vn, og, m.closed[3], 0., limit)
        assert got[1:] == want[1:]
        if np.isinf(want[0]):
            assert np.isinf(got[0])
        else:
            found += 1
            assert abs(got[0] - want[0]) <= 1e-12 * want[0], (limit, got, want)
        assert root.from_free(B, dev(pf, gpu_device), limit) == got, limit
    if amp == 1.5:
        assert found >= 2        # roots inside the bump at the larger limits
    pn = pf.copy()
    pn[5] = np.nan
    out = root.from_free_launch(B, dev(pn, gpu_device), 1.)
    assert float(out[2]) == 1.


def test_prescreen_bitwise(gpu_device):
    from seq_vp_3d_amd.prescreen import prescreen_flags, prescreen_lower_bounds
    import oracle.build
    oracle.build.build()
    print("case prescreen", flush=True)
    prob, og, m = case()
    D = m.closed[3]
    for seed, amp in ((80, .2), (81, .45), (82, .6)):
        u = rough_field(og.shape, amp, seed) + smooth_field(og.shape, .3, seed + 100)
        flat = u.reshape(-1, 3)
        idx = Q.cell_indices(og)
        flat[idx[17]] = counterexample_cell(-.05) * np.array(SPACING) + flat[idx[17][0]]
        u[[0, -1]] = 0.
        u[:, [0, -1]] = 0.
        u[:, :, [0, -1]] = 0.
        U = dev(u, gpu_device)
        got = prescreen_lower_bounds(U, og, D).cpu().numpy()
        want = ops.lower_bounds_c(u, og, D)
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)) or np.array_equal(got, want, equal_nan=True)
        cells = np.array([0, 5, 17, og.n_cells - 1], dtype=np.int64)
        got_l = prescreen_lower_bounds(U, og, D, cells).cpu().numpy()
        assert np.array_equal(got_l, want[cells], equal_nan=True)
        flags, count = prescreen_flags(U, og, D, 0.)
        proven = want > 0
        assert np.array_equal(flags.cpu().numpy().astype(bool), proven)
        assert int(count.item()) == int((~proven).sum())
        flags, count = prescreen_flags(U, og, D, .01)
        with np.errstate(invalid="ignore"):
            proven = np.nextafter(want - .01, -np.inf) > 0
        assert np.array_equal(flags.cpu().numpy().astype(bool), proven)


def test_spline_warp_and_points(gpu_device):
    from seq_vp_3d_amd.spline import DeviceSpline, build_image_sampler
    from seq_vp_3d_amd.problem import make_grid_3d
    print("case spline", flush=True)
    shape, spacing = (12, 13, 14), (1., 1.5, .7)
    grid = make_grid_3d(nx=14, ny=13, nz=12, hx=1., hy=1.5, hz=.7, x0=2., y0=-1., z0=.5)
    og = Q.Grid(shape, spacing, (2., -1., .5))
    img = np.random.default_rng(90).normal(size=shape)
    sp = DeviceSpline(build_image_sampler(img, grid), grid, gpu_device)
    ref = osp.build(img, og)
    u = smooth_field(shape, .4, 91)
    val, grad, bad = sp.warp(dev(u, gpu_device))
    assert int(bad.item()) == 0
    wv, wg = osp.sample(ref, og.identity() + u)
    assert rel(val.cpu().numpy(), wv) <= 1e-12 and rel(grad.cpu().numpy(), wg) <= 1e-12
    # out-of-support free node is counted
    u2 = u.copy()
    u2[5, 5, 1, 0] = -5.
    _, _, bad = sp.warp(dev(u2, gpu_device))
    assert int(bad.item()) == 1
    pts = og.identity()[1:-1, 1:-1, 1:-1].reshape(-1, 3) + .3
    v, g = sp.sample(dev(pts, gpu_device))
    wv, wg = osp.sample(ref, pts)
    assert rel(v.cpu().numpy(), wv) <= 1e-12 and rel(g.cpu().numpy(), wg) <= 1e-12
    # points exactly on nodes (knots), on the support ends and a hair inside them: the span search's edge cases
    ax = og.axes
    rng = np.random.default_rng(92)
    cand = [np.concatenate([a, [a[0], a[-1], np.nextafter(a[0], np.inf), np.nextafter(a[-1], -np.inf)]])
            for a in ax]
    pts = np.stack([rng.choice(c, 600) for c in cand], axis=-1)
    v, g = sp.sample(dev(np.ascontiguousarray(pts), gpu_device))
    wv, wg = osp.sample(ref, pts)
    assert rel(v.cpu().numpy(), wv) <= 1e-12 and rel(g.cpu().numpy(), wg) <= 1e-12
    # TorchSpline.sample's keywords (backends/spline.py:81-96): chunks give the same values; a strict mask refuses
    # a masked point on the support's end and accepts it unmasked
    v2, g2 = sp.sample(dev(np.ascontiguousarray(pts), gpu_device), chunk_size=97)
    assert torch.equal(v2, v) and torch.equal(g2, g)
    edge = np.ascontiguousarray(np.stack([np.full(4, ax[0][0]), np.full(4, ax[1][3]), np.full(4, ax[2][2])], -1))
    mask = torch.tensor([False, False, True, False], device=gpu_device)
    with pytest.raises(Exception, match="outside permitted support"):
        sp.sample(dev(edge, gpu_device), chunk_size=3, strict=mask)
    sp.sample(dev(edge, gpu_device), strict=torch.zeros(4, dtype=torch.bool, device=gpu_device))
    with pytest.raises(Exception, match="chunk"):
        sp.sample(dev(edge, gpu_device), chunk_size=0)


@pytest.mark.parametrize("fp32", [0, 1])
def test_mfma_layout_probe(gpu_device, fp32):
    from seq_vp_3d_amd import _lib
    rng = np.random.default_rng(95)
    a, b = rng.normal(size=(16, 4)), rng.normal(size=(4, 16))
    d = torch.zeros(256, dtype=torch.float64, device=gpu_device)
    A, B = dev(a, gpu_device), dev(b, gpu_device)        # keep both alive until the launch has run
    _lib.call("seqvp_mfma_probe", _lib.ptr(A), _lib.ptr(B), fp32, _lib.ptr(d), _lib.stream_handle(gpu_device))
    torch.cuda.synchronize()
    tol = 1e-5 if fp32 else 1e-14
    np.testing.assert_allclose(d.cpu().numpy().reshape(16, 16), a @ b, rtol=tol, atol=tol)


# interior lengths: odd and even, one wave pair (<= 128) and two (129..256) of the folded passes, and the plain
# 64x64 tiling beyond 256 (x = 258)
@pytest.mark.parametrize("shape", [(9, 10, 11), (20, 23, 37), (135, 5, 140), (4, 200, 7), (5, 6, 260)])
def test_dst_poisson(gpu_device, shape):
    from seq_vp_3d_amd.poisson import DevicePoisson, build_poisson_inverse
    from seq_vp_3d_amd.problem import make_grid_3d
    print("case dst", shape, flush=True)
    nz, ny, nx = shape
    grid = make_grid_3d(nx=nx, ny=ny, nz=nz, hx=1., hy=1.3, hz=.8)
    og = Q.Grid(shape, (1., 1.3, .8))
    inv = build_poisson_inverse(grid, a=.3, b=.04)
    den = opo.denominator(og, .3, .04)
    assert np.array_equal(inv.denominator, den)
    P = DevicePoisson(inv, gpu_device)
    rng = np.random.default_rng(96)
    g = rng.normal(size=3 * og.n_free_scalar)
    s = rng.uniform(.5, 2., size=g.size)
    got = P.apply(dev(g, gpu_device)).cpu().numpy()
    assert rel(got, opo.apply(g, den)) <= 1e-12
    got = P.equilibrated_direction(dev(s, gpu_device), dev(g, gpu_device)).cpu().numpy()
    assert rel(got, -(s * opo.apply(s * g, den))) <= 1e-12
    got32 = P.equilibrated_direction(dev(s, gpu_device), dev(g, gpu_device), fp32=True).cpu().numpy()
    assert rel(got32, -(s * opo.apply(s * g, den))) <= 2e-5
    # the reference's host-array protocol (hybrid.py:224-286) with TorchPoisson's argument order (torch_ops.py:107)
    from seq_vp_3d_amd.poisson import TorchPoisson
    T = TorchPoisson(gpu_device, inv)
    seen = []
    d, ok = T.apply_equilibrated(T.prepare_equilibrated(s, g), between=lambda: seen.append(1))
    assert ok is True and seen == [1] and isinstance(d, np.ndarray)
    assert rel(d, -(s * opo.apply(s * g, den))) <= 1e-12
    bad = g.copy()
    bad[5] = np.inf
    tag, x = T.apply_equilibrated(T.prepare_equilibrated(s, bad))
    assert tag == "x" and isinstance(x, np.ndarray) and x.shape == g.shape


def test_vector_ops(gpu_device):
    from seq_vp_3d_amd import _lib
    shape = (9, 10, 11)
    nz, ny, nx = shape
    og = Q.Grid(shape)
    rng = np.random.default_rng(97)
    theta = rng.normal(size=3 * og.n_free_scalar)
    p = rng.normal(size=theta.size)
    T, Pd = dev(theta, gpu_device), dev(p, gpu_device)
    u = torch.empty(shape + (3,), dtype=torch.float64, device=gpu_device)
    _lib.call("seqvp_inject", _lib.ptr(T), nx, ny, nz, _lib.ptr(u), _lib.stream_handle(gpu_device))
    assert np.array_equal(u.cpu().numpy(), Q.inject_free(theta, og))
    back = torch.empty_like(T)
    _lib.call("seqvp_restrict", _lib.ptr(u), nx, ny, nz, _lib.ptr(back), _lib.stream_handle(gpu_device))
    assert np.array_equal(back.cpu().numpy(), theta)
    tout = torch.empty_like(T)
    _lib.call("seqvp_axpy_inject", _lib.ptr(T), _lib.ptr(Pd), .37, nx, ny, nz, _lib.ptr(tout), _lib.ptr(u),
              _lib.stream_handle(gpu_device))
    assert np.array_equal(tout.cpu().numpy(), theta + .37 * p)
    assert np.array_equal(u.cpu().numpy(), Q.inject_free(theta + .37 * p, og))
    import ctypes
    xs = (ctypes.c_void_p * 2)(T.data_ptr(), Pd.data_ptr())
    ys = (ctypes.c_void_p * 2)(Pd.data_ptr(), Pd.data_ptr())
    out = torch.empty(2, dtype=torch.float64, device=gpu_device)
    ws = torch.empty(_lib.workspace_bytes(7, shape), dtype=torch.uint8, device=gpu_device)
    _lib.call("seqvp_dots", xs, ys, 2, theta.size, _lib.ptr(out), _lib.ptr(ws), ws.numel(),
              _lib.stream_handle(gpu_device))
    want = [float(np.sum(theta.astype(np.longdouble) * p)), float(np.sum(p.astype(np.longdouble) ** 2))]
    np.testing.assert_allclose(out.cpu().numpy(), want, rtol=1e-15)


@pytest.mark.parametrize("shape", [(9, 10, 11), (40, 33, 57)])
def test_direction_scalars_are_the_separate_entry_points(gpu_device, shape):
    """seqvp_direction_scalars (one pass) against seqvp_dots, seqvp_bb2_products, seqvp_step_bound and
    seqvp_max_abs_diff, bit for bit, with and without the previous state, with a NaN maximum case too."""
    import ctypes
    from seq_vp_3d_amd import _lib
    print("case direction scalars", shape, flush=True)
    nz, ny, nx = shape
    og = Q.Grid(shape, (1., 1.3, .8))
    rng = np.random.default_rng(98)
    n = 3 * og.n_free_scalar
    th, th0, g, g0, p0, p = (dev(rng.normal(size=n), gpu_device) for _ in range(6))
    u = dev(rng.normal(size=shape + (3,)) * .1, gpu_device)
    ws = torch.empty(_lib.workspace_bytes(7, shape), dtype=torch.uint8, device=gpu_device)
    s = _lib.stream_handle(gpu_device)
    ox, oy, oz = 0., 0., 0.
    hx, hy, hz = 1., 1.3, .8

    def separate(prev):
        out = torch.full((12,), np.nan, dtype=torch.float64, device=gpu_device)
        xs = (ctypes.c_void_p * 2)(g.data_ptr(), g.data_ptr())
        ys = (ctypes.c_void_p * 2)(p.data_ptr(), g.data_ptr())
        _lib.call("seqvp_dots", xs, ys, 2, n, _lib.ptr(out[0:2]), _lib.ptr(ws), ws.numel(), s)
        if prev:
            _lib.call("seqvp_bb2_products", _lib.ptr(th), _lib.ptr(th0), _lib.ptr(g), _lib.ptr(g0), _lib.ptr(p0),
                      _lib.ptr(p), n, _lib.ptr(out[4:8]), _lib.ptr(ws), ws.numel(), s)
            _lib.call("seqvp_max_abs_diff", _lib.ptr(th), _lib.ptr(th0), n, _lib.ptr(out[9:11]), _lib.ptr(ws),
                      ws.numel(), s)
        _lib.call("seqvp_step_bound", _lib.ptr(p), _lib.ptr(u), ox, oy, oz, hx, hy, hz, nx, ny, nz, _lib.ptr(out[8:9]),
                  _lib.ptr(ws), ws.numel(), s)
        return out.cpu().numpy()

    def fused(prev):
        out = torch.full((12,), np.nan, dtype=torch.float64, device=gpu_device)
        a = (th0, g0, p0) if prev else (None, None, None)
        _lib.call("seqvp_direction_scalars", _lib.ptr(th), _lib.ptr(a[0]), _lib.ptr(g), _lib.ptr(a[1]), _lib.ptr(a[2]),
                  _lib.ptr(p), _lib.ptr(u), ox, oy, oz, hx, hy, hz, nx, ny, nz, _lib.ptr(out), _lib.ptr(ws),
                  ws.numel(), s)
        return out.cpu().numpy()

    for prev in (True, False):
        a, b = separate(prev), fused(prev)
        assert np.array_equal(a, b, equal_nan=True), (prev, a, b)
        assert np.isfinite(b[8]) and (np.isfinite(b[4:8]).all() if prev else np.isnan(b[4:8]).all())
    th0[7] = float("nan")
    a, b = separate(True), fused(True)
    assert np.array_equal(a, b, equal_nan=True)


def test_node_energy(gpu_device):
    from seq_vp_3d_amd.q1 import FusedQ1
    prob, og, m = case()
    ref, base, u = fields(og.shape, 110)
    rng = np.random.default_rng(111)
    fixed, warped, bw = rng.uniform(size=og.shape), rng.uniform(size=og.shape), rng.uniform(size=og.shape)
    m.fixed = fixed
    q = FusedQ1(prob, gpu_device)
    out = q.node_energy(dev(warped, gpu_device), dev(fixed, gpu_device), dev(u, gpu_device), dev(ref, gpu_device),
                        m.lam, base_warped=dev(bw, gpu_device), base_u=dev(base, gpu_device)).cpu().numpy()
    image, prox = m.node_energies(u, ref, warped)
    ib, pb = m.node_energies(base, ref, bw)
    np.testing.assert_allclose(out[:2], [image, prox], rtol=1e-13)
    np.testing.assert_allclose(out[2], image - ib, rtol=1e-10)
    np.testing.assert_allclose(out[3], prox - pb, rtol=1e-10)


def test_node_terms(gpu_device):
    from seq_vp_3d_amd.q1 import FusedQ1
    prob, og, m = case()
    ref, _, u = fields(og.shape, 120)
    rng = np.random.default_rng(121)
    fixed, warped = rng.uniform(size=og.shape), rng.uniform(size=og.shape)
    b = rng.normal(size=og.shape + (3,))
    dw = rng.uniform(.5, 1., size=og.shape)
    m.fixed = fixed
    m.qd = m.q * dw
    q = FusedQ1(prob, gpu_device)
    img, tot, ggn = (torch.empty(og.shape + (3,), dtype=torch.float64, device=gpu_device) for _ in range(3))
    q.node_terms(dev(warped, gpu_device), dev(fixed, gpu_device), dev(b, gpu_device), dev(u, gpu_device),
                 dev(ref, gpu_device), m.lam, data_weight=dev(dw, gpu_device), image=img, total=tot, ggn=ggn)
    want = m.node_gradient_base(u, ref, warped, b)
    np.testing.assert_allclose(tot.cpu().numpy(), want, rtol=1e-13, atol=1e-15 * np.abs(want).max())
    np.testing.assert_allclose(img.cpu().numpy(), m.node_gradient_base(u, ref, warped, b, total=False), rtol=1e-14)
    np.testing.assert_allclose(ggn.cpu().numpy(), m.ggn_diagonal_base(b), rtol=1e-14)


@pytest.mark.parametrize("weighted", [False, True])
def test_gradient_free_is_the_restricted_node_gradient(gpu_device, weighted):
    """seqvp_q1_gradient_free (node terms, VJP gather and restriction in one pass) against the oracle's node
    gradient restricted to the free nodes, and against the three-call device path."""
    from seq_vp_3d_amd.q1 import FusedQ1
    prob, og, m = case()
    ref, _, u = fields(og.shape, 130)
    rng = np.random.default_rng(131)
    fixed, warped = rng.uniform(size=og.shape), rng.uniform(size=og.shape)
    b = rng.normal(size=og.shape + (3,))
    dw = rng.uniform(.5, 1., size=og.shape) if weighted else None
    m.fixed = fixed
    m.qd = m.q * dw if weighted else m.q
    q = FusedQ1(prob, gpu_device)
    D = lambda a: None if a is None else dev(a, gpu_device)
    print("case gradient_free", og.shape, "weighted", weighted, flush=True)
    g = q.gradient_free(D(warped), D(fixed), D(b), D(u), D(ref), m.lam, data_weight=D(dw)).cpu().numpy()
    want = Q.restrict_free(m.gradient_nodes(m.node_gradient_base(u, ref, warped, b), u, ref), og)
    np.testing.assert_allclose(g, want, rtol=1e-12, atol=1e-13 * np.abs(want).max())
    tot = torch.empty(og.shape + (3,), dtype=torch.float64, device=gpu_device)
    q.node_terms(D(warped), D(fixed), D(b), D(u), D(ref), m.lam, data_weight=D(dw), total=tot)
    nodes = q.gradient_nodes(tot, D(u), D(ref)).cpu().numpy()
    three = Q.restrict_free(nodes, og)
    np.testing.assert_allclose(g, three, rtol=1e-14, atol=1e-15 * np.abs(three).max())


def test_joint_free_is_the_restricted_joint(gpu_device):
    """seqvp_q1_joint_free (node terms, joint gathers and the three restrictions in one pass) against the oracle's
    image, outer and total node gradients restricted to the free nodes, and against the multi-call device path."""
    from seq_vp_3d_amd.q1 import FusedQ1
    prob, og, m = case(activation=.9)
    ref, _, u = fields(og.shape, 140)
    rng = np.random.default_rng(141)
    fixed, warped = rng.uniform(size=og.shape), rng.uniform(size=og.shape)
    b = rng.normal(size=og.shape + (3,))
    m.fixed = fixed
    m.qd = m.q
    q = FusedQ1(prob, gpu_device)
    D = lambda a: dev(a, gpu_device)
    print("case joint_free", og.shape, flush=True)
    gd, gf, ge = (t.cpu().numpy() for t in q.joint_free(D(warped), D(fixed), D(b), D(u), D(ref), m.lam))
    img = m.node_gradient_base(u, ref, warped, b, total=False)
    tot = m.node_gradient_base(u, ref, warped, b)
    R = lambda a: Q.restrict_free(a, og)
    for got, want in ((gd, R(img)), (gf, R(m.gradient_nodes(img, u, ref, soft=False, barrier=True))),
                      (ge, R(m.gradient_nodes(tot, u, ref, soft=True, barrier=True)))):
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-13 * np.abs(want).max())
    im_d, tot_d = (torch.empty(og.shape + (3,), dtype=torch.float64, device=gpu_device) for _ in range(2))
    q.node_terms(D(warped), D(fixed), D(b), D(u), D(ref), m.lam, image=im_d, total=tot_d)
    outer, total = q.joint(im_d, tot_d, D(u), D(ref))
    for got, want in ((gd, R(im_d.cpu().numpy())), (gf, R(outer.cpu().numpy())), (ge, R(total.cpu().numpy()))):
        np.testing.assert_allclose(got, want, rtol=1e-14, atol=1e-15 * np.abs(want).max())


@pytest.mark.parametrize("activation", [.05, .9])
def test_fp32_screen_energies_within_band(gpu_device, activation):
    """The FP32 screening energies stay within the screen's band of the FP64 values
    (64 ulps of binary32 for JD/curl, 4096 for the barrier; accelerated.py:67-76)."""
    from seq_vp_3d_amd.q1 import FusedQ1
    print("case fp32 screen", activation, flush=True)
    prob, og, m = case(activation)
    ref, base, u = fields(og.shape, 130)
    q = FusedQ1(prob, gpu_device)
    U, R, B = dev(u, gpu_device), dev(ref, gpu_device), dev(base, gpu_device)
    s64, s32 = q.soft_energy(U, R, B).cpu().numpy(), q.soft_energy_f32(U, R, B).cpu().numpy()
    b64, b32 = q.barrier_energy(U, B).cpu().numpy(), q.barrier_energy_f32(U, B).cpu().numpy()
    eps = 2. ** -24
    assert abs(s32[0] - s64[0]) <= 64 * eps * s64[0] and abs(s32[1] - s64[1]) <= 64 * eps * s64[1]
    assert abs(b32[0] - b64[0]) <= 4096 * eps * max(b64[0], 1e-300) + 1e-300
    assert abs(b32[1] - b64[1]) <= 8 * eps * abs(b64[1])
    jb, cb, _ = q.soft_energy(B, R).cpu().numpy()
    bb = q.barrier_energy(B).cpu().numpy()[0]
    band = eps * (64 * (s64[0] + s64[1] + jb + cb) + 4096 * (b64[0] + bb))
    w = prob.barrier_weight
    assert abs((s32[2] + w * b32[2]) - (s64[2] + w * b64[2])) <= band


@pytest.mark.parametrize("metric", ["q1", "seven_point"])
def test_device_denominator_matches_host(gpu_device, metric):
    from seq_vp_3d_amd.poisson import DevicePoisson, SpectralMetric, build_poisson_inverse
    from seq_vp_3d_amd.problem import make_grid_3d
    grid = make_grid_3d(nx=23, ny=19, nz=17, hx=1., hy=1.3, hz=.8)
    host = build_poisson_inverse(grid, a=.21, b=.037, metric=metric).denominator
    devp = DevicePoisson(SpectralMetric(grid, .21, .037, metric), gpu_device)
    np.testing.assert_allclose(devp._denominator.cpu().numpy(), host, rtol=2e-16, atol=0)


def test_fp64_ceiling_probe(gpu_device):
    """The roofline probe's streams compute what they claim: the VALU stream is the recurrence x <- x m + c per
    accumulator, the MFMA stream is linear in the number of rounds; bad arguments are refused."""
    from seq_vp_3d_amd import _lib
    out = torch.empty(2 * 256, dtype=torch.float64, device=gpu_device)
    s = _lib.stream_handle(gpu_device)
    _lib.call("seqvp_fp64_ceiling_probe", 1, 2, 5, _lib.ptr(out), s)
    got = out.cpu().numpy()
    lanes = np.arange(256)
    want = np.zeros(256)
    for j in range(8):
        x = 1.0 + (lanes + j) * 1e-9
        for _ in range(5):
            x = x * 0.999999999 + 1e-9
        want += x
    np.testing.assert_allclose(got[:256], want, rtol=1e-14)
    np.testing.assert_array_equal(got[256:], got[:256])
    r = []
    for iters in (3, 6):
        _lib.call("seqvp_fp64_ceiling_probe", 0, 2, iters, _lib.ptr(out), s)
        r.append(out.cpu().numpy().copy())
    assert np.all(np.isfinite(r[0])) and np.all(r[0] > 0)
    np.testing.assert_allclose(r[1], 2 * r[0], rtol=1e-13)
    assert _lib.raw("seqvp_fp64_ceiling_probe", 2, 2, 5, _lib.ptr(out), s) != 0
    assert _lib.raw("seqvp_fp64_ceiling_probe", 0, 0, 5, _lib.ptr(out), s) != 0
