"""Every sweep at the native size 160x192x224 (6,772,287 cells), against the oracle and by size-independent properties.

* node-local results (gradient, GGN diagonal, T69 node JD measure, spline warp) are compared with the oracle run on a
  z-slab sub-block at the nodes whose eight cells lie inside the block;
* the energies are additive over z slabs of cells: the native value equals the sum of the slabs' values, and one
  slab equals the oracle;
* the T66 root is the least root over the slabs (bit for bit) and one slab's root equals the oracle's;
* the DST-I inverse satisfies M (M^-1 x) = x with the oracle's independent operator action (poisson.py:68-83);
* the gradient's directional derivative matches central stable differences of the energy.
"""
import numpy as np
import pytest

from conftest import rough_field, smooth_field
from oracle import jd_affine as oja, jd_root as ojr, poisson as opo, q1 as Q, spline as osp

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

SHAPE = (160, 192, 224)
SPACING = (1., 1., 1.)


@pytest.fixture(scope="module")
def native(gpu_device):
    from seq_vp_3d_amd.problem import Q1Problem, Weights3D, make_grid_3d
    from seq_vp_3d_amd.q1 import FusedQ1
    print("case native 160x192x224 fields", flush=True)
    nz, ny, nx = SHAPE
    grid = make_grid_3d(nx=nx, ny=ny, nz=nz, hx=1., hy=1., hz=1.)
    w = Weights3D(lam=2.8e-6, mu=.03, nu=.03, kappa=.002)
    ref = smooth_field(SHAPE, 3., 1)
    base = ref + smooth_field(SHAPE, 4., 2) + rough_field(SHAPE, .02, 3)
    u = base + smooth_field(SHAPE, 2., 4) + rough_field(SHAPE, .02, 5)
    dev = lambda a: torch.tensor(a, dtype=torch.float64, device=gpu_device)
    out = dict(grid=grid, w=w, u=u, ref=ref, base=base, U=dev(u), R=dev(ref), B=dev(base), dev=dev)
    for act in (.05, .9):
        out[act] = FusedQ1(Q1Problem(grid, w, 0., act), gpu_device)
    return out


def _sub(a, z0, z1):
    return np.ascontiguousarray(a[z0:z1 + 1])


def _model(shape, w, act):
    g = Q.Grid(shape, SPACING)
    return g, Q.Model(g, lam=w.lam, mu=w.mu, nu=w.nu, kappa=w.kappa, delta=0., activation=act)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("act", [.05, .9])
def test_native_energies_are_slab_additive_and_match_the_oracle(native, act):
    print("case native energies", act, flush=True)
    q = native[act]
    U, R, B = native["U"], native["R"], native["B"]
    soft = q.soft_energy(U, R, B).tolist()
    bar = q.barrier_energy(U, B).tolist()
    assert bar[1] > 0 and soft[0] > 0
    if act == .9:
        assert bar[0] > 0
    cuts = [0, 40, 80, 120, SHAPE[0] - 1]
    parts_s, parts_b, mins = np.zeros(3), np.zeros(2), []
    from seq_vp_3d_amd.problem import Q1Problem, make_grid_3d
    from seq_vp_3d_amd.q1 import FusedQ1
    for z0, z1 in zip(cuts, cuts[1:]):
        g = make_grid_3d(nx=SHAPE[2], ny=SHAPE[1], nz=z1 - z0 + 1, hx=1., hy=1., hz=1.)
        qs = FusedQ1(Q1Problem(g, native["w"], 0., act), U.device)
        sl = lambda t: t[z0:z1 + 1].contiguous()
        s = qs.soft_energy(sl(U), sl(R), sl(B)).tolist()
        b = qs.barrier_energy(sl(U), sl(B)).tolist()
        parts_s += s
        parts_b += (b[0], b[2])
        mins.append(b[1])
    for k in range(3):
        assert abs(soft[k] - parts_s[k]) <= 1e-12 * (soft[0] + soft[1]), k
    assert abs(bar[0] - parts_b[0]) <= 1e-12 * max(bar[0], 1e-300)
    assert abs(bar[2] - parts_b[1]) <= 1e-12 * max(abs(bar[2]), abs(bar[0]), 1e-300)
    assert bar[1] == min(mins)
    # one slab against the oracle
    z0, z1 = 40, 42
    g, m = _model((z1 - z0 + 1,) + SHAPE[1:], native["w"], act)
    jd, curl, diff = m.soft_energy(_sub(native["u"], z0, z1), _sub(native["ref"], z0, z1), _sub(native["base"], z0, z1))
    gs = make_grid_3d(nx=SHAPE[2], ny=SHAPE[1], nz=z1 - z0 + 1, hx=1., hy=1., hz=1.)
    qs = FusedQ1(Q1Problem(gs, native["w"], 0., act), U.device)
    s = qs.soft_energy(U[z0:z1 + 1].contiguous(), R[z0:z1 + 1].contiguous(), B[z0:z1 + 1].contiguous()).tolist()
    assert abs(s[0] - jd) <= 1e-12 * jd and abs(s[1] - curl) <= 1e-12 * curl and abs(s[2] - diff) <= 1e-12 * (jd + curl)
    val, jmin, chg, _ = m.barrier_energy(_sub(native["u"], z0, z1), _sub(native["base"], z0, z1))
    b = qs.barrier_energy(U[z0:z1 + 1].contiguous(), B[z0:z1 + 1].contiguous()).tolist()
    assert abs(b[0] - val) <= 1e-12 * max(val, 1e-300) and abs(b[1] - jmin) <= 1e-13 * jmin


def test_native_gradient_ggn_and_node_measure_against_the_oracle(native):
    print("case native gradient/GGN/node measure", flush=True)
    q = native[.9]
    U, R = native["U"], native["R"]
    grad = q.gradient_nodes(torch.zeros_like(U), U, R)
    base_g = torch.tensor(np.random.default_rng(7).normal(size=SHAPE + (3,)), device=U.device)
    diag = q.diagonal_nodes(base_g, U)
    node = native[.05].node_min_jd(U, extension_activation=.25)
    z0, z1 = 100, 104                       # node planes 101..103 have all their cells in the block
    g, m = _model((z1 - z0 + 1,) + SHAPE[1:], native["w"], .9)
    us, rs = _sub(native["u"], z0, z1), _sub(native["ref"], z0, z1)
    want = m.gradient_nodes(np.zeros_like(us), us, rs)
    got = grad[z0 + 1:z1].cpu().numpy()
    assert rel(got, want[1:-1]) <= 1e-12
    bs = _sub(base_g.cpu().numpy(), z0, z1)
    wdiag = m.ggn_diagonal_nodes(bs, us)
    assert rel(diag[z0 + 1:z1].cpu().numpy() - bs[1:-1], wdiag[1:-1] - bs[1:-1]) <= 1e-12
    wnode, _ = oja.node_min_jd(us, g, Q.q1_tables(g, "closed27_uniform")[3], extension_activation=.25)
    np.testing.assert_allclose(node[z0 + 1:z1].cpu().numpy(), wnode[1:-1], rtol=1e-12, atol=0)


def test_native_root_is_the_least_slab_root(native):
    from seq_vp_3d_amd.problem import Q1Problem, make_grid_3d
    from seq_vp_3d_amd.q1 import FusedJDRoot, FusedQ1
    print("case native T66 root", flush=True)
    q = native[.05]
    U = native["U"]
    v = smooth_field(SHAPE, 6., 8) + rough_field(SHAPE, .3, 9)
    V = native["dev"](v)
    root = FusedJDRoot(None, q1=q)
    for limit in (1., 8.):
        s, invalid, points = root(U, V, limit)
        assert points == 27 * np.prod([n - 1 for n in SHAPE])
        cuts = [0, 50, 100, SHAPE[0] - 1]
        slabs = []
        for z0, z1 in zip(cuts, cuts[1:]):
            g = make_grid_3d(nx=SHAPE[2], ny=SHAPE[1], nz=z1 - z0 + 1, hx=1., hy=1., hz=1.)
            qs = FusedQ1(Q1Problem(g, native["w"], 0., .05), U.device)
            slabs.append(FusedJDRoot(None, q1=qs)(U[z0:z1 + 1].contiguous(), V[z0:z1 + 1].contiguous(), limit))
        assert s == min(r[0] for r in slabs) and invalid == sum(r[1] for r in slabs)
    # a slab against the oracle
    z0, z1 = 50, 53
    g, m = _model((z1 - z0 + 1,) + SHAPE[1:], native["w"], .05)
    want = ojr.global_root(_sub(native["u"], z0, z1), _sub(v, z0, z1), g, m.closed[3], 0., 8.)
    gs = make_grid_3d(nx=SHAPE[2], ny=SHAPE[1], nz=z1 - z0 + 1, hx=1., hy=1., hz=1.)
    qs = FusedQ1(Q1Problem(gs, native["w"], 0., .05), U.device)
    got = FusedJDRoot(None, q1=qs)(U[z0:z1 + 1].contiguous(), V[z0:z1 + 1].contiguous(), 8.)
    assert got[1:] == want[1:]
    assert (np.isinf(got[0]) and np.isinf(want[0])) or abs(got[0] - want[0]) <= 1e-12 * want[0]


def test_native_dst_inverse_and_spline_warp(native):
    from seq_vp_3d_amd.poisson import DevicePoisson, SpectralMetric
    from seq_vp_3d_amd.spline import DeviceSpline
    print("case native DST and spline", flush=True)
    grid = native["grid"]
    og = Q.Grid(SHAPE, SPACING)
    rng = np.random.default_rng(10)
    x = rng.normal(size=3 * og.n_free_scalar)
    P = DevicePoisson(SpectralMetric(grid, .1, .03), native["U"].device)
    y = P.apply(native["dev"](x)).cpu().numpy()
    back = opo.matvec(y, og, .1, .03)
    assert rel(back, x) <= 1e-12
    # the spline warp at native size against SciPy's spline at 20,000 nodes
    img = smooth_field(SHAPE, 1., 11)[..., 0] + .05 * rng.normal(size=SHAPE)
    sp = DeviceSpline.from_image(native["dev"](img), grid)
    uw = smooth_field(SHAPE, 2., 12)
    value, gradient, bad = sp.warp(native["dev"](uw))
    assert int(bad.item()) == 0
    ref = osp.build(img, og)
    idx = tuple(rng.integers(0, n, 20000) for n in SHAPE)
    pts = og.identity()[idx] + uw[idx]
    wv, wg = osp.sample(ref, pts)
    assert rel(value.cpu().numpy()[idx], wv) <= 1e-12 and rel(gradient.cpu().numpy()[idx], wg) <= 1e-12


def test_native_gradient_is_the_derivative_of_the_energy(native):
    """g . v = d/dh [soft + barrier](u + h v) at h = 0, from central stable differences (no reference formula)."""
    print("case native directional derivative", flush=True)
    q = native[.9]
    U, R = native["U"], native["R"]
    v = native["dev"](smooth_field(SHAPE, 1., 13))
    grad = q.gradient_nodes(torch.zeros_like(U), U, R)
    slope = float((grad * v).sum())
    h = 1e-4
    up, um = (U + h * v).contiguous(), (U - h * v).contiguous()
    d_soft = q.soft_energy(up, R, um).tolist()[2]           # E(u + hv) - E(u - hv), stable
    d_bar = q.barrier_energy(up, um).tolist()[2] * q.barrier_weight
    fd = (d_soft + d_bar) / (2 * h)
    assert abs(fd - slope) <= 1e-6 * abs(slope), (fd, slope)


def test_native_pair_in_a_batched_trial_launch(native):
    """The native fields in one multi-pair persistent launch with two uneven pairs (96^3 and 128x96x80): the native
    pair's 7 trial results equal its single-pair sweep bit for bit, for a persistent and a one-item-per-workgroup
    launch shape."""
    from seq_vp_3d_amd import _lib, batch as B
    from seq_vp_3d_amd.problem import Q1Problem, make_grid_3d
    from seq_vp_3d_amd.q1 import OP_TRIAL, FusedQ1
    print("case native batched trial launch", flush=True)
    q = native[.05]
    dev = native["dev"]
    items = [(q, native["U"], native["R"], native["B"])]
    for i, shape in enumerate([(96, 96, 96), (80, 96, 128)]):
        nz, ny, nx = shape
        g = make_grid_3d(nx=nx, ny=ny, nz=nz, hx=1., hy=1., hz=1.)
        qs = FusedQ1(Q1Problem(g, native["w"], 0., .05), q.device)
        r = smooth_field(shape, 2., 10 + i)
        b = r + smooth_field(shape, 1., 20 + i)
        items.append((qs, dev(b + rough_field(shape, .01, 30 + i)), dev(r), dev(b)))
    want = [qq.trial_energy(u, r, b) for qq, u, r, b in items]
    handle = _lib.stream_handle(q.device)
    for tile, grid in ((1, 0), (2, 1024)):
        outs = [torch.full((7,), float("nan"), dtype=torch.float64, device=q.device) for _ in items]
        wss = [torch.empty(_lib.workspace_bytes(OP_TRIAL, qq.shape), dtype=torch.uint8, device=q.device)
               for qq, *_ in items]
        B.trial_energy_batch([B.entry(qq, u, r, b, o, ws) for (qq, u, r, b), o, ws in zip(items, outs, wss)], tile, grid,
                             handle)
        for o, w_ in zip(outs, want):
            assert torch.equal(o, w_), (tile, grid)
