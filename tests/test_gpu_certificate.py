"""GPU certificate screen (edge hull + prescreen) and the device/CPU certificate against the oracle."""
import numpy as np
import pytest

from conftest import counterexample_cell, rough_field, smooth_field
from oracle import build as obuild, certificate as oc, prescreen as ops
from seq_vp_3d_amd.problem import make_grid_3d, q1_coefficients

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _oracle_built():
    obuild.build()


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def _zero_boundary(u):
    u[[0, -1]] = 0.
    u[:, [0, -1]] = 0.
    u[:, :, [0, -1]] = 0.
    return u


def folding_field(grid, seed=84):
    """A field exercising every branch of the per-cell interval arithmetic: smooth and rough parts, constant
    blocks (exact zero differences), subnormal cells, T05 cells (J <= 0 inside with positive corners) and one
    node with a huge displacement (overflow: flagged cells)."""
    rng = np.random.default_rng(seed)
    h = np.array(grid.spacing)
    nz, ny, nx = grid.shape
    zz, yy, xx = np.meshgrid(*(np.linspace(0., 1., n) for n in grid.shape), indexing="ij")
    u = (.08 * np.sin(np.pi * xx) * np.sin(2 * np.pi * yy) * np.sin(np.pi * zz))[..., None] * h
    u = u + rng.normal(size=grid.shape + (3,)) * (.02 * h)
    for k in range(12):
        z, y, x = (int(rng.integers(1, n - 3)) for n in grid.shape)
        if k % 4 == 0:
            u[z:z + 2, y:y + 2, x:x + 2] += rng.normal(size=(2, 2, 2, 3)) * (.6 * h)
        elif k % 4 == 1:
            u[z:z + 2, y:y + 2, x:x + 2] = u[z, y, x]
        elif k % 4 == 2:
            u[z:z + 2, y:y + 2, x:x + 2] = rng.normal(size=(2, 2, 2, 3)) * 5e-320
        else:
            v = counterexample_cell(-.05 if k % 8 == 3 else 0.) * h
            for a in range(8):
                u[z + (a >> 2), y + ((a >> 1) & 1), x + (a & 1)] = v[a]
    u[nz // 2, ny // 2, nx // 2] = (1e300, -1e300, 1e300)
    return np.ascontiguousarray(_zero_boundary(u))


def test_edge_hull_margins_are_bitwise_the_oracle(gpu_device):
    import torch
    from seq_vp_3d_amd.certificate import edge_hull
    grid = make_grid_3d(nx=22, ny=21, nz=20, hx=1.0, hy=1.2, hz=.8)
    for delta in (0., .01):
        for name, u in (("folding", folding_field(grid)), ("smooth", _zero_boundary(smooth_field(grid.shape, .8, 1))),
                        ("rough", rough_field(grid.shape, .3, 2))):
            print("case edge_hull", name, delta, flush=True)
            m, f = edge_hull(torch.tensor(u, device=gpu_device), grid, delta)
            want_m, want_f = oc.edge_hull_margins(oc.cell_vertices(u, grid.shape), grid.spacing, delta)
            f = f.cpu().numpy().astype(bool)
            assert np.array_equal(f, want_f), name
            ok = ~want_f
            assert np.array_equal(_bits(m.cpu().numpy()[ok]), _bits(want_m[ok])), name
            if name == "folding":
                assert want_f.any() and (want_m[ok] <= 0).any() and (want_m[ok] > 0).any()


def _codes_oracle(u, grid, D, delta, tile):
    verts = oc.cell_vertices(u, grid.shape)
    margin, flag = oc.edge_hull_margins(verts, grid.spacing, delta)
    n = grid.n_cells
    hull = np.zeros(n, dtype=bool)
    for s in range(0, n, tile):
        if not flag[s:s + tile].any():
            hull[s:s + tile] = margin[s:s + tile] > 0
    low = ops.lower_bounds_c(u, grid, D)
    with np.errstate(invalid="ignore"):
        pm = low if delta == 0 else np.nextafter(low - delta, -np.inf)
    pre = ~hull & (pm > 0)
    codes = np.where(hull, 0, np.where(pre, 1, 2)).astype(np.uint8)
    return codes, margin, pm


@pytest.mark.parametrize("tile", [8192, 301])
def test_screen_codes_and_summary(gpu_device, tile):
    import torch
    from seq_vp_3d_amd.certificate import CertificateConfig, DeviceCertificate
    grid = make_grid_3d(nx=22, ny=21, nz=20, hx=1.0, hy=1.2, hz=.8)
    D = q1_coefficients(grid.spacing, "closed27_uniform")[0]
    u = folding_field(grid, seed=5)
    u[grid.shape[0] // 2, grid.shape[1] // 2, grid.shape[2] // 2] = 0.     # keep the domain valid
    for delta in (0., .01):
        print("case screen", tile, delta, flush=True)
        cert = DeviceCertificate(grid, gpu_device, CertificateConfig(delta=delta, tile_size=tile), keep_codes=True)
        cert.screen(torch.tensor(u, device=gpu_device))
        s = cert.summary.cpu().numpy()
        codes, margin, pm = _codes_oracle(u, grid, D, delta, tile)
        assert np.array_equal(cert.codes.cpu().numpy(), codes)
        assert s[2] == (codes == 0).sum() and s[3] == (codes == 1).sum() and s[4] == (codes == 2).sum()
        assert _bits([s[0]])[0] == _bits([margin[codes == 0].min()])[0]
        if (codes == 1).any():
            assert _bits([s[1]])[0] == _bits([pm[codes == 1].min()])[0]
        listed = np.sort(cert.undecided[:int(s[4])].cpu().numpy())
        assert np.array_equal(listed, np.flatnonzero(codes == 2))
        assert (codes == 2).sum() > 0 and (codes == 1).sum() > 0


def _device_cases():
    grid = make_grid_3d(nx=14, ny=13, nz=12, hx=1.0, hy=1.1, hz=.9)
    cases = [(f"smooth{a}", grid, _zero_boundary(smooth_field(grid.shape, a, 3))) for a in (.3, .7, 1.0)]
    cases.append(("rough", grid, rough_field(grid.shape, .12, 4)))
    g4 = make_grid_3d(nx=4, ny=4, nz=4, hx=1., hy=1., hz=1.)
    z, y, x = np.meshgrid(*(np.arange(4, dtype=float),) * 3, indexing="ij")
    for a in (.8, 1.05):
        u = np.zeros(g4.shape + (3,))
        u[..., 1] = 2 * a * (x - 1.5) * (z - 1.5)
        u[..., 2] = -2 * a * (x - 1.5) * (y - 1.5)
        cases.append((f"shear{a}", g4, _zero_boundary(u)))
    g2 = make_grid_3d(nx=9, ny=8, nz=7, hx=1.0, hy=1e-308, hz=1.0)
    u = np.zeros(g2.shape + (3,))
    u[3, 3:5, 4, 0] = 2.
    cases.append(("flagged", g2, u))
    return cases


@pytest.mark.parametrize("case", _device_cases(), ids=lambda c: c[0])
def test_device_certificate_is_the_cpu_certificate(gpu_device, case):
    import torch
    from seq_vp_3d_amd.certificate import CertificateConfig, DeviceCertificate, certify_map
    name, grid, u = case
    for delta, tile in ((.01, 8192), (0., 97)):
        print("case device_certificate", name, delta, tile, flush=True)
        cfg = CertificateConfig(delta=delta, tile_size=tile)
        got = DeviceCertificate(grid, gpu_device, cfg, capacity=2)(torch.tensor(u, device=gpu_device))
        want = certify_map(u, grid, cfg)
        assert got == want, (name, got, want)
        D = None
        try:
            D = q1_coefficients(grid.spacing, "closed27_uniform")[0]
        except FloatingPointError:
            pass
        o = oc.certify_map(u, grid.shape, grid.spacing, grid.axes, delta, D=D, tile=tile)
        assert (got.status, got.reason, got.n_boxes, got.n_fast_cells, got.n_prescreen_cells) == \
               (o["status"], o["reason"], o["n_boxes"], o["n_fast_cells"], o["n_prescreen_cells"])


def test_native_size_certificate_matches_the_cpu(gpu_device):
    """160x192x224: the device certificate of a folding candidate equals the CPU one field for field."""
    import torch
    from seq_vp_3d_amd.certificate import CertificateConfig, DeviceCertificate, certify_map
    grid = make_grid_3d(nx=224, ny=192, nz=160, hx=1., hy=1., hz=1.)
    u = _zero_boundary(smooth_field(grid.shape, 9., 6, modes=4))
    cfg = CertificateConfig(delta=0.)
    print("case native certificate", flush=True)
    got = DeviceCertificate(grid, gpu_device, cfg)(torch.tensor(u, device=gpu_device))
    want = certify_map(u, grid, cfg, workers=16)
    assert got == want
    assert got.n_fast_cells > 0 and got.n_cells_checked == grid.n_cells
