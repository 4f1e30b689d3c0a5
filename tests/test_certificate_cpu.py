"""The C++ certificate of libseqvp_hip.so (host entry points, no GPU) against the oracle, bit for bit."""
import numpy as np
import pytest

from conftest import counterexample_cell, rough_field, smooth_field
from oracle import build as obuild, certificate as oc
from seq_vp_3d_amd import certificate as C
from seq_vp_3d_amd.problem import make_grid_3d, q1_coefficients

H = (1.0, 1.25, 0.75)


@pytest.fixture(scope="module", autouse=True)
def _oracle_built():
    obuild.build()


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def _cells():
    rng = np.random.default_rng(11)
    sets = [rng.normal(scale=s, size=(150, 8, 3)) for s in (.02, .15, .3, .6)]
    sets.append(rng.normal(size=(20, 8, 3)) * 5e-320)                   # subnormal differences
    sets.append(np.repeat(rng.normal(size=(10, 1, 3)), 8, axis=1))       # constant cells: exact zero differences
    big = rng.normal(size=(10, 8, 3))
    big[:, 3] *= 1e200                                                   # overflowing products
    sets.append(big)
    sets.append(np.stack([counterexample_cell(d) for d in (-.05, 0., .3, .6)]))
    return np.concatenate(sets)


CELLS = _cells()


@pytest.mark.parametrize("delta", [0., .01])
def test_fallback_arrays_are_bitwise_the_oracle(delta):
    plo, phi, clo, chi, s, ok = C.fallback_arrays(CELLS, H, delta)
    good = CELLS[:-14]                      # every cell before the overflow block is finite throughout
    p, c, s_o, ok_o = oc.bernstein_arrays(good, H, delta)
    assert ok_o.all() and not ok            # the overflow block makes the whole call fail, as one job
    n = len(good)
    for got, want in ((plo, p[0]), (phi, p[1]), (clo, c[0]), (chi, c[1]), (s, s_o)):
        assert np.array_equal(_bits(got[:n]), _bits(want))
    *_, ok_good = C.fallback_arrays(good, H, delta)
    assert ok_good


@pytest.mark.parametrize("delta", [0., .01])
def test_cell_classification_is_the_oracle_field_for_field(delta):
    cells = np.concatenate([CELLS[:-14], CELLS[-4:]])
    p, c, _, ok = oc.bernstein_arrays(cells, H, delta)
    seen = set()
    cfg = C.CertificateConfig(delta=delta)
    for k in range(len(cells)):
        want = oc.classify((p[0][k], p[1][k]), (c[0][k], c[1][k]), delta)
        got = C.certify_cell(cells[k], H, cfg)
        seen.add(want.reason)
        assert (got.status, got.reason, got.n_boxes, got.max_depth, got.witness_reference_point) == \
               (want.status, want.reason, want.boxes, want.depth, want.witness), k
        assert _bits([got.min_margin_lower_bound])[0] == _bits([want.margin])[0]
        assert _bits([got.min_jd_lower_bound])[0] == _bits([want.lower])[0]
    assert {"positive_bernstein_cover", "sample_witness"} <= seen


def test_budgets_and_subdivision_match():
    v = counterexample_cell(.3)
    for kw in (dict(max_boxes=1), dict(max_depth=0), dict(max_depth=1), dict()):
        p, c, _, _ = oc.bernstein_arrays(v[None], (1., 1., 1.), 0.)
        want = oc.classify((p[0][0], p[1][0]), (c[0][0], c[1][0]), 0., **kw)
        got = C.certify_cell(v, (1., 1., 1.), C.CertificateConfig(delta=0., **kw))
        assert (got.status, got.reason, got.n_boxes, got.max_depth) == (want.status, want.reason, want.boxes, want.depth)
    v = counterexample_cell(0.)
    got = C.certify_cell(v, (1., 1., 1.), C.CertificateConfig(delta=0.))
    assert got.status == "uncertain" and got.reason == "box_budget" and got.n_boxes == 255


def _map_cases():
    grid = make_grid_3d(nx=14, ny=13, nz=12, hx=1.0, hy=1.1, hz=.9)
    shape = grid.shape
    cases = []
    for amp in (.3, .55, .7, 1.0, 2.5):
        u = smooth_field(shape, amp, 3)
        u[[0, -1]] = 0.
        u[:, [0, -1]] = 0.
        u[:, :, [0, -1]] = 0.
        cases.append((f"smooth{amp}", grid, u))
    cases.append(("rough", grid, rough_field(shape, .12, 4)))
    u = smooth_field(shape, .3, 3)
    u[2, 3, 3, 1] = np.nan
    cases.append(("nonfinite", grid, u))
    u = np.zeros(shape + (3,))
    u[0, 4, 4] = .1
    cases.append(("boundary", grid, u))
    # anisotropic spacing: x-displacement differences divided by h_y = 1e-308 overflow (flagged tiles; the
    # fallback's arrays fail: arithmetic_uncertain)
    g2 = make_grid_3d(nx=9, ny=8, nz=7, hx=1.0, hy=1e-308, hz=1.0)
    u = np.zeros(g2.shape + (3,))
    u[3, 3:5, 4, 0] = 2.
    cases.append(("flagged", g2, u))
    # a shear whose middle cell has J = 1 + 4a^2 (x - 3/2)^2 with a negative root Bernstein coefficient
    # (subdivision in the fallback) and folding neighbours
    g4 = make_grid_3d(nx=4, ny=4, nz=4, hx=1., hy=1., hz=1.)
    z, y, x = np.meshgrid(*(np.arange(4, dtype=float),) * 3, indexing="ij")
    for a in (.8, 1.05):
        u = np.zeros(g4.shape + (3,))
        u[..., 1] = 2 * a * (x - 1.5) * (z - 1.5)
        u[..., 2] = -2 * a * (x - 1.5) * (y - 1.5)
        u[[0, -1]] = 0.
        u[:, [0, -1]] = 0.
        u[:, :, [0, -1]] = 0.
        cases.append((f"shear{a}", g4, u))
    return cases


@pytest.mark.parametrize("case", _map_cases(), ids=lambda c: c[0])
@pytest.mark.parametrize("delta,tile,prescreen", [(.01, 8192, True), (0., 97, True), (.01, 211, False)])
def test_map_certificate_is_the_oracle_field_for_field(case, delta, tile, prescreen):
    name, grid, u = case
    try:
        D = q1_coefficients(grid.spacing, "closed27_uniform")[0] if prescreen else None
    except FloatingPointError:
        D = None
    want = oc.certify_map(u, grid.shape, grid.spacing, grid.axes, delta, D=D, tile=tile)
    got = C.certify_map(u, grid, C.CertificateConfig(delta=delta, tile_size=tile, prescreen=prescreen), workers=4)
    for key, value in want.items():
        g = getattr(got, key)
        if isinstance(value, float):
            assert _bits([g])[0] == _bits([value])[0], key
        else:
            assert g == value, (key, g, value)


def test_map_cases_cover_every_stage():
    """The cases above reach the hull, the prescreen, the fallback's subdivision and its witnesses, and the
    flagged-tile rule (so the field-for-field test is not vacuous)."""
    seen = set()
    for name, grid, u in _map_cases():
        try:
            D = q1_coefficients(grid.spacing, "closed27_uniform")[0]
        except FloatingPointError:
            D = None
        r = oc.certify_map(u, grid.shape, grid.spacing, grid.axes, .01, D=D)
        seen.add(r["reason"])
        if r["n_fast_cells"]:
            seen.add("hull")
        if r["n_prescreen_cells"]:
            seen.add("prescreen")
        if r["n_boxes"] > r["n_cells_checked"]:
            seen.add("subdivision")
    assert {"hull", "prescreen", "subdivision", "JACOBIAN_VIOLATION", "arithmetic_uncertain",
            "NONFINITE_STATE", "NONZERO_BOUNDARY", "INVALID_DOMAIN", "positive_enclosure_cover"} <= seen, seen


@pytest.mark.parametrize("delta", [0., .01])
def test_host_edge_hull_is_the_oracle_per_cell(delta):
    """The host edge hull (four cells per AVX2 evaluation where the CPU has it, the row ends one by one) against
    the oracle, cell by cell: every unflagged margin bitwise, every flag equal. Rows of 13 cells mix both paths;
    the field mixes smooth, rough, zero, huge, tiny and nonfinite vertex values."""
    from seq_vp_3d_amd import _lib
    rng = np.random.default_rng(11)
    shape = (9, 11, 14)
    spacing = np.array([1.1, .7, 1.3])
    u = smooth_field(shape, .8, 5) + rough_field(shape, .2, 6)
    flat = u.reshape(-1, 3)
    pick = rng.choice(flat.shape[0], 60, replace=False)
    flat[pick[:10]] = 0.
    flat[pick[10:20]] *= 1e300
    flat[pick[20:30]] *= 1e-300
    flat[pick[30:33], 1] = np.nan
    flat[pick[33:36], 2] = np.inf
    u = np.ascontiguousarray(u)
    nz, ny, nx = shape
    ncell = (nx - 1) * (ny - 1) * (nz - 1)
    margin = np.empty(ncell)
    flag = np.empty(ncell, dtype=np.uint8)
    _lib.call("seqvp_edge_hull_host", _lib.hptr(u), nx, ny, nz, _lib.hptr(spacing), delta, 3, _lib.hptr(margin),
              _lib.hptr(flag))
    want, bad = oc.edge_hull_margins(oc.cell_vertices(u, shape), spacing, delta)
    assert np.array_equal(flag.astype(bool), bad)
    assert bad.any() and (~bad).sum() > ncell // 2
    np.testing.assert_array_equal(_bits(margin[~bad]), _bits(want[~bad]))
