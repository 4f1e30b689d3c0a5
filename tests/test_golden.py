"""The oracle against the reference's own outputs (tests/golden/reference_vectors.*, written by
tests/golden/make_golden.py from jd_curl.three_d's CPU functions).

Covers the T62/T69 node JD measure and its estimator rows (solvers/jd_affine.py:96-225), the T66 global root
(solvers/global_jd_step.py:107-129), the witness prediction and first_root_before (solvers/jd_step.py:13-37,
177-239), the Gauss-point geometry (geometry.py:60-89), the DST-I Poisson inverse (poisson.py:99-131) and the
whole-cell certificate with EnclosureCertificateConfig (feasibility.py:393-506).
"""
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import certificate as oc, jd_affine as oja, jd_root as ojr, jd_step as ojs, poisson as opo, q1 as Q

GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def golden():
    arrays = np.load(GOLDEN / "reference_vectors.npz")
    meta = json.loads((GOLDEN / "reference_vectors.json").read_text())
    return {k: arrays[k] for k in arrays.files}, meta


def _grid(meta):
    return Q.Grid(tuple(meta["shape"]), tuple(meta["spacing"]))


def _hex(x):
    return None if x is None else float.fromhex(x)


@pytest.mark.parametrize("name", ["mild", "folded", "dense"])
def test_node_min_jd_t62_t69(golden, name):
    arrays, meta = golden
    g = _grid(meta)
    D = Q.q1_tables(g, "closed27_uniform")[3]
    u = arrays[f"u_{name}"]
    node, _ = oja.node_min_jd(u, g, D)
    np.testing.assert_allclose(node, arrays[f"t62_{name}"], rtol=1e-13, atol=0)
    node, stats = oja.node_min_jd(u, g, D, extension_activation=.25)
    np.testing.assert_allclose(node, arrays[f"t69_{name}"], rtol=1e-12, atol=0)
    want = meta["t69_stats"][name]
    for key in ("candidates", "lowered_cells"):
        assert stats[key] == want[key], key
    # lowered_cells_any counts mu_c < m_c "including rounding" (jd_affine.py:196-197): a cell whose estimate
    # equals its closed27/edge minimum to rounding may count on one side and not the other
    _, _, J, _ = Q.geometry(u, g, D)
    m = oja.cell_minima(J)
    cand = oja.candidates(J, m, .25)
    mu = oja.face_interior_minima(J[cand])
    ambiguous = int(np.sum(np.abs(mu - m[cand]) <= 8 * np.finfo(float).eps * np.abs(m[cand])))
    assert abs(stats["lowered_cells_any"] - want["lowered_cells_any"]) <= ambiguous
    assert abs(stats["mu_min"] - want["mu_min"]) <= 1e-12 * abs(want["mu_min"])
    # the estimator on the reference's own rows
    rows = arrays[f"t69_rows_{name}"]
    mu = oja.face_interior_minima(rows)
    np.testing.assert_allclose(mu, arrays[f"t69_mu_{name}"], rtol=1e-12, atol=1e-15)


def test_t66_global_root(golden):
    arrays, meta = golden
    g = _grid(meta)
    D = Q.q1_tables(g, "closed27_uniform")[3]
    v = Q.inject_free(arrays["t66_direction"], g)
    for rec in meta["t66"]:
        s, invalid, points = ojr.global_root(arrays["t66_base"], v, g, D, 0., rec["limit"])
        assert (invalid, points) == (rec["invalid_base_points"], rec["points"])
        want = _hex(rec["root"])
        if want is None:
            assert np.isinf(s)
        else:
            assert abs(s * rec["limit"] - want) <= 1e-12 * want, (rec["limit"], s * rec["limit"], want)


def test_witness_prediction_and_first_roots(golden):
    arrays, meta = golden
    g = _grid(meta)
    mem = ojs.Memory()
    for cell, point in meta["witnesses"]:
        mem.witness(cell, point)
    v = Q.inject_free(arrays["t66_direction"], g)
    for rec in meta["predictions"]:
        tau, root, fallback = mem.predict(arrays["t66_base"], v, g, 0., rec["limit"], .9)
        assert fallback == rec["fallback"]
        assert abs(tau - _hex(rec["tau"])) <= 1e-12 * _hex(rec["tau"])
        assert (root is None) == (rec["root"] is None)
        if root is not None:
            assert abs(root - _hex(rec["root"])) <= 1e-12 * _hex(rec["root"])
    for c, row in zip(arrays["first_root_coefficients"], meta["first_roots"]):
        for limit, want in zip((.3, 1., 5.), row):
            got = ojs.first_root_before(c, limit)
            assert (got is None) == (want is None)
            if got is not None:
                assert got == _hex(want), (c, limit)


def test_gauss_geometry(golden):
    arrays, meta = golden
    g = _grid(meta)
    D = Q.q1_tables(g, "gauss27")[3]
    _, _, J, curl = Q.geometry(arrays["u_folded"], g, D, np.arange(100, 160))
    np.testing.assert_allclose(J, arrays["gauss_J_folded"], rtol=0, atol=1e-14)
    np.testing.assert_allclose(curl, arrays["gauss_curl_folded"], rtol=0, atol=1e-14)


def test_poisson_inverse(golden):
    arrays, meta = golden
    g = _grid(meta)
    for i, (a, b) in enumerate(meta["poisson_ab"]):
        got = opo.apply(arrays["poisson_rhs"], opo.denominator(g, a, b))
        want = arrays[f"poisson_out_{i}"]
        assert np.linalg.norm(got - want) <= 1e-14 * np.linalg.norm(want), (a, b)


CERT_KEYS = ("status", "reason", "min_jd_lower_bound", "min_margin_lower_bound", "sampled_min_jd", "n_cells_checked",
             "n_uncertain", "n_infeasible", "n_boxes", "max_depth", "witness_cell", "witness_reference_point",
             "n_fast_cells", "n_fallback_cells")


def test_certificate_records(golden):
    """The oracle without the prescreen is the reference's EnclosureCertificateConfig route: every record field
    equal (bounds bit for bit)."""
    arrays, meta = golden
    shape = tuple(meta["cert_shape"])
    g = Q.Grid(shape, tuple(meta["spacing"]))
    for key, want in meta["certificates"].items():
        name, delta = key.split("/")
        got = oc.certify_map(arrays[f"cert_u_{name}"], shape, g.spacing, g.axes, float(delta))
        for k in CERT_KEYS:
            w = want[k]
            if k in ("min_jd_lower_bound", "min_margin_lower_bound", "sampled_min_jd"):
                w = _hex(w)
            if k == "witness_reference_point" and w is not None:
                w = tuple(w)
                assert tuple(float(x) for x in got[k]) == w, (key, k)
                continue
            assert got[k] == w, (key, k, got[k], w)


def test_host_certificate_records(golden):
    """The library's C++ CPU certificate without the prescreen is the same route: every field equal, bounds bit for
    bit (boundary and sampling-domain refusals included)."""
    from seq_vp_3d_amd.certificate import CertificateConfig, certify_map
    from seq_vp_3d_amd.problem import make_grid_3d
    arrays, meta = golden
    nz, ny, nx = meta["cert_shape"]
    hx, hy, hz = meta["spacing"]
    grid = make_grid_3d(nx=nx, ny=ny, nz=nz, hx=hx, hy=hy, hz=hz)
    for key, want in meta["certificates"].items():
        name, delta = key.split("/")
        got = certify_map(arrays[f"cert_u_{name}"], grid, CertificateConfig(delta=float(delta), prescreen=False))
        for k in CERT_KEYS:
            w = want[k]
            if k in ("min_jd_lower_bound", "min_margin_lower_bound", "sampled_min_jd"):
                w = _hex(w)
            g = getattr(got, k)
            if k == "witness_reference_point" and w is not None:
                w, g = tuple(w), tuple(float(x) for x in g)
            assert g == w, (key, k, g, w)


def test_device_memory_predict_records(golden):
    """The library's witness memory (host logic of the device predictor), fed the oracle's cubics, makes the
    reference's records: tau, least root, fallback and the point count (cells finished before a failed cubic)."""
    from seq_vp_3d_amd.jd_step import JDStepMemory
    from seq_vp_3d_amd.problem import make_grid_3d
    arrays, meta = golden
    g = _grid(meta)
    nz, ny, nx = g.shape
    grid = make_grid_3d(nx=nx, ny=ny, nz=nz, hx=g.spacing[0], hy=g.spacing[1], hz=g.spacing[2])
    u, v = arrays["t66_base"], Q.inject_free(arrays["t66_direction"], g)

    def coefficients(cells, points):
        rows = []
        for cell in np.unique(cells):
            pts = points[cells == cell]
            A = ojs.point_gradients(u, g, cell, pts)
            V = ojs.point_gradients(v, g, cell, pts) - np.eye(3)
            rows.append(ojs.determinant_coefficients(A, V, 0.))
        return np.concatenate(rows)

    mem = JDStepMemory()
    mem.bind(grid)
    for cell, point in meta["witnesses"]:
        mem.witness(cell, point)
    for rec in meta["predictions"]:
        tau, info = mem.predict(coefficients, rec["limit"], .9)
        assert info["fallback"] == rec["fallback"] and info["points"] == rec["points"], (rec, info)
        assert abs(tau - _hex(rec["tau"])) <= 1e-12 * _hex(rec["tau"])
        assert (info["root"] is None) == (rec["root"] is None)
