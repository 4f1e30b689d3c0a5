"""Device entry points against the reference's own outputs (tests/golden/reference_vectors.*, written by
tests/golden/make_golden.py from jd_curl.three_d's CPU functions): T62/T69 node JD measure
(solvers/jd_affine.py:154-225), T66 global root (solvers/global_jd_step.py:107-129), the witness prediction
(solvers/jd_step.py:177-239), the DST-I Poisson inverse (poisson.py:99-131) and the whole-cell certificate
(feasibility.py:393-506, EnclosureCertificateConfig)."""
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import jd_affine as oja, q1 as Q

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def golden():
    arrays = np.load(GOLDEN / "reference_vectors.npz")
    meta = json.loads((GOLDEN / "reference_vectors.json").read_text())
    return {k: arrays[k] for k in arrays.files}, meta


def _hex(x):
    return None if x is None else float.fromhex(x)


def _setup(meta, device, shape=None):
    from seq_vp_3d_amd.problem import Q1Problem, Weights3D, make_grid_3d
    from seq_vp_3d_amd.q1 import FusedQ1
    nz, ny, nx = shape or meta["shape"]
    hx, hy, hz = meta["spacing"]
    grid = make_grid_3d(nx=nx, ny=ny, nz=nz, hx=hx, hy=hy, hz=hz)
    prob = Q1Problem(grid, Weights3D(lam=1e-3, mu=.03, nu=.03, kappa=.002), 0., .05)
    return grid, prob, FusedQ1(prob, device)


def _t(a, device):
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device=device)


@pytest.mark.parametrize("name", ["mild", "folded", "dense"])
def test_node_min_jd_against_reference(gpu_device, golden, name):
    arrays, meta = golden
    print("case golden node_min_jd", name, flush=True)
    grid, prob, q = _setup(meta, gpu_device)
    U = _t(arrays[f"u_{name}"], gpu_device)
    np.testing.assert_allclose(q.node_min_jd(U).cpu().numpy(), arrays[f"t62_{name}"], rtol=1e-13, atol=0)
    stats = torch.zeros(4, dtype=torch.float64, device=gpu_device)
    node = q.node_min_jd(U, extension_activation=.25, stats=stats).cpu().numpy()
    np.testing.assert_allclose(node, arrays[f"t69_{name}"], rtol=1e-12, atol=0)
    cand, lowered, lowered_any, mu_min = stats.tolist()
    want = meta["t69_stats"][name]
    assert (int(cand), int(lowered)) == (want["candidates"], want["lowered_cells"])
    # lowered_cells_any includes cells lowered by rounding only (jd_affine.py:196-197): it may differ by the cells
    # whose estimate equals their closed27/edge minimum to rounding
    og = Q.Grid(tuple(meta["shape"]), tuple(meta["spacing"]))
    _, _, J, _ = Q.geometry(arrays[f"u_{name}"], og, Q.q1_tables(og, "closed27_uniform")[3])
    m = oja.cell_minima(J)
    c = oja.candidates(J, m, .25)
    mu = oja.face_interior_minima(J[c])
    ambiguous = int(np.sum(np.abs(mu - m[c]) <= 8 * np.finfo(float).eps * np.abs(m[c])))
    assert abs(int(lowered_any) - want["lowered_cells_any"]) <= ambiguous
    assert abs(mu_min - want["mu_min"]) <= 1e-12 * abs(want["mu_min"])


def test_jd_root_against_reference(gpu_device, golden):
    from seq_vp_3d_amd.q1 import FusedJDRoot
    arrays, meta = golden
    print("case golden T66 root", flush=True)
    grid, prob, q = _setup(meta, gpu_device)
    og = Q.Grid(tuple(meta["shape"]), tuple(meta["spacing"]))
    root = FusedJDRoot(prob, q1=q)
    U = _t(arrays["t66_base"], gpu_device)
    V = _t(Q.inject_free(arrays["t66_direction"], og), gpu_device)
    P = _t(arrays["t66_direction"], gpu_device)
    for rec in meta["t66"]:
        s, invalid, points = root(U, V, rec["limit"])
        assert (invalid, points) == (rec["invalid_base_points"], rec["points"])
        assert root.from_free(U, P, rec["limit"]) == (s, invalid, points)
        want = _hex(rec["root"])
        if want is None:
            assert np.isinf(s)
        else:
            assert abs(s * rec["limit"] - want) <= 1e-12 * want, (rec["limit"], s * rec["limit"], want)


def test_witness_prediction_against_reference(gpu_device, golden):
    from seq_vp_3d_amd import _lib
    from seq_vp_3d_amd.jd_step import JDStepMemory
    arrays, meta = golden
    print("case golden witness prediction", flush=True)
    grid, prob, q = _setup(meta, gpu_device)
    og = Q.Grid(tuple(meta["shape"]), tuple(meta["spacing"]))
    U = _t(arrays["t66_base"], gpu_device)
    V = _t(Q.inject_free(arrays["t66_direction"], og), gpu_device)
    nz, ny, nx = grid.shape

    def coefficients(cells, points):
        c = torch.as_tensor(np.ascontiguousarray(cells, dtype=np.int64)).to(gpu_device)
        p = torch.as_tensor(np.ascontiguousarray(points, dtype=np.float64)).to(gpu_device)
        out = torch.empty((len(cells), 4), dtype=torch.float64, device=gpu_device)
        _lib.call("seqvp_witness_coefficients", _lib.ptr(U), _lib.ptr(V), _lib.ptr(q.k), nx, ny, nz, _lib.ptr(c),
                  _lib.ptr(p), len(cells), _lib.ptr(out), _lib.stream_handle(gpu_device))
        return out.cpu().numpy()

    mem = JDStepMemory()
    mem.bind(grid)
    for cell, point in meta["witnesses"]:
        mem.witness(cell, point)
    for rec in meta["predictions"]:
        tau, info = mem.predict(coefficients, rec["limit"], .9)
        assert info["fallback"] == rec["fallback"] and info["points"] == rec["points"]
        assert abs(tau - _hex(rec["tau"])) <= 1e-12 * _hex(rec["tau"]), rec["limit"]
        assert (info["root"] is None) == (rec["root"] is None)
        if rec["root"] is not None:
            assert abs(info["root"] - _hex(rec["root"])) <= 1e-12 * _hex(rec["root"])


def test_poisson_against_reference(gpu_device, golden):
    from seq_vp_3d_amd.poisson import DevicePoisson, build_poisson_inverse
    arrays, meta = golden
    grid, _, _ = _setup(meta, gpu_device)
    for i, (a, b) in enumerate(meta["poisson_ab"]):
        got = DevicePoisson(build_poisson_inverse(grid, a=a, b=b), gpu_device).apply(
            _t(arrays["poisson_rhs"], gpu_device)).cpu().numpy()
        want = arrays[f"poisson_out_{i}"]
        assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want), (a, b)


CERT_KEYS = ("status", "reason", "min_jd_lower_bound", "min_margin_lower_bound", "sampled_min_jd", "n_cells_checked",
             "n_uncertain", "n_infeasible", "n_boxes", "max_depth", "witness_cell", "witness_reference_point",
             "n_fast_cells", "n_fallback_cells")


def test_certificate_against_reference(gpu_device, golden):
    """Without the prescreen the device certificate is the reference's EnclosureCertificateConfig route: every
    record field equal, bounds bit for bit. With the prescreen (the default) the prescreen only adds proofs:
    status, reason, witness and the fast cells are the same and every cell is accounted for."""
    from seq_vp_3d_amd.certificate import CertificateConfig, DeviceCertificate
    arrays, meta = golden
    print("case golden certificate", flush=True)
    grid, _, _ = _setup(meta, gpu_device, shape=meta["cert_shape"])
    for key, want in meta["certificates"].items():
        if want["reason"] in ("INVALID_DOMAIN", "NONZERO_BOUNDARY"):
            continue        # the device screen certifies in-domain trial maps; the host route checks the domain
        name, delta = key.split("/")
        U = _t(arrays[f"cert_u_{name}"], gpu_device)
        got = DeviceCertificate(grid, gpu_device, CertificateConfig(delta=float(delta), prescreen=False))(U)
        for k in CERT_KEYS:
            w = want[k]
            if k in ("min_jd_lower_bound", "min_margin_lower_bound", "sampled_min_jd"):
                w = _hex(w)
            g = getattr(got, k)
            if k == "witness_reference_point" and w is not None:
                w, g = tuple(w), tuple(float(x) for x in g)
            assert g == w, (key, k, g, w)
        pre = DeviceCertificate(grid, gpu_device, CertificateConfig(delta=float(delta)))(U)
        assert (pre.status, pre.reason, pre.witness_cell, pre.n_fast_cells, pre.n_infeasible) == \
               (want["status"], want["reason"], want["witness_cell"], want["n_fast_cells"], want["n_infeasible"])
        assert pre.n_prescreen_cells + pre.n_fallback_cells == want["n_fallback_cells"], key
