"""JD witnesses: the host point enclosure (bitwise = oracle), the device nomination and the device cubic
coefficients of the witness predictor, against the oracle (oracle/jd_step.py)."""
import numpy as np
import pytest

from conftest import counterexample_cell, rough_field, smooth_field
from oracle import jd_step as js
from seq_vp_3d_amd import _lib


@pytest.mark.skipif(not _lib.LIB_PATH.exists(), reason="library not built")
def test_host_point_enclosure_is_bitwise_the_oracle():
    rng = np.random.default_rng(0)
    spacing = np.array([1.1, .9, 1.3])
    out = np.empty(2)
    for i in range(2000):
        v = np.ascontiguousarray(rng.normal(size=(8, 3)) * rng.choice([.05, .4, 1.5, 1e300]))
        t = np.ascontiguousarray(rng.uniform(0, 1, 3) if i % 2 else rng.choice([0., .5, 1.], size=3))
        lo, hi, ok = js.point_jd_interval(v, spacing, t)
        _lib.call("seqvp_point_jd_interval_host", _lib.hptr(v), _lib.hptr(spacing), _lib.hptr(t), _lib.hptr(out))
        if ok:
            assert out[0] == lo and out[1] == hi
        else:
            assert np.isnan(out).all()
    v = np.ascontiguousarray(counterexample_cell(-.2))
    _lib.call("seqvp_point_jd_interval_host", _lib.hptr(v), _lib.hptr(np.ones(3)), _lib.hptr(np.full(3, .5)),
              _lib.hptr(out))
    assert out[1] < 0
    with pytest.raises(_lib.SeqvpError):       # a point outside the cell
        _lib.call("seqvp_point_jd_interval_host", _lib.hptr(v), _lib.hptr(np.ones(3)), _lib.hptr(np.full(3, 1.5)),
                  _lib.hptr(out))


def test_first_root_before_is_the_oracles():
    from seq_vp_3d_amd.jd_step import first_root_before
    rng = np.random.default_rng(1)
    for _ in range(500):
        c = np.concatenate(([rng.uniform(.01, 1.)], rng.normal(size=3) * rng.choice([.1, 3., 30.])))
        limit = rng.uniform(.1, 3.)
        assert first_root_before(c, limit) == js.first_root_before(c, limit)
    for bad in ([0., 1., 0., 0.], [np.nan, 0., 0., 0.]):
        with pytest.raises(FloatingPointError):
            first_root_before(bad, 1.)


def test_memory_layout_and_capacity():
    from seq_vp_3d_amd.jd_step import CAPACITY, CLOSED_POINTS, JDStepMemory
    from seq_vp_3d_amd.problem import make_grid_3d
    g = make_grid_3d(nx=6, ny=7, nz=8, hx=1., hy=1., hz=1.)
    m = JDStepMemory()
    m.bind(g)
    m.witness(9, (.25, .5, .75))
    m.witness(2, (0., .5, 1.))          # a closed point adds no extra point
    m.witness(9, (.1, .2, .3))
    cells, pts = m.layout()
    assert list(np.unique(cells)) == [2, 9] and np.sum(cells == 2) == 27 and np.sum(cells == 9) == 29
    np.testing.assert_array_equal(pts[:27], np.array(CLOSED_POINTS))
    with pytest.raises(ValueError):
        m.witness(g.n_cells, (0., 0., 0.))
    for k in range(CAPACITY + 5):
        m.witness(k % g.n_cells, (k / 1000, 0., 0.))
    assert len(m.points) == CAPACITY
    m.certificate(dict(status="uncertain", witness_cell=None, witness_reference_point=None))
    m.certificate(dict(status="infeasible", witness_cell=3, witness_reference_point=(.5, .5, .5)))
    assert m.uncertain == 1 and list(m.points)[-1] == (3, (.5, .5, .5))
    m.bind(make_grid_3d(nx=7, ny=7, nz=8, hx=1., hy=1., hz=1.))        # another grid: a fresh memory
    assert not m.points and m.uncertain == 0


@pytest.mark.gpu
def test_device_nomination_and_coefficients(gpu_device):
    import torch
    from seq_vp_3d_amd.jd_step import nominate_witness
    from seq_vp_3d_amd.q1 import FusedQ1
    from test_gpu_kernels import case
    prob, og, m = case(shape=(12, 13, 14))
    u = smooth_field(og.shape, .1, 6)
    q = FusedQ1(prob, gpu_device)
    T = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device=gpu_device)
    print("case witness nomination", flush=True)
    ud = T(u)
    assert nominate_witness(ud, q.k, prob.grid, 0., gpu_device) is None
    folded = u.copy()
    folded[6, 6, 7] += np.array([3., 0., 0.])
    folded[3, 9, 4] += np.array([0., -2.5, 0.])
    fd = T(folded)
    for tile in (8192, 64, 7):
        got = nominate_witness(fd, q.k, prob.grid, 0., gpu_device, tile=tile)
        assert got == js.nominate_witness(folded, og, m.closed[3], 0., tile=tile), tile
    # cubic coefficients of det(A + s V) at closed and off-grid points of several cells
    v = rough_field(og.shape, .7, 8)
    rng = np.random.default_rng(9)
    cells = np.repeat(rng.choice(og.n_cells, 6, replace=False), 30)
    pts = np.concatenate([np.concatenate([js.CLOSED, rng.uniform(0, 1, (3, 3))]) for _ in range(6)])
    out = torch.empty((len(cells), 4), dtype=torch.float64, device=gpu_device)
    nz, ny, nx = og.shape
    print("case witness coefficients", flush=True)
    ud, vd, cd, pd = T(u), T(v), torch.tensor(cells, device=gpu_device), T(pts)      # held across the calls
    _lib.call("seqvp_witness_coefficients", _lib.ptr(ud), _lib.ptr(vd), _lib.ptr(q.k), nx, ny, nz, _lib.ptr(cd),
              _lib.ptr(pd), len(cells), _lib.ptr(out), _lib.stream_handle(gpu_device))
    got = out.cpu().numpy()
    for i, cell in enumerate(cells):
        A = js.point_gradients(u, og, cell, pts[i:i + 1])
        V = js.point_gradients(v, og, cell, pts[i:i + 1]) - np.eye(3)
        want = js.determinant_coefficients(A, V, 0.)[0]
        np.testing.assert_allclose(got[i], want, rtol=1e-12, atol=1e-13 * np.abs(want).max())
    # a cell id outside the grid yields NaN and touches nothing outside the field
    outside = torch.tensor([og.n_cells], device=gpu_device)
    _lib.call("seqvp_witness_coefficients", _lib.ptr(ud), _lib.ptr(vd), _lib.ptr(q.k), nx, ny, nz, _lib.ptr(outside),
              _lib.ptr(pd), 1, _lib.ptr(out), _lib.stream_handle(gpu_device))
    assert torch.isnan(out[0]).all()
