"""T69 face/interior extension of the node JD measure: the C++ estimator (host entry point, CPU) and the device
entry points (GPU) against the oracle (oracle/jd_affine.py)."""
import numpy as np
import pytest

from conftest import rough_field
from oracle import jd_affine as ja
from oracle import q1 as Q
from seq_vp_3d_amd import _lib
from test_oracle_jd_affine import bowl, sample27, spd
from test_oracle_q1 import make

# The estimate rounds differently from the oracle's einsum/BLAS contractions; where the start selection and the
# Newton decisions agree the values agree to rounding.
TOL = 1e-12


def rows(seed, n=2000):
    rng = np.random.default_rng(seed)
    out = [1. + .6 * rng.normal(size=(n // 2, 27)), rng.normal(size=(n // 4, 27))]
    bowls = []
    for _ in range(n - sum(len(o) for o in out)):
        s = rng.uniform(-.3, 1.3, 3)
        bowls.append(sample27(bowl(rng.uniform(-.5, .8), s, spd(rng))))
    return np.concatenate(out + [np.array(bowls)])


def close(got, want):
    return np.abs(got - want) <= TOL * np.maximum(1., np.abs(want))


@pytest.mark.skipif(not _lib.LIB_PATH.exists(), reason="library not built")
def test_host_estimator_matches_the_oracle():
    J = rows(0)
    want = ja.face_interior_minima(J)
    got = np.empty(len(J))
    _lib.call("seqvp_face_interior_minima_host", _lib.hptr(J), len(J), _lib.hptr(got))
    assert np.all(close(got, want)), np.flatnonzero(~close(got, want))[:5]
    # empty input, and a NULL pointer with rows
    _lib.call("seqvp_face_interior_minima_host", None, 0, None)
    with pytest.raises(_lib.SeqvpError):
        _lib.call("seqvp_face_interior_minima_host", None, 3, _lib.hptr(got))


@pytest.mark.gpu
def test_device_estimator_matches_the_oracle(gpu_device):
    import torch
    from seq_vp_3d_amd.q1 import FusedQ1
    J = rows(1, 20000)
    want = ja.face_interior_minima(J)
    print("case face_interior_minima 20000 rows", flush=True)
    got = FusedQ1.face_interior_minima(torch.tensor(J, device=gpu_device)).cpu().numpy()
    ok = close(got, want)
    assert np.all(ok), (np.sum(~ok), np.flatnonzero(~ok)[:5])


@pytest.mark.gpu
@pytest.mark.parametrize("amp,activation", [(.2, 1.), (.25, .25), (.05, .25)])
def test_extended_node_measure_matches_the_oracle(gpu_device, amp, activation):
    import torch
    from seq_vp_3d_amd.q1 import FusedQ1
    from test_gpu_kernels import case
    prob, og, m = case(shape=(12, 13, 14))
    u = rough_field(og.shape, amp, 3)
    want, wstats = ja.node_min_jd(u, og, m.closed[3], extension_activation=activation)
    q = FusedQ1(prob, gpu_device)
    stats = torch.full((4,), -1., dtype=torch.float64, device=gpu_device)
    print(f"case node_min_jd extended amp={amp} activation={activation}", flush=True)
    got = q.node_min_jd(torch.tensor(u, device=gpu_device), extension_activation=activation, stats=stats)
    got = got.cpu().numpy()
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14)
    cand, lowered, lowered_any, mu_min = stats.tolist()
    assert int(cand) == wstats["candidates"]
    assert int(lowered) == wstats["lowered_cells"]
    # "lowered at all" counts cells whose estimate falls below m_c by any amount, rounding included (a minimum at
    # a closed point is reached again up to rounding): it is pinned between the counts at -+ 1e-12 relative
    _, _, J, _ = Q.geometry(u, og, m.closed[3])
    minima = ja.cell_minima(J)
    cand_ids = ja.candidates(J, minima, activation)
    mu = ja.face_interior_minima(J[cand_ids])
    old = minima[cand_ids]
    slack = 1e-12 * np.maximum(1., np.abs(old))
    assert np.sum(mu < old - slack) <= int(lowered_any) <= np.sum(mu < old + slack)
    if wstats["candidates"]:
        assert abs(mu_min - wstats["mu_min"]) <= 1e-12 * max(1., abs(wstats["mu_min"]))
    else:
        assert mu_min == np.inf
    # without the extension the entry point is the T62 measure
    plain = q.node_min_jd(torch.tensor(u, device=gpu_device)).cpu().numpy()
    np.testing.assert_allclose(plain, ja.node_min_jd(u, og, m.closed[3])[0], rtol=1e-12, atol=1e-14)
    assert np.all(got <= plain)
