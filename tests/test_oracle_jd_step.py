"""Pins the oracle's JD witness and step-prediction functions (oracle/jd_step.py): exact rational enclosure of
J at a point, the cubic identity of det(A + s V), cubics with known roots against a dense sign scan, and the
prediction against a brute-force scan of det(A + s V) over s."""
from fractions import Fraction

import numpy as np
import pytest

from conftest import counterexample_cell, rough_field, smooth_field
from oracle import jd_step as js
from oracle import q1 as Q
from test_oracle_q1 import make


def exact_jd(vertices, spacing, point):
    """J at the point in exact rational arithmetic."""
    v = [[Fraction(float(x)) for x in row] for row in vertices]
    t = [Fraction(float(x)) for x in point]
    h = [Fraction(float(x)) for x in spacing]
    F = [[Fraction(int(r == j)) for j in range(3)] for r in range(3)]
    for a in range(8):
        bits = [a >> j & 1 for j in range(3)]
        f = [t[k] if bits[k] else 1 - t[k] for k in range(3)]
        for j in range(3):
            o = [k for k in range(3) if k != j]
            w = (1 if bits[j] else -1) / h[j] * f[o[0]] * f[o[1]]
            for r in range(3):
                F[r][j] += w * v[a][r]
    return (F[0][0] * (F[1][1] * F[2][2] - F[1][2] * F[2][1]) - F[0][1] * (F[1][0] * F[2][2] - F[1][2] * F[2][0])
            + F[0][2] * (F[1][0] * F[2][1] - F[1][1] * F[2][0]))


def test_point_enclosure_contains_the_exact_jacobian():
    rng = np.random.default_rng(0)
    spacing = (1.1, .9, 1.3)
    for _ in range(150):
        v = rng.normal(size=(8, 3)) * rng.choice([.05, .4, 1.5])
        t = rng.choice([0., .5, 1.], size=3) if rng.random() < .5 else rng.uniform(0, 1, 3)
        lo, hi, ok = js.point_jd_interval(v, spacing, t)
        assert ok
        e = exact_jd(v, spacing, t)
        assert Fraction(lo) <= e <= Fraction(hi)
        assert hi - lo <= 1e-12 * max(1., abs(float(e)))
    # the T05 cell with d < 0 folds at its centre (J = d there); with d > 0 the centre stays positive
    for d in (-.2, .3):
        lo, hi, ok = js.point_jd_interval(counterexample_cell(d), (1., 1., 1.), (.5, .5, .5))
        assert ok and lo <= d <= hi and hi - lo < 1e-14


def test_cubic_identity_of_the_determinant_coefficients():
    rng = np.random.default_rng(1)
    A = np.eye(3) + .3 * rng.normal(size=(20, 3, 3))
    V = rng.normal(size=(20, 3, 3))
    c = js.determinant_coefficients(A, V, .1)
    for s in (-1.3, .2, .7, 2.):
        np.testing.assert_allclose(c[:, 0] + s * c[:, 1] + s * s * c[:, 2] + s ** 3 * c[:, 3],
                                   np.linalg.det(A + s * V) - .1, rtol=1e-12, atol=1e-12)


def test_first_root_before_known_factors_and_a_dense_scan():
    # (1 - s/r1)(1 - s/r2)(1 - s/r3): the least root in (0, limit]
    for roots, limit, want in (((.3, .8, 2.), 1., .3), ((.3, .8, 2.), .25, None), ((1.5, 3., -1.), 2., 1.5),
                               ((.6, -.2, 5.), 1., .6)):
        r1, r2, r3 = roots
        c = np.array([1., -(1 / r1 + 1 / r2 + 1 / r3), 1 / (r1 * r2) + 1 / (r1 * r3) + 1 / (r2 * r3), -1 / (r1 * r2 * r3)])
        got = js.first_root_before(c, limit)
        assert (got is None) if want is None else abs(got - want) <= 1e-12
    with pytest.raises(FloatingPointError):
        js.first_root_before([0., 1., 0., 0.], 1.)
    # random cubics: a returned root is the first sign change of a dense scan
    rng = np.random.default_rng(2)
    s = np.linspace(0, 1, 200001)
    seen = 0
    for _ in range(300):
        c = np.concatenate(([rng.uniform(.05, 1.)], rng.normal(size=3) * 3))
        got = js.first_root_before(c, 1.)
        vals = c[0] + s * (c[1] + s * (c[2] + s * c[3]))
        neg = np.flatnonzero(vals <= 0)
        if got is None:
            assert neg.size == 0 or vals[neg].max() > -1e-9         # no crossing (a touch at rounding level at most)
        else:
            seen += 1
            assert neg.size and abs(s[neg[0]] - got) <= 1e-4
    assert seen > 50


def test_prediction_is_the_least_first_root_over_the_witness_points():
    g, m = make()
    rng = np.random.default_rng(3)
    u = smooth_field(g.shape, .2, 4)
    v = rough_field(g.shape, .8, 5)
    mem = js.Memory()
    assert mem.predict(u, v, g, 0., 1.) == (1., None, None)
    mem.witness(7, (.25, .5, .75))
    mem.witness(30, (0., 0., 0.))                  # a closed point: not added twice
    mem.witness(7, (.1, .9, .3))
    tau, least, fallback = mem.predict(u, v, g, 0., 1.)
    assert fallback is None
    # brute force: the first s in (0, 1] where det(I + grad(u + s v)) <= 0 at any point of the two cells
    s = np.linspace(0, 1, 4001)[1:]
    first = np.inf
    for cell, extra in ((7, [(.25, .5, .75), (.1, .9, .3)]), (30, [])):
        pts = np.concatenate([js.CLOSED, np.array(extra).reshape(-1, 3)])
        A = js.point_gradients(u, g, cell, pts)
        V = js.point_gradients(v, g, cell, pts) - np.eye(3)
        det = np.linalg.det(A[None] + s[:, None, None, None] * V[None])
        hit = np.flatnonzero((det <= 0).any(axis=1))
        if hit.size:
            first = min(first, s[hit[0]])
    if np.isfinite(first):
        assert least is not None and abs(least - first) <= 1 / 4000 and tau == pytest.approx(.9 * least)
    else:
        assert least is None and tau == 1.
    # capacity: the least recent witnesses go first
    for k in range(130):
        mem.witness(k % g.n_cells, (k / 200, 0., 0.))
    assert len(mem.points) == js.CAPACITY
    mem.certificate("uncertain")
    mem.certificate("infeasible", 3, (.5, .5, .5))
    assert mem.uncertain == 1 and list(mem.points)[-1] == (3, (.5, .5, .5))


def test_nomination_finds_a_planted_fold():
    g, m = make(shape=(8, 9, 10))
    u = smooth_field(g.shape, .1, 6)
    assert js.nominate_witness(u, g, m.closed[3], 0.) is None
    # push one interior node far enough to fold its cells
    u[4, 4, 5] += np.array([3., 0., 0.])
    w = js.nominate_witness(u, g, m.closed[3], 0.)
    assert w is not None
    cell, point = w
    nodes = u.reshape(-1, 3)[Q.cell_indices(g)[cell]]
    assert exact_jd(nodes, g.spacing, point) <= 0
    # it is the first minimum of the tile's closed27 values
    _, _, J, _ = Q.geometry(u, g, m.closed[3])
    flat = np.where(np.isfinite(J), J, np.inf).ravel()
    assert cell * 27 + js.CLOSED.tolist().index(list(point)) == int(np.argmin(flat))
