"""Pins the oracle's T69 face/interior node JD estimator (oracle/jd_affine.py) by what the mathematics fixes:
quadratic bowls (interior and face minima in closed form), a bounded quasi-Newton minimiser on random
triquadratic cells, and the Bernstein hull property."""
import numpy as np
import pytest
from scipy.optimize import minimize

from conftest import smooth_field
from oracle import jd_affine as ja
from oracle import q1 as Q
from test_oracle_q1 import make

CLOSED = np.array([0., .5, 1.])


def lagrange_eval(J27, pts):
    """The tensor Lagrange interpolant of the 27 closed values at points (m, 3) = (x, y, z)."""
    v = np.asarray(J27).reshape(3, 3, 3)
    L = [ja.lagrange(pts[:, a]) for a in range(3)]
    return np.einsum("zyx,mz,my,mx->m", v, L[2], L[1], L[0])


def sample27(f):
    """Closed27 values (index 9 iz + 3 iy + ix) of a function of (x, y, z)."""
    z, y, x = np.meshgrid(CLOSED, CLOSED, CLOSED, indexing="ij")
    return f(np.stack([x.ravel(), y.ravel(), z.ravel()], -1))


def bowl(c, s, A):
    return lambda p: c + np.einsum("mi,ij,mj->m", p - s, A, p - s)


def spd(rng, lo=.5, hi=3.):
    Qm, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    return Qm @ np.diag(rng.uniform(lo, hi, 3)) @ Qm.T


def test_interior_bowl_minimum_is_exact():
    rng = np.random.default_rng(0)
    rows, want = [], []
    for _ in range(40):
        c, s, A = rng.uniform(-.5, .8), rng.uniform(.15, .85, 3), spd(rng)
        rows.append(sample27(bowl(c, s, A)))
        want.append(c)
    mu = ja.face_interior_minima(np.array(rows))
    np.testing.assert_allclose(mu, want, rtol=0, atol=2e-14)
    # the grid alone misses the minimum: the Newton part is what reaches it
    grid = ja.grid_values(np.array(rows)).min(axis=1)
    assert np.all(grid >= np.array(want) - 1e-15) and np.mean(grid > np.array(want) + 1e-6) > .9


def test_face_minimum_with_the_outward_coordinate_fixed():
    rng = np.random.default_rng(1)
    for axis in range(3):
        for _ in range(10):
            A = spd(rng, 1., 2.)
            s = rng.uniform(.3, .7, 3)
            s[axis] = rng.choice([-.3, 1.3])
            face = 0. if s[axis] < 0 else 1.
            # constrained minimum on the face: Schur complement of the fixed coordinate, checked to lie inside
            o = [i for i in range(3) if i != axis]
            d = face - s[axis]
            shift = -np.linalg.solve(A[np.ix_(o, o)], A[o, axis] * d)
            if np.any(s[o] + shift <= 0) or np.any(s[o] + shift >= 1):
                continue
            c = .2
            want = c + d * d * (A[axis, axis] - A[axis, o] @ np.linalg.solve(A[np.ix_(o, o)], A[o, axis]))
            mu = ja.face_interior_minima(sample27(bowl(c, s, A))[None])[0]
            assert abs(mu - want) <= 1e-14, (axis, mu, want)


def test_random_triquadratic_cells_against_a_bounded_quasi_newton_minimiser():
    rng = np.random.default_rng(2)
    J = 1. + .6 * rng.normal(size=(60, 27))
    mu = ja.face_interior_minima(J)
    starts = rng.uniform(0, 1, size=(48, 3))
    agree = 0
    for n in range(len(J)):
        f = lambda p, n=n: lagrange_eval(J[n], p[None])[0]
        best = min(minimize(f, x0, method="L-BFGS-B", bounds=[(0, 1)] * 3,
                            options=dict(ftol=1e-15, gtol=1e-12, maxiter=500)).fun for x0 in starts)
        best = min(best, lagrange_eval(J[n], np.array([[a, b, c] for a in CLOSED for b in CLOSED for c in CLOSED])).min())
        # mu is a value J attains in the cell: never below its minimum
        assert mu[n] >= best - 1e-9 * max(1., abs(best))
        agree += abs(mu[n] - best) <= 1e-9 * max(1., abs(best))
    assert agree >= .9 * len(J), agree


def test_value_and_derivatives_of_the_monomial_form():
    rng = np.random.default_rng(3)
    J = rng.normal(size=(5, 27))
    C = ja.monomial(J)
    x = rng.uniform(0, 1, size=(5, 3))
    np.testing.assert_allclose(ja.value(C, x), [lagrange_eval(J[n], x[n:n + 1])[0] for n in range(5)], atol=1e-13)
    g, H = ja.derivatives(C, x)
    h = 1e-5
    for i in range(3):
        e = np.zeros(3)
        e[i] = h
        fd = (ja.value(C, x + e) - ja.value(C, x - e)) / (2 * h)
        np.testing.assert_allclose(g[:, i], fd, atol=1e-8)
        gp, _ = ja.derivatives(C, x + e)
        gm, _ = ja.derivatives(C, x - e)
        np.testing.assert_allclose(H[:, :, i], (gp - gm) / (2 * h), atol=1e-7)


def test_bernstein_hull_reproduces_and_bounds_the_cell():
    rng = np.random.default_rng(4)
    J = rng.normal(size=(20, 27))
    B = ja.bernstein(J).reshape(-1, 3, 3, 3)
    pts = rng.uniform(0, 1, size=(200, 3))
    basis = lambda t: np.stack([(1 - t) ** 2, 2 * t * (1 - t), t * t], -1)
    for n in range(len(J)):
        got = np.einsum("zyx,mz,my,mx->m", B[n], basis(pts[:, 2]), basis(pts[:, 1]), basis(pts[:, 0]))
        np.testing.assert_allclose(got, lagrange_eval(J[n], pts), atol=1e-13)
        assert B[n].min() <= lagrange_eval(J[n], pts).min() + 1e-15


def test_edge_minima_match_dense_sampling():
    rng = np.random.default_rng(5)
    f = rng.normal(size=(200, 3))
    got = ja.edge_minima(f[:, 0], f[:, 1], f[:, 2])
    t = np.linspace(0, 1, 200001)
    q = f[:, :1] * 2 * (t - .5) * (t - 1) - f[:, 1:2] * 4 * t * (t - 1) + f[:, 2:] * 2 * t * (t - .5)
    dense = q.min(axis=1)
    inner = np.isfinite(got)
    assert inner.sum() > 20
    np.testing.assert_allclose(got[inner], dense[inner], atol=1e-9)
    # no interior minimum: the dense minimum sits at an endpoint
    np.testing.assert_allclose(dense[~inner], np.minimum(f[~inner, 0], f[~inner, 2]), atol=1e-12)


@pytest.mark.parametrize("amp", [.15, .35])
def test_extended_node_measure_on_a_deformed_map(amp):
    g, m = make()
    u = smooth_field(g.shape, amp, 17)
    node0, none = ja.node_min_jd(u, g, m.closed[3])
    assert none is None
    np.testing.assert_array_equal(node0, m.node_min_jd(u))
    node, stats = ja.node_min_jd(u, g, m.closed[3], extension_activation=.25 if amp < .3 else 1.)
    assert np.all(node <= node0)
    _, _, J, _ = Q.geometry(u, g, m.closed[3])
    minima = ja.cell_minima(J)
    act = .25 if amp < .3 else 1.
    cand = ja.candidates(J, minima, act)
    assert stats["candidates"] == cand.size
    # outside the candidates the Bernstein hull, hence J itself, stays above min(m_c, activation)
    rest = np.setdiff1d(np.arange(len(J)), cand)
    pts = np.random.default_rng(6).uniform(0, 1, size=(64, 3))
    for c in rest[:50]:
        assert lagrange_eval(J[c], pts).min() >= min(minima[c], act) - 1e-14
    if cand.size:
        assert stats["mu_min"] <= minima[cand].min() + 1e-15
