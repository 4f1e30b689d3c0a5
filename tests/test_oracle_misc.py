"""Pins the Poisson, JD-root, prescreen and spline oracles against closed forms and brute force."""
from fractions import Fraction

import numpy as np
import pytest

from conftest import counterexample_cell, rough_field, smooth_field
from oracle import jd_root, poisson, prescreen, q1 as Q, spline


# ---- DST-I Poisson inverse ---------------------------------------------------------------------
@pytest.mark.parametrize("metric", ["q1", "seven_point"])
def test_poisson_inverse_inverts_the_operator(metric):
    g = Q.Grid((7, 9, 8), (1., 1.3, .8))
    a, b = .37, .015
    den = poisson.denominator(g, a, b, metric)
    x = np.random.default_rng(0).normal(size=3 * g.n_free_scalar)
    y = poisson.apply(x, den)
    np.testing.assert_allclose(poisson.matvec(y, g, a, b, metric), x, rtol=1e-11, atol=1e-11)


def test_poisson_matches_dense_kronecker_solve():
    g = Q.Grid((5, 6, 7), (1., 1., 1.))
    a, b = .2, .03
    n = g.n_free_scalar
    M = np.stack([poisson.matvec(np.eye(3 * n)[:, i], g, a, b)[:n] for i in range(n)], axis=1)
    x = np.random.default_rng(1).normal(size=n)
    den = poisson.denominator(g, a, b)
    y = poisson.apply(np.concatenate([x, x, x]), den)[:n]
    np.testing.assert_allclose(y, np.linalg.solve(M, x), rtol=1e-11)


# ---- T66 first JD root ---------------------------------------------------------------------------
def test_first_roots_match_numpy_roots():
    rng = np.random.default_rng(2)
    b0 = rng.uniform(.1, 1., 4000)
    b1, b2, b3 = (rng.normal(size=4000) * 2 for _ in range(3))
    got = jd_root.first_roots(b0, b1, b2, b3)
    for i in range(4000):
        r = np.roots([b3[i], b2[i], b1[i], b0[i]])
        real = sorted(x.real for x in r if abs(x.imag) < 1e-9 and 0 < x.real <= 1)
        q = lambda s: b0[i] + s * (b1[i] + s * (b2[i] + s * b3[i]))
        if not real:
            # no root in (0,1]: inf, or a near-tangent touch where |q| is within rounding
            assert got[i] == np.inf or abs(q(got[i])) < 1e-9
        else:
            assert got[i] <= real[0] + 1e-9 and got[i] >= real[0] - 1e-6, (i, got[i], real)


def test_global_root_counts_and_closed_form():
    g = Q.Grid((5, 5, 5), (1., 1., 1.))
    D = Q.q1_tables(g, "closed27_uniform")[3]
    u = np.zeros(g.shape + (3,))
    # v = -X on the free nodes is a uniform compression where the cells are interior: J(tau) = (1-tau)^3
    X = g.identity()
    v = np.zeros_like(u)
    v[2, 2, 2] = (-.5, -.5, -.5)
    s, invalid, points = jd_root.global_root(u, v, g, D, 0., 4.)
    assert invalid == 0 and points == 27 * g.n_cells
    assert np.isfinite(s) and 0 < s <= 1
    # at the root the minimum closed27 J of u + s*L*v is ~0 and positive just before it
    J = Q.geometry(u + s * 4. * v, g, D)[2]
    assert J.min() > 0 and J.min() < 1e-6
    del X


# ---- T75 prescreen -------------------------------------------------------------------------------
def _exact_min_beta(u, g, D, c):
    """Exact rational min Bernstein coefficient of J on cell c, with exact 1/h."""
    nz, ny, nx = g.shape
    cz, cy, cx = g.cell_shape
    x0, y0, z0 = c % cx, (c // cx) % cy, c // (cx * cy)
    flat = u.reshape(-1, 3)
    base = (z0 * ny + y0) * nx + x0
    verts = [base + (z * ny + y) * nx + x for x, y, z in Q.VERTICES]
    ih = [1 / Fraction(h) for h in g.spacing]
    pts = [(x, y, z) for z in (0, Fraction(1, 2), 1) for y in (0, Fraction(1, 2), 1) for x in (0, Fraction(1, 2), 1)]
    b = []
    for p in pts:
        F = [[Fraction(0)] * 3 for _ in range(3)]
        for a, vert in enumerate(Q.VERTICES):
            for j in range(3):
                o = [k for k in range(3) if k != j]
                f = [(p[k] if vert[k] else 1 - p[k]) for k in range(3)]
                w = (1 if vert[j] else -1) * ih[j] * f[o[0]] * f[o[1]]
                for r in range(3):
                    F[r][j] += Fraction(float(flat[verts[a], r])) * w
        for r in range(3):
            F[r][r] += 1
        b.append(F[0][0] * (F[1][1] * F[2][2] - F[1][2] * F[2][1]) + F[0][1] * (F[1][2] * F[2][0] - F[1][0] * F[2][2])
                 + F[0][2] * (F[1][0] * F[2][1] - F[1][1] * F[2][0]))
    for stride in (1, 3, 9):
        for q in range(27):
            if (q // stride) % 3 == 1:
                b[q] = 2 * b[q] - (b[q - stride] + b[q + stride]) / 2
    return min(b)


def test_prescreen_c_matches_python_bitwise():
    g = Q.Grid((5, 6, 7), (1., 1.3, .8))
    D = Q.q1_tables(g, "closed27_uniform")[3]
    u = smooth_field(g.shape, .2, 3) + rough_field(g.shape, .05, 4)
    c_out = prescreen.lower_bounds_c(u, g, D)
    for c in range(0, g.n_cells, 7):
        assert c_out[c] == prescreen.lower_bound_py(u, g, D, c) or (np.isnan(c_out[c]))


@pytest.mark.parametrize("spacing", [(1., 1., 1.), (1., 1.3, .8)])
def test_prescreen_bound_contains_exact_bernstein_minimum(spacing):
    g = Q.Grid((4, 4, 4), spacing)
    D = Q.q1_tables(g, "closed27_uniform")[3]
    rng = np.random.default_rng(5)
    for trial in range(6):
        u = rough_field(g.shape, .25 if trial % 2 else .45, 10 + trial)
        low = prescreen.lower_bounds_c(u, g, D)
        for c in range(g.n_cells):
            exact = _exact_min_beta(u, g, D, c)
            assert Fraction(float(low[c])) <= exact, (trial, c)
            # not vacuous: the a-priori bound stays within 1e-9 of the exact minimum for O(1) cells
            assert float(exact) - low[c] < 1e-9
    del rng


def test_prescreen_never_proves_a_folding_cell():
    g = Q.Grid((2, 2, 2), (1., 1., 1.))
    D = Q.q1_tables(g, "closed27_uniform")[3]
    for d in (-.05, 0., .3):
        u = counterexample_cell(d).reshape(2, 2, 2, 3)
        low = prescreen.lower_bounds_c(u, g, D)
        assert not low[0] > 0, d
        assert prescreen.lower_bound_py(u, g, D, 0) == low[0]


# ---- spline --------------------------------------------------------------------------------------
def test_spline_reproduces_cubics_and_interpolates():
    g = Q.Grid((8, 9, 10), (1., 1.5, .7), (2., -1., .5))
    X = g.identity()
    p = lambda x, y, z: .3 * x ** 3 - x * y * z + .2 * y ** 2 * z - z ** 3 / 7 + 1.5
    img = p(X[..., 0], X[..., 1], X[..., 2])
    s = spline.build(img, g)
    rng = np.random.default_rng(6)
    lo = np.array([a[0] for a in g.axes])
    hi = np.array([a[-1] for a in g.axes])
    pts = lo + rng.uniform(0, 1, size=(200, 3)) * (hi - lo)
    v, grad = spline.sample(s, pts)
    x, y, z = pts.T
    np.testing.assert_allclose(v, p(x, y, z), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(grad[:, 0], .9 * x ** 2 - y * z, rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(grad[:, 1], -x * z + .4 * y * z, rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(grad[:, 2], -x * y + .2 * y ** 2 - 3 * z ** 2 / 7, rtol=1e-10, atol=1e-10)
    rough = np.random.default_rng(7).normal(size=g.shape)
    s2 = spline.build(rough, g)
    v2, _ = spline.sample(s2, X, with_gradient=False)
    np.testing.assert_allclose(v2, rough, atol=1e-12)


def test_span_guess_and_walk_is_searchsorted():
    """The device sampler finds a knot span by a proportional guess and two walks instead of a bisection
    (csrc/spline.hip cubic_basis); on the not-a-knot knot vectors of uniform grids that procedure returns
    searchsorted(t, x, right=True) - 1 for every x (knots, ends, between, outside, NaN)."""
    from scipy.interpolate import make_interp_spline

    def walk(t, x):
        nt = len(t)
        if not (x >= t[0]):
            return -1
        if x >= t[-1]:
            return nt - 1
        f = (x - t[0]) / (t[-1] - t[0]) * (nt - 1)
        s = 0 if f < 0 else (nt - 2 if f > nt - 2 else int(f))
        while t[s + 1] <= x:
            s += 1
        while t[s] > x:
            s -= 1
        return s

    rng = np.random.default_rng(3)
    for n, h, o in ((8, 1., 0.), (23, .7, -3.), (160, 1., 0.), (224, 1.1, 2.5)):
        xs = o + h * np.arange(n)
        t = make_interp_spline(xs, np.zeros(n), k=3).t
        probes = np.concatenate([t, np.nextafter(t, np.inf), np.nextafter(t, -np.inf), rng.uniform(t[0] - 1, t[-1] + 1, 500),
                                 [np.nan, -np.inf, np.inf]])
        for x in probes:
            want = np.searchsorted(t, x, side="right") - 1 if not np.isnan(x) else -1
            assert walk(t, x) == want, (n, x)
