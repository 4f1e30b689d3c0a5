"""Pins the values of the Q1 oracle's energies and barrier derivatives by closed forms and by high-precision
evaluation of the model's definitions (mpmath), independently of the formulas the oracle uses.

* psi, psi', psi'': the localized barrier b(J) = -r^2 log t, t = (J-delta)/(a-delta), r = 1 - t on 0 < t < 1 and
  0 above (docs/method.md:48-52), differentiated by mpmath at 40 digits, on both log branches (log1p when r < 1/2)
  and on both sides of r = 1/2; the second derivative also by central differences of the first.
* the Gauss-Legendre and closed 27-point rules: weights sum to the cell volume, and the Gauss rule integrates
  x^i y^j z^k exactly for i, j, k <= 5 (cell_q1.py:101-129).
* affine maps u = (M - I) X: F = M at every point, so every term has a closed form in det M and curl M
  (residuals.py:166-186, energy_comparison.py:196-240, barrier.py:19-36) and the node terms reduce to the
  trapezoid mass sum (grid.py:37-39).
"""
import mpmath as mp
import numpy as np
import pytest

from oracle import q1 as Q

mp.mp.dps = 40


def psi_mp(J, delta, activation):
    """The barrier's definition in 40-digit arithmetic (docs/method.md:48-52)."""
    J, d, a = mp.mpf(J), mp.mpf(delta), mp.mpf(activation)
    if J >= a:
        return mp.mpf(0)
    t = (J - d) / (a - d)
    return -(1 - t) ** 2 * mp.log(t)


def _points(delta, activation):
    # r = (a - J)/(a - delta): both branches (r < 1/2 takes log1p) and both sides of r = 1/2
    rs = (.02, .2, .45, .4999, .5, .5001, .55, .8, .93)
    return [float(mp.mpf(delta) + (mp.mpf(activation) - delta) * (1 - mp.mpf(r))) for r in rs]


@pytest.mark.parametrize("delta,activation", [(0., .05), (0., .9), (.01, .05), (.1, .6)])
def test_barrier_value_and_derivatives_against_the_definition(delta, activation):
    J = np.array(_points(delta, activation))
    w = .37
    value, first, second = Q.barrier_terms(J, delta, w, activation)
    for j, v, f, s in zip(J, value, first, second):
        exact = [w * mp.diff(lambda x: psi_mp(x, delta, activation), mp.mpf(j), n) for n in (0, 1, 2)]
        for got, want, tol in ((v, exact[0], 4e-15), (f, exact[1], 4e-15), (s, exact[2], 4e-15)):
            want = float(want)
            assert abs(got - want) <= tol * abs(want), (delta, activation, j, got, want)


@pytest.mark.parametrize("delta,activation", [(0., .05), (.01, .05)])
def test_barrier_second_derivative_is_the_derivative_of_the_first(delta, activation):
    """Central differences of psi' (the gradient's weight) reproduce psi'' (the GGN weight)."""
    for j in _points(delta, activation):
        h = 1e-6 * (activation - delta)
        fp = Q.barrier_terms([j + h], delta, 1., activation)[1][0]
        fm = Q.barrier_terms([j - h], delta, 1., activation)[1][0]
        s = Q.barrier_terms([j], delta, 1., activation)[2][0]
        assert abs((fp - fm) / (2 * h) - s) <= 1e-6 * abs(s), (j, (fp - fm) / (2 * h), s)


def test_barrier_closed_form_at_half_activation():
    """At r = t = 1/2, -r^2 log t gives psi = log(2)/4, psi' = (-log 2 - 1/2)/(a - delta) and
    psi'' = (2 log 2 + 5)/(a - delta)^2."""
    for delta, a in ((0., .05), (.02, .3)):
        j = delta + (a - delta) / 2
        v, f, s = (x[0] for x in Q.barrier_terms([j], delta, 1., a))
        assert abs(v - np.log(2) / 4) <= 2e-16
        # d/dJ (-r^2 log t) = (2 r log t - r^2 / t)/(a-delta) = (-log 2 - 1/2)/(a-delta) at r = t = 1/2
        assert abs(f - (-np.log(2) - .5) / (a - delta)) <= 4e-16 / (a - delta)
        # d2/dJ2 = (-2 log t + 4 r/t + (r/t)^2)/(a-delta)^2 = (2 log 2 + 5)/(a-delta)^2
        assert abs(s - (2 * np.log(2) + 5) / (a - delta) ** 2) <= 4e-15 / (a - delta) ** 2


def test_quadrature_rules_are_exact():
    g = Q.Grid((4, 5, 6), (1.1, .7, 1.3))
    points, w, _, _ = Q.q1_tables(g, "gauss27")
    assert abs(w.sum() - g.omega) <= 1e-15 * g.omega
    for i in range(6):
        for j in range(6):
            for k in range(6):
                got = np.sum(w * points[:, 0] ** i * points[:, 1] ** j * points[:, 2] ** k)
                assert abs(got - g.omega / ((i + 1) * (j + 1) * (k + 1))) <= 1e-15 * g.omega, (i, j, k)
    points, w, _, _ = Q.q1_tables(g, "closed27_uniform")
    assert np.array_equal(w, np.full(27, g.omega / 27))
    assert sorted(set(points.ravel())) == [0., .5, 1.]
    # point index p = 9 iz + 3 iy + ix (cell_q1.py:114)
    assert tuple(points[9 * 2 + 3 * 1 + 0]) == (0., .5, 1.)


def _curl(M):
    return np.array([M[2, 1] - M[1, 2], M[0, 2] - M[2, 0], M[1, 0] - M[0, 1]])


def affine_case(shape=(5, 6, 7), spacing=(1., 1.3, .8), activation=.9, seed=0):
    rng = np.random.default_rng(seed)
    g = Q.Grid(shape, spacing, (2., -1., .5))
    m = Q.Model(g, lam=2e-3, mu=.03, nu=.05, kappa=.002, delta=0., activation=activation)
    X = g.identity()
    # det M and det M_b stay below the activation so the barrier is active in every cell
    mats = [s * (np.eye(3) + rng.uniform(-.15, .15, size=(3, 3))) for s in (.85, 1., .88)]
    fields = [X @ (M - np.eye(3)).T for M in mats]
    return g, m, mats, fields


def test_affine_maps_closed_forms():
    """u = (M - I)X, ref = (M_r - I)X, base = (M_b - I)X (all Q1-exact): with C cells of volume omega,
    E_jd = mu/2 C omega (det M - det M_r)^2, E_curl = nu/2 C omega |curl M - curl M_r|^2,
    B = kappa C omega psi(det M), and the stable differences are the differences of those."""
    g, m, (M, Mr, Mb), (u, ref, base) = affine_case()
    C, om = g.n_cells, g.omega
    dM, dR, dB = (np.linalg.det(x) for x in (M, Mr, Mb))
    assert max(dM, dB) < m.activation
    jd, curl, diff = m.soft_energy(u, ref, base)
    assert abs(jd - .5 * m.mu * C * om * (dM - dR) ** 2) <= 1e-12 * jd
    cm, cr, cb = _curl(M), _curl(Mr), _curl(Mb)
    assert abs(curl - .5 * m.nu * C * om * np.sum((cm - cr) ** 2)) <= 1e-12 * curl
    want = .5 * C * om * (m.mu * ((dM - dR) ** 2 - (dB - dR) ** 2) + m.nu * (np.sum((cm - cr) ** 2)
                                                                          - np.sum((cb - cr) ** 2)))
    assert abs(diff - want) <= 1e-12 * (jd + curl)
    val, jmin, chg, bmin = m.barrier_energy(u, base)
    psiM, psiB = (float(psi_mp(d, 0., m.activation)) for d in (dM, dB))
    assert abs(val - m.kappa * C * om * psiM) <= 1e-12 * val
    assert abs(jmin - dM) <= 1e-14 and abs(bmin - dB) <= 1e-14
    assert abs(chg - m.kappa * C * om * (psiM - psiB)) <= 1e-11 * val
    # image and proximity node terms: the trapezoid masses sum to C * omega
    fixed = np.zeros(g.shape)
    m.fixed, m.qd = fixed, m.q
    image, prox = m.node_energies(u, ref, fixed + .3)
    assert abs(image - .5 * .09 * C * om) <= 1e-13 * image
    v = np.array([.1, -.2, .05])
    image, prox = m.node_energies(ref + v, ref, fixed)
    assert abs(prox - .5 * m.lam * C * om * (v @ v)) <= 1e-13 * prox and image == 0.


def test_affine_gradient_is_the_boundary_flux():
    """For an affine map the cell-wise constant fluxes cancel between neighbouring cells, so the soft and barrier
    gradient vanishes at every interior node (a discrete divergence of a constant field)."""
    g, m, (M, Mr, _), (u, ref, _) = affine_case()
    out = m.gradient_nodes(np.zeros_like(u), u, ref)
    interior = out[1:-1, 1:-1, 1:-1]
    scale = np.abs(out).max()
    assert scale > 1e-6
    assert np.abs(interior).max() <= 1e-13 * scale
