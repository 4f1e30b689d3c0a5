"""Pins the Q1 oracle (oracle/q1.py) against closed forms, finite differences and brute force."""
import numpy as np
import pytest

from conftest import rough_field, smooth_field
from oracle import q1 as Q


def make(shape=(5, 6, 7), spacing=(1., 1.3, .8), activation=.9):
    g = Q.Grid(shape, spacing)
    m = Q.Model(g, lam=1e-3, mu=.03, nu=.05, kappa=.002, delta=0., activation=activation)
    return g, m


def test_identity_map_closed_forms():
    g, m = make()
    u = np.zeros(g.shape + (3,))
    for D in (m.gauss[3], m.closed[3]):
        F, A, J, C = Q.geometry(u, g, D)
        assert np.array_equal(J, np.ones_like(J))
        assert np.array_equal(C, np.zeros_like(C))
    # zero change energy and zero barrier above the activation
    assert m.soft_energy(u, u, u) == (0., 0., 0.)
    assert m.barrier_energy(u)[0] == 0. and m.barrier_energy(u)[1] == 1.


def test_affine_map_gives_constant_F():
    g, m = make()
    M = np.array([[1.1, .2, -.1], [.05, .9, .1], [-.2, .1, 1.2]])
    X = g.identity()
    u = X @ (M - np.eye(3)).T
    for D in (m.gauss[3], m.closed[3]):
        F, A, J, C = Q.geometry(u, g, D)
        np.testing.assert_allclose(F, np.broadcast_to(M, F.shape), rtol=0, atol=1e-13)
        np.testing.assert_allclose(J, np.linalg.det(M), rtol=1e-13)
        np.testing.assert_allclose(C[..., 0], M[2, 1] - M[1, 2], atol=1e-13)
        # cofactor rows: A = det(M) M^-T
        np.testing.assert_allclose(A, np.broadcast_to(np.linalg.det(M) * np.linalg.inv(M).T, A.shape), atol=1e-12)


def test_determinant_along_a_line_is_the_cubic():
    """det(F + t V) evaluated by geometry() equals the cubic c0 + c1 t + c2 t^2 + c3 t^3 (method.md:32)."""
    from oracle.jd_root import point_coefficients
    g, m = make()
    u, v = smooth_field(g.shape, .3, 1), smooth_field(g.shape, .4, 2)
    D = m.closed[3]
    F, _, _, _ = Q.geometry(u, g, D)
    V = Q.geometry(v, g, D)[0] - np.eye(3)
    c0, c1, c2, c3 = point_coefficients(F, V, 0.)
    for t in (.3, -.7, 1.9):
        J = Q.geometry(u + t * v, g, D)[2]
        np.testing.assert_allclose(J, c0 + t * (c1 + t * (c2 + t * c3)), rtol=1e-12, atol=1e-12)


def _total_energy(m, u, ref):
    jd, curl, _ = m.soft_energy(u, ref)
    return jd + curl + m.barrier_energy(u)[0]


@pytest.mark.parametrize("activation", [.9, .05])
def test_gradient_matches_central_differences(activation):
    g, m = make(activation=activation)
    ref = smooth_field(g.shape, .15, 3)
    u = ref + rough_field(g.shape, .05, 4) + smooth_field(g.shape, .25, 5)
    assert Q.geometry(u, g, m.closed[3])[2].min() > .1
    if activation == .9:
        assert Q.geometry(u, g, m.closed[3])[2].min() < .9   # barrier active somewhere
    grad = m.gradient_nodes(np.zeros_like(u), u, ref)
    rng = np.random.default_rng(0)
    free = np.argwhere(np.ones(tuple(n - 2 for n in g.shape), dtype=bool)) + 1
    for z, y, x in free[rng.choice(len(free), 12, replace=False)]:
        for c in range(3):
            h = 1e-6
            up, um = u.copy(), u.copy()
            up[z, y, x, c] += h
            um[z, y, x, c] -= h
            fd = (_total_energy(m, up, ref) - _total_energy(m, um, ref)) / (2 * h)
            assert abs(fd - grad[z, y, x, c]) <= 1e-6 * max(1e-6, abs(fd)) + 1e-11, (z, y, x, c, fd, grad[z, y, x, c])


def test_soft_gradient_vanishes_at_reference():
    g, m = make()
    ref = smooth_field(g.shape, .2, 6)
    out = m.gradient_nodes(np.zeros_like(ref), ref, ref, soft=True, barrier=False)
    assert np.array_equal(out, np.zeros_like(out))


def test_ggn_diagonal_by_brute_force():
    """diag of J_r^T J_r (soft residual Jacobian by finite differences) + sum psi''(J) (dJ/du_i)^2 + node terms."""
    g, m = make(shape=(4, 4, 5))
    u = smooth_field(g.shape, .1, 7)
    b = np.random.default_rng(8).normal(size=g.shape + (3,))
    _, qw, _, D = m.gauss
    Dc = m.closed[3]

    def residual(uu):
        _, _, J, C = Q.geometry(uu, g, D)
        return np.concatenate([(np.sqrt(m.mu * qw) * J).ravel(), (np.sqrt(m.nu * qw)[None, :, None] * C).ravel()])

    Jc0 = Q.geometry(u, g, Dc)[2]
    second = Q.barrier_terms(Jc0, m.delta, m.barrier_weight, m.activation)[2]
    base = m.ggn_diagonal_base(b)
    diag = m.ggn_diagonal_nodes(base, u)
    for z in range(1, g.shape[0] - 1):
        for y in range(1, g.shape[1] - 1):
            for x in range(1, g.shape[2] - 1):
                for c in range(3):
                    h = 1e-6
                    up, um = u.copy(), u.copy()
                    up[z, y, x, c] += h
                    um[z, y, x, c] -= h
                    dr = (residual(up) - residual(um)) / (2 * h)
                    dJ = (Q.geometry(up, g, Dc)[2] - Q.geometry(um, g, Dc)[2]) / (2 * h)
                    expect = base[z, y, x, c] + dr @ dr + np.sum(second * dJ ** 2)
                    assert abs(diag[z, y, x, c] - expect) <= 1e-7 * abs(expect), (z, y, x, c)


def test_ggn_diagonal_is_diagonal_of_hvp():
    g, m = make(shape=(4, 5, 4))
    u = smooth_field(g.shape, .1, 9)
    b = np.random.default_rng(10).normal(size=g.shape + (3,))
    diag = m.ggn_diagonal_nodes(m.ggn_diagonal_base(b), u)
    for (z, y, x, c) in [(1, 1, 1, 0), (2, 3, 2, 2), (1, 2, 2, 1)]:
        e = np.zeros_like(u)
        e[z, y, x, c] = 1.
        np.testing.assert_allclose(m.hvp_nodes(u, e, b)[z, y, x, c], diag[z, y, x, c], rtol=1e-13)


def test_stable_differences_equal_direct_differences():
    g, m = make()
    ref = smooth_field(g.shape, .1, 11)
    base = smooth_field(g.shape, .25, 12)
    u = base + smooth_field(g.shape, .05, 13)
    jd, curl, diff = m.soft_energy(u, ref, base)
    jb, cb, _ = m.soft_energy(base, ref)
    np.testing.assert_allclose(diff, (jd + curl) - (jb + cb), rtol=1e-10)
    val, _, chg, _ = m.barrier_energy(u, base)
    vb = m.barrier_energy(base)[0]
    np.testing.assert_allclose(chg, val - vb, rtol=1e-9)


def test_barrier_branches_and_continuity():
    """psi, psi', psi'' vanish at J = activation (C2, method.md:52) and psi'' > 0 inside."""
    J = np.array([.05 - 1e-9, .05, .2, .049, .01, .03])
    v, f, s = Q.barrier_terms(J, 0., 1., .05)
    assert v[1] == f[1] == s[1] == 0.
    assert abs(v[0]) < 1e-12 and abs(f[0]) < 1e-6
    assert np.all(s[3:] > 0) and np.all(v[3:] > 0)
    # first derivative matches a finite difference of the value on both log branches
    for j in (.01, .04):
        h = 1e-7
        fd = (Q.barrier_terms([j + h], 0., 1., .05)[0][0] - Q.barrier_terms([j - h], 0., 1., .05)[0][0]) / (2 * h)
        np.testing.assert_allclose(Q.barrier_terms([j], 0., 1., .05)[1][0], fd, rtol=1e-6)


def test_node_min_jd_bounds_sampled_jd():
    g, m = make()
    u = smooth_field(g.shape, .3, 14)
    node = m.node_min_jd(u)
    J = Q.geometry(u, g, m.closed[3])[2]
    idx = Q.cell_indices(g)
    for c in range(g.n_cells):
        # cell minimum <= every closed27 value and >= the Bernstein lower hull would allow
        assert np.all(node.ravel()[idx[c]] <= J[c].min() + 1e-15)
    # brute force along one cell edge: the exact quadratic minimum is the dense-sample minimum
    f0, f1, f2 = J[0, 0], J[0, 1], J[0, 2]
    t = np.linspace(0, 1, 20001)
    q = f0 * 2 * (t - .5) * (t - 1) - f1 * 4 * t * (t - 1) + f2 * 2 * t * (t - .5)
    a = 2 * (f0 - 2 * f1 + f2)
    b = -3 * f0 + 4 * f1 - f2
    if a > 0 and b < 0 and -b < 2 * a:
        np.testing.assert_allclose(f0 - b * b / (4 * a), q.min(), atol=1e-8)
