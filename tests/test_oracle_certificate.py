"""Pins the certificate oracle (oracle/certificate.py) by exact rational arithmetic and closed forms.

Every check computes the exact value independently (fractions.Fraction on the binary64 inputs), so a
dropped outward step, a wrong edge weight, a transposed cofactor or a wrong Bernstein transform makes an
enclosure miss the exact value or a decision contradict it.
"""
from fractions import Fraction
from itertools import product

import numpy as np
import pytest

from conftest import counterexample_cell, smooth_field
from oracle import certificate as oc

H = (1.0, 1.25, 0.75)


def exact_F(v, h, s):
    """Exact I + grad u at reference point s = (x, y, z) (Fractions) of the Q1 cell with vertex values v."""
    F = [[Fraction(int(r == j)) for j in range(3)] for r in range(3)]
    for j in range(3):
        for a in range(8):
            if a & (1 << j):
                continue
            b = a | (1 << j)
            w = Fraction(1)
            for k in range(3):
                if k != j:
                    w *= s[k] if (a >> k) & 1 else 1 - s[k]
            for r in range(3):
                F[r][j] += (Fraction(v[b][r]) - Fraction(v[a][r])) / Fraction(h[j]) * w
    return F


def exact_J(v, h, s):
    F = exact_F(v, h, s)
    return (F[0][0] * (F[1][1] * F[2][2] - F[1][2] * F[2][1]) - F[0][1] * (F[1][0] * F[2][2] - F[1][2] * F[2][0])
            + F[0][2] * (F[1][0] * F[2][1] - F[1][1] * F[2][0]))


POINTS = [(Fraction(ix, 2), Fraction(iy, 2), Fraction(iz, 2)) for iz in range(3) for iy in range(3) for ix in range(3)]


def exact_bernstein(vals):
    """Lagrange values at {0,1/2,1}^3 (index 9iz+3iy+ix) -> tensor degree-2 Bernstein coefficients, exactly."""
    b = list(vals)
    for st in (1, 3, 9):
        for m in range(27):
            if (m // st) % 3 == 1:
                b[m] = 2 * b[m] - (b[m - st] + b[m + st]) / 2
    return b


def cells(n, scale, seed):
    return np.random.default_rng(seed).normal(scale=scale, size=(n, 8, 3))


def test_outward_operations_contain_the_exact_result():
    rng = np.random.default_rng(0)
    vals = np.concatenate([rng.normal(size=200), rng.normal(size=50) * 1e-310, rng.normal(size=50) * 1e150])
    a = (np.minimum(vals[:150], vals[150:]), np.maximum(vals[:150], vals[150:]))
    b = (np.minimum(vals[::2][:150], vals[1::2][:150]), np.maximum(vals[::2][:150], vals[1::2][:150]))
    for op, f in ((oc.iadd, lambda x, y: x + y), (oc.isub, lambda x, y: x - y), (oc.imul, lambda x, y: x * y)):
        lo, hi, ok = op(a, b)
        for i in range(150):
            if not ok[i]:
                continue
            ends = [f(Fraction(x), Fraction(y)) for x in (a[0][i], a[1][i]) for y in (b[0][i], b[1][i])]
            assert Fraction(lo[i]) < min(ends) and max(ends) < Fraction(hi[i])


@pytest.mark.parametrize("scale", [.05, .3])
def test_interval_jacobians_and_bernstein_coefficients_contain_exact_values(scale):
    v = cells(12, scale, 1)
    delta = .01
    p, c, samples, ok = oc.bernstein_arrays(v, H, delta)
    assert ok.all()
    for k in range(len(v)):
        exact = [exact_J(v[k], H, s) - Fraction(delta) for s in POINTS]
        bern = exact_bernstein(exact)
        for i in range(27):
            assert Fraction(p[0][k, i]) <= exact[i] <= Fraction(p[1][k, i])
            assert Fraction(c[0][k, i]) <= bern[i] <= Fraction(c[1][k, i])
            # the float64 samples are the ordinary rounding of J (relative error of a few ulps of the terms)
            assert abs(Fraction(samples[k, i]) - (exact[i] + Fraction(delta))) < Fraction(1, 10 ** 12)


def test_edge_hull_margin_bounds_exact_J_inside_the_cell():
    rng = np.random.default_rng(3)
    v = cells(20, .08, 2)
    margin, flag = oc.edge_hull_margins(v, H, .01)
    assert not flag.any()
    for k in range(len(v)):
        pts = POINTS + [tuple(Fraction(int(t), 64) for t in rng.integers(0, 65, 3)) for _ in range(10)]
        m = min(exact_J(v[k], H, s) for s in pts) - Fraction(.01)
        assert Fraction(margin[k]) <= m


def test_identity_and_affine_maps_are_certified_with_the_exact_margin():
    v0 = np.zeros((8, 3))
    p, c, _, ok = oc.bernstein_arrays(v0[None], H, .01)
    r = oc.classify((p[0][0], p[1][0]), (c[0][0], c[1][0]), .01)
    assert r.status == "certified_feasible" and r.boxes == 1
    assert r.margin <= .99 and r.margin > .99 - 1e-12     # a few hundred outward ulps
    # affine u = A X: J = det(I + A) everywhere
    A = np.array([[.1, .05, -.02], [.03, -.2, .04], [.01, .02, .15]])
    X = np.array([(x * H[0], y * H[1], z * H[2]) for z in (0, 1) for y in (0, 1) for x in (0, 1)], dtype=float)
    v = X @ A.T
    J = np.linalg.det(np.eye(3) + A)
    p, c, _, ok = oc.bernstein_arrays(v[None], H, 0.)
    r = oc.classify((p[0][0], p[1][0]), (c[0][0], c[1][0]), 0.)
    assert r.status == "certified_feasible"
    assert r.margin <= J and r.margin > J - 1e-12
    m, _ = oc.edge_hull_margins(v[None], H, 0.)
    assert m[0] <= J and m[0] > J - 1e-12


def test_t05_cell_with_positive_d_is_certified_by_subdivision():
    # J = d + 4 (x - 1/2)^2: the root Bernstein box dips to d - 1 < 0, one split at x = 1/2 proves min J = d
    v = counterexample_cell(.3)
    p, c, _, ok = oc.bernstein_arrays(v[None], (1., 1., 1.), 0.)
    assert c[0][0].min() < 0
    r = oc.classify((p[0][0], p[1][0]), (c[0][0], c[1][0]), 0.)
    assert r.status == "certified_feasible" and r.boxes == 3 and r.depth == 1
    assert r.margin <= .3 and r.margin > .3 - 1e-12


@pytest.mark.parametrize("d", [-.05, 0.])
def test_t05_folding_cell_is_never_certified(d):
    v = counterexample_cell(d)
    h = (1., 1., 1.)
    p, c, _, ok = oc.bernstein_arrays(v[None], h, 0.)
    r = oc.classify((p[0][0], p[1][0]), (c[0][0], c[1][0]), 0.)
    assert r.status != "certified_feasible"
    m, _ = oc.edge_hull_margins(v[None], h, 0.)
    assert not m[0] > 0
    if r.status == "infeasible":
        assert exact_J(v, h, tuple(Fraction(x) for x in r.witness)) <= 0


def _subdivided_cells():
    """Cells whose root Bernstein box is not positive but which subdivision certifies (searched)."""
    found = []
    for seed in range(40):
        v = cells(400, .22, 100 + seed)
        p, c, _, ok = oc.bernstein_arrays(v, H, .01)
        for k in np.flatnonzero(ok & (c[0].min(axis=1) <= 0) & (p[1].min(axis=1) > 0)):
            r = oc.classify((p[0][k], p[1][k]), (c[0][k], c[1][k]), .01)
            if r.status == "certified_feasible" and r.boxes > 1:
                found.append(v[k])
            if len(found) >= 6:
                return np.array(found)
    return np.array(found)


SUBDIVIDED = _subdivided_cells()


def test_subdivision_certificates_are_sound_and_witnesses_are_real():
    assert len(SUBDIVIDED) >= 3
    lattice = [tuple(Fraction(i, 6) for i in t) for t in product(range(7), repeat=3)]
    for v in SUBDIVIDED:
        p, c, _, _ = oc.bernstein_arrays(v[None], H, .01)
        r = oc.classify((p[0][0], p[1][0]), (c[0][0], c[1][0]), .01)
        low = min(exact_J(v, H, s) for s in lattice) - Fraction(.01)
        assert Fraction(r.margin) <= low and low > 0
    # witnesses: every infeasible decision names a point where the exact J <= delta
    v = cells(300, .3, 7)
    p, c, _, ok = oc.bernstein_arrays(v, H, .01)
    n_witness = 0
    for k in np.flatnonzero(ok):
        r = oc.classify((p[0][k], p[1][k]), (c[0][k], c[1][k]), .01)
        if r.status == "infeasible":
            n_witness += 1
            assert exact_J(v[k], H, tuple(Fraction(x) for x in r.witness)) <= Fraction(.01)
        elif r.status == "certified_feasible":
            assert min(exact_J(v[k], H, s) for s in POINTS) > Fraction(.01)
    assert n_witness > 20


def test_budgets_end_in_uncertain():
    v = SUBDIVIDED[0]
    p, c, _, _ = oc.bernstein_arrays(v[None], H, .01)
    args = ((p[0][0], p[1][0]), (c[0][0], c[1][0]), .01)
    assert oc.classify(*args, max_boxes=1).reason == "box_budget"
    assert oc.classify(*args, max_depth=0).reason == "depth_budget"


def test_map_certificate_closed_forms():
    shape = (6, 7, 8)
    axes = [np.arange(n, dtype=float) for n in shape[::-1]]
    u = np.zeros(shape + (3,))
    r = oc.certify_map(u, shape, (1., 1., 1.), axes, .01)
    assert r["status"] == "certified_feasible" and r["reason"] == "positive_enclosure_cover"
    assert r["n_fast_cells"] == 5 * 6 * 7 and r["min_margin_lower_bound"] <= .99
    bad = u.copy()
    bad[0, 2, 2, 0] = .1
    assert oc.certify_map(bad, shape, (1., 1., 1.), axes, .01)["reason"] == "NONZERO_BOUNDARY"
    bad = smooth_field(shape, 40., 0)
    bad[[0, -1]] = 0.
    bad[:, [0, -1]] = 0.
    bad[:, :, [0, -1]] = 0.
    assert oc.certify_map(bad, shape, (1., 1., 1.), axes, .01)["reason"] == "INVALID_DOMAIN"
    bad = u.copy()
    bad[2, 2, 2, 1] = np.nan
    assert oc.certify_map(bad, shape, (1., 1., 1.), axes, .01)["reason"] == "NONFINITE_STATE"
