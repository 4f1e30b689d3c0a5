"""The per-cell bound by which the T66 root sweep skips cells (root_free_cell in q1_sweeps.hip) against the cubic
coefficients the oracle computes at the 27 closed points (global_jd_step.py:90-97): every point gradient lies within
its cell's edge spread about the edges' mean and within its largest edge, J at every point is at least the
centre-expansion bound, and |c1|, |c2|, |c3| obey the Frobenius-norm bounds; a cell the predicate passes has no
root and no c0 <= 0 point."""
import numpy as np
import pytest

from conftest import rough_field, smooth_field
from oracle import jd_root as ojr, q1 as Q

SQRT3 = np.sqrt(3.)


def cell_edges(f, spacing):
    """(C, 3, 3, 4) [r, j, k]: d f_r / d x_j along the four parallel edges of each cell, cells in z, y, x order."""
    out = []
    for j, h in enumerate(spacing):           # j = 0, 1, 2: x, y, z (axes 2, 1, 0 of the node array)
        ax = 2 - j
        d = np.diff(f, axis=ax) / h
        parts = []
        for o1 in (0, 1):
            for o2 in (0, 1):
                idx = []
                for a in range(3):
                    if a == ax:
                        idx.append(slice(None))
                    else:
                        first = (a == min(b for b in range(3) if b != ax))
                        o = o1 if first else o2
                        idx.append(slice(o, d.shape[a] - 1 + o))
                parts.append(d[tuple(idx)])
        out.append(np.stack([q.reshape(-1, 3) for q in parts], axis=-1))     # (C, 3 components, 4 edges)
    return np.stack(out, axis=2)                                               # (C, r, j, k)


def edge_maxima(f, spacing):
    """(C, 3, 3) [r, j]: the largest |d f_r / d x_j| over the four parallel edges of each cell."""
    return np.abs(cell_edges(f, spacing)).max(axis=-1)


@pytest.mark.parametrize("amp,seed", [(.05, 1), (.4, 2), (1.5, 3), (4., 4)])
def test_cell_bound_holds_at_every_closed_point(amp, seed):
    shape, spacing = (7, 8, 9), (1., 1.3, .8)
    g = Q.Grid(shape, spacing)
    D = Q.q1_tables(g, "closed27_uniform")[3]
    u = smooth_field(shape, .3, seed) + rough_field(shape, amp * .1, seed + 10)
    v = smooth_field(shape, amp, seed + 20) + rough_field(shape, amp, seed + 30)
    F, _, _, _ = Q.geometry(u, g, D)
    V = Q.geometry(v, g, D)[0] - np.eye(3)
    gu, gv = edge_maxima(u, spacing), edge_maxima(v, spacing)
    # every point gradient within its cell's edge maxima (the blend weights are >= 0 and sum to 1); the oracle forms
    # the gradient as sum_a D u_a rather than from edge differences, which rounds at the scale of |u| / h
    tol_u = 1e-14 * np.abs(u).max() / min(spacing)
    tol_v = 1e-14 * np.abs(v).max() / min(spacing)
    assert np.all(np.abs(F - np.eye(3)) <= gu[:, None] * (1 + 1e-14) + tol_u)
    assert np.all(np.abs(V) <= gv[:, None] * (1 + 1e-14) + tol_v)
    nf2 = np.sum((np.eye(3) + gu) ** 2, axis=(1, 2))
    nv2 = np.sum(gv ** 2, axis=(1, 2))
    c0, c1, c2, c3 = ojr.point_coefficients(F, V, 0.)
    b1 = nf2 / SQRT3 * np.sqrt(nv2)
    b2 = np.sqrt(nf2) * nv2 / SQRT3
    b3 = nv2 * np.sqrt(nv2) / (3 * SQRT3)
    for c, b in ((c1, b1), (c2, b2), (c3, b3)):
        assert np.all(np.abs(c) <= b[:, None] * (1 + 1e-12) + 1e-12 * b.max())
    # the bounds are not vacuous: within a small factor of the largest coefficient of some cell
    assert np.max(np.abs(c1).max(axis=1) / b1) > .2


def test_cell_bound_is_tight_for_uniform_maps():
    """F = a I, V = b I: |c1| = 3 a^2 |b|, |c2| = 3 |a| b^2, |c3| = |b|^3, each bound's equality case."""
    a, b = 1.3, -.2
    F = a * np.eye(3)[None, None]
    V = b * np.eye(3)[None, None]
    c0, c1, c2, c3 = ojr.point_coefficients(F, V, 0.)
    nf2, nv2 = 3 * a * a, 3 * b * b
    assert np.isclose(abs(c1[0, 0]), nf2 / SQRT3 * np.sqrt(nv2))
    assert np.isclose(abs(c2[0, 0]), np.sqrt(nf2) * nv2 / SQRT3)
    assert np.isclose(abs(c3[0, 0]), nv2 * np.sqrt(nv2) / (3 * SQRT3))


def skip_predicate(eu, ev, delta, L):
    """root_free_cell's test in NumPy on cell edges (C, r, j, k): (skip, Jlb, NF, NV)."""
    m = eu.mean(axis=-1)
    spread = np.abs(eu - m[..., None]).max(axis=-1)
    Fc = np.eye(3) + m
    A = Q.cofactor(Fc)
    det = np.sum(Fc[:, 0, :] * A[:, 0, :], axis=-1)
    S = np.sqrt(np.sum(spread ** 2, axis=(1, 2)))
    nfc = np.sqrt(np.sum(Fc ** 2, axis=(1, 2)))
    na = np.sqrt(np.sum(A ** 2, axis=(1, 2)))
    nv = np.sqrt(np.sum(np.abs(ev).max(axis=-1) ** 2, axis=(1, 2)))
    jlb = det - (na * S + nfc * S * S / SQRT3 + S ** 3 / (3 * SQRT3))
    NF = nfc + S
    bound = L * (NF ** 2 / SQRT3 * nv + L * (NF * nv ** 2 / SQRT3 + L * nv ** 3 / (3 * SQRT3)))
    return (jlb - delta) > bound * (1 + 1e-9) + 1e-12 * (1 + NF ** 3), jlb, NF, nv


@pytest.mark.parametrize("amp,limit,seed", [(.05, 1., 1), (.4, .5, 2), (1.5, 2., 3), (.4, 10., 5)])
def test_skipped_cells_have_no_root(amp, limit, seed):
    shape, spacing = (7, 8, 9), (1., 1.3, .8)
    g = Q.Grid(shape, spacing)
    D = Q.q1_tables(g, "closed27_uniform")[3]
    u = smooth_field(shape, .3, seed) + rough_field(shape, amp * .1, seed + 10)
    v = smooth_field(shape, amp * .1, seed + 20) + rough_field(shape, amp * .05, seed + 30)
    F, _, J, _ = Q.geometry(u, g, D)
    V = Q.geometry(v, g, D)[0] - np.eye(3)
    eu, ev = cell_edges(u, spacing), cell_edges(v, spacing)
    skip, jlb, NF, nv = skip_predicate(eu, ev, 0., limit)
    # the centre expansion bounds J and the gradients' norms at every closed point
    assert np.all(J >= jlb[:, None] - 1e-12)
    assert np.all(np.sqrt(np.sum(F ** 2, axis=(2, 3))) <= NF[:, None] * (1 + 1e-12))
    assert np.all(np.sqrt(np.sum(V ** 2, axis=(2, 3))) <= nv[:, None] * (1 + 1e-12) + 1e-14)
    c0, c1, c2, c3 = ojr.point_coefficients(F, V, 0.)
    b = np.abs(c1) * limit + np.abs(c2) * limit ** 2 + np.abs(c3) * limit ** 3
    # a skipped cell: every point valid and no point whose cubic may have a root (the sweep's own screen)
    assert np.all(c0[skip] > 0) and np.all(c0[skip] > b[skip])
    print("skipped fraction", skip.mean())
    if amp == .05:
        assert skip.mean() > .5          # the bound is not vacuous for small steps


def test_index_division_by_multiply_high():
    """The cell index split of the sweeps (div_magic and cell_coords in q1_common.h): floor(n / v) as the high 64
    bits of M n, M = floor((2^64 - 1) / v) + 1, for 32-bit n and 2 <= v < 2^32, checked exactly with Python integers
    on the divisors of the native grids and others, at the multiples of v, their neighbours, the top of the range and
    random indices."""
    rng = np.random.default_rng(9)
    divisors = [2, 3, 5, 7, 157, 158, 159, 189, 190, 191, 221, 222, 223, 224, 1 << 16, (1 << 16) + 1, 65535,
                (1 << 31) - 1, 1 << 31, (1 << 32) - 1] + [int(x) for x in rng.integers(2, 1 << 32, size=40)]
    top = (1 << 32) - 1
    for v in divisors:
        m = ((1 << 64) - 1) // v + 1
        assert 0 < m < (1 << 64)
        ns = {0, 1, v - 1, v, v + 1, top, top - 1, top // v * v, top // v * v - 1}
        ns |= {k * v + o for k in (1, 2, 3, 1000, top // v - 1) for o in (-1, 0, 1) if 0 <= k * v + o <= top}
        ns |= {int(x) for x in rng.integers(0, top, size=2000, endpoint=True)}
        for n in ns:
            assert (m * n) >> 64 == n // v, (v, n)
    # the node grids' cell split: x, y, z of cell c for a 160x192x224 node grid, by multiplies, at both ends and
    # random cells
    cx, cy, cz = 223, 191, 159
    mcx, mcy = ((1 << 64) - 1) // cx + 1, ((1 << 64) - 1) // cy + 1
    for c in [0, cx - 1, cx, cx * cy - 1, cx * cy, cx * cy * cz - 1] + [int(x) for x in rng.integers(0, cx * cy * cz,
                                                                                                        size=2000)]:
        q = (mcx * c) >> 64
        q2 = (mcy * q) >> 64
        assert (c - q * cx, q - q2 * cy, q2) == (c % cx, (c // cx) % cy, c // (cx * cy))
