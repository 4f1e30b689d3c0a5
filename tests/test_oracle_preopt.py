"""Pins the pre-optimiser oracle (oracle/preopt.py) by finite differences, SciPy and closed forms."""
import numpy as np
import pytest
from scipy import ndimage

from oracle import preopt as op


def _problem(shape=(7, 8, 9), seed=0, amp=.3):
    rng = np.random.default_rng(seed)
    D, H, W = shape
    z, y, x = np.meshgrid(*(np.linspace(0, 1, n) for n in shape), indexing="ij")
    F = np.sin(3 * x) * np.cos(2 * y) + .5 * z * z
    M = np.sin(3 * x + .2) * np.cos(2 * y - .1) + .5 * (z + .05) ** 2
    u = rng.normal(size=(3, D - 2, H - 2, W - 2)) * amp
    return u, F, M


@pytest.mark.parametrize("alpha,j0", [(0., .2), (1e-2, 1.05)])
def test_gradient_matches_central_differences(alpha, j0):
    # j0 = 1.05 puts most cells in the relu floor, so its derivative is exercised too
    u, F, M = _problem()
    h = (1., 1.3, .8)
    args = (F, M, h, .7, alpha, 3., j0)
    terms, g = op.energy_grad(u, *args)
    rng = np.random.default_rng(1)
    worst = 0.
    for _ in range(40):
        idx = tuple(int(rng.integers(0, n)) for n in u.shape)
        e = np.zeros_like(u)
        e[idx] = 1e-6
        fd = (op.energy_grad(u + e, *args)[0]["total"] - op.energy_grad(u - e, *args)[0]["total"]) / 2e-6
        worst = max(worst, abs(fd - g[idx]) / max(1., abs(g[idx])))
    assert worst < 1e-6
    if j0 > 1:
        assert terms["floor"] > 0


def test_trilinear_matches_scipy_map_coordinates_with_border_clamping():
    rng = np.random.default_rng(2)
    M = rng.normal(size=(6, 7, 8))
    pts = rng.uniform(-1.5, 9., size=(3, 500))          # many outside: border clamping
    val, _ = op.trilinear(M, pts[0], pts[1], pts[2])
    want = ndimage.map_coordinates(M, np.stack([np.clip(pts[2], 0, 5), np.clip(pts[1], 0, 6), np.clip(pts[0], 0, 7)]),
                                   order=1, mode="nearest")
    np.testing.assert_allclose(val, want, rtol=1e-13, atol=1e-13)


def test_smoothing_matches_scipy_correlate1d():
    rng = np.random.default_rng(3)
    f = rng.normal(size=(3, 5, 6, 7))
    w, r = op.gauss_kernel(1.)
    got = op.smooth(f, w)
    want = f
    for axis in (1, 2, 3):
        want = ndimage.correlate1d(want, w, axis=axis, mode="nearest")
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-14)


def test_prolongation_reproduces_affine_fields_and_keeps_zero_faces():
    src_shape, tgt_shape = (5, 6, 7), (9, 11, 13)
    h_src = (1.5, 1.2, 1.)          # x, y, z spacing of the source
    axes_src = [np.arange(n) * hh for n, hh in zip(src_shape[::-1], h_src)]          # x, y, z
    h_tgt = op.coarse_spacing(src_shape, h_src, tgt_shape)
    axes_tgt = [np.arange(n) * hh for n, hh in zip(tgt_shape[::-1], h_tgt)]
    A = np.array([[.1, -.2, .05], [.3, .0, -.1], [.02, .04, .2]])
    def field(axes):
        x, y, z = axes
        X = np.stack(np.meshgrid(z, y, x, indexing="ij")[::-1], axis=-1)            # (nz, ny, nx, (x, y, z))
        return X @ A.T + np.array([.3, -.1, .2])
    got = op.prolong(field(axes_src), axes_src[::-1], axes_tgt[::-1], h_src[::-1])
    np.testing.assert_allclose(got, field(axes_tgt), rtol=0, atol=1e-13)
    u = np.zeros(src_shape + (3,))
    u[1:-1, 1:-1, 1:-1] = np.random.default_rng(4).normal(size=(3, 4, 5, 3))
    out = op.prolong(u, axes_src[::-1], axes_tgt[::-1], h_src[::-1])
    assert not out[[0, -1]].any() and not out[:, [0, -1]].any() and not out[:, :, [0, -1]].any()


def test_adam_step_respects_the_cap_and_descends():
    u, F, M = _problem(amp=.05)
    args = (F, M, (1., 1., 1.), 1., 0., 10., .2)
    terms, g = op.energy_grad(u, *args)
    w, _ = op.gauss_kernel(1.)
    m = v = np.zeros_like(u)
    # the first Adam step is lr * g / (|g| + eps): with lr = 1 its largest entry is ~1 and the cap scales it to .5
    u1, m, v, capped = op.adam_step(u, m, v, g, 1, 1., .5, .9, .99, 1e-8, w_grad=w)
    assert capped and abs(np.max(np.abs(u1 - u)) - .5) < 1e-12
    u2, *_ = op.adam_step(u, np.zeros_like(u), np.zeros_like(u), g, 1, 1e-3, .5, .9, .99, 1e-8)
    assert op.energy_grad(u2, *args)[0]["total"] < terms["total"]
