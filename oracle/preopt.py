"""Oracle: the pre-optimiser's objective, gradient, update and pyramid transfers in float64 NumPy (test
infrastructure only).

Follows jd_curl/three_d/preopt.py:157-233 (_pre_energy_tensors, the Adam/cap loop of _run_level_default),
preopt_grad.py:1-20, 120-205 (the analytic gradient: data r dT chi, cell G scattered with +-1/4, alpha h Lap(h Lap u)),
multiresolution.py:45-84 (four_level_grids, prolong_displacement) and nifti.py:112-149 (coarse_grid, downsample_array,
which is SciPy's gaussian_filter + map_coordinates; the oracle calls SciPy itself there).

Pinned in tests/test_oracle_preopt.py: the gradient by central finite differences of the energy, the trilinear
sampler by scipy.ndimage.map_coordinates(order=1, mode="nearest"), the smoothing by scipy.ndimage.correlate1d
(mode "nearest"), the prolongation by exact reproduction of affine fields.
"""
import numpy as np
from scipy import ndimage


def pad(u_int):
    """(3, D-2, H-2, W-2) -> (3, D, H, W) with zero faces."""
    return np.pad(u_int, ((0, 0), (1, 1), (1, 1), (1, 1)))


def trilinear(M, ix, iy, iz):
    """Value and (d/dix, d/diy, d/diz) of the trilinear interpolation of M at clamped voxel coordinates."""
    D, H, W = M.shape
    ix, iy, iz = (np.clip(a, 0, n - 1) for a, n in ((ix, W), (iy, H), (iz, D)))
    x0, y0, z0 = np.floor(ix), np.floor(iy), np.floor(iz)
    wx1, wy1, wz1 = ix - x0, iy - y0, iz - z0
    wx0, wy0, wz0 = 1 - wx1, 1 - wy1, 1 - wz1
    x0, y0, z0 = x0.astype(int), y0.astype(int), z0.astype(int)

    def c(dz, dy, dx):
        z, y, x = z0 + dz, y0 + dy, x0 + dx
        inside = (z <= D - 1) & (y <= H - 1) & (x <= W - 1)
        return np.where(inside, M[np.minimum(z, D - 1), np.minimum(y, H - 1), np.minimum(x, W - 1)], 0.)
    C = {(a, b, e): c(a, b, e) for a in (0, 1) for b in (0, 1) for e in (0, 1)}
    wz, wy, wx = (wz0, wz1), (wy0, wy1), (wx0, wx1)
    val = sum(C[(a, b, e)] * wx[e] * wy[b] * wz[a] for a in (0, 1) for b in (0, 1) for e in (0, 1))
    dtx = sum((C[(a, b, 1)] - C[(a, b, 0)]) * wy[b] * wz[a] for a in (0, 1) for b in (0, 1))
    dty = sum((C[(a, 1, e)] - C[(a, 0, e)]) * wx[e] * wz[a] for a in (0, 1) for e in (0, 1))
    dtz = sum((C[(1, b, e)] - C[(0, b, e)]) * wx[e] * wy[b] for b in (0, 1) for e in (0, 1))
    return val, (dtx, dty, dtz)


def cell_derivatives(u):
    """Voxel derivatives along x, y, z at the cell centres: (3 axes)[comp] arrays (D-1, H-1, W-1)."""
    out = []
    for axis in (2, 1, 0):                      # x, y, z
        g = np.diff(u, axis=1 + axis)
        for ax in range(3):
            if ax != axis:
                g = .5 * (np.take(g, range(g.shape[1 + ax] - 1), axis=1 + ax) +
                          np.take(g, range(1, g.shape[1 + ax]), axis=1 + ax))
        out.append(g)
    return out                                  # out[d][c]


def laplacian_interior(f, h):
    hx, hy, hz = h
    ui = f[:, 1:-1, 1:-1, 1:-1]
    return ((f[:, 2:, 1:-1, 1:-1] - 2 * ui + f[:, :-2, 1:-1, 1:-1]) / (hz * hz)
            + (f[:, 1:-1, 2:, 1:-1] - 2 * ui + f[:, 1:-1, :-2, 1:-1]) / (hy * hy)
            + (f[:, 1:-1, 1:-1, 2:] - 2 * ui + f[:, 1:-1, 1:-1, :-2]) / (hx * hx))


def energy_grad(u_int, F, M, h, mu, alpha, kappa, j0):
    """(terms dict, grad (3, D-2, H-2, W-2)) of P at u_int (voxel units, float64)."""
    u = pad(np.asarray(u_int, dtype=np.float64))
    D, H, W = F.shape
    zz, yy, xx = np.meshgrid(np.arange(D), np.arange(H), np.arange(W), indexing="ij")
    ix, iy, iz = xx + u[0], yy + u[1], zz + u[2]
    val, dts = trilinear(M, ix, iy, iz)
    r = val - F
    data = .5 * np.sum(r * r)
    chi = ((ix > 0) & (ix < W - 1), (iy > 0) & (iy < H - 1), (iz > 0) & (iz < D - 1))
    grad = np.stack([np.where(chi[a], r * dts[a], 0.) for a in range(3)])
    der = cell_derivatives(u)
    R = [[h[c] / h[d] for d in range(3)] for c in range(3)]
    F_ = [[(1. if c == d else 0.) + der[d][c] * R[c][d] for d in range(3)] for c in range(3)]
    (F00, F01, F02), (F10, F11, F12), (F20, F21, F22) = F_
    cof = [[F11 * F22 - F12 * F21, -(F10 * F22 - F12 * F20), F10 * F21 - F11 * F20],
           [-(F01 * F22 - F02 * F21), F00 * F22 - F02 * F20, -(F00 * F21 - F01 * F20)],
           [F01 * F12 - F02 * F11, -(F00 * F12 - F02 * F10), F00 * F11 - F01 * F10]]
    det = F00 * cof[0][0] + F01 * cof[0][1] + F02 * cof[0][2]
    curl = (F21 - F12, F02 - F20, F10 - F01)
    rl = np.maximum(j0 - det, 0.)
    A = mu * (det - 1) - 2 * kappa * rl
    z = np.zeros_like(det)
    adj = [[z, -curl[2], curl[1]], [curl[2], z, -curl[0]], [-curl[1], curl[0], z]]
    G = [[(A * cof[c][d] + mu * adj[c][d]) * R[c][d] for d in range(3)] for c in range(3)]
    for oz in (0, 1):
        for oy in (0, 1):
            for ox in (0, 1):
                s = ((1. if ox else -1.), (1. if oy else -1.), (1. if oz else -1.))
                for c in range(3):
                    grad[c, oz:oz + D - 1, oy:oy + H - 1, ox:ox + W - 1] += .25 * sum(s[d] * G[c][d] for d in range(3))
    geometry = .5 * mu * (np.sum((det - 1) ** 2) + sum(np.sum(k * k) for k in curl))
    floor = kappa * np.sum(rl * rl)
    curvature = 0.
    if alpha > 0:
        hv = np.array(h)[:, None, None, None]
        lap = laplacian_interior(u, h) * hv
        curvature = .5 * alpha * np.sum(lap * lap)
        grad[:, 1:-1, 1:-1, 1:-1] += alpha * hv * laplacian_interior(pad(lap), h)
    total = data + geometry + floor + curvature
    terms = dict(total=total, data=data, geometry=geometry, floor=floor, curvature=curvature, min_det=float(det.min()))
    return terms, grad[:, 1:-1, 1:-1, 1:-1]


def gauss_kernel(sigma):
    """The update's smoothing weights: radius ceil(3 sigma), normalised (preopt._gaussian_1d)."""
    r = int(np.ceil(3 * sigma))
    x = np.arange(-r, r + 1, dtype=np.float64)
    k = np.exp(-.5 * (x / sigma) ** 2)
    return k / k.sum(), r


def smooth(field, w):
    """Separable correlation of a (3, D', H', W') array along z, y, x with replicate padding."""
    r = (len(w) - 1) // 2
    out = field
    for axis in (1, 2, 3):
        n = out.shape[axis]
        acc = np.zeros_like(out)
        for k in range(-r, r + 1):
            idx = np.clip(np.arange(n) + k, 0, n - 1)
            acc = acc + w[k + r] * np.take(out, idx, axis=axis)
        out = acc
    return out


def adam_step(u, m, v, g, t, lr, cap, b1, b2, eps, w_grad=None, w_field=None):
    """One update (preopt.py:219-233): g smoothed, Adam moments, step scaled to max |step| <= cap, optional field
    smoothing. Returns (u, m, v, capped)."""
    gs = smooth(g, w_grad) if w_grad is not None else g
    m = b1 * m + (1 - b1) * gs
    v = b2 * v + (1 - b2) * gs * gs
    step = lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    largest = float(np.max(np.abs(step)))
    scale = 1. if largest <= cap else cap / largest
    u = u - step * scale
    if w_field is not None:
        u = smooth(u, w_field)
    return u, m, v, scale < 1


def coarse_spacing(shape, spacing, coarse):
    """(hx, hy, hz) of the coarse grid with the same node-centre FOV (nifti.py:112-120)."""
    return tuple((n - 1) * h / (m - 1) for n, h, m in zip(shape[::-1], spacing, coarse[::-1]))


def downsample(values, shape_zyx):
    """SciPy's gaussian_filter (truncate 4, nearest) + order-1 map_coordinates at the linspace node positions."""
    src = np.asarray(values.shape)
    ratios = (src - 1) / (np.asarray(shape_zyx) - 1)
    sigma = .5 * np.sqrt(ratios ** 2 - 1)
    smooth_ = ndimage.gaussian_filter(values, sigma=sigma, mode="nearest", truncate=4.)
    axes = [np.linspace(0, old - 1, new) for old, new in zip(values.shape, shape_zyx)]
    coords = np.array(np.meshgrid(*axes, indexing="ij"))
    return ndimage.map_coordinates(smooth_, coords, order=1, mode="nearest", prefilter=False)


def prolong(u, src_axes_zyx, tgt_axes_zyx, src_spacing_zyx):
    """Separable physical Q1 interpolation of a (nz, ny, nx, 3) field (multiresolution.py:55-84)."""
    result = u
    for axis, (old, new, h) in enumerate(zip(src_axes_zyx, tgt_axes_zyx, src_spacing_zyx)):
        t = (new - old[0]) / h
        t[0], t[-1] = 0., len(old) - 1
        left = np.floor(t).astype(np.intp)
        left[-1] = len(old) - 2
        weight = t - left
        shape = [1] * 4
        shape[axis] = len(new)
        weight = weight.reshape(shape)
        result = (1 - weight) * np.take(result, left, axis=axis) + weight * np.take(result, left + 1, axis=axis)
    return result
